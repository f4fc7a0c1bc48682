"""Pins the CPU oracle against facts fixed independently of its own code:
closed-form gamma tables, the free-field momentum-space symbol (with
boundary-phase-shifted momenta), an explicit single-site hop formula on a
non-uniform lattice, dense brute force (adjoint, gamma5-hermiticity, LU),
gauge covariance and the pseudofermion identity."""

import itertools

import numpy as np
import pytest

from jaxqft_amd import fields
from oracle import wilson_oracle as wo

RNG = np.random.default_rng(1234)


# ----------------------------------------------------------------- gamma
def test_gamma_closed_form_and_clifford():
    g = wo.build_euclidean_gamma(4)
    s1 = np.array([[0, 1], [1, 0]], complex)
    s2 = np.array([[0, -1j], [1j, 0]], complex)
    s3 = np.array([[1, 0], [0, -1]], complex)
    I2 = np.eye(2)
    for mu, ref in enumerate([np.kron(s1, s1), np.kron(s1, s2), np.kron(s2, I2), np.kron(s3, I2)]):
        np.testing.assert_array_equal(g[mu], ref)
    for mu in range(4):
        np.testing.assert_array_equal(g[mu], np.conj(g[mu].T))
        for nu in range(4):
            anti = g[mu] @ g[nu] + g[nu] @ g[mu]
            np.testing.assert_allclose(anti, 2 * np.eye(4) * (mu == nu), atol=0)
    g5 = wo.gamma5(g)
    np.testing.assert_array_equal(g5, np.kron(s1, s3))
    for mu in range(4):
        np.testing.assert_allclose(g5 @ g[mu] + g[mu] @ g5, 0, atol=0)
    np.testing.assert_allclose(g5 @ g5, np.eye(4), atol=0)


def test_gamma_sparse_tables_match_closed_form():
    perm, phase = wo.sparse_gamma_tables(wo.build_euclidean_gamma(4))
    np.testing.assert_array_equal(perm, [[3, 2, 1, 0], [3, 2, 1, 0], [2, 3, 0, 1], [0, 1, 2, 3]])
    np.testing.assert_array_equal(phase, [[1, 1, 1, 1], [-1j, 1j, -1j, 1j], [-1j, -1j, 1j, 1j], [1, 1, -1, -1]])


def test_gamma_other_dimensions():
    for d in (2, 3, 5, 6):
        g = wo.build_euclidean_gamma(d)
        for mu in range(d):
            for nu in range(d):
                anti = g[mu] @ g[nu] + g[nu] @ g[mu]
                np.testing.assert_allclose(anti, 2 * np.eye(g.shape[-1]) * (mu == nu), atol=1e-14)


# --------------------------------------------------------- free field
def identity_links(L, nc):
    if nc == 1:
        return np.ones((1, 4, *L), complex)
    return np.broadcast_to(np.eye(3, dtype=complex), (1, 4, *L, 3, 3)).copy()


@pytest.mark.parametrize("nc", [3, 1])
@pytest.mark.parametrize("kernel", ["reference", "optimized"])
@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (1, 1, 1, -1), (np.exp(0.3j), 1, -1, np.exp(-1.1j))])
def test_free_field_plane_waves(nc, kernel, bc):
    L = (3, 4, 5, 6)
    ora = wo.WilsonOracle(L, nc=nc, mass=0.07, fermion_bc=bc, kernel=kernel)
    theta = [np.angle(complex(b)) for b in bc]
    for n in [(0, 0, 0, 0), (1, 2, 3, 4), (2, 1, 4, 5)]:
        p = [(2 * np.pi * n[mu] + theta[mu]) / L[mu] for mu in range(4)]
        u = RNG.standard_normal((4, nc)) + 1j * RNG.standard_normal((4, nc))
        psi = wo.plane_wave(L, p, u)
        sym = wo.free_field_symbol(p, 0.07)
        expect = wo.plane_wave(L, p, sym @ u)
        got = ora.apply_D(identity_links(L, nc), psi)
        np.testing.assert_allclose(got, expect, atol=1e-12)
        # D^dag has the Hermitian-conjugate symbol
        got_d = ora.apply_Ddag(identity_links(L, nc), psi)
        np.testing.assert_allclose(got_d, wo.plane_wave(L, p, np.conj(sym.T) @ u), atol=1e-12)


# ------------------------------------------- explicit single-site hops
@pytest.mark.parametrize("nc", [3, 1])
def test_point_source_hops_explicit(nc):
    """(D delta_y u)(x): (m+4)u at x=y; -1/2 (1 - g_mu) U~_mu(x) u at x=y-mu;
    -1/2 (1 + g_mu) U~_mu(y)^dag u at x=y+mu (U~ carries p_mu where x_mu=L-1).
    Written with explicit coordinate arithmetic, no rolls."""
    L = (3, 4, 5, 6)
    bc = (np.exp(0.4j), 1.0, -1.0, -1.0)
    mass = 0.13
    rng = fields.rng_for(7)
    if nc == 3:
        U = fields.su3_hot_start(rng, 1, L, 0.5)
    else:
        U = fields.u1_hot_start(rng, 1, L, 0.5)
    g = wo.build_euclidean_gamma(4)
    ora = wo.WilsonOracle(L, nc=nc, mass=mass, fermion_bc=bc)
    for y in [(0, 0, 0, 0), (2, 3, 4, 5), (1, 0, 4, 2)]:
        u = RNG.standard_normal((4, nc)) + 1j * RNG.standard_normal((4, nc))
        psi = np.zeros((1, *L, 4, nc), complex)
        psi[(0, *y)] = u
        out = ora.apply_D(U, psi)
        expect = np.zeros_like(out)
        expect[(0, *y)] += (mass + 4) * u
        for mu in range(4):
            xm = list(y)
            xm[mu] = (y[mu] - 1) % L[mu]  # x with x + mu = y
            link = U[(0, mu, *xm)]
            if nc == 1:
                link = np.array([[link]])
            ph = bc[mu] if xm[mu] == L[mu] - 1 else 1.0
            expect[(0, *xm)] += -0.5 * (np.eye(4) - g[mu]) @ (ph * (link @ u.T)).T
            xp = list(y)
            xp[mu] = (y[mu] + 1) % L[mu]  # x with x - mu = y
            link_y = U[(0, mu, *y)]
            if nc == 1:
                link_y = np.array([[link_y]])
            ph = np.conj(bc[mu]) if y[mu] == L[mu] - 1 else 1.0
            expect[(0, *xp)] += -0.5 * (np.eye(4) + g[mu]) @ (ph * (np.conj(link_y.T) @ u.T)).T
        np.testing.assert_allclose(out, expect, atol=1e-13)


# ------------------------------------------------ stencil variants agree
@pytest.mark.parametrize("nc", [3, 1])
def test_reference_and_optimized_variants_agree(nc):
    L = (4, 2, 3, 4)
    rng = fields.rng_for(3)
    U = fields.su3_hot_start(rng, 2, L, 0.7) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.7)
    psi = fields.random_fermion(rng, (2, *L, 4, nc))
    a = wo.WilsonOracle(L, nc, 0.05, fermion_bc=(1, 1, 1, -1), kernel="reference")
    b = wo.WilsonOracle(L, nc, 0.05, fermion_bc=(1, 1, 1, -1), kernel="optimized")
    for f in ("apply_D", "apply_Ddag", "apply_normal"):
        x, y = getattr(a, f)(U, psi), getattr(b, f)(U, psi)
        assert np.linalg.norm(x - y) / np.linalg.norm(x) < 1e-14


# --------------------------------------------------------- brute force
@pytest.fixture(scope="module")
def dense_su3():
    L = (2, 2, 2, 4)
    rng = fields.rng_for(11)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    ora = wo.WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1))
    shape = (1, *L, 4, 3)
    D = wo.dense_matrix(lambda v: ora.apply_D(U, v), shape)
    Dd = wo.dense_matrix(lambda v: ora.apply_Ddag(U, v), shape)
    return L, U, ora, D, Dd


def test_dense_adjoint_and_gamma5_hermiticity(dense_su3):
    L, U, ora, D, Dd = dense_su3
    np.testing.assert_allclose(Dd, np.conj(D.T), atol=1e-14)
    g5 = wo.gamma5(wo.build_euclidean_gamma(4))
    G5 = np.kron(np.eye(int(np.prod(L))), np.kron(g5, np.eye(3)))
    np.testing.assert_allclose(G5 @ D @ G5, Dd, atol=1e-14)
    N = wo.dense_matrix(lambda v: ora.apply_normal(U, v), (1, *L, 4, 3))
    np.testing.assert_allclose(N, Dd @ D, atol=1e-12)
    assert np.linalg.eigvalsh(0.5 * (N + np.conj(N.T))).min() > 0


def test_cg_matches_dense_lu(dense_su3):
    L, U, ora, D, Dd = dense_su3
    rng = fields.rng_for(12)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    x_lu = np.linalg.solve(Dd @ D, b.reshape(-1)).reshape(b.shape)
    x, k = wo.cg(lambda v: ora.apply_normal(U, v), b, tol=1e-12, maxiter=2000)
    assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 1e-9
    assert np.linalg.norm(ora.apply_normal(U, x) - b) / np.linalg.norm(b) <= 1e-12
    # maxiter caps the iteration count like the reference call
    _, k2 = wo.cg(lambda v: ora.apply_normal(U, v), b, tol=1e-12, maxiter=5)
    assert k2 == 5


def test_pseudofermion_identity():
    L = (4, 4, 4, 4)
    rng = fields.rng_for(13)
    U = fields.su3_hot_start(rng, 1, L, 0.3)
    ora = wo.WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1))
    eta = fields.random_fermion(rng, (1, *L, 4, 3))
    phi = ora.apply_Ddag(U, eta)
    x, _ = wo.cg(lambda v: ora.apply_normal(U, v), phi, tol=1e-12, maxiter=2000)
    S = np.vdot(phi, x).real
    assert abs(S - np.vdot(eta, eta).real) / np.vdot(eta, eta).real < 1e-10


def test_gauge_covariance():
    L = (2, 4, 3, 4)
    rng = fields.rng_for(14)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    Om = fields.expm_su3_algebra(fields.project_su3_algebra(
        0.7 * (rng.standard_normal((1, *L, 3, 3)) + 1j * rng.standard_normal((1, *L, 3, 3)))))
    Up = wo.gauge_transform_links(U, Om)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    ora = wo.WilsonOracle(L, 3, 0.05, fermion_bc=(1, 1, 1, -1))
    lhs = ora.apply_D(Up, wo.gauge_transform_fermion(psi, Om))
    rhs = wo.gauge_transform_fermion(ora.apply_D(U, psi), Om)
    assert np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs) < 1e-14


def test_layouts_equivalent():
    L = (2, 3, 2, 4)
    rng = fields.rng_for(15)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    U_b = np.moveaxis(U, 1, 5).copy()  # BXYMIJ
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    a = wo.WilsonOracle(L, 3, 0.05, layout="BMXYIJ").apply_D(U, psi)
    b = wo.WilsonOracle(L, 3, 0.05, layout="BXYMIJ").apply_D(U_b, psi)
    np.testing.assert_array_equal(a, b)


def test_bc_parsing():
    assert wo.normalize_fermion_bc(None, 4) == (1, 1, 1, 1)
    assert wo.normalize_fermion_bc(None, 4, default_antiperiodic_t=True) == (1, 1, 1, -1)
    assert wo.normalize_fermion_bc("antiperiodic-t", 4) == (1, 1, 1, -1)
    assert wo.normalize_fermion_bc("1,1,1,-1", 4) == (1, 1, 1, -1)
    assert wo.normalize_fermion_bc("1,1,1,1i", 4)[3] == 1j
    with pytest.raises(ValueError):
        wo.normalize_fermion_bc((1, 1), 4)


# --------------------------------------------------------- index tables
@pytest.mark.parametrize("L", [(4, 4, 4, 4), (4, 4, 4, 8), (2, 3, 4, 5)])
def test_index_tables_by_coordinates(L):
    t = wo.checkerboard_neighbor_maps(L)
    V = int(np.prod(L))
    coords = list(itertools.product(*[range(n) for n in L]))  # C-order = flat index order
    par = np.array([sum(c) % 2 for c in coords])
    even = [i for i in range(V) if par[i] == 0]
    odd = [i for i in range(V) if par[i] == 1]
    np.testing.assert_array_equal(t["even_idx"], even)
    np.testing.assert_array_equal(t["odd_idx"], odd)
    cinv = {c: i for i, c in enumerate(coords)}
    einv = {s: k for k, s in enumerate(even)}
    oinv = {s: k for k, s in enumerate(odd)}
    for mu in range(4):
        for k, s in enumerate(odd):
            c = list(coords[s])
            cp = list(c)
            cp[mu] = (c[mu] + 1) % L[mu]
            cm = list(c)
            cm[mu] = (c[mu] - 1) % L[mu]
            sp, sm = cinv[tuple(cp)], cinv[tuple(cm)]
            assert t["plus_oe"][mu, k] == einv.get(sp, -1)
            assert t["minus_oe"][mu, k] == einv.get(sm, -1)
            assert t["bwd_oe"][mu, k] == sm
        for k, s in enumerate(even):
            c = list(coords[s])
            cp = list(c)
            cp[mu] = (c[mu] + 1) % L[mu]
            cm = list(c)
            cm[mu] = (c[mu] - 1) % L[mu]
            assert t["plus_eo"][mu, k] == oinv.get(cinv[tuple(cp)], -1)
            assert t["minus_eo"][mu, k] == oinv.get(cinv[tuple(cm)], -1)
            assert t["bwd_eo"][mu, k] == cinv[tuple(cm)]
