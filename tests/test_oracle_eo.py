"""Pinning of the oracle's even-odd restatement (CPU, NumPy).

The oracle's parity blocks and Schur operators are checked against what the
mathematics fixes independently of them: the dense hopping matrix built column
by column (nearest-neighbour hopping couples opposite parities only, so its
parity blocks are K_oe and K_eo), the Schur complement built from those dense
blocks, the dense LU solution of D x = b (the even part of x solves the Schur
system), adjointness, and the literal compact-index loop of the reference
against the masked full-lattice form.  The closed-form compact index used by
the GPU kernels (full index / 2 for L3 even) is checked against the restated
index tables."""

import numpy as np
import pytest

from jaxqft_amd import fields
from oracle.wilson_oracle import WilsonOracle, checkerboard_indices, dense_matrix, parity_field


def _setup(L, nc, bc, seed):
    rng = fields.rng_for(seed)
    U = fields.su3_hot_start(rng, 1, L, 0.7) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.7)
    ora = WilsonOracle(L, nc, 0.08, fermion_bc=bc)
    return rng, U, ora


def _dense(ora, U, fn):
    shape = (1, *ora.lattice_shape, 4, ora.nc)
    return dense_matrix(lambda v: fn(U, v), shape)


def _parity_dof(L, nc):
    par = parity_field(L).reshape(-1) % 2
    return np.repeat(par, 4 * nc)


@pytest.mark.parametrize("nc,bc", [(3, (1, 1, 1, -1)), (1, (np.exp(0.3j), 1, -1, 1))])
def test_hopping_blocks_and_schur_match_dense(nc, bc):
    L = (2, 2, 2, 4)
    rng, U, ora = _setup(L, nc, bc, 500 + nc)
    K = _dense(ora, U, ora.apply_offdiag)
    dof_par = _parity_dof(L, nc)
    e, o = dof_par == 0, dof_par == 1
    # nearest-neighbour hopping never connects two sites of the same parity
    assert np.abs(K[np.ix_(e, e)]).max() < 1e-14 and np.abs(K[np.ix_(o, o)]).max() < 1e-14
    psi = fields.random_fermion(rng, (1, *L, 4, nc))
    v = psi.reshape(-1)
    Koe = np.zeros_like(K)
    Koe[np.ix_(o, e)] = K[np.ix_(o, e)]
    Keo = np.zeros_like(K)
    Keo[np.ix_(e, o)] = K[np.ix_(e, o)]
    np.testing.assert_allclose(ora.apply_K_oe(U, psi).reshape(-1), Koe @ v, atol=1e-13)
    np.testing.assert_allclose(ora.apply_K_eo(U, psi).reshape(-1), Keo @ v, atol=1e-13)
    np.testing.assert_allclose(ora.apply_Kdag_oe(U, psi).reshape(-1), Keo.conj().T @ v, atol=1e-13)
    np.testing.assert_allclose(ora.apply_Kdag_eo(U, psi).reshape(-1), Koe.conj().T @ v, atol=1e-13)
    m0 = ora.m0
    Pe = np.diag(e.astype(float))
    Po = np.diag(o.astype(float))
    Se = m0 * Pe - Keo @ Koe / m0
    So = m0 * Po - Koe @ Keo / m0
    np.testing.assert_allclose(ora.apply_eo_schur_even(U, psi).reshape(-1), Se @ v, atol=1e-12)
    np.testing.assert_allclose(ora.apply_eo_schur_even(U, psi, normalized=True).reshape(-1), Se @ v / m0, atol=1e-12)
    np.testing.assert_allclose(ora.apply_eo_schur_even_dagger(U, psi).reshape(-1), Se.conj().T @ v, atol=1e-12)
    np.testing.assert_allclose(ora.apply_eo_schur_odd(U, psi).reshape(-1), So @ v, atol=1e-12)


@pytest.mark.parametrize("nc", [3, 1])
def test_schur_solution_is_even_part_of_full_solution(nc):
    """D x = b  =>  S_e x_e = b_e - K_eo b_o / m0; the oracle's even-odd normal
    solve, fed S_e^dag (that) / m0 in the normalized form, returns x_e of the
    dense LU solution."""
    L = (2, 2, 4, 2)
    rng, U, ora = _setup(L, nc, (1, 1, 1, -1), 520 + nc)
    D = _dense(ora, U, ora.apply_D)
    b = fields.random_fermion(rng, (1, *L, 4, nc))
    x = np.linalg.solve(D, b.reshape(-1)).reshape(b.shape)
    c = ora.project_even(b) - ora.apply_K_eo(U, ora.project_odd(b)) / ora.m0
    rhs = ora.apply_eo_schur_even_dagger(U, c / ora.m0, normalized=True)
    xe = ora.solve_eop_even_normal(U, rhs, tol=1e-13, maxiter=2000)
    np.testing.assert_allclose(xe, ora.project_even(x), atol=1e-10)
    # reconstruction of the odd part
    xo = (ora.project_odd(b) - ora.apply_K_oe(U, xe)) / ora.m0
    np.testing.assert_allclose(xe + xo, x, atol=1e-10)


@pytest.mark.parametrize("daggered", [False, True])
@pytest.mark.parametrize("normalized", [False, True])
@pytest.mark.parametrize("nc", [3, 1])
def test_compact_loop_matches_masked_form(daggered, normalized, nc):
    """The reference's literal compact-index loop (neighbour tables) equals the
    masked full-lattice Schur operator."""
    L = (4, 2, 2, 4)
    rng, U, ora = _setup(L, nc, (-1, 1, np.exp(0.5j), -1), 540 + nc)
    psi = fields.random_fermion(rng, (1, *L, 4, nc))
    full = (ora.apply_eo_schur_even_dagger if daggered else ora.apply_eo_schur_even)(U, psi, normalized=normalized)
    comp = ora.apply_eo_schur_even_compact_one(U[0], ora.pack_parity(psi, 0)[0], daggered=daggered,
                                               normalized=normalized)
    np.testing.assert_allclose(ora.pack_parity(full, 0)[0], comp, atol=1e-13)
    np.testing.assert_allclose(ora.unpack_parity(comp[None], 0), full, atol=1e-13)


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (2, 4, 6, 8), (8, 2, 2, 2), (6, 6, 4, 2)])
def test_compact_index_is_half_the_full_index(L):
    """With L3 even, the rank of a site among the sites of its parity in
    ascending flat order (su3_wilson_nf2.py:831-840) is its flat index // 2,
    and compact index i of parity q sits at x3 = 2 (i mod L3/2) + (x0+x1+x2+q) mod 2;
    the GPU kernels and pack routines rely on both."""
    even, odd = checkerboard_indices(L)
    for q, idx in ((0, even), (1, odd)):
        np.testing.assert_array_equal(idx // 2, np.arange(idx.size))
        i = np.arange(idx.size)
        L3h = L[3] // 2
        row, j = i // L3h, i % L3h
        x2 = row % L[2]
        x1 = (row // L[2]) % L[1]
        x0 = row // (L[2] * L[1])
        flat = row * L[3] + 2 * j + ((x0 + x1 + x2 + q) & 1)
        np.testing.assert_array_equal(flat, idx)
