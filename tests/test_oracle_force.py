"""Pinning of the oracle's analytic pseudofermion force (CPU, NumPy) by the
calculus identity it must satisfy: with S(U) = phi^dag (D^dag D)^-1 phi,
X = (D^dag D)^-1 phi, Y = D X and the link U_mu(z) -> exp(eps H) U_mu(z) for
a random traceless anti-Hermitian H,
    dS/d eps |_0 = Re tr(H F_mu(z)),
checked against a central finite difference of S computed by dense LU.  The
same for the even-odd action phi_e^dag (S^dag S)^-1 phi_e, and for U(1)
(H = i h, h real)."""

import numpy as np
import pytest
import scipy.linalg

from jaxqft_amd import fields
from oracle.wilson_oracle import WilsonOracle, dense_matrix


def _action(ora, U, phi, eo=False):
    shape = (1, *ora.lattice_shape, 4, ora.nc)
    if eo:
        S = dense_matrix(lambda v: ora.apply_eo_schur_even(U, v, normalized=True), shape)
        even = np.repeat((np.indices(ora.lattice_shape).sum(axis=0) % 2 == 0).reshape(-1), 4 * ora.nc)
        Se = S[np.ix_(even, even)]
        pe = ora.project_even(phi).reshape(-1)[even]
        return float(np.real(np.vdot(pe, np.linalg.solve(Se.conj().T @ Se, pe))))
    D = dense_matrix(lambda v: ora.apply_D(U, v), shape)
    p = phi.reshape(-1)
    return float(np.real(np.vdot(p, np.linalg.solve(D.conj().T @ D, p))))


def _perturbed(U, nc, mu, z, H, eps):
    V = U.copy()
    idx = (0, mu, *z)
    if nc == 3:
        V[idx] = scipy.linalg.expm(eps * H) @ U[idx]
    else:
        V[idx] = np.exp(eps * H) * U[idx]
    return V


@pytest.mark.parametrize("nc,bc", [(3, (1, 1, 1, -1)), (3, (np.exp(0.4j), 1, 1, -1)), (1, (1, -1, 1, -1))])
@pytest.mark.parametrize("eo", [False, True])
def test_force_is_the_derivative_of_the_action(nc, bc, eo):
    L = (2, 2, 2, 4)
    rng = fields.rng_for(1700 + nc + (10 if eo else 0))
    U = fields.su3_hot_start(rng, 1, L, 0.6) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.6)
    phi = fields.random_fermion(rng, (1, *L, 4, nc))
    ora = WilsonOracle(L, nc, 0.15, fermion_bc=bc)
    shape = (1, *L, 4, nc)
    if eo:
        S = dense_matrix(lambda v: ora.apply_eo_schur_even(U, v, normalized=True), shape)
        even = np.repeat((np.indices(L).sum(axis=0) % 2 == 0).reshape(-1), 4 * nc)
        Se = S[np.ix_(even, even)]
        pe = ora.project_even(phi).reshape(-1)[even]
        xe = np.zeros(S.shape[0], dtype=np.complex128)
        xe[even] = np.linalg.solve(Se.conj().T @ Se, pe)
        Xe = xe.reshape(shape)
        Ye = ora.apply_eo_schur_even(U, Xe, normalized=True)
        F = ora.eop_pseudofermion_force_from_solution(U, Xe, Ye)
    else:
        D = dense_matrix(lambda v: ora.apply_D(U, v), shape)
        X = np.linalg.solve(D.conj().T @ D, phi.reshape(-1)).reshape(shape)
        F = ora.pseudofermion_force_from_solution(U, X, ora.apply_D(U, X))
    eps = 1e-5
    for mu, z in ((0, (1, 0, 1, 3)), (3, (0, 1, 1, 3)), (2, (1, 1, 0, 0))):
        if nc == 3:
            a = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
            H = fields.project_su3_algebra(a)
        else:
            H = 1j * rng.standard_normal()
        fd = (_action(ora, _perturbed(U, nc, mu, z, H, eps), phi, eo)
              - _action(ora, _perturbed(U, nc, mu, z, H, -eps), phi, eo)) / (2 * eps)
        Fz = F[(0, mu, *z)]
        an = float(np.real(np.trace(H @ Fz))) if nc == 3 else float(np.real(H * Fz))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (mu, z, fd, an)


def test_force_is_in_the_algebra():
    L = (2, 2, 2, 4)
    rng = fields.rng_for(1750)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    X = fields.random_fermion(rng, (1, *L, 4, 3))
    Y = fields.random_fermion(rng, (1, *L, 4, 3))
    F = WilsonOracle(L, 3, 0.1).pseudofermion_force_from_solution(U, X, Y)
    assert np.abs(F + np.conj(np.swapaxes(F, -1, -2))).max() < 1e-14
    assert np.abs(np.trace(F, axis1=-2, axis2=-1)).max() < 1e-14
    Fx = WilsonOracle(L, 3, 0.1, layout="BXYMIJ").pseudofermion_force_from_solution(np.moveaxis(U, 1, 5), X, Y)
    np.testing.assert_allclose(np.moveaxis(Fx, 5, 1), F, atol=1e-15)
