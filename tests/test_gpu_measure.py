"""GPU: the point-source propagator driver (core/measurements.py:277-371) as one
multi-RHS solve, against the oracle's sequential loop and, at 8^4, by
D S = point sources; and the chronological guesses (su3_wilson_nf2.py:85-194)
on device tensors driving warm starts along a slowly varying gauge field."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def theory(L, mass, **kw):
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    return SU3WilsonNf2(lattice_shape=L, mass=mass, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, **kw)


def test_propagator_matches_oracle_loop():
    print("[gpu] propagator 4^4", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.measurements import full_point_source_propagator
    from oracle.wilson_oracle import WilsonOracle, dense_matrix

    L = (4, 4, 4, 4)
    rng = fields.rng_for(1900)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    th = theory(L, 0.1, cg_tol=1e-12, cg_maxiter=3000)
    src = (1, 2, 3, 0)
    prop, info = full_point_source_propagator(U, th, src)
    assert prop.shape == (1, *L, 4, 3, 4, 3) and info["inv_n_solves"] == 12.0
    ora = WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1))
    D = dense_matrix(lambda v: ora.apply_D(U, v), (1, *L, 4, 3))
    for s in range(4):
        for c in range(3):
            e = np.zeros((1, *L, 4, 3), dtype=np.complex128)
            e[(0, *src, s, c)] = 1.0
            ref = np.linalg.solve(D, e.reshape(-1)).reshape(e.shape)
            got = prop[..., s, c]
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-9, (s, c)


def test_propagator_solves_D_8x4():
    print("[gpu] propagator 8^4", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.measurements import full_point_source_propagator

    L = (8, 8, 8, 8)
    rng = fields.rng_for(1910)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    th = theory(L, 0.05, cg_tol=1e-12, cg_maxiter=4000, solver_precision="f16", solver_recon=12)
    rq = th.resident_gauge(U)
    src = (7, 0, 5, 3)
    prop, _ = full_point_source_propagator(rq, th, src)
    for s in range(4):
        for c in range(3):
            e = np.zeros((1, *L, 4, 3), dtype=np.complex128)
            e[(0, *src, s, c)] = 1.0
            r = np.asarray(th.apply_D(rq, np.ascontiguousarray(prop[..., s, c]))) - e
            assert np.linalg.norm(r) <= 1e-9, (s, c, np.linalg.norm(r))


def test_chrono_guesses_cut_iterations():
    """Pseudofermion-style solves along U_k = exp(k eps H) U with phi_k = D_k^dag eta:
    MRE guesses from (x_i, D_i x_i) with target eta, and polynomial guesses,
    both reach the same true residual in fewer iterations than a cold start."""
    print("[gpu] chrono", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.chrono import chrono_predict_from_history, chrono_predict_mre

    L = (8, 8, 8, 8)
    rng = fields.rng_for(1920)
    U0 = torch.from_numpy(fields.su3_hot_start(rng, 1, L, 0.5)).cuda()
    a = torch.from_numpy(rng.standard_normal((1, 4, *L, 3, 3)) + 1j * rng.standard_normal((1, 4, *L, 3, 3))).cuda()
    H = 0.5 * (a - a.conj().transpose(-1, -2))
    H = H - (torch.diagonal(H, dim1=-2, dim2=-1).sum(-1) / 3.0)[..., None, None] * torch.eye(3, dtype=H.dtype,
                                                                                           device="cuda")
    eta = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    th = theory(L, 0.1, cg_tol=1e-10, cg_maxiter=4000)
    xs, ys, cold, mre, poly = [], [], [], [], []
    for k in range(6):
        Uk = torch.linalg.matrix_exp((0.01 * k) * H.reshape(-1, 3, 3)).reshape(H.shape) @ U0
        rq = th.resident_gauge(Uk)
        phi = th.apply_Ddag(rq, eta)
        th.solve_normal(rq, phi)
        cold.append(th.last_solve_stats[0]["iters"])
        if xs:
            g = chrono_predict_mre(xs, ys, eta, 4)
            xm = th.solve_normal_with_guess(rq, phi, g)
            mre.append(th.last_solve_stats[0]["iters"])
            assert th.last_solve_stats[0]["status"] == "converged"
            gp = chrono_predict_from_history(xs, "poly", 3)
            th.solve_normal_with_guess(rq, phi, gp)
            poly.append(th.last_solve_stats[0]["iters"])
        else:
            xm = th.solve_normal(rq, phi)
        xs.append(xm)
        ys.append(th.apply_D(rq, xm))
    print(f"[gpu] iterations cold {cold} mre {mre} poly {poly}", flush=True)
    assert np.mean(mre) < 0.8 * np.mean(cold[1:]) and np.mean(poly) < np.mean(cold[1:])
