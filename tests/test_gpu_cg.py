"""GPU parity of the normal-equation CG (plain FP64 and reliable-update mixed
precision) against the oracle: every solve must reach a true residual
||b - D^dag D x|| / ||b|| <= 1e-10 evaluated by the FP64 oracle operator, and
agree with the oracle's converged solution."""

from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


def true_resid(L, nc, mass, bc, U, b, x):
    from oracle.wilson_oracle import WilsonOracle

    ora = WilsonOracle(L, nc, mass, fermion_bc=bc)
    r = ora.apply_normal(U, x) - b
    return float(np.linalg.norm(r) / np.linalg.norm(b))


@pytest.mark.parametrize("policy", [
    dict(solver_precision="f64"),
    dict(solver_precision="f64", solver_stop="iterated"),
    dict(solver_precision="f64", use_graph=True),
    dict(solver_precision="f32"),
    dict(solver_precision="f32", solver_recon=12),
    dict(solver_precision="f16"),
    dict(solver_precision="f16", solver_recon=12, use_graph=True),
    dict(solver_precision="f16", solver_recon=12, solver_accum="f16", reliable_delta=0.01),
    dict(solver_precision="f32", solver_accum="f64", use_graph=True),
    dict(solver_precision="f16", solver_recon=12, solver_accum="f16", reliable_delta=0.01, solver_search="xpay"),
    dict(solver_precision="f32", solver_search="xpay", use_graph=True),
    dict(solver_precision="f16", solver_recon=12, solver_stop="iterated"),
])
def test_config1_cg_4x4x4x4(policy):
    print(f"[gpu] cg config-1 {policy}", flush=True)
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    z = np.load(GOLD / "cg_su3_4x4x4x4.npz")
    L = (4, 4, 4, 4)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=3000, **policy)
    x = th.solve_normal(z["U"], z["b"])
    st = th.last_solve_stats[0]
    assert st["status"] == "converged", st
    res = true_resid(L, 3, 0.1, (1, 1, 1, -1), z["U"], z["b"], x)
    if policy.get("solver_stop") == "iterated":
        assert res <= 2e-10, res  # reference rule stops on the iterated residual
    else:
        assert res <= 1e-10, (res, st)
        assert abs(res - st["true_resid"]) <= 1e-3 * res + 1e-14
    assert np.linalg.norm(x - z["x"]) / np.linalg.norm(z["x"]) < 1e-8


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("nc", [3, 1])
def test_batch_and_guess(nc, prec):
    """Two configurations per call, cold and from a guess near the solution,
    SU(3) and U(1), plain FP64 and mixed precision: every column meets 1e-10
    by the oracle and the guess cuts the iterations."""
    print(f"[gpu] cg batch/guess nc={nc} {prec}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    L = (4, 4, 2, 8)
    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    th = cls(lattice_shape=L, mass=0.05, batch_size=2, dtype=np.complex128, cg_tol=1e-10, cg_maxiter=4000,
             solver_precision=prec, reliable_delta=0.01)
    rng = fields.rng_for(500 + nc)
    U = fields.su3_hot_start(rng, 2, L, 0.5) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.5)
    b = fields.random_fermion(rng, th.fermion_shape())
    x = th.solve_normal(U, b)
    it_cold = [s["iters"] for s in th.last_solve_stats]
    for k in range(2):
        assert true_resid(L, nc, 0.05, th.fermion_bc, U[k:k + 1], b[k:k + 1], x[k:k + 1]) <= 1e-10
    x2 = th.solve_normal_with_guess(U, b, x + 1e-6 * b)
    it_warm = [s["iters"] for s in th.last_solve_stats]
    assert all(w < c for w, c in zip(it_warm, it_cold)), (it_warm, it_cold)
    for k in range(2):
        assert true_resid(L, nc, 0.05, th.fermion_bc, U[k:k + 1], b[k:k + 1], x2[k:k + 1]) <= 1e-10


def test_pseudofermion_identity_on_gpu():
    print("[gpu] pfid", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 4)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-12,
                      cg_maxiter=3000, seed=9)
    U = th.hot_start(scale=0.3)
    eta = th.random_fermion()
    phi = th.apply_Ddag(U, eta)
    S = th.pseudofermion_action(U, phi)
    assert abs(S[0] - np.vdot(eta, eta).real) / np.vdot(eta, eta).real < 1e-10
    st = th.normal_residual_stats(U, phi, th.solve_normal(U, phi))
    assert st["max_rel_residual"] <= 1e-12
    # device tensors: the action is reduced on the GPU and agrees with the host sum
    import torch

    St = th.pseudofermion_action(torch.from_numpy(U).cuda(), torch.from_numpy(phi).cuda())
    assert isinstance(St, np.ndarray) and St.shape == (1,)
    assert abs(St[0] - S[0]) <= 1e-12 * abs(S[0])
    Se = th.eop_pseudofermion_action(U, phi)
    Set = th.eop_pseudofermion_action(torch.from_numpy(U).cuda(), torch.from_numpy(phi).cuda())
    assert abs(Set[0] - Se[0]) <= 1e-12 * abs(Se[0])


def test_status_maxiter_and_nan():
    print("[gpu] cg status", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 4)
    rng = fields.rng_for(600)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, dtype=np.complex128, cg_tol=1e-10, cg_maxiter=7)
    th.solve_normal(U, b)
    assert th.last_solve_stats[0]["status"] == "maxiter"
    assert th.last_solve_stats[0]["iters"] == 7
    bad = b.copy()
    bad[0, 1, 2, 3, 0, 2, 1] = np.nan
    th2 = SU3WilsonNf2(lattice_shape=L, mass=0.1, dtype=np.complex128, cg_tol=1e-10, cg_maxiter=100)
    th2.solve_normal(U, bad)
    assert th2.last_solve_stats[0]["status"] == "nan"
    # zero right-hand side converges immediately to x = 0
    x0 = th2.solve_normal(U, np.zeros_like(b))
    assert th2.last_solve_stats[0]["status"] == "converged" and np.all(x0 == 0)


def test_gauge_upload_waits_for_the_producing_stream():
    """Links computed on a non-default torch stream and uploaded at once: the
    upload (which takes no stream) reads them only after that work finishes,
    so D on the resident links matches the oracle on the same links and the
    mixed solve converges (a pack racing the producer would read partial
    links)."""
    print("[gpu] gauge upload ordering", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from oracle.wilson_oracle import WilsonOracle

    L = (16, 16, 16, 16)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(670)
    U0h = fields.su3_hot_start(rng, 1, L, 0.5)
    ah = rng.standard_normal((1, 4, *L, 3, 3)) + 1j * rng.standard_normal((1, 4, *L, 3, 3))
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10, cg_maxiter=3000,
                      solver_precision="f16", solver_recon=12, reliable_delta=0.01, use_graph=True, poll_every=16)
    ora = WilsonOracle(L, 3, 0.05, fermion_bc=bc)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        U0 = torch.from_numpy(U0h).cuda()
        a = torch.from_numpy(ah).cuda()
        H = 0.5 * (a - a.conj().transpose(-1, -2))
        H = H - (torch.diagonal(H, dim1=-2, dim2=-1).sum(-1) / 3.0)[..., None, None] * torch.eye(
            3, dtype=H.dtype, device="cuda")
        pt = torch.from_numpy(psi).cuda()
        for k in (1, 2):
            Uk = torch.linalg.matrix_exp((0.01 * k) * H.reshape(-1, 3, 3)).reshape(H.shape) @ U0
            rq = th.resident_gauge(Uk)
            got = th.apply_D(rq, pt)
            x = th.solve_normal(rq, pt)
            st = th.last_solve_stats[0]
            side.synchronize()
            Ukh = Uk.cpu().numpy()
            ref = ora.apply_D(Ukh, psi)
            got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
            err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            assert err < 1e-13, (k, err)
            assert st["status"] == "converged", (k, st)
            del rq


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_mixed_solve_independent_of_block_length(prec):
    """A mixed solve decides its reliable updates on the device and alternates
    its two search-direction buffers by iteration: the iterates, and so the
    solution and the iteration count, are the same bit for bit whatever the
    host's block length (odd requests round up to even blocks) and with or
    without graphs; updates fall after iterations of both parities (delta 0.1
    triggers many), and every solve meets 1e-10 by the oracle."""
    print(f"[gpu] block-length independence {prec}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(660)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, 3, *L, 4, 3))
    ref = None
    for poll, graph in ((1, False), (3, True), (16, True), (16, False)):
        th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10,
                          cg_maxiter=3000, solver_precision=prec, solver_recon=12, reliable_delta=0.1,
                          poll_every=poll, use_graph=graph)
        x = np.asarray(th.solve_normal_multi(U, b))
        st = th.last_solve_stats
        assert all(c["status"] == "converged" for c in st), st
        assert max(c["reliable_updates"] for c in st) >= 3
        got = (x, [c["iters"] for c in st])
        if ref is None:
            ref = got
            for j in range(3):
                assert true_resid(L, 3, 0.05, bc, U, b[:, j], x[:, j]) <= 1e-10
        else:
            assert got[1] == ref[1] and np.array_equal(got[0], ref[0]), (poll, graph)


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_search_direction_forms_agree(prec):
    """The mixed solve's two search-direction forms -- by recurrence (t = D r
    + beta t_prev, p = r + beta p_prev formed in the D^dag launch; the
    default) and in a pass of its own (p = r + beta p, t = D p) -- are the same
    CG in exact arithmetic.  At 8^4 with three right-hand sides each form
    meets the true residual 1e-10 by the oracle on every column, the two
    solutions agree to solver accuracy, and their iteration counts differ by
    rounding only.  The C ABI rejects any other form."""
    print(f"[gpu] search-direction forms {prec}", flush=True)
    import ctypes

    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (8, 8, 8, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(690)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, 3, *L, 4, 3))
    got = {}
    for form in ("recurrence", "xpay"):
        th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10,
                          cg_maxiter=3000, solver_precision=prec, solver_recon=12, reliable_delta=0.01,
                          solver_accum="f16" if prec == "f16" else "auto", use_graph=True, solver_search=form)
        x = np.asarray(th.solve_normal_multi(U, b))
        st = th.last_solve_stats
        assert all(c["status"] == "converged" for c in st), (form, st)
        for j in range(3):
            assert true_resid(L, 3, 0.1, bc, U, b[:, j], x[:, j]) <= 1e-10, (form, j)
        got[form] = (x, [c["iters"] for c in st])
    (xr, ir), (xp, ip) = got["recurrence"], got["xpay"]
    assert np.linalg.norm(xr - xp) / np.linalg.norm(xp) < 1e-8
    assert all(abs(a - c) <= 3 for a, c in zip(ir, ip)), (ir, ip)
    # an undefined form is an argument error, reported before any device work
    lib = nat.load()
    ctx = nat.Context(L, 3, 1, 0.1, 1.0, bc)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fb, fx = nat.Field(ctx, nat.WL_F64, 1), nat.Field(ctx, nat.WL_F64, 1)
    fb.import_(torch.from_numpy(b[:, 0]).cuda())
    sp, keep = nat.solve_params(1e-10, 100, inner_prec=nat.PREC[prec])
    sp.search_update = 7
    st = (nat.WlSolveStats * 1)()
    rc = lib.wl_cg_normal_field(ctx.h, g.h, fb.h, fx.h, 0, 1, ctypes.byref(sp), st, None)
    assert rc == nat.WL_ERR_BAD_ARG and "search_update" in lib.wl_last_error().decode()
    torch.cuda.synchronize()


@pytest.mark.parametrize("perturb", [0.0, 1e-8])
@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_true_residual_on_links_as_given(perturb, prec):
    """The reliable updates and the final true residual use 12-real FP64
    links only when those rebuild the stored third rows to 1e-14: on exact
    SU(3) links and on links pushed 1e-8 off the group (where a rebuilt third
    row would change the operator by ~1e-8), every solve meets 1e-10 by the
    oracle applied to the links exactly as given."""
    print(f"[gpu] true residual on given links perturb={perturb} {prec}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(650)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    if perturb:
        U = U + perturb * (rng.standard_normal(U.shape) + 1j * rng.standard_normal(U.shape))
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    kw = dict(solver_precision=prec, solver_recon=12) if prec != "f64" else dict(solver_precision="f64")
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=3000, **kw)
    x = np.asarray(th.solve_normal(U, b))
    assert th.last_solve_stats[0]["status"] == "converged", th.last_solve_stats
    res = true_resid(L, 3, 0.05, bc, U, b, x)
    assert res <= 1e-10, res
    assert abs(res - th.last_solve_stats[0]["true_resid"]) <= 1e-3 * res + 1e-15


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_zero_guess_start_matches_explicit_zero_guess(prec):
    """A solve without a guess starts from r = b with |r|^2 = |b|^2 and no
    operator application; the same solve given an explicit x0 = 0 (which
    evaluates b - A x0 in FP64) takes the same number of iterations (within
    one: |b|^2 from the norm kernel and from the residual kernel may differ
    in the last bit) and reaches the same solution to solver precision, and
    a zero column among them still converges at once."""
    print(f"[gpu] zero-guess start {prec}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(640)
    U = fields.su3_hot_start(rng, 2, L, 0.5)
    b = fields.random_fermion(rng, (2, *L, 4, 3))
    b[1] = 0.0
    kw = dict(solver_precision=prec, solver_recon=12) if prec != "f64" else dict(solver_precision="f64")
    th = SU3WilsonNf2(lattice_shape=L, mass=0.08, fermion_bc=bc, batch_size=2, dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=3000, **kw)
    x = np.asarray(th.solve_normal(U, b))
    st = th.last_solve_stats
    xg = np.asarray(th.solve_normal_with_guess(U, b, np.zeros_like(b)))
    stg = th.last_solve_stats
    assert st[1]["status"] == "converged" and st[1]["iters"] == 0 and np.all(x[1] == 0)
    assert stg[1]["status"] == "converged" and stg[1]["iters"] == 0
    assert st[0]["status"] == stg[0]["status"] == "converged"
    assert abs(st[0]["iters"] - stg[0]["iters"]) <= 1
    assert true_resid(L, 3, 0.08, bc, U[:1], b[:1], x[:1]) <= 1e-10
    assert np.linalg.norm(x[0] - xg[0]) / np.linalg.norm(xg[0]) < 1e-8


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_multi_rhs_same_alone_and_in_the_mix(prec):
    """Per-column masses and convergence masking: every column of a batched
    multi-RHS solve is bit-identical to the same column solved alone."""
    print(f"[gpu] multi-rhs {prec}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    rng = fields.rng_for(700)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    k = 5
    rhs = fields.random_fermion(rng, (1, k, *L, 4, 3))
    rhs[0, 2] = fields.point_sources(L, (1, 2, 3, 4))[7]
    masses = [0.02, 0.2, 0.05, 0.1, 0.03]
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=5000, solver_precision=prec)
    X = th.solve_normal_multi(U, rhs, masses=masses)
    stats = list(th.last_solve_stats)
    assert all(s["status"] == "converged" for s in stats), stats
    for j in range(k):
        xj = th.solve_normal_multi(U, rhs[:, j:j + 1], masses=[masses[j]])
        assert np.array_equal(xj[:, 0], X[:, j]), j
        assert th.last_solve_stats[0]["iters"] == stats[j]["iters"]
        assert true_resid(L, 3, masses[j], (1, 1, 1, -1), U, rhs[:, j], X[:, j]) <= 1e-10


@pytest.mark.parametrize("search", ["recurrence", "xpay"])
def test_mixed_maxiter_reports_the_true_residual(search):
    """A mixed solve stopped by cg_maxiter ends with one FP64 true-residual
    evaluation: status maxiter after exactly maxiter iterations, the x it
    returns holds every iteration's correction, and the reported true
    residual is the oracle's residual of that x."""
    print(f"[gpu] mixed maxiter {search}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(720)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    for maxiter in (7, 12):
        th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10,
                          cg_maxiter=maxiter, solver_precision="f16", solver_recon=12, reliable_delta=0.01,
                          solver_search=search)
        x = th.solve_normal(U, b)
        st = th.last_solve_stats[0]
        assert st["status"] == "maxiter" and st["iters"] == maxiter, st
        res = true_resid(L, 3, 0.05, bc, U, b, x)
        assert 1e-10 < res < 1.0
        assert abs(res - st["true_resid"]) <= 1e-6 * res, (res, st)


@pytest.mark.parametrize("search", ["recurrence", "xpay"])
@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_mixed_batch_with_zero_and_nan_columns(prec, search):
    """Columns that stop at once share a mixed solve with ordinary ones: a zero
    source converges at its first true residual (x = 0 exactly, no search
    direction is ever formed), a source holding a NaN stops with status nan,
    and the two Gaussian columns beside them are bit-identical to the same
    sources solved alone and meet the oracle's true residual 1e-10."""
    print(f"[gpu] zero and nan columns {prec} {search}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(710)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    rhs = fields.random_fermion(rng, (1, 4, *L, 4, 3))
    rhs[0, 1] = 0.0
    rhs[0, 2, 1, 2, 3, 4, 0, 2] = np.nan
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=3000, solver_precision=prec, solver_recon=12, reliable_delta=0.01,
                      solver_search=search, use_graph=True)
    X = np.asarray(th.solve_normal_multi(U, rhs))
    st = list(th.last_solve_stats)
    assert [s["status"] for s in st] == ["converged", "converged", "nan", "converged"], st
    assert np.all(X[:, 1] == 0) and st[1]["iters"] == 0
    for j in (0, 3):
        xj = np.asarray(th.solve_normal_multi(U, rhs[:, j:j + 1]))
        assert np.array_equal(xj[:, 0], X[:, j]), j
        assert th.last_solve_stats[0]["iters"] == st[j]["iters"]
        assert true_resid(L, 3, 0.05, bc, U, rhs[:, j], X[:, j]) <= 1e-10


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("march,slab,box", [(2, 0, 0), (8, 0, 0), (1, 2, 0), (1, 0, 1)])
def test_cg_with_marching_stencil(prec, march, slab, box):
    """The fused CG epilogues of the marching kernel and of the x1-slab site
    order drive CG to the same FP64 true residual."""
    print(f"[gpu] cg march={march} slab={slab} box={box} {prec}", flush=True)
    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (8, 4, 4, 8)
    rng = fields.rng_for(1000)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=4000, solver_precision=prec)
    for p in (nat.WL_F64, nat.WL_F32, nat.WL_F16):
        th._context().set_tuning(p, False, march, 0, True, slab, False, 0, box)
    x = th.solve_normal(U, b)
    assert th.last_solve_stats[0]["status"] == "converged"
    assert true_resid(L, 3, 0.05, (1, 1, 1, -1), U, b, x) <= 1e-10


@pytest.mark.parametrize("prec,nc", [("f16", 3), ("f32", 1)])
def test_theory_tune_then_solve(prec, nc):
    """th.tune installs launch shapes for the inner precision and the FP64
    update stencil; the operator after it agrees with the one before to
    rounding and the solves, one and two configurations, meet the FP64 true
    residual; a 4-RHS tuning leaves the one-RHS shapes alone and the 4-RHS
    solve meets the true residual too."""
    print(f"[gpu] theory tune {prec} nc={nc}", flush=True)
    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    L = (8, 8, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1090 + nc)
    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    U = fields.su3_hot_start(rng, 2, L, 0.5) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.5)
    b = fields.random_fermion(rng, (2, *L, 4, nc))
    for bs in (1, 2):
        th = cls(lattice_shape=L, mass=0.05, fermion_bc=bc, batch_size=bs, dtype=np.complex128, cg_tol=1e-10,
                 cg_maxiter=4000, solver_precision=prec, solver_recon=12 if nc == 3 else 18)
        D0 = th.apply_D(U[:bs], b[:bs])
        chosen = th.tune(U[:bs], nrhs=1)
        assert set(chosen) == {prec, "f64"}
        for name, t in chosen.items():
            got = th._context().get_tuning(nat.PREC[name], 1)
            assert all(got[k] == t[k] for k in ("march", "slab", "cols_outer", "lds_box")), (name, got, t)
        D1 = th.apply_D(U[:bs], b[:bs])
        assert np.linalg.norm(D1 - D0) <= 1e-14 * np.linalg.norm(D0)
        x = th.solve_normal(U[:bs], b[:bs])
        for j in range(bs):
            assert th.last_solve_stats[j]["status"] == "converged"
            assert true_resid(L, nc, 0.05, bc, U[j:j + 1], b[j:j + 1], x[j:j + 1]) <= 1e-10
    one = {p: th._context().get_tuning(nat.PREC[p], 1) for p in chosen}
    th.tune(U[:bs], nrhs=4)
    assert {p: th._context().get_tuning(nat.PREC[p], 1) for p in chosen} == one
    rhs = fields.random_fermion(rng, (bs, 4, *L, 4, nc))
    X = th.solve_normal_multi(U[:bs], rhs)
    for j in range(4):
        assert true_resid(L, nc, 0.05, bc, U[:bs], rhs[:, j], X[:, j]) <= 1e-10


@pytest.mark.parametrize("delta", [1e-4, 1e-3])
def test_small_delta_still_converges(delta):
    """A reliable-update trigger below what FP16 storage can follow is raised
    to the precision floor and backed by a forced update every 100 inner
    iterations, so the solve converges instead of running to maxiter."""
    print(f"[gpu] small delta {delta}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (8, 8, 8, 8)
    rng = fields.rng_for(1600)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.02, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=3000, solver_precision="f16", solver_recon=12, reliable_delta=delta)
    x = th.solve_normal(U, b)
    st = th.last_solve_stats[0]
    assert st["status"] == "converged", st
    assert st["iters"] < 1500, st
    assert true_resid(L, 3, 0.02, (1, 1, 1, -1), U, b, x) <= 1e-10



@pytest.mark.parametrize("march", [2, 4])
def test_multi_rhs_cg_with_marching_f16_stencil(march):
    """The fused CG epilogues of the FP16 two-RHS marching kernel (norm on D,
    the CG step on D^dag) drive every column of a batched mixed solve to the
    FP64 true residual."""
    print(f"[gpu] multi-rhs cg march={march}", flush=True)
    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (8, 4, 6, 4)
    rng = fields.rng_for(1700)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    k = 4
    rhs = fields.random_fermion(rng, (1, k, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=4000, solver_precision="f16", solver_recon=12, reliable_delta=0.01)
    th._context().set_tuning(nat.WL_F16, True, march, 2, True, 0)
    X = th.solve_normal_multi(U, rhs)
    assert all(s["status"] == "converged" for s in th.last_solve_stats), th.last_solve_stats
    for j in range(k):
        assert true_resid(L, 3, 0.05, (1, 1, 1, -1), U, rhs[:, j], X[:, j]) <= 1e-10


@pytest.mark.parametrize("plane,k", [(2, 1), (4, 3)])
def test_cg_with_plane_kernel(plane, k):
    """The fused CG epilogues of the LDS-staged plane kernel drive a mixed
    FP16 solve (one or several right-hand sides) to the FP64 true residual."""
    print(f"[gpu] cg plane={plane} k={k}", flush=True)
    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 8, 8, 16)
    rng = fields.rng_for(1800)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    rhs = fields.random_fermion(rng, (1, k, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=4000, solver_precision="f16", solver_recon=12, reliable_delta=0.01)
    th._context().set_tuning(nat.WL_F16, k > 1, 1, 1 if k > 1 else 0, True, 0, 0, 0, 0, plane)
    X = th.solve_normal_multi(U, rhs)
    assert all(s["status"] == "converged" for s in th.last_solve_stats), th.last_solve_stats
    for j in range(k):
        assert true_resid(L, 3, 0.05, (1, 1, 1, -1), U, rhs[:, j], X[:, j]) <= 1e-10


def test_mixed_multi_rhs_at_32x4():
    """The headline configuration at full size (32^4, eps 0.6, FP16 12-real
    inner, FP16 accumulator, delta 0.01), 12 right-hand sides with four
    different masses: every column converges, its true residual recomputed
    with the FP64 operator (apply_normal, independently of the solver's own
    statistics) is <= 1e-10, and a column solved alone agrees with its
    batched solution to solver accuracy."""
    print("[gpu] mixed 12-RHS at 32^4", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (32, 32, 32, 32)
    rng = fields.rng_for(680)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.02, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=5000, solver_precision="f16", solver_recon=12, reliable_delta=0.01, solver_accum="f16",
                      use_graph=True, poll_every=16)
    G = th.resident_gauge(fields.su3_hot_start(rng, 1, L, 0.6))
    b = torch.from_numpy(fields.random_fermion(rng, (1, 12, *L, 4, 3))).cuda()
    masses = [0.02, 0.05, 0.1, 0.2] * 3
    X = th.solve_normal_multi(G, b, masses=masses)
    st = th.last_solve_stats
    assert all(c["status"] == "converged" for c in st), st
    assert len({c["iters"] for c in st}) > 1
    # true residuals with the FP64 operator of each column's mass
    for m in sorted(set(masses)):
        th.mass = m
        cols = [j for j, mm in enumerate(masses) if mm == m]
        for j in cols:
            r = b[:, j] - th.apply_normal(G, X[:, j])
            res = float(torch.linalg.norm(r) / torch.linalg.norm(b[:, j]))
            assert res <= 1e-10, (j, m, res)
    th.mass = 0.02
    xa = th.solve_normal_multi(G, b[:, 3:4], masses=[masses[3]])
    assert float(torch.linalg.norm(xa[:, 0] - X[:, 3]) / torch.linalg.norm(X[:, 3])) < 1e-8


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_complex64_inputs_solve_in_fp64(prec):
    """complex64 links and sources (the reference's default dtype): the solve
    runs on the FP64 promotion of the given values and returns complex64 of
    the FP64 solution, so it equals the oracle's FP64 solve of the same
    (complex64) system to complex64 rounding, for the full and the even-odd
    normal equations."""
    print(f"[gpu] complex64 solve {prec}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(700)
    U = fields.su3_hot_start(rng, 1, L, 0.5).astype(np.complex64)
    b = fields.random_fermion(rng, (1, *L, 4, 3)).astype(np.complex64)
    kw = dict(solver_precision=prec, solver_recon=12) if prec != "f64" else dict(solver_precision="f64")
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=bc, dtype=np.complex64, cg_tol=1e-10, cg_maxiter=3000,
                      **kw)
    ora = WilsonOracle(L, 3, 0.1, fermion_bc=bc)
    U64, b64 = U.astype(np.complex128), b.astype(np.complex128)
    x = np.asarray(th.solve_normal(U, b))
    assert x.dtype == np.complex64 and th.last_solve_stats[0]["status"] == "converged"
    from oracle.wilson_oracle import cg

    xr, _ = cg(lambda v: ora.apply_normal(U64, v), b64, tol=1e-12, maxiter=3000)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-6
    xe = np.asarray(th.solve_eop_even_normal(U, b))
    assert xe.dtype == np.complex64 and th.last_solve_stats[0]["status"] == "converged"
    xer = ora.solve_eop_even_normal(U64, b64, tol=1e-12, maxiter=3000)
    assert np.linalg.norm(xe - xer) / np.linalg.norm(xer) < 1e-6


@pytest.mark.parametrize("L", [(24, 24, 24, 48), (32, 32, 32, 32)])
def test_u1_full_size_operator_and_mixed_solve(L):
    """U(1) at the benchmark sizes: D and D^dag match the FP64 oracle
    (1e-13), gamma5-hermiticity holds, and a batched FP16-inner 4-RHS solve
    (two masses, antiperiodic in x3, the U(1) default) reaches an oracle true
    residual <= 1e-10 in every column."""
    print(f"[gpu] u1 full size L={L}", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2
    from oracle.wilson_oracle import WilsonOracle

    rng = fields.rng_for(1400)
    bc = (1, 1, 1, -1)
    th = U1WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10, cg_maxiter=6000,
                     solver_precision="f16", reliable_delta=0.01, use_graph=True, poll_every=16)
    U = fields.u1_hot_start(rng, 1, L, 0.6)
    psi = fields.random_fermion(rng, (1, *L, 4, 1))
    ora = WilsonOracle(L, 1, 0.05, fermion_bc=bc)
    asnp = lambda v: v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)
    G = th.resident_gauge(U)
    for got, ref in ((th.apply_D(G, psi), ora.apply_D(U, psi)), (th.apply_Ddag(G, psi), ora.apply_Ddag(U, psi))):
        got = asnp(got)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13
    g5 = lambda v: np.stack([v[..., 2, :], -v[..., 3, :], v[..., 0, :], -v[..., 1, :]], axis=-2)
    ddag = asnp(th.apply_Ddag(G, psi))
    assert np.linalg.norm(ddag - g5(asnp(th.apply_D(G, g5(psi))))) / np.linalg.norm(ddag) < 1e-14
    b = fields.random_fermion(rng, (1, 4, *L, 4, 1))
    masses = [0.05, 0.2, 0.05, 0.2]
    X = asnp(th.solve_normal_multi(G, torch.from_numpy(b).cuda(), masses=masses))
    assert all(c["status"] == "converged" for c in th.last_solve_stats), th.last_solve_stats
    for j, m in enumerate(masses):
        assert true_resid(L, 1, m, bc, U, b[:, j], X[:, j]) <= 1e-10, (j, m)
