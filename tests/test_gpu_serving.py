"""GPU tests of the continuous-batching server: every result reaches the FP64
true residual 1e-10 (oracle operator), and equals the same request solved
alone (relative difference <= 1e-7: both solutions satisfy the 1e-10
residual bound, so they agree to about cond(D^dag D) * 2e-10)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_serving_same_alone_and_in_the_mix():
    print("[gpu] serving mix vs alone", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.serving import Request, Server, fp64_queue
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from oracle.wilson_oracle import WilsonOracle

    bc = (1, 1, 1, -1)
    lattices = [(4, 4, 4, 4), (4, 4, 4, 8)]
    U = {}
    for i, L in enumerate(lattices):
        for f in range(2):
            U[(L, f)] = fields.su3_hot_start(fields.rng_for(1200 + 10 * i + f), 1, L, 0.5)
    rng = np.random.default_rng(5)
    reqs = []
    for rid in range(12):
        key = list(U)[int(rng.integers(len(U)))]
        k = int(rng.integers(1, 5))
        rhs = torch.from_numpy(fields.random_fermion(rng, (k, *key[0], 4, 3))).cuda()
        reqs.append(Request(rid=rid, key=key, mass=float(rng.uniform(0.02, 0.2)), rhs=rhs))
    srv = Server(U, capacity=6, inner="f16", recon=12, tol=1e-10, delta=0.01, poll_every=4)
    res = srv.run(reqs, realtime=False)
    assert res["requests"] == 12 and res["all_converged"], res
    for r in reqs:
        L = r.key[0]
        ora = WilsonOracle(L, 3, r.mass, fermion_bc=bc)
        x = r.x.cpu().numpy()
        b = r.rhs.cpu().numpy()
        for j in range(r.nrhs):
            res_j = np.linalg.norm(ora.apply_normal(U[r.key], x[j:j + 1]) - b[j:j + 1]) / np.linalg.norm(b[j])
            assert res_j <= 1e-10, (r.rid, j, res_j)
        th = SU3WilsonNf2(lattice_shape=L, mass=r.mass, fermion_bc=bc, dtype=np.complex128, cg_tol=1e-10,
                          cg_maxiter=5000, solver_precision="f16", solver_recon=12, reliable_delta=0.01)
        alone = th.solve_normal_multi(U[r.key], b[None])[0]
        assert np.linalg.norm(alone - x) / np.linalg.norm(alone) <= 1e-7, r.rid
    # the unbatched FP64 queue gives the same solutions
    ctxs = srv.backend.ctxs
    gauges = srv.backend.gauges
    base = [Request(rid=r.rid, key=r.key, mass=r.mass, rhs=r.rhs) for r in reqs]
    out = fp64_queue(base, gauges, ctxs, realtime=False)
    assert out["all_converged"]
    for r, q in zip(reqs, base):
        assert float(torch.linalg.norm(r.x - q.x) / torch.linalg.norm(q.x)) <= 1e-7


def test_serving_budget_admits_one_lane_at_a_time():
    print("[gpu] serving budget", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.serving import Request, Server

    L = (4, 4, 4, 4)
    U = {(L, f): fields.su3_hot_start(fields.rng_for(1300 + f), 1, L, 0.5) for f in range(3)}
    probe = Server({(L, 0): U[(L, 0)]}, capacity=4)
    one_lane = probe.lane_cost(L)
    budget = probe.resident_bytes * 3 + 1.5 * one_lane
    srv = Server(U, capacity=4, hbm_budget_bytes=budget, inner="f32", recon=18)
    rng = np.random.default_rng(6)
    reqs = [Request(rid=i, key=(L, i % 3), mass=0.1,
                    rhs=torch.from_numpy(fields.random_fermion(rng, (2, *L, 4, 3))).cuda()) for i in range(9)]
    res = srv.run(reqs, realtime=False)
    assert res["requests"] == 9 and res["all_converged"]
    assert res["true_resid_max"] <= 1e-10


def test_batch_slot_rules_and_readmission():
    """The batch engine through the C ABI: admitting into an occupied slot,
    retiring a free one, an out-of-range slot and a bad dtype are refused
    with an error; a slot retired and at once re-admitted with another mass
    and source solves the new system (true residual <= 1e-10 by the oracle
    at the new mass) while the other slot keeps its own."""
    print("[gpu] batch slot rules", flush=True)
    import torch
    from ctypes import byref, c_int, c_void_p

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1300)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = torch.from_numpy(fields.random_fermion(rng, (3, *L, 4, 3))).cuda()
    ctx = nat.Context(L, 3, 1, 0.1, 1.0, bc)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    lib = ctx._lib
    sp, keep = nat.solve_params(1e-10, 5000, nat.WL_STOP_TRUE_RESIDUAL, nat.WL_F16, 12, 0.01, 4, True)
    x = torch.zeros_like(b)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())  # b and x are ready before the batch stream uses them
    h = c_void_p()
    nat.check(lib.wl_batch_create(ctx.h, g.h, 2, byref(sp), c_void_p(s.cuda_stream), byref(h)))
    stats = (nat.WlSolveStats * 2)()
    state = (c_int * 2)()
    ptr = lambda t: c_void_p(t.data_ptr())
    try:
        nat.check(lib.wl_batch_admit(h, 0, ptr(b[0]), nat.WL_C128, 0.1))
        nat.check(lib.wl_batch_admit(h, 1, ptr(b[1]), nat.WL_C128, 0.2))
        assert lib.wl_batch_admit(h, 0, ptr(b[2]), nat.WL_C128, 0.1) == nat.WL_ERR_BAD_ARG
        assert b"occupied" in lib.wl_last_error()
        assert lib.wl_batch_admit(h, 2, ptr(b[2]), nat.WL_C128, 0.1) == nat.WL_ERR_BAD_ARG
        assert lib.wl_batch_admit(h, -1, ptr(b[2]), nat.WL_C128, 0.1) == nat.WL_ERR_BAD_ARG
        assert lib.wl_batch_admit(h, 0, ptr(b[2]), 99, 0.1) == nat.WL_ERR_BAD_ARG

        def run_until(slot):
            for _ in range(2000):
                nat.check(lib.wl_batch_step(h))
                nat.check(lib.wl_batch_query(h, stats, state))
                if state[slot] == 2:
                    return
            raise AssertionError(f"slot {slot} did not finish")

        run_until(0)
        nat.check(lib.wl_batch_retire(h, 0, ptr(x[0]), nat.WL_C128))
        assert lib.wl_batch_retire(h, 0, ptr(x[0]), nat.WL_C128) == nat.WL_ERR_BAD_ARG
        assert b"free" in lib.wl_last_error()
        # re-admit the slot just retired, with another mass and source
        nat.check(lib.wl_batch_admit(h, 0, ptr(b[2]), nat.WL_C128, 0.3))
        run_until(0)
        run_until(1)
        nat.check(lib.wl_batch_retire(h, 0, ptr(x[2]), nat.WL_C128))
        nat.check(lib.wl_batch_retire(h, 1, ptr(x[1]), nat.WL_C128))
        nat.check(lib.wl_batch_query(h, stats, state))
        assert list(state) == [0, 0]
        s.synchronize()
    finally:
        lib.wl_batch_destroy(h)
        del keep
    xs, bs = x.cpu().numpy(), b.cpu().numpy()
    for j, m in ((0, 0.1), (1, 0.2), (2, 0.3)):
        ora = WilsonOracle(L, 3, m, fermion_bc=bc)
        r = np.linalg.norm(ora.apply_normal(U, xs[j:j + 1]) - bs[j:j + 1]) / np.linalg.norm(bs[j])
        assert r <= 1e-10, (j, m, r)
