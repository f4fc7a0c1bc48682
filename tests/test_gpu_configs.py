"""GPU tests of the benchmark configurations at their full sizes, checked
against the FP64 operator (wl_apply D^dag D in FP64, independent of the
solver's own statistics):

* mixed precision at 32^4 (eps 0.6, m 0.02, delta 0.1): FP32 and FP16 inner
  operators with and without 12-real links, each solve's true residual
  <= 1e-10 and its error against the plain FP64 solve;
* batched solves at 24^3x48 (eps 0.6, m 0.03, FP32 inner): k in {1, 4, 12,
  24}, k = 12 being D^dag applied to the 12 unit point sources at one site
  (core/measurements.py:333-345), every column to 1e-10 and a column solved
  alone equal to its batched solution;
* serving over {8^4, 16^4, 24^3x48, 32^4} (4 fields each, eps 0.5, m in
  [0.02, 0.2], 1-12 right-hand sides): every result to 1e-10 and the same
  alone and in the mix.

Tolerance of "the same solution": two solves that both meet the relative
true residual 1e-10 differ by at most cond(D^dag D) 1e-10 in principle;
measured differences are ~1e-11 (mixed vs FP64) or exactly 0 (a column
alone vs batched with the same kernels), so 1e-8 catches any wrong column
while leaving room for rounding.

Fields come from the seeded device generator (fields.su3_hot_start_device),
the same distribution as the NumPy generator."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    import torch

    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _true_resid(th, G, b, x):
    """||b - D^dag D x|| / ||b|| with the FP64 operator on the GPU."""
    import torch

    r = b - th.apply_normal(G, x)
    return float(torch.linalg.norm(r) / torch.linalg.norm(b))


@pytest.fixture(scope="module")
def config3():
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (32, 32, 32, 32)
    U = fields.su3_hot_start_device(fields.SEED_BASE + 30, 1, L, 0.6)
    gen = torch.Generator(device="cuda").manual_seed(31)
    b = torch.complex(torch.randn((1, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64),
                      torch.randn((1, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64)) / np.sqrt(2)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.02, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=5000, solver_precision="f64", use_graph=True, poll_every=16)
    G = th.resident_gauge(U)
    x64 = th.solve_normal(G, b)
    assert th.last_solve_stats[0]["status"] == "converged"
    assert _true_resid(th, G, b, x64) <= 1e-10
    return th, G, b, x64


@pytest.mark.parametrize("prec,recon", [("f32", 18), ("f32", 12), ("f16", 18), ("f16", 12)])
def test_config3_mixed_precision_32x32x32x32(config3, prec, recon):
    print(f"[gpu] config3 32^4 {prec} r{recon} delta 0.1", flush=True)
    th, G, b, x64 = config3
    from jaxqft_amd import _native as nat

    x = th.solve_normal_multi(G, b[:, None], inner_prec=nat.PREC[prec], inner_recon=recon, delta=0.1)[:, 0]
    st = th.last_solve_stats[0]
    assert st["status"] == "converged", st
    assert st["reliable_updates"] >= 2, st
    res = _true_resid(th, G, b, x)
    assert res <= 1e-10, res
    err = _rel(x, x64)
    print(f"[gpu] config3 {prec} r{recon}: {st['iters']} iterations, {st['reliable_updates']} updates, "
          f"true residual {res:.2e}, error vs FP64 {err:.2e}", flush=True)
    assert err <= 1e-8, err


@pytest.fixture(scope="module")
def config4():
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (24, 24, 24, 48)
    U = fields.su3_hot_start_device(fields.SEED_BASE + 40, 1, L, 0.6)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.03, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, cg_tol=1e-10,
                      cg_maxiter=5000, solver_precision="f32", solver_recon=12, reliable_delta=0.01,
                      use_graph=True, poll_every=16)
    return th, th.resident_gauge(U), L


@pytest.mark.parametrize("k", [1, 4, 12, 24])
def test_config4_batched_24x24x24x48(config4, k):
    print(f"[gpu] config4 24^3x48 k={k}", flush=True)
    import torch

    from jaxqft_amd import fields

    th, G, L = config4
    if k == 12:
        site = (5, 17, 2, 33)
        src = torch.from_numpy(fields.point_sources(L, site)).cuda()
        # phi = D^dag e_(s,c) for the 12 (spin, colour) sources, measurements.py:338-345
        b = torch.stack([th.apply_Ddag(G, src[j:j + 1])[0] for j in range(12)])[None]
    else:
        gen = torch.Generator(device="cuda").manual_seed(400 + k)
        b = torch.complex(torch.randn((1, k, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64),
                          torch.randn((1, k, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64)) / np.sqrt(2)
    X = th.solve_normal_multi(G, b)
    st = list(th.last_solve_stats)
    assert len(st) == k and all(c["status"] == "converged" for c in st), st
    worst = 0.0
    for j in range(k):
        res = _true_resid(th, G, b[:, j], X[:, j])
        worst = max(worst, res)
        assert res <= 1e-10, (j, res)
    if k == 12:
        # D x = e_(s,c): the propagator columns solve the Dirac equation itself
        for j in (0, 7, 11):
            r = float(torch.linalg.norm(th.apply_D(G, X[:, j]) - src[j:j + 1]) / torch.linalg.norm(src[j]))
            assert r <= 1e-7, (j, r)
    j = k - 1
    alone = th.solve_normal_multi(G, b[:, j:j + 1].contiguous())
    assert th.last_solve_stats[0]["status"] == "converged"
    d = _rel(alone[:, 0], X[:, j])
    print(f"[gpu] config4 k={k}: iterations {[c['iters'] for c in st][:4]}..., worst true residual {worst:.2e}, "
          f"column {j} alone vs batched {d:.2e}", flush=True)
    assert d <= 1e-8, d


def test_config5_serving_full_lattices():
    """Continuous batching over the four serving lattices with resident fields
    (realtime=False: every request is offered at t = 0, so lanes fill and
    retire columns continuously)."""
    print("[gpu] config5 serving over 8^4, 16^4, 24^3x48, 32^4", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.serving import Server, poisson_requests
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    lattices = [(8, 8, 8, 8), (16, 16, 16, 16), (24, 24, 24, 48), (32, 32, 32, 32)]
    U = {}
    for i, L in enumerate(lattices):
        for f in range(4):
            U[(L, f)] = fields.su3_hot_start_device(fields.SEED_BASE + 500 + 10 * i + f, 1, L, 0.5)
    torch.cuda.synchronize()
    free, _ = torch.cuda.mem_get_info()
    reqs = poisson_requests(200.0, 16, lattices, 4, seed=fields.SEED_BASE + 55)
    assert len({r.key[0] for r in reqs}) == 4 and min(r.nrhs for r in reqs) >= 1 and max(r.nrhs for r in reqs) <= 12
    srv = Server(U, capacity=12, inner="f16", recon=12, tol=1e-10, delta=0.01, poll_every=8,
                 hbm_budget_bytes=0.5 * free)
    res = srv.run(reqs, realtime=False)
    assert res["requests"] == len(reqs) and res["all_converged"], res
    print(f"[gpu] config5 served {res['rhs']} right-hand sides at {res['rhs_per_s']} RHS/s", flush=True)
    theories = {}
    gauges = {}
    alone_done = set()
    for r in reqs:
        L = r.key[0]
        th = theories.get(L)
        if th is None:
            th = theories[L] = SU3WilsonNf2(lattice_shape=L, mass=r.mass, fermion_bc=(1, 1, 1, -1),
                                            dtype=np.complex128, cg_tol=1e-10, cg_maxiter=5000,
                                            solver_precision="f16", solver_recon=12, reliable_delta=0.01,
                                            use_graph=True, poll_every=8)
        th.mass = r.mass
        if r.key not in gauges:
            gauges[r.key] = th.resident_gauge(U[r.key])
        G = gauges[r.key]
        for j in range(r.nrhs):
            res_j = _true_resid(th, G, r.rhs[j:j + 1], r.x[j:j + 1])
            assert res_j <= 1e-10, (r.rid, L, j, res_j)
        if L not in alone_done:
            # one request per lattice solved on its own through the theory object
            alone_done.add(L)
            xa = th.solve_normal_multi(G, r.rhs[None], masses=[r.mass] * r.nrhs)[0]
            assert all(c["status"] == "converged" for c in th.last_solve_stats)
            d = _rel(xa, r.x)
            assert d <= 1e-8, (r.rid, L, d)
    assert alone_done == set(lattices)


def test_calls_on_two_streams_are_ordered():
    """Calls on one context share its device workspace; issued on different
    streams with no synchronization between them, they still run in the
    order issued (each call waits for the previous call's end): apply, solve
    and apply again alternate between two streams and give the results of
    the same calls made one after another on one stream."""
    print("[gpu] calls on two streams", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from ctypes import byref, c_void_p

    L = (16, 16, 16, 16)
    U = fields.su3_hot_start_device(fields.SEED_BASE + 60, 1, L, 0.5)
    gen = torch.Generator(device="cuda").manual_seed(61)
    psi = torch.complex(torch.randn((2, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64),
                        torch.randn((2, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64))
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, (1, 1, 1, -1))
    gauge = nat.Gauge(ctx, U, nat.WL_LAYOUT_BMXYIJ)
    lib = ctx._lib
    sp, keep = nat.solve_params(1e-10, 3000, nat.WL_STOP_TRUE_RESIDUAL, nat.WL_F64, 18, 0.0, 8, False)

    def run(streams):
        sa, sb = streams
        d1, x, d2 = torch.empty_like(psi), torch.empty_like(psi), torch.empty_like(psi)
        st = (nat.WlSolveStats * 2)()
        torch.cuda.synchronize()
        nat.check(lib.wl_apply(ctx.h, gauge.h, nat.WL_OP_DDAGD, c_void_p(psi.data_ptr()), c_void_p(d1.data_ptr()), 2,
                               nat.WL_C128, c_void_p(sa.cuda_stream)))
        nat.check(lib.wl_cg_normal(ctx.h, gauge.h, c_void_p(psi.data_ptr()), c_void_p(x.data_ptr()), None, 2,
                                   nat.WL_C128, byref(sp), st, c_void_p(sb.cuda_stream)))
        nat.check(lib.wl_apply(ctx.h, gauge.h, nat.WL_OP_DDAGD, c_void_p(x.data_ptr()), c_void_p(d2.data_ptr()), 2,
                               nat.WL_C128, c_void_p(sa.cuda_stream)))
        nat.check(lib.wl_apply(ctx.h, gauge.h, nat.WL_OP_D, c_void_p(psi.data_ptr()), c_void_p(d1.data_ptr()), 2,
                               nat.WL_C128, c_void_p(sb.cuda_stream)))
        torch.cuda.synchronize()
        assert all(s["status"] == "converged" for s in nat.stats_list(st, 2))
        return d1, x, d2

    one = torch.cuda.Stream()
    ref = run((one, one))
    got = run((torch.cuda.Stream(), torch.cuda.Stream()))
    for a, b in zip(ref, got):
        assert torch.equal(a, b)
    assert _rel(ref[2], psi) <= 1e-10
    del keep


def test_field_import_and_export_are_ordered_across_streams():
    """wl_field_import and wl_field_export join the context's call order: a
    field imported on one stream, run through a 12-RHS FP64 stencil (about
    2 ms at 32^4) on a second stream and exported at once on a third, with no
    synchronization between the calls, equals the same three calls on one
    stream, and an import on another stream waits for the stencil still
    reading that field."""
    print("[gpu] field import/export across streams", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    L, k = (32, 32, 32, 32), 12
    U = fields.su3_hot_start_device(fields.SEED_BASE + 62, 1, L, 0.6)
    gen = torch.Generator(device="cuda").manual_seed(63)
    psi = torch.complex(torch.randn((k, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64),
                        torch.randn((k, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64))
    ctx = nat.Context(L, 3, 1, 0.02, 1.0, (1, 1, 1, -1))
    gauge = nat.Gauge(ctx, U, nat.WL_LAYOUT_BMXYIJ)
    gauge.prepare(nat.WL_F64, 18)
    fi, fo = nat.Field(ctx, nat.WL_F64, k), nat.Field(ctx, nat.WL_F64, k)
    lib = ctx._lib
    from ctypes import c_void_p

    def run(sa, sb, sc):
        out = torch.empty_like(psi)
        torch.cuda.synchronize()
        nat.check(lib.wl_field_import(fi.h, c_void_p(psi.data_ptr()), nat.WL_C128, c_void_p(sa.cuda_stream)))
        nat.check(lib.wl_dslash(ctx.h, gauge.h, nat.WL_OP_D, 18, fi.h, fo.h, k, c_void_p(sb.cuda_stream)))
        nat.check(lib.wl_field_export(fo.h, c_void_p(out.data_ptr()), nat.WL_C128, c_void_p(sc.cuda_stream)))
        # overwrite the stencil's input from yet another stream: must not reach the stencil
        nat.check(lib.wl_field_import(fi.h, c_void_p(out.data_ptr()), nat.WL_C128, c_void_p(sa.cuda_stream)))
        torch.cuda.synchronize()
        return out

    one = torch.cuda.Stream()
    ref = run(one, one, one)
    got = run(torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Stream())
    assert torch.equal(ref, got)
    assert float(torch.linalg.norm(ref)) > 0.0


def test_profiled_solve_matches_graph_solve():
    """Kernel timing runs the iterations host-driven with events between the
    launches: the solve is bit-identical to the graph solve, and every
    iteration of both parities is timed."""
    print("[gpu] profiled solve", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from ctypes import byref

    L = (8, 8, 8, 16)
    U = fields.su3_hot_start_device(fields.SEED_BASE + 70, 1, L, 0.5)
    gen = torch.Generator(device="cuda").manual_seed(71)
    b = torch.complex(torch.randn((4, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64),
                      torch.randn((4, *L, 4, 3), generator=gen, device="cuda", dtype=torch.float64))
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, (1, 1, 1, -1))
    gauge = nat.Gauge(ctx, U, nat.WL_LAYOUT_BMXYIJ)
    stream = torch.cuda.Stream()
    out = []
    for prof in (False, True):
        ctx.set_profiling(prof)
        sp, keep = nat.solve_params(1e-10, 3000, nat.WL_STOP_TRUE_RESIDUAL, nat.WL_F16, 12, 0.01, 16, True,
                                    accum="f16")
        bf, xf = nat.Field(ctx, nat.WL_F64, 4), nat.Field(ctx, nat.WL_F64, 4)
        with torch.cuda.stream(stream):
            bf.import_(b)
            st = (nat.WlSolveStats * 4)()
            nat.check(ctx._lib.wl_cg_normal_field(ctx.h, gauge.h, bf.h, xf.h, 0, 4, byref(sp), st,
                                                  nat.c_void_p(stream.cuda_stream)))
            x = torch.empty_like(b)
            xf.export(x)
        stream.synchronize()
        stats = nat.stats_list(st, 4)
        assert all(s["status"] == "converged" for s in stats)
        out.append((x, [s["iters"] for s in stats], ctx.profile()))
        del keep
    ctx.set_profiling(False)
    assert torch.equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
    p = out[1][2]
    # every queued iteration is timed (blocks may run a few past the last column's convergence)
    assert max(out[1][1]) <= sum(p["iterations"]) <= max(out[1][1]) + 16 and min(p["iterations"]) > 0, p
    assert min(p["us_op"]) > 0 and min(p["us_op_dag"]) > 0 and p["us_xpay"] > 0, p
    assert sum(out[0][2]["iterations"]) == 0  # profiling off: nothing timed


def test_stream_probe_kernels():
    """The HBM streaming probes: the copy reproduces its source bit for bit,
    the write fills its pattern, the read leaves its source unchanged, for
    every unroll, both access hints, a single pass and a grid-stride loop,
    and a size that is not a multiple of one workgroup's step."""
    print("[gpu] stream probes", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from ctypes import c_void_p

    lib = nat.load()
    n = (3 << 20) + 48  # bytes: not a multiple of 4 KiB
    src = torch.randint(0, 2 ** 31 - 1, (n // 4,), dtype=torch.int32, device="cuda")
    keep = src.clone()
    pattern = torch.tensor([0x3f800000, 0, 0x3f800000, 0], dtype=torch.int32, device="cuda").repeat(n // 16)
    for ntm in (0, 1):
        for unroll in (1, 2, 4):
            for blocks in (0, 7):
                dst = torch.zeros_like(src)
                nat.check(lib.wl_stream_probe(nat.WL_STREAM_COPY, ntm, unroll, c_void_p(src.data_ptr()),
                                              c_void_p(dst.data_ptr()), n, blocks, None))
                torch.cuda.synchronize()
                assert torch.equal(dst, src), (ntm, unroll, blocks)
                nat.check(lib.wl_stream_probe(nat.WL_STREAM_WRITE, ntm, unroll, None, c_void_p(dst.data_ptr()), n,
                                              blocks, None))
                torch.cuda.synchronize()
                assert torch.equal(dst, pattern), (ntm, unroll, blocks)
                sink = torch.zeros(4, dtype=torch.int32, device="cuda")
                nat.check(lib.wl_stream_probe(nat.WL_STREAM_READ, ntm, unroll, c_void_p(src.data_ptr()),
                                              c_void_p(sink.data_ptr()), n, blocks, None))
                torch.cuda.synchronize()
                assert torch.equal(src, keep)
