"""GPU tests of the LDS box kernel (FP16 storage, SU(3), 12-real links): a
workgroup owns one 256-site box of one configuration and walks all of its
columns two per pass, with the box's forward links staged in LDS and in-box
neighbours read from LDS.

* every operator against the oracle, on boxes that wrap a whole axis inside
  themselves (the box spans L_mu), on box grids split into 8 XCD regions and
  on grids too small to split, with an odd number of columns (a pass with one
  column), two configurations and complex boundary phases in every direction;
* 18-real links keep the other kernels (same results as with the box unset);
* mixed-precision CG through the box kernel with FP16, FP32 and FP64
  accumulators and per-column masses: every column meets the FP64 true
  residual by the oracle, and a column solved alone equals its batched
  solution bit for bit (a column's arithmetic never depends on its partner);
* presets that do not divide the lattice, or a U(1) context, are refused.

Tolerance: FP16 storage against the FP64 oracle, 3e-3 relative (half
precision rounds each stored component to 2^-11 of the site's largest)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BC = (np.exp(0.4j), np.exp(-0.7j), -1, np.exp(1.1j))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("L,box,nrhs,B", [
    ((4, 4, 4, 16), 1, 3, 2),   # 4 boxes per configuration: no region split
    ((4, 4, 8, 16), 1, 4, 1),   # 8 boxes, split (2, 2, 2, 1)
    ((2, 4, 4, 8), 2, 2, 1),    # one box spanning the lattice: every axis wraps inside it
    ((4, 8, 8, 16), 3, 5, 1),   # x0 spans the box (wraps), split over x1..x3
    ((8, 8, 8, 16), 1, 2, 1),   # 32 boxes, split (2, 2, 2, 1)
])
def test_box_kernel_against_oracle(L, box, nrhs, B):
    print(f"[gpu] box kernel L={L} box={box} nrhs={nrhs} B={B}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    rng = fields.rng_for(1700 + box)
    U = fields.su3_hot_start(rng, B, L, 0.5)
    psi = fields.random_fermion(rng, (B * nrhs, *L, 4, 3))
    ctx = nat.Context(L, 3, B, 0.04, 1.0, BC)
    pc = nat.WL_F16
    ctx.set_tuning(pc, nrhs > 1, 1, 2 if nrhs > 1 else 0, True, lds_box=box)
    assert ctx.get_tuning(pc, nrhs)["lds_box"] == box
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, pc, B * nrhs), nat.Field(ctx, pc, B * nrhs)
    fi.import_(torch.from_numpy(psi).cuda())
    out = torch.empty((B * nrhs, *L, 4, 3), dtype=torch.complex128, device="cuda")
    ora = WilsonOracle(L, 3, 0.04, fermion_bc=BC)
    for op, fn in ((nat.WL_OP_D, ora.apply_D), (nat.WL_OP_DDAG, ora.apply_Ddag), (nat.WL_OP_G5D, ora.apply_g5D),
                   (nat.WL_OP_K, ora.apply_offdiag), (nat.WL_OP_KDAG, ora.apply_offdiag_dagger)):
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, 12, fi.h, fo.h, nrhs, nat.c_void_p(nat.current_stream())))
        fo.export(out)
        got = out.cpu().numpy()
        for c in range(B * nrhs):
            cfg = c // nrhs
            ref = fn(U[cfg:cfg + 1], psi[c:c + 1])
            assert rel(got[c:c + 1], ref) < 3e-3, (op, c, rel(got[c:c + 1], ref))


def test_box_kernel_leaves_18_real_links_to_the_other_kernels():
    """With 18-real links the box setting does not apply: results equal
    those with the box unset, bit for bit."""
    print("[gpu] box kernel, 18-real links", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    L = (4, 4, 8, 16)
    rng = fields.rng_for(1710)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (4, *L, 4, 3))
    outs = []
    for box in (0, 1):
        ctx = nat.Context(L, 3, 1, 0.04, 1.0, BC)
        ctx.set_tuning(nat.WL_F16, True, 1, 2, True, lds_box=box)
        g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
        fi, fo = nat.Field(ctx, nat.WL_F16, 4), nat.Field(ctx, nat.WL_F16, 4)
        fi.import_(torch.from_numpy(psi).cuda())
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, nat.WL_OP_D, 18, fi.h, fo.h, 4, nat.c_void_p(nat.current_stream())))
        out = torch.empty((4, *L, 4, 3), dtype=torch.complex128, device="cuda")
        fo.export(out)
        outs.append(out)
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("accum", ["f16", "f32", "f64"])
def test_box_kernel_mixed_cg(accum):
    """Mixed-precision CG (FP16 12-real inner, FP64 reliable updates) with
    the inner stencils on the box kernel: five columns with three masses
    (a pass with one column, columns converging at different iterations so
    the gate drops them pass by pass) all meet the FP64 true residual, by the
    oracle; the last column solved alone equals its batched solution bit for
    bit."""
    print(f"[gpu] box kernel mixed CG accum={accum}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 8, 16)
    bc = (1, 1, 1, -1)
    k = 5
    masses = [0.1, 0.3, 0.1, 0.6, 0.3]
    rng = fields.rng_for(1720)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (k, *L, 4, 3))
    ctx = nat.Context(L, 3, 1, 0.1, 1.0, bc)
    for multi in (False, True):
        ctx.set_tuning(nat.WL_F16, multi, 1, 2 if multi else 0, True, lds_box=1)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    side = torch.cuda.Stream()  # the solves capture HIP graphs

    def solve(bb, ms):
        n = bb.shape[0]
        sp, keep = nat.solve_params(1e-10, 3000, nat.WL_STOP_TRUE_RESIDUAL, nat.WL_F16, 12, 0.1, 8, True,
                                    masses=ms, accum=accum)
        bt = torch.from_numpy(bb).cuda()
        xt = torch.empty_like(bt)
        st = (nat.WlSolveStats * n)()
        side.wait_stream(torch.cuda.current_stream())
        nat.check(ctx._lib.wl_cg_normal(ctx.h, g.h, nat.c_void_p(bt.data_ptr()), nat.c_void_p(xt.data_ptr()), None, n,
                                        nat.WL_C128, nat.byref(sp), st, nat.c_void_p(side.cuda_stream)))
        side.synchronize()
        del keep
        return xt.cpu().numpy(), nat.stats_list(st, n)

    x, stats = solve(b, masses)
    assert all(s["status"] == "converged" for s in stats), stats
    for j in range(k):
        ora = WilsonOracle(L, 3, masses[j], fermion_bc=bc)
        r = b[j:j + 1] - ora.apply_normal(U, x[j:j + 1])
        assert np.linalg.norm(r) / np.linalg.norm(b[j]) <= 1e-10, (j, stats[j])
    xa, sa = solve(b[k - 1:k], masses[k - 1:k])
    assert sa[0]["iters"] == stats[k - 1]["iters"]
    assert np.array_equal(xa[0], x[k - 1])


def test_box_presets_refused_where_they_do_not_fit():
    from jaxqft_amd import _native as nat

    ctx = nat.Context((4, 4, 4, 8), 3, 1, 0.1, 1.0, (1, 1, 1, -1))
    with pytest.raises(nat.NativeError, match="lds_box"):
        ctx.set_tuning(nat.WL_F16, True, 1, 2, True, lds_box=1)  # x3 extent 16 does not divide 8
    with pytest.raises(nat.NativeError, match="lds_box"):
        ctx.set_tuning(nat.WL_F16, True, 1, 2, True, lds_box=4)
    ctx.set_tuning(nat.WL_F16, True, 1, 2, True, lds_box=2)  # (2, 4, 4, 8) fits
    assert ctx.get_tuning(nat.WL_F16, 2)["lds_box"] == 2
    assert ctx.get_tuning(nat.WL_F64, 2)["lds_box"] == 0
    u1 = nat.Context((4, 4, 4, 16), 1, 1, 0.1, 1.0, (1, 1, 1, -1))
    with pytest.raises(nat.NativeError, match="lds_box"):
        u1.set_tuning(nat.WL_F16, True, 1, 2, True, lds_box=1)


def test_tuner_times_the_box_presets():
    """The stencil tuner includes the box kernel among the FP16 12-real
    multi-RHS candidates (each preset that divides the lattice) and installs
    whichever is fastest."""
    print("[gpu] tuner with box presets", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.autotune import tune

    L = (8, 8, 8, 16)
    rng = fields.rng_for(1730)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (4, *L, 4, 3))
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, (1, 1, 1, -1))
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, nat.WL_F16, 4), nat.Field(ctx, nat.WL_F16, 4)
    fi.import_(torch.from_numpy(psi).cuda())
    t = tune(ctx, g, nat.WL_F16, 12, fi, fo, 4, reps=3)
    boxes = sorted(c["lds_box"] for c in t["ranked"] if c["lds_box"])
    assert boxes == [1, 2, 3], boxes
    assert ctx.get_tuning(nat.WL_F16, 4)["lds_box"] == t["lds_box"]
