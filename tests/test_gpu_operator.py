"""GPU parity of the Wilson operator (through the C ABI) against the oracle's
golden fixtures and against the oracle directly, plus size-independent
properties at the benchmark sizes.  FP64 tolerance: 1e-13 relative to |D psi|."""

from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
TOL = 1e-13


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def theory_for(L, nc, mass, bc, **kw):
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    return cls(lattice_shape=L, mass=mass, fermion_bc=bc, dtype=np.complex128, **kw)


@pytest.mark.parametrize("name", ["op_su3_2x2x2x2", "op_su3_4x4x4x4", "op_su3_2x3x4x5", "op_u1_2x2x2x2",
                                  "op_u1_4x4x4x4"])
def test_operator_matches_golden(name):
    print(f"[gpu] operator golden {name}", flush=True)
    z = np.load(GOLD / f"{name}.npz")
    L = tuple(int(v) for v in z["L"])
    nc = z["psi"].shape[-1]
    th = theory_for(L, nc, float(z["mass"]), tuple(z["bc"]))
    U, psi = z["U"], z["psi"]
    assert rel(th.apply_D(U, psi), z["D"]) < TOL
    assert rel(th.apply_Ddag(U, psi), z["Ddag"]) < TOL
    assert rel(th.apply_g5D(U, psi), z["g5D"]) < TOL
    assert rel(th.apply_normal(U, psi), z["DdagD"]) < TOL
    assert rel(th.apply_offdiag(U, psi), z["K"]) < TOL
    assert rel(th.apply_offdiag_dagger(U, psi), z["Kdag"]) < TOL


@pytest.mark.parametrize("nc", [3, 1])
@pytest.mark.parametrize("L", [(8, 8, 8, 8), (4, 6, 2, 10), (1, 2, 3, 4)])
def test_operator_matches_oracle_random(nc, L):
    print(f"[gpu] operator oracle nc={nc} L={L}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    bc = (np.exp(0.7j), -1, 1, -1)
    rng = fields.rng_for(200 + nc)
    th = theory_for(L, nc, 0.05, bc, batch_size=2)
    U = fields.su3_hot_start(rng, 2, L, 0.6) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.6)
    psi = fields.random_fermion(rng, th.fermion_shape())
    ora = WilsonOracle(L, nc, 0.05, fermion_bc=bc)
    assert rel(th.apply_D(U, psi), ora.apply_D(U, psi)) < TOL
    assert rel(th.apply_Ddag(U, psi), ora.apply_Ddag(U, psi)) < TOL
    assert rel(th.apply_normal(U, psi), ora.apply_normal(U, psi)) < TOL


def test_layout_bxymij_and_complex64():
    print("[gpu] layout/complex64", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 2, 6)
    rng = fields.rng_for(300)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    ref = WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1)).apply_D(U, psi)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=(1, 1, 1, -1), layout="BXYMIJ", dtype=np.complex128)
    assert rel(th.apply_D(np.moveaxis(U, 1, 5).copy(), psi), ref) < TOL
    th32 = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=(1, 1, 1, -1))
    out = th32.apply_D(U.astype(np.complex64), psi.astype(np.complex64))
    assert out.dtype == np.complex64
    ref32 = WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1)).apply_D(U.astype(np.complex64), psi.astype(np.complex64))
    assert rel(out, ref32) < 1e-6


def test_torch_inputs_stay_on_device():
    print("[gpu] torch inputs", flush=True)
    import torch

    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 4)
    rng = fields.rng_for(301)
    U = fields.su3_hot_start(rng, 1, L, 0.4)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    th = theory_for(L, 3, 0.2, (1, 1, 1, -1))
    out = th.apply_Ddag(torch.from_numpy(U).cuda(), torch.from_numpy(psi).cuda())
    assert isinstance(out, torch.Tensor) and out.is_cuda
    assert rel(out.cpu().numpy(), WilsonOracle(L, 3, 0.2, fermion_bc=(1, 1, 1, -1)).apply_Ddag(U, psi)) < TOL


def test_l2_operator_with_prephased_links():
    """WilsonDiracOperator receives links with the phase baked in, like the
    reference's L2 operator (wilson.py:171-192)."""
    print("[gpu] L2 operator", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.fermions import WilsonDiracOperator
    from oracle import wilson_oracle as wo

    L = (4, 4, 4, 8)
    rng = fields.rng_for(302)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    Ul = wo.links_with_fermion_bc(U, (1, 1, 1, -1))
    op = WilsonDiracOperator(ndim=4, mass=0.3)
    m0 = 4.3
    assert rel(op.apply(Ul, psi, sign=-1), m0 * psi + wo.dslash_reference(Ul, psi, -1)) < TOL
    assert rel(op.apply_dagger(Ul, psi), m0 * psi + wo.dslash_reference(Ul, psi, +1)) < TOL
    assert rel(op.apply_dslash(Ul, psi, sign=-1), wo.dslash_optimized(Ul, psi, -1)) < TOL


@pytest.mark.parametrize("L", [(16, 16, 16, 16), (24, 24, 24, 48), (32, 32, 32, 32)])
def test_properties_at_benchmark_sizes(L):
    """gamma5-hermiticity, adjointness, linearity and gauge covariance at full
    size (no oracle: these hold exactly for the true operator)."""
    print(f"[gpu] properties L={L}", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    rng = fields.rng_for(400)
    th = SU3WilsonNf2(lattice_shape=L, mass=0.02, fermion_bc=(1, 1, 1, -1), dtype=np.complex128)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    G = th.resident_gauge(U)
    psi = torch.from_numpy(fields.random_fermion(rng, th.fermion_shape())).cuda()
    chi = torch.from_numpy(fields.random_fermion(rng, th.fermion_shape())).cuda()
    g5 = lambda v: torch.stack([v[..., 2, :], -v[..., 3, :], v[..., 0, :], -v[..., 1, :]], dim=-2)
    Dpsi = th.apply_D(G, psi)
    Ddag_psi = th.apply_Ddag(G, psi)
    assert float(torch.linalg.norm(Ddag_psi - g5(th.apply_D(G, g5(psi)))) / torch.linalg.norm(Ddag_psi)) < 1e-14
    lhs = torch.vdot(chi.flatten(), Dpsi.flatten())
    rhs = torch.vdot(th.apply_Ddag(G, chi).flatten(), psi.flatten())
    assert float(abs(lhs - rhs) / abs(lhs)) < 1e-12
    a, b = 0.3 - 1.2j, 2.1 + 0.4j
    lin = th.apply_D(G, a * psi + b * chi) - (a * Dpsi + b * th.apply_D(G, chi))
    assert float(torch.linalg.norm(lin) / torch.linalg.norm(Dpsi)) < 1e-14
    n = th.apply_normal(G, psi)
    q = torch.vdot(psi.flatten(), n.flatten())
    assert abs(float(q.imag)) < 1e-10 * float(q.real) and float(q.real) > 0


@pytest.mark.parametrize("march,slab", [(1, 0), (1, 1), (1, 2), (2, 0), (4, 0), (8, 0)])
@pytest.mark.parametrize("nc", [3, 1])
def test_march_variants_match_oracle(march, slab, nc):
    """Every stencil launch shape (x0 sites per thread; flat or x1-slab site
    order, whose slabs of 24 and 48 sites leave most of each tile padding)
    gives the oracle's D and D^dag."""
    print(f"[gpu] march={march} slab={slab} nc={nc}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (8, 4, 6, 4)
    bc = (-1, np.exp(0.3j), 1, -1)
    rng = fields.rng_for(800 + march)
    th = theory_for(L, nc, 0.07, bc, batch_size=2)
    th._context().set_tuning(nat.WL_F64, False, march, 0, True, slab)
    assert th._context().get_tuning(nat.WL_F64)["march"] == march
    assert th._context().get_tuning(nat.WL_F64)["slab"] == slab
    U = fields.su3_hot_start(rng, 2, L, 0.6) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.6)
    psi = fields.random_fermion(rng, th.fermion_shape())
    ora = WilsonOracle(L, nc, 0.07, fermion_bc=bc)
    assert rel(th.apply_D(U, psi), ora.apply_D(U, psi)) < TOL
    assert rel(th.apply_Ddag(U, psi), ora.apply_Ddag(U, psi)) < TOL
    assert rel(th.apply_g5D(U, psi), ora.apply_g5D(U, psi)) < TOL
    assert rel(th.apply_offdiag(U, psi), ora.apply_offdiag(U, psi)) < TOL


@pytest.mark.parametrize("box", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f16", 3e-3)])
def test_box_orders_match_oracle(box, prec, tol):
    """Every 4D workgroup box preset (16 x 8 x 4 x 16 lattice, so all fit) gives
    the oracle's D and D^dag, single-RHS and three right-hand sides."""
    print(f"[gpu] box={box} {prec}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (16, 8, 4, 16)
    bc = (1, -1, 1, -1)
    rng = fields.rng_for(1400 + box)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (3, *L, 4, 3))
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, bc)
    pc = nat.PREC[prec]
    ctx.set_tuning(pc, False, 1, 0, True, 0, False, 0, box)
    ctx.set_tuning(pc, True, 1, 1, True, 0, False, 0, box)
    assert ctx.get_tuning(pc, 1)["block_shape"] == box and ctx.get_tuning(pc, 3)["block_shape"] == box
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    ora = WilsonOracle(L, 3, 0.05, fermion_bc=bc)
    for nrhs in (1, 3):
        fi, fo = nat.Field(ctx, pc, nrhs), nat.Field(ctx, pc, nrhs)
        fi.import_(torch.from_numpy(psi[:nrhs].copy()).cuda())
        out = torch.empty((nrhs, *L, 4, 3), dtype=torch.complex128, device="cuda")
        for op, fn in ((nat.WL_OP_D, ora.apply_D), (nat.WL_OP_DDAG, ora.apply_Ddag)):
            nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, 18, fi.h, fo.h, nrhs, nat.c_void_p(nat.current_stream())))
            fo.export(out)
            got = out.cpu().numpy()
            for j in range(nrhs):
                assert rel(got[j:j + 1], fn(U, psi[j:j + 1])) < tol, (op, nrhs, j)


def test_box_shape_rejected_when_it_does_not_fit():
    """A box preset whose extents do not divide the lattice is refused."""
    from jaxqft_amd import _native as nat

    ctx = nat.Context((8, 4, 6, 4), 3, 1, 0.05, 1.0, (1, 1, 1, -1))
    with pytest.raises(RuntimeError, match="block_shape"):
        ctx.set_tuning(nat.WL_F64, False, 1, 0, True, 0, False, 0, 1)


@pytest.mark.parametrize("prec,recon,tol", [("f64", 12, 1e-13), ("f32", 18, 2e-6), ("f32", 12, 2e-6),
                                            ("f16", 18, 3e-3), ("f16", 12, 3e-3)])
@pytest.mark.parametrize("march", [1, 4])
def test_low_precision_stencils(prec, recon, tol, march):
    """Resident-field stencils in every storage precision against the FP64
    oracle: FP32 within 2e-6 and FP16 (per-site scaled) within 3e-3 relative;
    FP64 with 12-real reconstruction within 1e-13 (links are exactly SU(3))."""
    print(f"[gpu] stencil {prec} r{recon} march={march}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (8, 8, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(900)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, bc)
    pc = nat.PREC[prec]
    ctx.set_march(pc, march)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, pc, 1), nat.Field(ctx, pc, 1)
    fi.import_(torch.from_numpy(psi).cuda())
    ora = WilsonOracle(L, 3, 0.05, fermion_bc=bc)
    out = torch.empty((1, *L, 4, 3), dtype=torch.complex128, device="cuda")
    for op, ref in ((nat.WL_OP_D, ora.apply_D(U, psi)), (nat.WL_OP_DDAG, ora.apply_Ddag(U, psi))):
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, recon, fi.h, fo.h, 1, nat.c_void_p(nat.current_stream())))
        fo.export(out)
        assert rel(out.cpu().numpy(), ref) < tol, (op, rel(out.cpu().numpy(), ref))


@pytest.mark.parametrize("nc,recon", [(3, 12), (3, 18), (1, 18)])
@pytest.mark.parametrize("prec,tol", [("f32", 2e-6), ("f16", 3e-3)])
@pytest.mark.parametrize("nrhs,per_thread", [(1, 0), (2, 2)])
def test_low_precision_complex_bc(nc, recon, prec, tol, nrhs, per_thread):
    """Twisted (complex) boundary phases in directions 0 and 3 and an
    antiperiodic direction 2 through the reduced-precision stencils (the F16
    path applies real phases as a factor and complex ones in FP32), SU(3) and
    U(1), one and two right-hand sides per thread, against the FP64 oracle."""
    print(f"[gpu] complex bc {prec} nc={nc} r{recon} nrhs={nrhs}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 6, 4, 8)
    bc = (np.exp(0.4j), 1, -1, np.exp(-0.7j))
    rng = fields.rng_for(1300 + nc)
    U = fields.su3_hot_start(rng, 1, L, 0.5) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (nrhs, *L, 4, nc))
    ctx = nat.Context(L, nc, 1, 0.06, 1.0, bc)
    pc = nat.PREC[prec]
    ctx.set_tuning(pc, nrhs > 1, 1, per_thread, True, 0)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, pc, nrhs), nat.Field(ctx, pc, nrhs)
    fi.import_(torch.from_numpy(psi).cuda())
    ora = WilsonOracle(L, nc, 0.06, fermion_bc=bc)
    out = torch.empty((nrhs, *L, 4, nc), dtype=torch.complex128, device="cuda")
    for op, fn in ((nat.WL_OP_D, ora.apply_D), (nat.WL_OP_DDAG, ora.apply_Ddag)):
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, recon, fi.h, fo.h, nrhs, nat.c_void_p(nat.current_stream())))
        fo.export(out)
        got = out.cpu().numpy()
        for j in range(nrhs):
            ref = fn(U, psi[j:j + 1])
            assert rel(got[j:j + 1], ref) < tol, (op, j, rel(got[j:j + 1], ref))


@pytest.mark.parametrize("prec,recon,tol", [("f64", 18, 1e-13), ("f32", 12, 2e-6), ("f16", 12, 3e-3)])
@pytest.mark.parametrize("per_thread,slab,cols_outer,slab_x2", [(1, 0, 0, 0), (2, 0, 0, 0), (4, 0, 0, 0), (2, 2, 0, 0),
                                                              (4, 4, 0, 0), (1, 2, 1, 0), (2, 0, 1, 0), (1, 4, 2, 0),
                                                              (1, 0, 4, 0), (2, 0, 3, 0), (2, 4, 2, 0), (2, 4, 0, 2),
                                                              (1, 2, 3, 1), (2, 8, 0, 2), (4, 8, 1, 1)])
def test_multi_rhs_stencil(prec, recon, tol, per_thread, slab, cols_outer, slab_x2):
    """k right-hand sides sharing each link load (1, 2 or 4 per thread; flat,
    x1-slab or (x1, x2)-tile site order, the 8 x 2 x 6 tiles being narrower
    than a workgroup; the columns of a tile back to back, or blocks of 1, 2,
    3 (lowered to 2) or 4 columns sweeping the lattice in turn), two gauge
    configurations, against the oracle column by column."""
    print(f"[gpu] multi-rhs stencil {prec} k/thread={per_thread} slab={slab} x2={slab_x2} cols_outer={cols_outer}",
          flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 8, 4, 6)
    bc = (1, -1, 1, -1)
    k, B = 4, 2
    rng = fields.rng_for(1100)
    U = fields.su3_hot_start(rng, B, L, 0.5)
    psi = fields.random_fermion(rng, (B * k, *L, 4, 3))
    ctx = nat.Context(L, 3, B, 0.04, 1.0, bc)
    pc = nat.PREC[prec]
    ctx.set_tuning(pc, True, 1, per_thread, True, slab, cols_outer, slab_x2=slab_x2)
    assert ctx.get_tuning(pc, k)["rhs_per_thread"] == per_thread
    assert ctx.get_tuning(pc, k)["slab"] == slab
    assert ctx.get_tuning(pc, k)["slab_x2"] == (slab_x2 if slab else 0)
    assert ctx.get_tuning(pc, k)["cols_outer"] == cols_outer
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, pc, B * k), nat.Field(ctx, pc, B * k)
    fi.import_(torch.from_numpy(psi).cuda())
    out = torch.empty((B * k, *L, 4, 3), dtype=torch.complex128, device="cuda")
    ora = WilsonOracle(L, 3, 0.04, fermion_bc=bc)
    for op, fn in ((nat.WL_OP_D, ora.apply_D), (nat.WL_OP_DDAG, ora.apply_Ddag), (nat.WL_OP_G5D, ora.apply_g5D)):
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, recon, fi.h, fo.h, k, nat.c_void_p(nat.current_stream())))
        fo.export(out)
        got = out.cpu().numpy()
        for c in range(B * k):
            cfg = c // k
            ref = fn(U[cfg:cfg + 1], psi[c:c + 1])
            assert rel(got[c:c + 1], ref) < tol, (op, c)


@pytest.mark.parametrize("recon", [12, 18])
@pytest.mark.parametrize("L,march,cols_outer", [((8, 4, 6, 4), 2, False), ((8, 4, 6, 4), 8, True),
                                                ((4, 6, 10, 6), 4, False), ((6, 6, 10, 6), 3, False)])
def test_multi_rhs_march_stencil_f16(recon, L, march, cols_outer):
    """The FP16 two-RHS marching kernel (x0 neighbours from registers, the
    backward x0 hop carried between sites) on lattices whose x1*x2*x3 columns
    fill a partial workgroup, with a complex x0 boundary phase, every operator,
    against the oracle column by column."""
    print(f"[gpu] multi-rhs march f16 r{recon} L={L} march={march} cols_outer={cols_outer}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    bc = (np.exp(0.3j), -1, 1, -1)
    k, B = 4, 2
    rng = fields.rng_for(1200)
    U = fields.su3_hot_start(rng, B, L, 0.5)
    psi = fields.random_fermion(rng, (B * k, *L, 4, 3))
    ctx = nat.Context(L, 3, B, 0.04, 1.0, bc)
    pc = nat.WL_F16
    ctx.set_tuning(pc, True, march, 2, True, 0, cols_outer)
    t = ctx.get_tuning(pc, k)
    assert (t["march"], t["rhs_per_thread"], t["cols_outer"]) == (march, 2, cols_outer)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, pc, B * k), nat.Field(ctx, pc, B * k)
    fi.import_(torch.from_numpy(psi).cuda())
    out = torch.empty((B * k, *L, 4, 3), dtype=torch.complex128, device="cuda")
    ora = WilsonOracle(L, 3, 0.04, fermion_bc=bc)
    for op, fn in ((nat.WL_OP_D, ora.apply_D), (nat.WL_OP_DDAG, ora.apply_Ddag), (nat.WL_OP_G5D, ora.apply_g5D),
                   (nat.WL_OP_K, ora.apply_offdiag), (nat.WL_OP_KDAG, ora.apply_offdiag_dagger)):
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, recon, fi.h, fo.h, k, nat.c_void_p(nat.current_stream())))
        fo.export(out)
        got = out.cpu().numpy()
        for c in range(B * k):
            cfg = c // k
            ref = fn(U[cfg:cfg + 1], psi[c:c + 1])
            assert rel(got[c:c + 1], ref) < 3e-3, (op, c, rel(got[c:c + 1], ref))


@pytest.mark.parametrize("L,plane,nrhs,cols_outer", [((4, 8, 6, 16), 2, 1, 0), ((6, 4, 8, 32), 4, 1, 0),
                                                     ((4, 6, 10, 12), 1, 1, 0), ((4, 6, 10, 12), 6, 3, 0),
                                                     ((8, 4, 4, 32), 2, 4, 2)])
def test_plane_kernel_f16(L, plane, nrhs, cols_outer):
    """The LDS-staged plane kernel (x2/x3 neighbours from a double-buffered LDS
    tile with periodic halo rows, x1 neighbours carried in registers): tiles
    with idle threads, tiles covering all of x2, complex boundary phases in
    every direction, one or several right-hand sides, every operator, against
    the oracle."""
    print(f"[gpu] plane f16 L={L} plane={plane} nrhs={nrhs} cols_outer={cols_outer}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    bc = (np.exp(0.4j), np.exp(-0.7j), -1, np.exp(1.1j))
    B = 2
    rng = fields.rng_for(1300)
    U = fields.su3_hot_start(rng, B, L, 0.5)
    psi = fields.random_fermion(rng, (B * nrhs, *L, 4, 3))
    ctx = nat.Context(L, 3, B, 0.04, 1.0, bc)
    pc = nat.WL_F16
    ctx.set_tuning(pc, nrhs > 1, 1, 1 if nrhs > 1 else 0, True, 0, cols_outer, 0, 0, plane)
    assert ctx.get_tuning(pc, nrhs)["plane"] == plane
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    fi, fo = nat.Field(ctx, pc, B * nrhs), nat.Field(ctx, pc, B * nrhs)
    fi.import_(torch.from_numpy(psi).cuda())
    out = torch.empty((B * nrhs, *L, 4, 3), dtype=torch.complex128, device="cuda")
    ora = WilsonOracle(L, 3, 0.04, fermion_bc=bc)
    for recon in (12, 18):
        for op, fn in ((nat.WL_OP_D, ora.apply_D), (nat.WL_OP_DDAG, ora.apply_Ddag), (nat.WL_OP_G5D, ora.apply_g5D),
                       (nat.WL_OP_K, ora.apply_offdiag), (nat.WL_OP_KDAG, ora.apply_offdiag_dagger)):
            nat.check(ctx._lib.wl_dslash(ctx.h, g.h, op, recon, fi.h, fo.h, nrhs, nat.c_void_p(nat.current_stream())))
            fo.export(out)
            got = out.cpu().numpy()
            for c in range(B * nrhs):
                cfg = c // nrhs
                ref = fn(U[cfg:cfg + 1], psi[c:c + 1])
                assert rel(got[c:c + 1], ref) < 3e-3, (recon, op, c, rel(got[c:c + 1], ref))


def test_plane_kernel_rejected_shapes_fall_back():
    """A plane march that does not divide L1 is rejected; a lattice whose x3
    rows do not fit a workgroup keeps the other kernels."""
    import torch

    from jaxqft_amd import _native as nat

    ctx = nat.Context((4, 6, 4, 8), 3, 1, 0.1, 1.0, (1, 1, 1, -1))
    with pytest.raises(nat.NativeError, match="plane"):
        ctx.set_tuning(nat.WL_F16, False, 1, 0, True, 0, 0, 0, 0, 4)
    # a tile x2 extent must divide L2; it is reported only with a slab order
    with pytest.raises(nat.NativeError, match="slab_x2"):
        ctx.set_tuning(nat.WL_F16, True, 1, 2, True, 2, 0, slab_x2=3)
    ctx.set_tuning(nat.WL_F16, True, 1, 2, True, 2, 0, slab_x2=2)
    assert ctx.get_tuning(nat.WL_F16, 4)["slab_x2"] == 2
    ctx.set_tuning(nat.WL_F16, True, 1, 2, True, 0, 0, slab_x2=2)
    assert ctx.get_tuning(nat.WL_F16, 4)["slab_x2"] == 0


def test_mass_change_reaches_raw_and_resident_links_alike():
    """After theory.mass changes, D on a resident gauge and D on the raw link
    array both use the new mass (and match the oracle at that mass)."""
    print("[gpu] mass change", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(690)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    th = SU3WilsonNf2(lattice_shape=L, mass=0.1, fermion_bc=bc, dtype=np.complex128)
    G = th.resident_gauge(U)
    th.apply_D(G, psi)
    th.mass = 0.3
    ref = WilsonOracle(L, 3, 0.3, fermion_bc=bc).apply_D(U, psi)
    a = np.asarray(th.apply_D(G, psi))
    b = np.asarray(th.apply_D(U, psi))
    assert np.linalg.norm(a - ref) / np.linalg.norm(ref) < 1e-13
    assert np.array_equal(a, b)


def test_link_encoding_built_on_one_stream_is_ready_on_another():
    """A low-precision link encoding built on a busy stream (queued behind
    long work) and used at once from an unrelated non-blocking stream: the
    build waits for its own kernels, so the stencil on the other stream reads
    finished links and matches the oracle (one racing the build would read
    unconverted memory)."""
    print("[gpu] link encoding across streams", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (16, 16, 16, 16)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(910)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, bc)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    ref = WilsonOracle(L, 3, 0.05, fermion_bc=bc).apply_D(U, psi)
    busy, other = torch.cuda.Stream(), torch.cuda.Stream()
    m = torch.randn(4096, 4096, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(busy):
        for _ in range(12):
            m = m @ m
            m = m / m.norm()
        g.prepare(nat.WL_F16, 12)
    with torch.cuda.stream(other):
        fi, fo = nat.Field(ctx, nat.WL_F16, 1), nat.Field(ctx, nat.WL_F16, 1)
        fi.import_(torch.from_numpy(psi).cuda())
        nat.check(ctx._lib.wl_dslash(ctx.h, g.h, nat.WL_OP_D, 12, fi.h, fo.h, 1, nat.c_void_p(nat.current_stream())))
        out = torch.empty((1, *L, 4, 3), dtype=torch.complex128, device="cuda")
        fo.export(out)
        other.synchronize()
    assert rel(out.cpu().numpy(), ref) < 3e-3


@pytest.mark.filterwarnings("ignore:The CUDA Graph is empty")
def test_link_encoding_not_built_under_graph_capture():
    """Building a link encoding inside a graph capture would only record its
    kernels: the library refuses and says to prepare it first; a prepared
    encoding is then used without a build."""
    print("[gpu] link encoding under capture", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    rng = fields.rng_for(911)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, (1, 1, 1, -1))
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    err = None
    with torch.cuda.stream(s):
        graph.capture_begin()
        try:
            g.prepare(nat.WL_F32, 12)
        except nat.NativeError as e:
            err = str(e)
        finally:
            graph.capture_end()
    torch.cuda.current_stream().wait_stream(s)
    assert err is not None and "wl_gauge_prepare" in err, err
    g.prepare(nat.WL_F32, 12)
    torch.cuda.synchronize()


def test_foreign_and_mismatched_handles_rejected():
    """Handles from another context, fields of the wrong precision or column
    count and in == out are refused with WL_ERR_BAD_ARG before any launch;
    the context stays usable afterwards."""
    print("[gpu] foreign handles", flush=True)
    import torch
    from ctypes import byref, c_void_p

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(920)
    U = torch.from_numpy(fields.su3_hot_start(rng, 1, L, 0.5)).cuda()
    psi = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    out = torch.empty_like(psi)
    c1, c2 = nat.Context(L, 3, 1, 0.1, 1.0, bc), nat.Context(L, 3, 1, 0.1, 1.0, bc)
    g1, g2 = nat.Gauge(c1, U, nat.WL_LAYOUT_BMXYIJ), nat.Gauge(c2, U, nat.WL_LAYOUT_BMXYIJ)
    lib = c1._lib
    B = nat.WL_ERR_BAD_ARG
    h = c_void_p(nat.current_stream())
    p = lambda t: c_void_p(t.data_ptr())
    sp, keep = nat.solve_params(1e-10, 100)
    st = (nat.WlSolveStats * 1)()
    assert lib.wl_apply(c1.h, g2.h, nat.WL_OP_D, p(psi), p(out), 1, nat.WL_C128, h) == B
    assert lib.wl_cg_normal(c1.h, g2.h, p(psi), p(out), None, 1, nat.WL_C128, byref(sp), st, h) == B
    assert lib.wl_eo_apply(c1.h, g2.h, nat.WL_EO_K_OE, 1, p(psi), p(out), 1, nat.WL_C128, h) == B
    assert lib.wl_pf_force(c1.h, g2.h, p(psi), p(psi), p(out), nat.WL_LAYOUT_BMXYIJ, nat.WL_C128, h) == B
    bt = c_void_p()
    assert lib.wl_batch_create(c1.h, g2.h, 2, byref(sp), h, byref(bt)) == B
    f64a, f64b = nat.Field(c1, nat.WL_F64, 1), nat.Field(c1, nat.WL_F64, 1)
    f32 = nat.Field(c1, nat.WL_F32, 1)
    f64x2 = nat.Field(c1, nat.WL_F64, 2)
    foreign = nat.Field(c2, nat.WL_F64, 1)
    assert lib.wl_dslash(c1.h, g1.h, nat.WL_OP_D, 18, f64a.h, f32.h, 1, h) == B      # precision
    assert lib.wl_dslash(c1.h, g1.h, nat.WL_OP_D, 18, f64a.h, f64x2.h, 1, h) == B    # columns
    assert lib.wl_dslash(c1.h, g1.h, nat.WL_OP_D, 18, f64a.h, f64a.h, 1, h) == B     # in == out
    assert lib.wl_dslash(c1.h, g1.h, nat.WL_OP_D, 18, f64a.h, foreign.h, 1, h) == B  # foreign field
    assert lib.wl_dslash(c1.h, g1.h, nat.WL_OP_D, 18, f64x2.h, f64x2.h, 2, h) == B
    assert lib.wl_dslash(c1.h, g1.h, nat.WL_OP_D, 18, f64a.h, f64b.h, 2, h) == B     # batch * nrhs
    assert lib.wl_dslash(c1.h, g1.h, 99, 18, f64a.h, f64b.h, 1, h) == B              # op
    assert lib.wl_cg_normal_field(c1.h, g1.h, f32.h, f64a.h, 0, 1, byref(sp), st, h) == B
    # the context still works
    nat.check(lib.wl_apply(c1.h, g1.h, nat.WL_OP_D, p(psi), p(out), 1, nat.WL_C128, h))
    ref = torch.empty_like(psi)
    nat.check(lib.wl_apply(c2.h, g2.h, nat.WL_OP_D, p(psi), p(ref), 1, nat.WL_C128, h))
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    del keep


def test_links_with_fermion_bc_on_device():
    """_links_with_fermion_bc of a device tensor equals the NumPy result, and
    stays on the device."""
    print("[gpu] links with fermion bc (torch)", flush=True)
    import torch

    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (4, 4, 4, 8)
    th = SU3WilsonNf2(lattice_shape=L, fermion_bc=(np.exp(0.3j), 1, 1, -1), batch_size=2)
    U = fields.su3_hot_start(fields.rng_for(1890), 2, L, 0.5)
    got = th._links_with_fermion_bc(torch.from_numpy(U).cuda())
    assert got.is_cuda
    np.testing.assert_allclose(got.cpu().numpy(), th._links_with_fermion_bc(U), rtol=1e-15, atol=0)
