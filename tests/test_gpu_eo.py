"""GPU parity of the even-odd (checkerboard) path against the oracle: the
parity hopping blocks and Schur operators (every op, normalized or not, SU(3)
and U(1), twisted boundary phases, several configurations and right-hand
sides), the even-odd normal CG in FP64 and mixed precision (true residual of
the Schur normal equations <= 1e-10 by the FP64 oracle), the per-column
independence of batched solves, and at 16^4 the reconstruction of the full
solution of D x = b from the even-odd solve (checked with the full operator)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-13


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def theory(L, nc, mass, bc, batch=1, **kw):
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    return cls(lattice_shape=L, mass=mass, fermion_bc=bc, batch_size=batch, dtype=np.complex128, **kw)


@pytest.mark.parametrize("nc", [3, 1])
def test_eo_operators_match_oracle(nc):
    print(f"[gpu] eo operators nc={nc}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 6, 4, 8)
    bc = (np.exp(0.4j), 1, -1, -1)
    rng = fields.rng_for(1500 + nc)
    U = fields.su3_hot_start(rng, 2, L, 0.6) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.6)
    psi = fields.random_fermion(rng, (2, *L, 4, nc))
    th = theory(L, nc, 0.07, bc, batch=2)
    ora = WilsonOracle(L, nc, 0.07, fermion_bc=bc)
    for name in ("apply_K_oe", "apply_K_eo", "apply_Kdag_oe", "apply_Kdag_eo"):
        assert rel(getattr(th, name)(U, psi), getattr(ora, name)(U, psi)) < TOL, name
    for norm in (False, True):
        for name in ("apply_eo_schur_even", "apply_eo_schur_even_dagger", "apply_eo_schur_odd"):
            got = getattr(th, name)(U, psi, normalized=norm)
            assert rel(got, getattr(ora, name)(U, psi, normalized=norm)) < TOL, (name, norm)
        got = th.apply_eo_schur_even_compact(U, psi, daggered=True, normalized=norm)
        assert rel(got, ora.apply_eo_schur_even_dagger(U, psi, normalized=norm)) < TOL
    # outputs vanish on the other parity
    odd = ora.parity_masks()[1]
    assert np.abs(np.asarray(th.apply_eo_schur_even(U, psi)) * odd).max() == 0.0


def test_eo_schur_odd_dagger_is_adjoint():
    """The odd Schur dagger (no reference counterpart) is the adjoint of S_o."""
    print("[gpu] eo odd dagger", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    rng = fields.rng_for(1510)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    th = theory(L, 3, 0.05, (1, 1, 1, -1))
    chi = fields.random_fermion(rng, (1, *L, 4, 3))
    psi = fields.random_fermion(rng, (1, *L, 4, 3))
    s_psi = np.asarray(th.apply_eo_schur_odd(U, psi, normalized=True))
    sd_chi = np.asarray(th._eo(U, chi, nat.WL_EO_SCHUR_O_DAG, True))
    from oracle.wilson_oracle import WilsonOracle

    ora = WilsonOracle(L, 3, 0.05, fermion_bc=(1, 1, 1, -1))
    co, po = ora.project_odd(chi), ora.project_odd(psi)
    lhs = np.vdot(co, s_psi)
    rhs = np.vdot(sd_chi, po)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("nrhs", [1, 3, 4])
def test_eo_apply_multi_rhs_columns(nrhs):
    """k right-hand sides per configuration through the C ABI, column by column
    (an even k shares each link load between two right-hand sides)."""
    print(f"[gpu] eo multi-rhs nrhs={nrhs}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 6, 4)
    bc = (1, -1, 1, -1)
    rng = fields.rng_for(1520 + nrhs)
    U = fields.su3_hot_start(rng, 2, L, 0.5)
    psi = fields.random_fermion(rng, (2 * nrhs, *L, 4, 3))
    ctx = nat.Context(L, 3, 2, 0.05, 1.0, bc)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    ora = WilsonOracle(L, 3, 0.05, fermion_bc=bc)
    pin = torch.from_numpy(psi).cuda()
    out = torch.empty_like(pin)
    nat.check(ctx._lib.wl_eo_apply(ctx.h, g.h, nat.WL_EO_SCHUR_E_DAG, 1, nat.c_void_p(pin.data_ptr()),
                                   nat.c_void_p(out.data_ptr()), nrhs, nat.WL_C128,
                                   nat.c_void_p(nat.current_stream())))
    got = out.cpu().numpy()
    for col in range(2 * nrhs):
        cfg = col // nrhs
        ref = ora.apply_eo_schur_even_dagger(U[cfg:cfg + 1], psi[col:col + 1], normalized=True)
        assert rel(got[col:col + 1], ref) < TOL, col


def _schur_normal_resid(ora, U, b, x):
    se = ora.apply_eo_schur_even(U, x, normalized=True)
    r = ora.apply_eo_schur_even_dagger(U, se, normalized=True) - ora.project_even(b)
    return float(np.linalg.norm(r) / np.linalg.norm(ora.project_even(b)))


@pytest.mark.parametrize("policy", [
    dict(solver_precision="f64"),
    dict(solver_precision="f64", use_graph=True),
    dict(solver_precision="f32", solver_recon=12),
    dict(solver_precision="f16", solver_recon=12, reliable_delta=0.01),
    dict(solver_precision="f16", solver_recon=18, solver_accum="f16"),
    dict(solver_precision="f16", solver_recon=12, reliable_delta=0.01, solver_search="xpay"),
])
@pytest.mark.parametrize("nc", [3, 1])
def test_eo_cg_matches_oracle(policy, nc):
    print(f"[gpu] eo cg nc={nc} {policy}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1530 + nc)
    U = fields.su3_hot_start(rng, 1, L, 0.5) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, *L, 4, nc))
    th = theory(L, nc, 0.06, bc, cg_tol=1e-10, cg_maxiter=3000, **policy)
    x = np.asarray(th.solve_eop_even_normal(U, b))
    st = th.last_solve_stats[0]
    assert st["status"] == "converged", st
    ora = WilsonOracle(L, nc, 0.06, fermion_bc=bc)
    res = _schur_normal_resid(ora, U, b, x)
    assert res <= 1e-10, (res, st)
    assert abs(res - st["true_resid"]) <= 1e-3 * res + 1e-14
    assert np.abs(x * ora.parity_masks()[1]).max() == 0.0
    xr = ora.solve_eop_even_normal(U, b, tol=1e-12, maxiter=3000)
    assert rel(x, xr) < 1e-8


def test_eo_cg_guess_and_batch():
    print("[gpu] eo cg guess/batch", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 4)
    rng = fields.rng_for(1540)
    U = fields.su3_hot_start(rng, 2, L, 0.5)
    b = fields.random_fermion(rng, (2, *L, 4, 3))
    th = theory(L, 3, 0.1, (1, 1, 1, -1), batch=2, cg_tol=1e-10, cg_maxiter=2000)
    x = np.asarray(th.solve_eop_even_normal(U, b))
    it0 = [s["iters"] for s in th.last_solve_stats]
    ora = WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1))
    for k in range(2):
        assert _schur_normal_resid(ora, U[k:k + 1], b[k:k + 1], x[k:k + 1]) <= 1e-10
    x2 = np.asarray(th.solve_eop_even_normal_with_guess(U, b, x))
    assert all(s["iters"] <= 2 for s in th.last_solve_stats), th.last_solve_stats
    assert rel(x2, x) < 1e-9 and max(it0) > 5


def test_eo_multi_rhs_alone_vs_mix_bitwise():
    """Columns of a batched even-odd solve with different masses converge
    independently: each equals its own single solve bit for bit."""
    print("[gpu] eo multi-rhs alone vs mix", flush=True)
    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    rng = fields.rng_for(1550)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    rhs = fields.random_fermion(rng, (1, 4, *L, 4, 3))
    masses = [0.05, 0.2, 0.1, 0.3]
    th = theory(L, 3, 0.1, (1, 1, 1, -1), cg_tol=1e-10, cg_maxiter=3000, solver_precision="f16", solver_recon=12)
    X = np.asarray(th.solve_eop_even_normal_multi(U, rhs, masses=masses))
    stats = th.last_solve_stats
    assert len({s["iters"] for s in stats}) > 1
    for j in range(4):
        xj = np.asarray(th.solve_eop_even_normal_multi(U, rhs[:, j:j + 1], masses=[masses[j]]))
        assert np.array_equal(xj[:, 0], X[:, j]), j
        assert th.last_solve_stats[0]["iters"] == stats[j]["iters"]


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_eo_reconstructs_full_solution_16x4(prec):
    """At 16^4: D x = b solved through the Schur system,
    x_e = (S^dag S)^-1 S^dag (b_e - K_eo b_o / m0) / m0 (normalized S), and
    x_o = (b_o - K_oe x_e) / m0, has |D x - b| / |b| <= 1e-9, with the full
    operator applied on the GPU; the even-odd solve needs fewer iterations
    than the full normal-equation solve of the same system."""
    print(f"[gpu] eo reconstruct 16^4 {prec}", flush=True)
    from jaxqft_amd import fields

    L = (16, 16, 16, 16)
    rng = fields.rng_for(1560)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    th = theory(L, 3, 0.05, (1, 1, 1, -1), cg_tol=1e-11, cg_maxiter=4000, solver_precision=prec, solver_recon=12)
    m0 = th.mass + 4.0
    par = (np.indices(L).sum(axis=0) % 2)[None, ..., None, None]
    be, bo = b * (par == 0), b * (par == 1)
    c = be - np.asarray(th.apply_K_eo(U, bo)) / m0
    rhs = np.asarray(th.apply_eo_schur_even_dagger(U, c / m0, normalized=True))
    xe = np.asarray(th.solve_eop_even_normal(U, rhs))
    it_eo = th.last_solve_stats[0]["iters"]
    xo = (bo - np.asarray(th.apply_K_oe(U, xe))) / m0
    x = xe + xo
    r = np.asarray(th.apply_D(U, x)) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-9
    th.solve_normal(U, np.asarray(th.apply_Ddag(U, b)))
    it_full = th.last_solve_stats[0]["iters"]
    print(f"[gpu] iterations eo {it_eo} full {it_full}", flush=True)
    assert it_eo < it_full


def _set_eo_slab_all(ctx, eo_slab):
    from jaxqft_amd import _native as nat
    from jaxqft_amd.autotune import set_eo_slab

    for pc in (nat.WL_F64, nat.WL_F32, nat.WL_F16):
        for multi in (False, True):
            set_eo_slab(ctx, pc, multi, eo_slab)
            assert ctx.get_tuning(pc, 2 if multi else 1)["eo_slab"] == eo_slab


@pytest.mark.parametrize("eo_slab", [1, 2, 3, 6])
def test_eo_slab_order_operators(eo_slab):
    """The even-odd hop walking slabs of x1 planes x0-fastest (slabs of 1, 2,
    3 and all 6 planes; 16 compact sites a plane, so most threads of a
    workgroup idle) gives the oracle's blocks and Schur operators."""
    print(f"[gpu] eo slab operators eo_slab={eo_slab}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 6, 4, 8)
    bc = (np.exp(0.4j), 1, -1, -1)
    rng = fields.rng_for(1560 + eo_slab)
    U = fields.su3_hot_start(rng, 2, L, 0.6)
    psi = fields.random_fermion(rng, (2, *L, 4, 3))
    th = theory(L, 3, 0.07, bc, batch=2)
    _set_eo_slab_all(th._context(), eo_slab)
    ora = WilsonOracle(L, 3, 0.07, fermion_bc=bc)
    for name in ("apply_K_oe", "apply_K_eo", "apply_Kdag_oe", "apply_Kdag_eo"):
        assert rel(getattr(th, name)(U, psi), getattr(ora, name)(U, psi)) < TOL, name
    for name in ("apply_eo_schur_even", "apply_eo_schur_even_dagger", "apply_eo_schur_odd"):
        assert rel(getattr(th, name)(U, psi), getattr(ora, name)(U, psi)) < TOL, name


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_eo_slab_order_multi_rhs_cg(prec):
    """Batched even-odd CG with the slab order of the hop: each column meets
    the true-residual target of its Schur normal equations and equals the
    flat-order solve to solver precision."""
    print(f"[gpu] eo slab cg {prec}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 8, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1570)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    rhs = fields.random_fermion(rng, (1, 4, *L, 4, 3))
    masses = [0.05, 0.1, 0.2, 0.3]
    kw = dict(solver_precision=prec, solver_recon=12) if prec != "f64" else dict(solver_precision="f64")
    th0 = theory(L, 3, 0.1, bc, cg_tol=1e-10, cg_maxiter=3000, **kw)
    X0 = np.asarray(th0.solve_eop_even_normal_multi(U, rhs, masses=masses))
    th = theory(L, 3, 0.1, bc, cg_tol=1e-10, cg_maxiter=3000, **kw)
    _set_eo_slab_all(th._context(), 4)
    X = np.asarray(th.solve_eop_even_normal_multi(U, rhs, masses=masses))
    for j, m in enumerate(masses):
        st = th.last_solve_stats[j]
        assert st["status"] == "converged", st
        ora = WilsonOracle(L, 3, m, fermion_bc=bc)
        res = _schur_normal_resid(ora, U, rhs[:, j], X[:, j])
        assert res <= 1e-10, (j, res, st)
        assert rel(X[:, j], X0[:, j]) < 1e-8, j


def test_eo_slab_must_divide_l1():
    print("[gpu] eo slab rejected", flush=True)
    from jaxqft_amd import _native as nat
    from jaxqft_amd.autotune import set_eo_slab

    th = theory((4, 6, 4, 8), 3, 0.1, (1, 1, 1, -1))
    ctx = th._context()
    with pytest.raises(nat.NativeError, match="eo_slab"):
        set_eo_slab(ctx, nat.WL_F64, False, 4)
    assert ctx.get_tuning(nat.WL_F64, 1)["eo_slab"] == 0


def test_tune_eo_installs_a_valid_slab():
    """The even-odd tuner times each slab that divides L1 on capped solves,
    installs the fastest, leaves the full-lattice launch shape alone, and a
    solve after it still meets the true-residual target."""
    print("[gpu] tune_eo", flush=True)
    import torch
    from jaxqft_amd import _native as nat, fields
    from jaxqft_amd.autotune import tune_eo
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 8, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1580)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, 4, *L, 4, 3))
    th = theory(L, 3, 0.1, bc, cg_tol=1e-10, cg_maxiter=3000, solver_precision="f16", solver_recon=12)
    ctx = th._context()
    before = ctx.get_tuning(nat.WL_F16, 4)
    before64 = ctx.get_tuning(nat.WL_F64, 4)
    g = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
    bt = torch.from_numpy(b.reshape(4, *L, 4, 3)).cuda()
    xt = torch.empty_like(bt)
    with pytest.raises(ValueError):
        tune_eo(ctx, g, bt[:2], xt, 4, nat.WL_F16, 12)
    t = tune_eo(ctx, g, bt, xt, 4, nat.WL_F16, 12, slabs=(0, 1, 2, 3, 4, 8), iters=6)
    assert set(t["ms"]) == {"0", "1", "2", "4", "8"}
    assert t["eo_slab"] in (0, 1, 2, 4, 8)
    after = ctx.get_tuning(nat.WL_F16, 4)
    assert after["eo_slab"] == t["eo_slab"]
    assert {k: v for k, v in after.items() if k != "eo_slab"} == {k: v for k, v in before.items() if k != "eo_slab"}
    # the FP64 hops of the reliable updates: tuned the same way, installed for FP64 at this nrhs
    assert set(t["update_ms"]) == {"0", "2", "4", "8"} and t["update_eo_slab"] in (0, 2, 4, 8)
    after64 = ctx.get_tuning(nat.WL_F64, 4)
    assert after64["eo_slab"] == t["update_eo_slab"]
    assert {k: v for k, v in after64.items() if k != "eo_slab"} == {k: v for k, v in before64.items() if k != "eo_slab"}
    assert "update_eo_slab" not in tune_eo(ctx, g, bt, xt, 4, nat.WL_F16, 12, slabs=(0,), iters=6, update_slabs=())
    assert ctx.get_tuning(nat.WL_F64, 4)["eo_slab"] == t["update_eo_slab"]
    X = np.asarray(th.solve_eop_even_normal_multi(U, b, masses=[0.1] * 4))
    ora = WilsonOracle(L, 3, 0.1, fermion_bc=bc)
    for j in range(4):
        assert _schur_normal_resid(ora, U, b[:, j], X[:, j]) <= 1e-10, j


@pytest.mark.parametrize("L", [(16, 16, 16, 16), (32, 32, 32, 32)])
def test_eo_properties_at_benchmark_sizes(L):
    """At full size (no oracle: these hold exactly for the true operators):
    K_eo and K^dag_oe, K_oe and K^dag_eo are adjoint pairs between the
    parities; S_e^dag is the adjoint of S_e and S_o^dag of S_o; the Schur
    complements are linear and vanish off their parity; and the even-odd
    factorisation D = [[m0, K_eo], [K_oe, m0]] reproduces D on the full
    lattice (D psi = m0 psi + K_eo psi_o + K_oe psi_e)."""
    print(f"[gpu] eo properties L={L}", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    rng = fields.rng_for(1590)
    th = theory(L, 3, 0.02, (1, 1, 1, -1))
    G = th.resident_gauge(fields.su3_hot_start(rng, 1, L, 0.6))
    psi = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    chi = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    x = torch.arange(L[0]).view(-1, 1, 1, 1) + torch.arange(L[1]).view(1, -1, 1, 1) + \
        torch.arange(L[2]).view(1, 1, -1, 1) + torch.arange(L[3]).view(1, 1, 1, -1)
    even = ((x % 2) == 0).to(torch.complex128).cuda().view(1, *L, 1, 1)
    odd = 1.0 - even

    def dot(a, b):
        return torch.vdot(a.flatten(), b.flatten())

    def close(a, b, tol):
        return float(abs(a - b) / max(abs(a), abs(b))) < tol

    pe, po, ce, co = even * psi, odd * psi, even * chi, odd * chi
    # <chi_o, K_oe psi_e> = <K^dag_eo chi_o, psi_e> and <chi_e, K_eo psi_o> = <K^dag_oe chi_e, psi_o>
    assert close(dot(co, th._eo(G, pe, nat.WL_EO_K_OE)), dot(th._eo(G, co, nat.WL_EO_KDAG_EO), pe), 1e-12)
    assert close(dot(ce, th._eo(G, po, nat.WL_EO_K_EO)), dot(th._eo(G, ce, nat.WL_EO_KDAG_OE), po), 1e-12)
    for norm in (False, True):
        se = th._eo(G, pe, nat.WL_EO_SCHUR_E, norm)
        assert close(dot(ce, se), dot(th._eo(G, ce, nat.WL_EO_SCHUR_E_DAG, norm), pe), 1e-12)
        so = th._eo(G, po, nat.WL_EO_SCHUR_O, norm)
        assert close(dot(co, so), dot(th._eo(G, co, nat.WL_EO_SCHUR_O_DAG, norm), po), 1e-12)
        assert float(torch.abs(se * odd).max()) == 0.0 and float(torch.abs(so * even).max()) == 0.0
        a, b = 0.3 - 1.2j, 2.1 + 0.4j
        lin = th._eo(G, a * pe + b * ce, nat.WL_EO_SCHUR_E, norm) - (a * se + b * th._eo(G, ce, nat.WL_EO_SCHUR_E, norm))
        assert float(torch.linalg.norm(lin) / torch.linalg.norm(se)) < 1e-14
    m0 = 0.02 + 4.0
    full = m0 * psi + th._eo(G, po, nat.WL_EO_K_EO) + th._eo(G, pe, nat.WL_EO_K_OE)
    d = th.apply_D(G, psi)
    assert float(torch.linalg.norm(full - d) / torch.linalg.norm(d)) < 1e-14


def test_eo_mixed_multi_rhs_at_32x4():
    """The even-odd solve at full size (32^4, eps 0.6, FP16 12-real inner,
    delta 0.005, slabs of 4 x1 planes in the hop), 12 right-hand sides with
    three masses: every column converges, its Schur normal-equation true
    residual |S^dag S x - P_e b| / |P_e b| recomputed with the FP64 even-odd
    operators (at the column's mass) is <= 1e-10, and x vanishes on the odd
    sites."""
    print("[gpu] eo mixed 12-RHS at 32^4", flush=True)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.autotune import set_eo_slab

    L = (32, 32, 32, 32)
    rng = fields.rng_for(1600)
    th = theory(L, 3, 0.02, (1, 1, 1, -1), cg_tol=1e-10, cg_maxiter=5000, solver_precision="f16", solver_recon=12,
                reliable_delta=0.005, solver_accum="f16", use_graph=True, poll_every=16)
    G = th.resident_gauge(fields.su3_hot_start(rng, 1, L, 0.6))
    for pc in (nat.WL_F64, nat.WL_F32, nat.WL_F16):
        set_eo_slab(th._context(), pc, True, 4)
    b = torch.from_numpy(fields.random_fermion(rng, (1, 12, *L, 4, 3))).cuda()
    masses = [0.02, 0.05, 0.1] * 4
    X = th.solve_eop_even_normal_multi(G, b, masses=masses)
    st = th.last_solve_stats
    assert all(c["status"] == "converged" for c in st), st
    x = torch.arange(L[0]).view(-1, 1, 1, 1) + torch.arange(L[1]).view(1, -1, 1, 1) + \
        torch.arange(L[2]).view(1, 1, -1, 1) + torch.arange(L[3]).view(1, 1, 1, -1)
    even = ((x % 2) == 0).to(torch.complex128).cuda().view(1, *L, 1, 1)
    assert float(torch.abs(X * (1.0 - even).unsqueeze(1)).max()) == 0.0
    for m in sorted(set(masses)):
        th.mass = m
        for j in [j for j, mm in enumerate(masses) if mm == m]:
            be = even * b[:, j]
            s = th.apply_eo_schur_even(G, X[:, j], normalized=True)
            r = th.apply_eo_schur_even_dagger(G, s, normalized=True) - be
            res = float(torch.linalg.norm(r) / torch.linalg.norm(be))
            assert res <= 1e-10, (j, m, res)
