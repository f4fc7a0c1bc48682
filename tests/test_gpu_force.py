"""GPU parity of the analytic pseudofermion force (su3_wilson_nf2.py:1556-1597)
against the oracle (itself pinned by finite differences of the action): both
layouts, SU(3) and U(1), twisted boundary phases, two configurations, complex64
and complex128 inputs, the even-odd force, and at 16^4 the identity
dS/d eps = sum_links Re tr(H F) for U -> exp(eps H) U with the action
evaluated by the GPU solver."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def theory(L, nc, mass, bc, batch=1, layout="BMXYIJ", **kw):
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    return cls(lattice_shape=L, mass=mass, fermion_bc=bc, batch_size=batch, dtype=np.complex128, layout=layout, **kw)


@pytest.mark.parametrize("layout", ["BMXYIJ", "BXYMIJ"])
@pytest.mark.parametrize("nc", [3, 1])
def test_force_matches_oracle(nc, layout):
    print(f"[gpu] force nc={nc} {layout}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 6, 4, 8)
    bc = (np.exp(0.3j), 1, -1, -1)
    rng = fields.rng_for(1800 + nc)
    U = fields.su3_hot_start(rng, 2, L, 0.6, layout=layout) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.6,
                                                                                                 layout=layout)
    X = fields.random_fermion(rng, (2, *L, 4, nc))
    Y = fields.random_fermion(rng, (2, *L, 4, nc))
    th = theory(L, nc, 0.05, bc, batch=2, layout=layout)
    ora = WilsonOracle(L, nc, 0.05, fermion_bc=bc, layout=layout)
    ref = ora.pseudofermion_force_from_solution(U, X, Y)
    assert rel(th.pseudofermion_force_from_solution(U, X, Y), ref) < 1e-13
    got64 = th.pseudofermion_force_from_solution(U, X.astype(np.complex64), Y.astype(np.complex64))
    assert rel(got64, ref) < 2e-6


@pytest.mark.parametrize("layout", ["BMXYIJ", "BXYMIJ"])
@pytest.mark.parametrize("L", [(3, 5, 2, 7), (2, 2, 2, 2), (4, 4, 4, 16)])
def test_force_ragged_blocks_and_wraps(L, layout):
    """The force over 64-site workgroups: a last workgroup with idle lanes
    (210 sites), a lattice smaller than one workgroup, odd and tiny extents
    and a 16-site x3 row, with twisted phases, all match the oracle."""
    print(f"[gpu] force ragged {L} {layout}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    bc = (1, np.exp(-0.7j), 1, -1)
    rng = fields.rng_for(1810 + sum(L))
    U = fields.su3_hot_start(rng, 2, L, 0.6, layout=layout)
    X = fields.random_fermion(rng, (2, *L, 4, 3))
    Y = fields.random_fermion(rng, (2, *L, 4, 3))
    th = theory(L, 3, 0.05, bc, batch=2, layout=layout)
    ora = WilsonOracle(L, 3, 0.05, fermion_bc=bc, layout=layout)
    assert rel(th.pseudofermion_force_from_solution(U, X, Y), ora.pseudofermion_force_from_solution(U, X, Y)) < 1e-13


@pytest.mark.parametrize("nc", [3, 1])
def test_eo_force_matches_oracle(nc):
    print(f"[gpu] eo force nc={nc}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1810 + nc)
    U = fields.su3_hot_start(rng, 1, L, 0.5) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.5)
    ora = WilsonOracle(L, nc, 0.07, fermion_bc=bc)
    Xe = ora.project_even(fields.random_fermion(rng, (1, *L, 4, nc)))
    Ye = ora.apply_eo_schur_even(U, Xe, normalized=True)
    th = theory(L, nc, 0.07, bc)
    got = th.eop_pseudofermion_force_from_solution(U, Xe, Ye)
    assert rel(got, ora.eop_pseudofermion_force_from_solution(U, Xe, Ye)) < 1e-12


@pytest.mark.parametrize("nc", [3, 1])
def test_eo_monomial_force(nc):
    """The even-odd monomial's force from phi (su3_wilson_nf2.py:566-579): x_e
    solves S^dag S x_e = phi_e to the solver tolerance (oracle operator), the
    force equals the oracle's force from (x_e, S x_e); the eo_split form gives
    the same force, and a second call from the chronological guess (a history
    separate from the standard monomial's) starts at the solution."""
    print(f"[gpu] eo monomial force nc={nc}", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    bc = (1, np.exp(0.4j), 1, -1)
    rng = fields.rng_for(1850 + nc)
    U = fields.su3_hot_start(rng, 1, L, 0.5) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.5)
    ora = WilsonOracle(L, nc, 0.07, fermion_bc=bc)
    phi = ora.project_even(fields.random_fermion(rng, (1, *L, 4, nc)))
    th = theory(L, nc, 0.07, bc, cg_tol=1e-12, cg_maxiter=3000)
    F = th.eop_pseudofermion_force(U, phi)
    Xe = th.solve_eop_even_normal(U, phi)
    Ye = ora.apply_eo_schur_even(U, Xe, normalized=True)
    res = ora.apply_eo_schur_even_dagger(U, Ye, normalized=True) - phi
    assert np.linalg.norm(res) <= 2e-12 * np.linalg.norm(phi)
    assert rel(F, ora.eop_pseudofermion_force_from_solution(U, Xe, Ye)) < 1e-12
    ths = theory(L, nc, 0.07, bc, cg_tol=1e-12, cg_maxiter=3000, solver_form="eo_split", solver_kind="gmres")
    assert rel(ths.eop_pseudofermion_force(U, phi), F) < 1e-12
    thg = theory(L, nc, 0.07, bc, cg_tol=1e-12, cg_maxiter=3000, use_solver_guess=True)
    assert rel(thg.eop_pseudofermion_force(U, phi), F) < 1e-12
    it_cold = thg.last_solve_stats[0]["iters"]
    assert it_cold > 10 and len(thg.eop_solver_guess.xs) == 1 and not thg.solver_guess.xs
    assert rel(thg.eop_pseudofermion_force(U, phi), F) < 1e-10
    assert thg.last_solve_stats[0]["iters"] <= 2
    thg.clear_solver_guess()
    assert not thg.eop_solver_guess.xs


def test_eo_monomial_force_batched_torch_resident():
    """The even-odd monomial force on a batch of two configurations given as
    torch tensors on the device, with the links resident (resident_gauge) and
    the launch shapes tuned (th.tune): each configuration's force equals the
    force computed for it alone from numpy inputs."""
    print("[gpu] eo monomial force batched torch", flush=True)
    import torch

    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1880)
    U = fields.su3_hot_start(rng, 2, L, 0.5)
    kw = dict(cg_tol=1e-12, cg_maxiter=3000, solver_precision="f32", solver_recon=12)
    th2 = theory(L, 3, 0.08, bc, batch=2, **kw)
    phi = th2.project_even(fields.random_fermion(rng, (2, *L, 4, 3)))
    G = th2.resident_gauge(torch.from_numpy(U).cuda())
    th2.tune(G, nrhs=1)
    F2 = th2.eop_pseudofermion_force(G, torch.from_numpy(phi).cuda())
    assert isinstance(F2, torch.Tensor) and F2.is_cuda
    for j in range(2):
        F1 = theory(L, 3, 0.08, bc, **kw).eop_pseudofermion_force(U[j:j + 1], phi[j:j + 1])
        assert rel(F2[j:j + 1].cpu().numpy(), F1) < 1e-9, j


def test_monomial_force_composition():
    print("[gpu] force from phi", flush=True)
    from jaxqft_amd import fields
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 4)
    rng = fields.rng_for(1820)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    phi = fields.random_fermion(rng, (1, *L, 4, 3))
    th = theory(L, 3, 0.1, (1, 1, 1, -1), cg_tol=1e-12, cg_maxiter=3000)
    F = th.pseudofermion_force(U, phi)
    ora = WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1))
    X = np.asarray(th.solve_normal(U, phi))
    ref = ora.pseudofermion_force_from_solution(U, X, ora.apply_D(U, X))
    assert rel(F, ref) < 1e-12


@pytest.mark.parametrize("layout", ["BMXYIJ", "BXYMIJ"])
@pytest.mark.parametrize("nc", [3, 1])
def test_force_on_resident_fields(nc, layout):
    """The monomial force runs solve, D and force on resident FP64 fields
    (wl_cg_normal_field, wl_dslash, wl_pf_force_field): in complex128 it equals
    solve_normal, apply_D and the force from the solution in turn bit for bit,
    from a zero start and from a chronological guess; wl_pf_force_field on the
    fields equals wl_pf_force on their exported arrays bit for bit, and checks
    its handles."""
    print(f"[gpu] resident force nc={nc} {layout}", flush=True)
    import ctypes

    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    L = (4, 6, 4, 8)
    bc = (np.exp(0.3j), 1, -1, -1)
    rng = fields.rng_for(1840 + nc)
    U = fields.su3_hot_start(rng, 2, L, 0.6, layout=layout) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.6,
                                                                                                 layout=layout)
    phi = fields.random_fermion(rng, (2, *L, 4, nc))
    th = theory(L, nc, 0.05, bc, batch=2, layout=layout, cg_tol=1e-12, cg_maxiter=3000)
    F = th.pseudofermion_force(U, phi)
    X = th.solve_normal(U, phi)
    Y = th.apply_D(U, X)
    assert np.array_equal(F, th.pseudofermion_force_from_solution(U, X, Y))
    # from a guess: X0 = X + a small perturbation, as a chronological predictor gives
    X0 = X + 1e-3 * fields.random_fermion(rng, X.shape)
    Fg = th._force_resident(U, phi, X0, keep_xy=True)
    Xg = th.solve_normal_with_guess(U, phi, X0)
    assert np.array_equal(Fg[1], Xg) and np.array_equal(Fg[2], th.apply_D(U, Xg))
    assert np.array_equal(Fg[0], th.pseudofermion_force_from_solution(U, Xg, Fg[2]))
    # the entry point itself: complex128 against wl_pf_force on the arrays;
    # complex64 output is the FP64 force rounded once at the store
    ctx, g = th._gauge(U)
    fx, fy = nat.Field(ctx, nat.WL_F64, 2), nat.Field(ctx, nat.WL_F64, 2)
    Xt, Yt = torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda()
    fx.import_(Xt)
    fy.import_(Yt)
    lay = th._layout_code()
    Ff = torch.empty(th.field_shape(), dtype=torch.complex128, device="cuda")
    Fa = torch.empty_like(Ff)
    Ff32 = torch.empty(th.field_shape(), dtype=torch.complex64, device="cuda")
    vp = ctypes.c_void_p
    nat.check(ctx._lib.wl_pf_force_field(ctx.h, g.h, fx.h, fy.h, vp(Ff.data_ptr()), lay, nat.WL_C128, None))
    nat.check(ctx._lib.wl_pf_force(ctx.h, g.h, vp(Xt.data_ptr()), vp(Yt.data_ptr()), vp(Fa.data_ptr()), lay,
                                   nat.WL_C128, None))
    nat.check(ctx._lib.wl_pf_force_field(ctx.h, g.h, fx.h, fy.h, vp(Ff32.data_ptr()), lay, nat.WL_C64, None))
    torch.cuda.synchronize()
    assert torch.equal(Ff, Fa)
    assert torch.equal(Ff32, Fa.to(torch.complex64))
    # handle checks: an FP32 field, a field with the wrong column count
    f32 = nat.Field(ctx, nat.WL_F32, 2)
    f3 = nat.Field(ctx, nat.WL_F64, 3)
    for a, b in ((f32, fy), (fx, f3)):
        rc = ctx._lib.wl_pf_force_field(ctx.h, g.h, a.h, b.h, vp(Ff.data_ptr()), lay, nat.WL_C128, None)
        assert rc == nat.WL_ERR_BAD_ARG
    assert "column" in ctx._lib.wl_last_error().decode()


def test_monomial_force_with_guess():
    """pseudofermion_force with use_solver_guess (su3_wilson_nf2.py:378-392)
    keeps (X, D X) in its history: with phi = D^dag eta the 'mre' guess fits
    eta, so a second call starts at the solution; the forces equal the cold
    one."""
    print("[gpu] force with guess", flush=True)
    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1860)
    U = fields.su3_hot_start(rng, 2, L, 0.5)
    eta = fields.random_fermion(rng, (2, *L, 4, 3))
    th = theory(L, 3, 0.1, bc, batch=2, cg_tol=1e-12, cg_maxiter=3000)
    phi = th.apply_Ddag(U, eta)
    F = th.pseudofermion_force(U, phi)
    thg = theory(L, 3, 0.1, bc, batch=2, cg_tol=1e-12, cg_maxiter=3000, use_solver_guess=True,
                 solver_guess_mode="mre", solver_guess_history=3)
    assert np.array_equal(thg.pseudofermion_force(U, phi, eta), F)
    cold = [st["iters"] for st in thg.last_solve_stats]
    X, Y = thg.solver_guess.xs[0], thg.solver_guess.ys[0]
    assert np.array_equal(Y, thg.apply_D(U, X)) and rel(Y, eta) < 1e-10
    assert rel(thg.pseudofermion_force(U, phi, eta), F) < 1e-10
    warm = [st["iters"] for st in thg.last_solve_stats]
    assert min(cold) > 10 and max(warm) <= 5, (cold, warm)
    assert len(thg.solver_guess.xs) == 2 and not thg.eop_solver_guess.xs


def test_resident_force_graph_and_complex64():
    """The resident-field force with HIP-graph iteration blocks (the solve
    then runs on a side stream when the caller is on the default stream)
    equals the force without graphs bit for bit; complex64 phi gives the
    complex128 force rounded once (X and Y stay FP64 inside), to within the
    complex64 rounding of phi itself."""
    print("[gpu] resident force graph / complex64", flush=True)
    import torch

    from jaxqft_amd import fields

    L = (4, 4, 4, 8)
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(1870)
    U = torch.from_numpy(fields.su3_hot_start(rng, 2, L, 0.5)).cuda()
    phi = torch.from_numpy(fields.random_fermion(rng, (2, *L, 4, 3))).cuda()
    kw = dict(batch=2, cg_tol=1e-12, cg_maxiter=3000, solver_precision="f16", solver_recon=12)
    F = theory(L, 3, 0.1, bc, **kw).pseudofermion_force(U, phi)
    Fg = theory(L, 3, 0.1, bc, use_graph=True, **kw).pseudofermion_force(U, phi)
    assert torch.equal(F, Fg)
    F64 = theory(L, 3, 0.1, bc, **kw).pseudofermion_force(U, phi.to(torch.complex64))
    assert F64.dtype == torch.complex64
    ref = theory(L, 3, 0.1, bc, **kw).pseudofermion_force(U, phi.to(torch.complex64).to(torch.complex128))
    assert torch.equal(F64, ref.to(torch.complex64))


def test_force_is_action_derivative_16x4():
    """At 16^4, with the action S = phi^dag (D^dag D)^-1 phi evaluated by the GPU
    solver (FP64, tol 1e-13), (S(eps) - S(-eps)) / 2 eps for U -> exp(eps H) U
    (H a random su(3) field on every link) equals sum Re tr(H F)."""
    print("[gpu] force derivative 16^4", flush=True)
    import torch

    from jaxqft_amd import fields

    L = (16, 16, 16, 16)
    rng = fields.rng_for(1830)
    U = torch.from_numpy(fields.su3_hot_start(rng, 1, L, 0.5)).cuda()
    phi = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    a = torch.from_numpy(rng.standard_normal((1, 4, *L, 3, 3)) + 1j * rng.standard_normal((1, 4, *L, 3, 3))).cuda()
    H = 0.5 * (a - a.conj().transpose(-1, -2))
    H = H - (torch.diagonal(H, dim1=-2, dim2=-1).sum(-1) / 3.0)[..., None, None] * torch.eye(3, dtype=H.dtype,
                                                                                           device=H.device)
    th = theory(L, 3, 0.1, (1, 1, 1, -1), cg_tol=1e-13, cg_maxiter=4000)

    def action(Uc):
        x = th.solve_normal(Uc, phi)
        return float(torch.vdot(phi.flatten(), x.flatten()).real)

    F = th.pseudofermion_force(U, phi)
    an = float(torch.einsum("...ij,...ji->", H, F).real)
    eps = 1e-4
    ex = torch.linalg.matrix_exp(eps * H.reshape(-1, 3, 3)).reshape(H.shape)
    exm = torch.linalg.matrix_exp(-eps * H.reshape(-1, 3, 3)).reshape(H.shape)
    fd = (action(ex @ U) - action(exm @ U)) / (2 * eps)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    # the force is traceless anti-Hermitian
    assert float((F + F.conj().transpose(-1, -2)).abs().max()) < 1e-13 * float(F.abs().max())


def test_force_properties_at_32x4():
    """At 32^4 (no oracle): every force matrix is traceless anti-Hermitian,
    the two output layouts hold the same matrices, and the force is additive
    in Y (it is real-bilinear in X and Y)."""
    print("[gpu] force properties 32^4", flush=True)
    import torch

    from jaxqft_amd import fields

    L = (32, 32, 32, 32)
    rng = fields.rng_for(1830)
    th = theory(L, 3, 0.02, (1, 1, 1, -1))
    thl = theory(L, 3, 0.02, (1, 1, 1, -1), layout="BXYMIJ")
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    G = th.resident_gauge(U)
    Gl = thl.resident_gauge(np.ascontiguousarray(np.moveaxis(U, 1, -3)))
    X = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    Y1 = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    Y2 = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
    F1 = th.pseudofermion_force_from_solution(G, X, Y1)
    F2 = th.pseudofermion_force_from_solution(G, X, Y2)
    F12 = th.pseudofermion_force_from_solution(G, X, Y1 + Y2)
    scale = float(torch.abs(F1).max())
    assert float(torch.abs(F1 + F1.conj().transpose(-1, -2)).max()) <= 1e-14 * scale
    tr = torch.diagonal(F1, dim1=-2, dim2=-1).sum(-1)
    assert float(torch.abs(tr).max()) <= 1e-14 * scale
    assert float(torch.linalg.norm(F12 - F1 - F2) / torch.linalg.norm(F12)) < 1e-13
    Fl = thl.pseudofermion_force_from_solution(Gl, X, Y1)
    assert torch.equal(torch.movedim(Fl, -3, 1), F1)
