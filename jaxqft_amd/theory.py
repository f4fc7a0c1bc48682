"""Theory-level drop-in for the Wilson-fermion operator and normal-equation CG.

Mirrors the fermion-operator surface of the reference theory objects
``SU3WilsonNf2`` (jaxqft/models/su3_wilson_nf2.py:594-1635) and
``U1WilsonNf2`` (jaxqft/models/u1_wilson_nf2.py:296-1025): the same dataclass
field names, shapes, layouts, boundary-phase and mass conventions and method
names (``apply_D``, ``apply_Ddag``, ``apply_normal``, ``apply_offdiag``,
``solve_normal``, ``solve_normal_with_guess``, ``sample_pseudofermion``,
``pseudofermion_action``, ``normal_residual_stats``, ``fermion_shape``, ...).

Every operator application and solve runs in libwilson_amd on the GPU.
Inputs may be NumPy arrays or torch tensors; outputs come back as the same
kind (and dtype as the fermion input).  Computation is FP64 regardless of the
input dtype (complex64 inputs are widened on the device).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as nat
from ._arrays import from_device, to_device
from .fermions.gamma import build_euclidean_gamma, gamma5
from .fermions.wilson import WilsonDiracOperator
from .chrono import SolverGuess
from . import fields as _fields


def _vdot_batch(a, b):
    """Re <a_b, b_b> for each batch element b (leading axis), as a NumPy array.
    Two tensors on one GPU are reduced there: only the B results cross to the
    host."""
    if isinstance(a, torch.Tensor) and isinstance(b, torch.Tensor) and a.is_cuda and b.device == a.device:
        return torch.sum(torch.real(torch.conj(a) * b), dim=tuple(range(1, a.ndim))).cpu().numpy()
    aa = np.asarray(a.cpu() if isinstance(a, torch.Tensor) else a)
    bb = np.asarray(b.cpu() if isinstance(b, torch.Tensor) else b)
    return np.real(np.sum(np.conj(aa) * bb, axis=tuple(range(1, aa.ndim))))


def _as_like(a, ref):
    """a as the same kind of array as ref (torch tensor on ref's device, or NumPy)."""
    if isinstance(ref, torch.Tensor):
        return a.to(ref.device) if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a), device=ref.device)
    return np.asarray(a.cpu()) if isinstance(a, torch.Tensor) else np.asarray(a)


def parse_fermion_bc(bc, nd: int, default_antiperiodic_t: bool) -> Tuple[complex, ...]:
    """su3_wilson_nf2.py:711-724 / :1645-1660 and u1_wilson_nf2.py:461-471."""
    if bc is None:
        vals = [1.0 + 0.0j] * nd
        if default_antiperiodic_t:
            vals[-1] = -1.0 + 0.0j
    elif isinstance(bc, str):
        txt = bc.strip().lower()
        if txt in ("", "periodic", "p"):
            vals = [1.0 + 0.0j] * nd
        elif txt in ("antiperiodic-t", "anti-t", "ap-t", "apt", "chroma"):
            vals = [1.0 + 0.0j] * nd
            vals[-1] = -1.0 + 0.0j
        else:
            vals = [complex(t.strip().replace("I", "j").replace("i", "j")) for t in bc.split(",") if t.strip()]
    else:
        vals = [complex(v) for v in bc]
    if len(vals) != nd:
        raise ValueError(f"fermion_bc must have Nd={nd} entries; got {len(vals)}")
    return tuple(vals)


class ResidentGauge:
    """Gauge configurations kept on the device; pass it wherever ``U`` is taken
    to skip the per-call upload.  Links are stored without the boundary phase."""

    def __init__(self, theory: "WilsonNf2Theory", U: Any):
        theory._check_links(U)
        Ut, _, _ = to_device(U)
        self.theory = theory
        self.ctx = theory._context()
        self.gauge = nat.Gauge(self.ctx, Ut, theory._layout_code())
        self.shape = tuple(Ut.shape)


@dataclass
class WilsonNf2Theory:
    lattice_shape: Tuple[int, ...]
    beta: float = 0.0
    batch_size: int = 1
    layout: str = "BMXYIJ"
    seed: int = 0
    dtype: Any = np.complex64
    mass: float = 0.0
    wilson_r: float = 1.0
    cg_tol: float = 1e-8
    cg_maxiter: int = 500
    solver_kind: str = "cg"
    solver_form: str = "normal"
    preconditioner_kind: str = "none"
    fermion_bc: Optional[Any] = None
    dirac_kernel: str = "optimized"
    # further reference constructor fields (su3_wilson_nf2.py:594-623, su3_ym.py:181-188,
    # u1_wilson_nf2.py:296-320), accepted so that a caller constructs the drop-in unchanged:
    # the solver-guess fields drive the chronological guesses of pseudofermion_action /
    # pseudofermion_force; the others select JAX jit, GMRES, half-spinor or HMC-monomial
    # options that have no meaning here and are validated, stored and otherwise unused
    exp_method: str = "expm"
    gmres_restart: int = 32
    gmres_solve_method: str = "batched"
    eo_use_half_spinor: bool = True
    include_gauge_monomial: bool = True
    include_fermion_monomial: bool = True
    fermion_monomial_kind: str = "unpreconditioned"
    gauge_timescale: int = 0
    fermion_timescale: int = 1
    pseudofermion_refresh: str = "heatbath"
    pseudofermion_gamma: Optional[float] = None
    pseudofermion_force_mode: str = "autodiff"
    smd_gamma: float = 0.3
    auto_refresh_pseudofermions: bool = True
    use_solver_guess: bool = False
    solver_guess_mode: str = "last"
    solver_guess_history: int = 1
    jit_dirac_kernels: bool = True
    jit_solvers: bool = True
    # native solver policy (no counterpart in the reference)
    solver_precision: str = "f64"   # inner precision: f64 (plain CG) | f32 | f16 (reliable updates)
    solver_recon: int = 18          # 18 | 12: link reconstruction of the inner operator
    reliable_delta: float = 0.1     # reliable-update trigger on the residual drop
    solver_stop: str = "true"       # "true": FP64 true residual <= cg_tol; "iterated": reference rule
    poll_every: int = 8             # iterations per host poll (rounded up to an even number)
    use_graph: bool = False
    solver_accum: str = "auto"      # inner solution accumulator storage: auto (f32) | f64 | f32 | f16
    solver_search: str = "recurrence"  # mixed precision: t = D r + beta t_prev (recurrence) | p = r + beta p (xpay)
    Nc: int = field(default=3, init=False)
    last_solve_stats: List[Dict[str, Any]] = field(default_factory=list, init=False, repr=False)

    _default_antiperiodic_t = False

    def __post_init__(self):
        self.lattice_shape = tuple(int(v) for v in self.lattice_shape)
        self.Nd = len(self.lattice_shape)
        self.Bs = int(self.batch_size)
        lay = str(self.layout).upper()
        lay = {"BM...IJ": "BMXYIJ", "B...MIJ": "BXYMIJ"}.get(lay, lay)
        if lay not in ("BMXYIJ", "BXYMIJ"):
            raise ValueError("layout must be one of: BMXYIJ, BXYMIJ, BM...IJ, B...MIJ")
        self.layout = lay
        self.solver_kind = str(self.solver_kind).lower()
        self.solver_form = str(self.solver_form).lower()
        self.preconditioner_kind = str(self.preconditioner_kind).lower()
        self.dirac_kernel = str(self.dirac_kernel).lower()
        # the reference's solver fields and checks (su3_wilson_nf2.py:634-639); every solve here is CG on
        # the normal equations (D^dag D or S^dag S), whose solutions are those of the reference's
        # non-Hermitian split solves (see "split solver forms" below)
        if self.solver_kind not in ("cg", "bicgstab", "gmres"):
            raise ValueError("solver_kind must be one of: cg, bicgstab, gmres")
        if self.solver_form not in ("normal", "split", "eo_split"):
            raise ValueError("solver_form must be one of: normal, split, eo_split")
        if self.solver_form in ("split", "eo_split") and self.solver_kind == "cg":
            raise ValueError("solver_form=split/eo_split requires a non-Hermitian solver (use solver_kind=bicgstab)")
        # A scalar Jacobi preconditioner 1/m0^2 leaves the CG iterates and the
        # residual-based stopping rule unchanged, so both kinds map to plain CG.
        if self.preconditioner_kind not in ("none", "jacobi"):
            raise ValueError("preconditioner_kind must be one of: none, jacobi")
        if self.dirac_kernel not in ("optimized", "reference"):
            raise ValueError("dirac_kernel must be one of: optimized, reference")
        # reference validation of the remaining constructor fields (su3_wilson_nf2.py:625-657, su3_ym.py:204-206)
        self.exp_method = str(self.exp_method).lower()
        if self.exp_method not in ("expm", "su3"):
            raise ValueError("exp_method must be one of: expm, su3")
        self.fermion_monomial_kind = str(self.fermion_monomial_kind).lower()
        if self.fermion_monomial_kind not in ("unpreconditioned", "eo_preconditioned"):
            raise ValueError("fermion_monomial_kind must be one of: unpreconditioned, eo_preconditioned")
        self.gmres_restart = int(self.gmres_restart)
        if self.gmres_restart <= 0:
            raise ValueError("gmres_restart must be positive")
        self.solver_guess_mode = str(self.solver_guess_mode).lower()
        self.solver_guess_history = int(self.solver_guess_history)
        self.solver_guess = SolverGuess(self.solver_guess_mode, self.solver_guess_history)
        # the even-odd monomial keeps a history of its own (su3_wilson_nf2.py:416-421)
        self.eop_solver_guess = SolverGuess(self.solver_guess_mode, self.solver_guess_history)
        if self.pseudofermion_gamma is None:
            refresh = str(self.pseudofermion_refresh).lower()
            self.pseudofermion_gamma = float(self.smd_gamma) if refresh == "ou" else 0.3
        if self.solver_precision not in nat.PREC:
            raise ValueError("solver_precision must be one of: f64, f32, f16")
        if int(self.solver_recon) not in (18, 12):
            raise ValueError("solver_recon must be 18 or 12")
        if self.solver_stop not in ("true", "iterated"):
            raise ValueError("solver_stop must be 'true' or 'iterated'")
        if self.solver_accum not in nat.ACCUM:
            raise ValueError("solver_accum must be one of: auto, f64, f32, f16")
        if self.solver_search not in nat.SEARCH:
            raise ValueError("solver_search must be 'recurrence' or 'xpay'")
        self.fermion_bc = parse_fermion_bc(self.fermion_bc, self.Nd, self._default_antiperiodic_t)
        self.dirac = WilsonDiracOperator(ndim=self.Nd, mass=self.mass, wilson_r=self.wilson_r, dtype=self.dtype)
        self.gamma = self.dirac.gamma
        self.Ns = self.dirac.Ns
        self._m0 = float(self.mass + self.wilson_r * self.Nd)
        self._rng = _fields.rng_for(self.seed)
        self._ctx: Optional[nat.Context] = None
        self._force_fields = None  # (ctx, b, x, y): resident FP64 fields of the force step
        # per-lattice constants built on first use: parity masks (NumPy, and per
        # torch device) and the boundary-phase factor of the links
        self._parity_masks: Dict[Any, Any] = {}

    # ------------------------------------------------------------------ shapes
    def fermion_shape(self) -> Tuple[int, ...]:
        return (self.Bs, *self.lattice_shape, self.Ns, self.Nc)

    def field_shape(self) -> Tuple[int, ...]:
        raise NotImplementedError

    def _layout_code(self) -> int:
        return nat.WL_LAYOUT_BMXYIJ if self.layout == "BMXYIJ" else nat.WL_LAYOUT_BXYMIJ

    def _check_links(self, U) -> None:
        if isinstance(U, ResidentGauge):
            return
        shp = tuple(U.shape)
        if shp != tuple(self.field_shape()):
            raise ValueError(f"link field has shape {shp}, expected {self.field_shape()}")

    def _check_fermion(self, psi, name="psi") -> None:
        if tuple(psi.shape) != tuple(self.fermion_shape()):
            raise ValueError(f"{name} has shape {tuple(psi.shape)}, expected {self.fermion_shape()}")

    # ------------------------------------------------------------ native plumbing
    def _context(self) -> nat.Context:
        if self._ctx is None:
            self._ctx = nat.Context(self.lattice_shape, self.Nc, self.Bs, self.mass, self.wilson_r, self.fermion_bc)
        if self._ctx.mass != float(self.mass):
            self._ctx.set_mass(float(self.mass))
        return self._ctx

    def _gauge(self, U) -> Tuple[nat.Context, Any]:
        if isinstance(U, ResidentGauge):
            ctx = self._context()  # also brings the context's mass up to date with self.mass
            if U.theory is not self and U.ctx is not ctx:
                raise ValueError("resident gauge belongs to another theory")
            return U.ctx, U.gauge
        self._check_links(U)
        ctx = self._context()
        Ut, _, _ = to_device(U)
        return ctx, nat.Gauge(ctx, Ut, self._layout_code())

    def resident_gauge(self, U) -> ResidentGauge:
        return ResidentGauge(self, U)

    def _links_with_fermion_bc(self, U):
        """U with the fermion boundary phase on the links U_mu(x) of the slice
        x_mu = L_mu - 1 (su3_wilson_nf2.py:751-759), for callers that build on
        the phased links (the conserved current, core/measurements.py:1620);
        the kernels apply the same phase on the hop that crosses the boundary.
        One configuration or a batch, NumPy or torch, in the theory's layout;
        U itself when every phase is 1."""
        if all(abs(complex(p) - 1.0) <= 1e-12 for p in self.fermion_bc):
            return U
        fac = self._parity_masks.get("bc_factor")
        if fac is None:
            nmat = len(self.field_shape()) - self.Nd - 2  # trailing colour axes: 2 for SU(3), 0 for U(1)
            mu_first = self.layout == "BMXYIJ"
            shape = ((self.Nd, *self.lattice_shape) if mu_first else (*self.lattice_shape, self.Nd)) + (1,) * nmat
            fac = np.ones(shape, dtype=np.complex128)
            for mu, p in enumerate(self.fermion_bc):
                if abs(complex(p) - 1.0) <= 1e-12:
                    continue
                sl = [slice(None)] * len(shape)
                sl[0 if mu_first else self.Nd] = mu
                sl[(1 + mu) if mu_first else mu] = int(self.lattice_shape[mu]) - 1
                fac[tuple(sl)] *= complex(p)
            self._parity_masks["bc_factor"] = fac
        if U.ndim == fac.ndim + 1:
            f = fac[None]
        elif U.ndim == fac.ndim:
            f = fac
        else:
            raise ValueError(f"Unexpected link tensor rank for fermion BC application: ndim={U.ndim}")
        if isinstance(U, torch.Tensor):
            return U * torch.as_tensor(f, device=U.device).to(U.dtype)
        U = np.asarray(U)
        return U * f.astype(U.dtype)

    def tune(self, U, nrhs: int = 1, reps: int = 3) -> Dict[str, Dict[str, object]]:
        """Pick this lattice's stencil launch shapes with jaxqft_amd.autotune.tune
        for the solver's inner precision and link encoding and for the FP64
        stencil of the reliable updates, at nrhs right-hand sides per
        configuration (timed on Gaussian spinors), and install them in the
        context: later operator calls and solves use them.  Launches with one
        right-hand side per configuration and with several are tuned
        separately (nrhs decides which).  Returns the chosen shape per
        precision."""
        from .autotune import tune as _tune

        ctx, g = self._gauge(U)
        ncol = self.Bs * int(nrhs)
        if ncol < 1:
            raise ValueError("nrhs must be positive")
        gen = torch.Generator(device="cuda").manual_seed(int(self.seed))
        shape = (ncol, *self.lattice_shape, self.Ns, self.Nc)
        src = torch.complex(torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen),
                            torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen))
        inner = nat.PREC[self.solver_precision]
        plan = [(inner, int(self.solver_recon))]
        if inner != nat.WL_F64:
            plan.append((nat.WL_F64, 12 if self.Nc == 3 else 18))
        chosen = {}
        for pc, recon in plan:
            fi, fo = nat.Field(ctx, pc, ncol), nat.Field(ctx, pc, ncol)
            fi.import_(src)
            t = _tune(ctx, g, pc, recon, fi, fo, nrhs=int(nrhs), reps=reps)
            chosen[self._prec_name(pc)] = {k: v for k, v in t.items() if k not in ("ranked", "us")}
        torch.cuda.synchronize()
        return chosen

    @staticmethod
    def _prec_name(pc: int) -> str:
        return {v: k for k, v in nat.PREC.items()}[pc]

    def _apply(self, U, psi, op: int):
        self._check_fermion(psi)
        ctx, g = self._gauge(U)
        pt, kind, orig = to_device(psi)
        out = torch.empty_like(pt)
        dt = nat.WL_C128 if pt.dtype == torch.complex128 else nat.WL_C64
        nat.check(ctx._lib.wl_apply(ctx.h, g.h, int(op), nat.c_void_p(pt.data_ptr()), nat.c_void_p(out.data_ptr()), 1,
                                    dt, nat.c_void_p(nat.current_stream())))
        return from_device(out, kind, orig)

    # -------------------------------------------------------------- operator API
    def apply_D(self, U, psi):
        return self._apply(U, psi, nat.WL_OP_D)

    def apply_Ddag(self, U, psi):
        return self._apply(U, psi, nat.WL_OP_DDAG)

    def apply_normal(self, U, psi):
        return self._apply(U, psi, nat.WL_OP_DDAGD)

    def apply_offdiag(self, U, psi):
        return self._apply(U, psi, nat.WL_OP_K)

    def apply_offdiag_dagger(self, U, psi):
        return self._apply(U, psi, nat.WL_OP_KDAG)

    def apply_g5D(self, U, psi):
        """gamma5 D (the composition used by the reference's convention test)."""
        return self._apply(U, psi, nat.WL_OP_G5D)

    def apply_D_with_kernel(self, U, psi, kernel: str, use_sparse_gamma: bool = True):
        if str(kernel).lower() not in ("optimized", "reference"):
            raise ValueError("kernel must be one of: optimized, reference")
        return self.apply_D(U, psi)

    def apply_normal_with_kernel(self, U, psi, kernel: str, use_sparse_gamma: bool = True):
        if str(kernel).lower() not in ("optimized", "reference"):
            raise ValueError("kernel must be one of: optimized, reference")
        return self.apply_normal(U, psi)

    # The reference's dense-gamma forms (su3_wilson_nf2.py:1157-1167) apply the
    # same operator with the gamma matrices multiplied out; the projector
    # kernels here compute it once.
    def apply_D_dense(self, U, psi):
        return self.apply_D(U, psi)

    def apply_Ddag_dense(self, U, psi):
        return self.apply_Ddag(U, psi)

    def apply_normal_dense(self, U, psi):
        return self.apply_normal(U, psi)

    # ------------------------------------------------- even-odd (checkerboard)
    def _eo(self, U, psi, op: int, normalized: bool = False, nrhs: int = 1):
        ctx, g = self._gauge(U)
        pt, kind, orig = to_device(psi)
        out = torch.empty_like(pt)
        dt = nat.WL_C128 if pt.dtype == torch.complex128 else nat.WL_C64
        nat.check(ctx._lib.wl_eo_apply(ctx.h, g.h, int(op), 1 if normalized else 0, nat.c_void_p(pt.data_ptr()),
                                       nat.c_void_p(out.data_ptr()), int(nrhs), dt, nat.c_void_p(nat.current_stream())))
        return from_device(out, kind, orig)

    def apply_K_oe(self, U, psi_even):
        """K_oe = P_o K P_e, even -> odd hopping block (su3_wilson_nf2.py:1057-1061, :1081-1084)."""
        self._check_fermion(psi_even, "psi_even")
        return self._eo(U, psi_even, nat.WL_EO_K_OE)

    def apply_K_eo(self, U, psi_odd):
        """K_eo = P_e K P_o (su3_wilson_nf2.py:1063-1067)."""
        self._check_fermion(psi_odd, "psi_odd")
        return self._eo(U, psi_odd, nat.WL_EO_K_EO)

    def apply_Kdag_oe(self, U, psi_even):
        """P_o K^dag P_e (su3_wilson_nf2.py:1069-1073)."""
        self._check_fermion(psi_even, "psi_even")
        return self._eo(U, psi_even, nat.WL_EO_KDAG_OE)

    def apply_Kdag_eo(self, U, psi_odd):
        """P_e K^dag P_o (su3_wilson_nf2.py:1075-1079)."""
        self._check_fermion(psi_odd, "psi_odd")
        return self._eo(U, psi_odd, nat.WL_EO_KDAG_EO)

    def apply_eo_schur_even(self, U, psi_even, normalized: bool = False):
        """S_e = m0 - K_eo K_oe / m0 on the even sites of psi_even (normalized:
        1 - K_eo K_oe / m0^2); full-lattice result, zero on odd sites
        (su3_wilson_nf2.py:1112-1121)."""
        self._check_fermion(psi_even, "psi_even")
        return self._eo(U, psi_even, nat.WL_EO_SCHUR_E, normalized)

    def apply_eo_schur_even_dagger(self, U, psi_even, normalized: bool = False):
        """S_e^dag (su3_wilson_nf2.py:1138-1146)."""
        self._check_fermion(psi_even, "psi_even")
        return self._eo(U, psi_even, nat.WL_EO_SCHUR_E_DAG, normalized)

    def apply_eo_schur_odd(self, U, psi_odd, normalized: bool = False):
        """S_o = m0 - K_oe K_eo / m0 (su3_wilson_nf2.py:1148-1156)."""
        self._check_fermion(psi_odd, "psi_odd")
        return self._eo(U, psi_odd, nat.WL_EO_SCHUR_O, normalized)

    def apply_eo_schur_even_compact(self, U, psi_even, daggered: bool = False, normalized: bool = True):
        """Compact-path Schur matvec with a full-lattice interface
        (su3_wilson_nf2.py:1123-1136); the same operator as apply_eo_schur_even."""
        self._check_fermion(psi_even, "psi_even")
        return self._eo(U, psi_even, nat.WL_EO_SCHUR_E_DAG if daggered else nat.WL_EO_SCHUR_E, normalized)

    def _parity_mask(self, parity: int, device=None):
        """Boolean mask of the sites of one parity, shaped (1, *L, 1, 1); built
        once per lattice (on the host, then once per torch device)."""
        key = (parity, None if device is None else str(device))
        m = self._parity_masks.get(key)
        if m is None:
            if device is None:
                even = np.indices(self.lattice_shape).sum(axis=0) % 2 == 0
                m = (even if parity == 0 else ~even).reshape((1, *self.lattice_shape, 1, 1))
            else:
                m = torch.as_tensor(self._parity_mask(parity), device=device)
            self._parity_masks[key] = m
        return m

    def _parity_project(self, psi, parity: int):
        """psi on the sites of one parity (sum of coordinates mod 2), zero on the
        others (su3_wilson_nf2.py:884-892)."""
        if isinstance(psi, torch.Tensor):
            return psi * self._parity_mask(parity, psi.device).to(psi.dtype)
        return np.asarray(psi) * self._parity_mask(parity).astype(np.asarray(psi).real.dtype)

    def project_even(self, psi):
        return self._parity_project(psi, 0)

    def project_odd(self, psi):
        return self._parity_project(psi, 1)

    # the reference's even-odd monomial calls the projections under these
    # names (su3_wilson_nf2.py:490-505, :884-892)
    _project_even = project_even
    _project_odd = project_odd

    def _solve_eo(self, U, rhs, x0, nrhs: int = 1, masses=None, **overrides):
        ctx, g = self._gauge(U)
        pt, kind, orig = to_device(rhs)
        x = torch.empty_like(pt)
        x0t = None
        if x0 is not None:
            x0t, _, _ = to_device(x0, dtype=pt.dtype)
        ncol = self.Bs * nrhs
        sp, keep = self._solve_params(masses=masses, **overrides)
        stats = (nat.WlSolveStats * ncol)()
        dt = nat.WL_C128 if pt.dtype == torch.complex128 else nat.WL_C64
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            nat.check(ctx._lib.wl_cg_eo_normal(ctx.h, g.h, nat.c_void_p(pt.data_ptr()), nat.c_void_p(x.data_ptr()),
                                               nat.c_void_p(x0t.data_ptr() if x0t is not None else 0), nrhs, dt,
                                               nat.byref(sp), stats, nat.c_void_p(side.cuda_stream)))
        torch.cuda.current_stream().wait_stream(side)
        del keep
        self.last_solve_stats = nat.stats_list(stats, ncol)
        return from_device(x, kind, orig)

    def solve_eop_even_normal(self, U, rhs):
        """x_e = (S_e^dag S_e)^-1 P_e rhs with the normalized even Schur complement
        (su3_wilson_nf2.py:1373-1389); full-lattice result, zero on odd sites."""
        self._check_fermion(rhs, "rhs")
        return self._solve_eo(U, rhs, None)

    def solve_eop_even_normal_with_guess(self, U, rhs, x0):
        """Same with initial guess P_e x0 (su3_wilson_nf2.py:1391-1392)."""
        self._check_fermion(rhs, "rhs")
        self._check_fermion(x0, "x0")
        return self._solve_eo(U, rhs, x0)

    def solve_eop_even(self, U, phi):
        """solver_form dispatch of su3_wilson_nf2.py:1402-1408."""
        if self.solver_form == "normal":
            return self.solve_eop_even_normal(U, phi)
        return self.solve_eop_even_split(U, phi)

    # ------------------------------------------------- split solver forms
    # The reference solves D x = b, D^dag x = b and their even-odd Schur forms
    # with a non-Hermitian Krylov method (bicgstab or gmres).  Here each is CG
    # on the normal equations, with the same solution: D x = b is D^dag D x =
    # D^dag b, and D^dag x = b is x = D z with D^dag D z = b (likewise with the
    # normalized even Schur complement S).  The stopping rule applies to the
    # normal system's residual (cg_tol on its FP64 true residual by default).
    def solve_direct(self, U, rhs):
        """x with D x = rhs (su3_wilson_nf2.py:1271-1283): x = (D^dag D)^-1 D^dag rhs."""
        self._check_fermion(rhs, "rhs")
        return self.solve_normal(U, self.apply_Ddag(U, rhs))

    def solve_dagger(self, U, rhs):
        """x with D^dag x = rhs (su3_wilson_nf2.py:1285-1297): x = D (D^dag D)^-1 rhs."""
        self._check_fermion(rhs, "rhs")
        return self.apply_D(U, self.solve_normal(U, rhs))

    def _solve_eo_reduced(self, U, rhs, dagger: bool):
        """D x = rhs (or D^dag x = rhs) by the even-odd reduction of
        su3_wilson_nf2.py:1299-1332: the even system (m0 - K_eo K_oe / m0) x_e =
        b_e - K_eo b_o / m0 (its dagger for D^dag) is m0 times the normalized
        Schur system, solved through the even-odd normal equations; then
        x_o = (b_o - K_oe x_e) / m0."""
        self._check_fermion(rhs, "rhs")
        m0 = self._m0_value()
        b_e, b_o = self.project_even(rhs), self.project_odd(rhs)
        k_eo = self.apply_Kdag_eo if dagger else self.apply_K_eo
        k_oe = self.apply_Kdag_oe if dagger else self.apply_K_oe
        c = (b_e - _as_like(k_eo(U, b_o), b_e) / m0) / m0
        if dagger:
            x_e = self.apply_eo_schur_even(U, self.solve_eop_even_normal(U, c), normalized=True)
        else:
            x_e = self.solve_eop_even_normal(U, self.apply_eo_schur_even_dagger(U, c, normalized=True))
        x_e = _as_like(x_e, b_e)
        x_o = (b_o - _as_like(k_oe(U, x_e), b_o)) / m0
        return x_e + x_o

    def solve_direct_eo(self, U, rhs):
        """x with D x = rhs through the even-odd reduction (su3_wilson_nf2.py:1334-1335)."""
        return self._solve_eo_reduced(U, rhs, dagger=False)

    def solve_dagger_eo(self, U, rhs):
        """x with D^dag x = rhs through the even-odd reduction (su3_wilson_nf2.py:1337-1338)."""
        return self._solve_eo_reduced(U, rhs, dagger=True)

    def solve_eop_even_direct(self, U, rhs):
        """x_e with S x_e = P_e rhs, S the normalized even Schur complement
        (su3_wilson_nf2.py:1340-1353): x_e = (S^dag S)^-1 S^dag P_e rhs."""
        self._check_fermion(rhs, "rhs")
        return self.solve_eop_even_normal(U, self.apply_eo_schur_even_dagger(U, self.project_even(rhs),
                                                                            normalized=True))

    def solve_eop_even_dagger(self, U, rhs):
        """x_e with S^dag x_e = P_e rhs (su3_wilson_nf2.py:1355-1368): x_e = S (S^dag S)^-1 P_e rhs."""
        self._check_fermion(rhs, "rhs")
        return self.apply_eo_schur_even(U, self.solve_eop_even_normal(U, rhs), normalized=True)

    def solve_eop_even_split_with_intermediate(self, U, phi):
        """y = S^-dag P_e phi and x = S^-1 y (su3_wilson_nf2.py:1394-1397): x = (S^dag S)^-1 P_e phi
        from one even-odd normal solve, y = S x."""
        self._check_fermion(phi, "phi")
        x = self.solve_eop_even_normal(U, phi)
        return x, self.apply_eo_schur_even(U, x, normalized=True)

    def solve_eop_even_split(self, U, phi):
        """su3_wilson_nf2.py:1399-1401."""
        return self.solve_eop_even_split_with_intermediate(U, phi)[0]

    def solve_eop_even_normal_multi(self, U, rhs, masses: Optional[Sequence[float]] = None, **overrides):
        """k right-hand sides per configuration through one even-odd solve.
        rhs: (B, k, *L, Ns, Nc)."""
        shp = tuple(rhs.shape)
        if len(shp) != self.Nd + 4 or shp[0] != self.Bs or shp[2:] != tuple(self.fermion_shape()[1:]):
            raise ValueError(f"rhs must have shape (B, k, *L, Ns, Nc); got {shp}")
        k = int(shp[1])
        x = self._solve_eo(U, rhs.reshape((self.Bs * k, *shp[2:])), None, nrhs=k, masses=masses, **overrides)
        return x.reshape(shp)

    # ------------------------------------------------------ pseudofermion force
    def pseudofermion_force_from_solution(self, U, X, Y):
        """Analytic force of phi^dag (D^dag D)^-1 phi from X = (D^dag D)^-1 phi and
        Y = D X, in the link layout (su3_wilson_nf2.py:1556-1576)."""
        self._check_fermion(X, "X")
        self._check_fermion(Y, "Y")
        ctx, g = self._gauge(U)
        xt, kind, orig = to_device(X)
        yt, _, _ = to_device(Y, dtype=xt.dtype)
        shp = tuple(self.field_shape())
        F = torch.empty(shp, dtype=xt.dtype, device=xt.device)
        dt = nat.WL_C128 if xt.dtype == torch.complex128 else nat.WL_C64
        nat.check(ctx._lib.wl_pf_force(ctx.h, g.h, nat.c_void_p(xt.data_ptr()), nat.c_void_p(yt.data_ptr()),
                                       nat.c_void_p(F.data_ptr()), self._layout_code(), dt,
                                       nat.c_void_p(nat.current_stream())))
        return from_device(F, kind, orig)

    def eop_pseudofermion_force_from_solution(self, U, X_e, Y_e):
        """Force of phi_e^dag (S^dag S)^-1 phi_e (normalized even Schur complement)
        from X_e and Y_e = S X_e: two standard force terms,
        F(K_oe X_e, -Y_e / m0^2) + F(X_e, -K^dag_oe Y_e / m0^2)
        (su3_wilson_nf2.py:1578-1597)."""
        alpha = 1.0 / (self._m0_value() ** 2)
        kx = self.apply_K_oe(U, X_e)
        ky = self.apply_Kdag_oe(U, Y_e)
        f1 = self.pseudofermion_force_from_solution(U, kx, -alpha * _as_like(Y_e, kx))
        f2 = self.pseudofermion_force_from_solution(U, X_e, -alpha * ky)
        return f1 + f2

    def _solve_with_history(self, U, phi, eta=None):
        """solve_normal, from a chronological guess when use_solver_guess is
        set (the monomial's action and force_analytic, su3_wilson_nf2.py:322-343,
        :378-392)."""
        if not self.use_solver_guess:
            return self.solve_normal(U, phi)
        x0 = self.solver_guess.predict(phi, eta)
        return self.solve_normal(U, phi) if x0 is None else self.solve_normal_with_guess(U, phi, x0)

    def clear_solver_guess(self) -> None:
        """Forget the solution histories (a pseudofermion refresh, su3_wilson_nf2.py:318-320, :432-435)."""
        self.solver_guess.clear()
        self.eop_solver_guess.clear()

    def eop_pseudofermion_force(self, U, phi, eta=None):
        """The even-odd monomial's analytic force (su3_wilson_nf2.py:566-579,
        u1_wilson_nf2.py:275-281): x_e = (S^dag S)^-1 phi_e with the normalized
        even Schur complement, y_e = S x_e, then
        eop_pseudofermion_force_from_solution.  With use_solver_guess the solve
        starts from the guess of eop_solver_guess, a history separate from the
        standard monomial's ('mre' fits eta, phi = S^dag eta, when given).  A
        split solver form takes (x_e, y_e) from
        solve_eop_even_split_with_intermediate, as the reference does."""
        self._check_fermion(phi, "phi")
        if self.solver_form in ("split", "eo_split"):
            Xe, Ye = self.solve_eop_even_split_with_intermediate(U, phi)
            return self.eop_pseudofermion_force_from_solution(U, Xe, Ye)
        x0 = self.eop_solver_guess.predict(phi, eta) if self.use_solver_guess else None
        Xe = self.solve_eop_even_normal(U, phi) if x0 is None else self.solve_eop_even_normal_with_guess(U, phi, x0)
        Ye = self.apply_eo_schur_even(U, Xe, normalized=True)
        if self.use_solver_guess:
            self.eop_solver_guess.update(Xe, Ye)
        return self.eop_pseudofermion_force_from_solution(U, Xe, Ye)

    def pseudofermion_force(self, U, phi, eta=None):
        """The analytic monomial force: X = solve_normal(U, phi), Y = D X, then
        the force from the solution (su3_wilson_nf2.py:378-392).  With
        use_solver_guess the solve starts from the guess of solver_guess_mode
        over the last solver_guess_history (X, Y) pairs ('mre' fits eta,
        phi = D^dag eta, when given).  A split solver form takes (X, Y) from
        solve_split_with_intermediate, as the reference does.

        The normal form runs on resident FP64 fields (_force_resident): X and
        Y stay in the library's layout from the solve through D to the force,
        with the same results as the three separate calls."""
        if self.solver_form in ("split", "eo_split"):
            X, Y = self.solve_split_with_intermediate(U, phi)
            return self.pseudofermion_force_from_solution(U, X, Y)
        x0 = self.solver_guess.predict(phi, eta) if self.use_solver_guess else None
        F, X, Y = self._force_resident(U, phi, x0, keep_xy=self.use_solver_guess)
        if self.use_solver_guess:
            self.solver_guess.update(X, Y)
        return F

    def _resident_fields(self, ctx):
        """The (b, x, y) FP64 fields of the force step, one column per
        configuration, kept across calls on the same context."""
        if self._force_fields is None or self._force_fields[0] is not ctx:
            self._force_fields = (ctx,) + tuple(nat.Field(ctx, nat.WL_F64, self.Bs) for _ in range(3))
        return self._force_fields[1:]

    def _force_resident(self, U, phi, x0, keep_xy: bool = False):
        """X = (D^dag D)^-1 phi (from x0 when given), Y = D X and the force from
        them on resident fields: wl_cg_normal_field, wl_dslash and
        wl_pf_force_field, so X and Y are never converted to the reference
        layout and back.  Returns (F, X, Y) with X and Y exported only when
        keep_xy (else None).  Same arithmetic as solve_normal, apply_D and
        pseudofermion_force_from_solution in turn."""
        self._check_fermion(phi, "phi")
        ctx, g = self._gauge(U)
        pt, kind, orig = to_device(phi)
        x0t = None
        if x0 is not None:
            self._check_fermion(x0, "x0")
            x0t, _, _ = to_device(x0, dtype=pt.dtype)
        fb, fx, fy = self._resident_fields(ctx)
        sp, keep = self._solve_params()
        stats = (nat.WlSolveStats * self.Bs)()
        dt = nat.WL_C128 if pt.dtype == torch.complex128 else nat.WL_C64
        # outputs are allocated on the caller's stream, which owns them afterwards
        F = torch.empty(tuple(self.field_shape()), dtype=pt.dtype, device=pt.device)
        xs, ys = (torch.empty_like(pt), torch.empty_like(pt)) if keep_xy else (None, None)
        lib = ctx._lib

        def run():
            h = nat.c_void_p(nat.current_stream())
            fb.import_(pt)
            if x0t is not None:
                fx.import_(x0t)
            nat.check(lib.wl_cg_normal_field(ctx.h, g.h, fb.h, fx.h, 1 if x0t is not None else 0, 1, nat.byref(sp),
                                             stats, h))
            nat.check(lib.wl_dslash(ctx.h, g.h, nat.WL_OP_D, 18, fx.h, fy.h, 1, h))
            nat.check(lib.wl_pf_force_field(ctx.h, g.h, fx.h, fy.h, nat.c_void_p(F.data_ptr()), self._layout_code(),
                                            dt, h))
            if keep_xy:
                fx.export(xs)
                fy.export(ys)

        if sp.use_graph and nat.current_stream() == 0:
            # graph capture needs a stream other than the default one
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                run()
            torch.cuda.current_stream().wait_stream(side)
        else:
            run()
        del keep
        self.last_solve_stats = nat.stats_list(stats, self.Bs)
        X = from_device(xs, kind, orig) if keep_xy else None
        Y = from_device(ys, kind, orig) if keep_xy else None
        return from_device(F, kind, orig), X, Y

    def _m0_value(self) -> float:
        return float(self.mass) + float(self.wilson_r) * self.Nd

    # ----------------------------------------------------------------- solvers
    def _solve_params(self, masses=None, **overrides):
        pol = dict(
            tol=self.cg_tol,
            maxiter=self.cg_maxiter,
            stop_rule=nat.WL_STOP_ITERATED if self.solver_stop == "iterated" else nat.WL_STOP_TRUE_RESIDUAL,
            inner_prec=nat.PREC[self.solver_precision],
            inner_recon=int(self.solver_recon),
            delta=self.reliable_delta,
            poll_every=self.poll_every,
            use_graph=self.use_graph,
            masses=masses,
            accum=self.solver_accum,
            search=self.solver_search,
        )
        pol.update(overrides)
        return nat.solve_params(**pol)

    def _solve(self, U, phi, x0, nrhs: int = 1, masses=None, **overrides):
        ctx, g = self._gauge(U)
        pt, kind, orig = to_device(phi)
        x = torch.empty_like(pt)
        x0t = None
        if x0 is not None:
            x0t, _, _ = to_device(x0, dtype=pt.dtype)
        ncol = self.Bs * nrhs
        sp, keep = self._solve_params(masses=masses, **overrides)
        stats = (nat.WlSolveStats * ncol)()
        dt = nat.WL_C128 if pt.dtype == torch.complex128 else nat.WL_C64
        stream = nat.current_stream()
        if sp.use_graph and stream == 0:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                stream = side.cuda_stream
                nat.check(ctx._lib.wl_cg_normal(ctx.h, g.h, nat.c_void_p(pt.data_ptr()), nat.c_void_p(x.data_ptr()),
                                                nat.c_void_p(x0t.data_ptr() if x0t is not None else 0), nrhs, dt,
                                                nat.byref(sp), stats, nat.c_void_p(stream)))
            torch.cuda.current_stream().wait_stream(side)
        else:
            nat.check(ctx._lib.wl_cg_normal(ctx.h, g.h, nat.c_void_p(pt.data_ptr()), nat.c_void_p(x.data_ptr()),
                                            nat.c_void_p(x0t.data_ptr() if x0t is not None else 0), nrhs, dt,
                                            nat.byref(sp), stats, nat.c_void_p(stream)))
        del keep
        self.last_solve_stats = nat.stats_list(stats, ncol)
        return from_device(x, kind, orig)

    def solve_normal(self, U, phi):
        """x = (D^dag D)^-1 phi per batch element (su3_wilson_nf2.py:1410-1429)."""
        self._check_fermion(phi, "phi")
        return self._solve(U, phi, None)

    def solve_normal_with_guess(self, U, phi, x0):
        """Same with initial guess x0 (su3_wilson_nf2.py:1431-1432)."""
        self._check_fermion(phi, "phi")
        self._check_fermion(x0, "x0")
        return self._solve(U, phi, x0)

    def solve_split_with_intermediate(self, U, phi):
        """y = D^-dag phi and x = D^-1 y (su3_wilson_nf2.py:1434-1441).  For
        solver_form 'eo_split' both solves go through the even-odd reduction, as
        in the reference; otherwise x = (D^dag D)^-1 phi from one normal solve
        and y = D x."""
        self._check_fermion(phi, "phi")
        if self.solver_form == "eo_split":
            y = self.solve_dagger_eo(U, phi)
            return self.solve_direct_eo(U, y), y
        x = self.solve_normal(U, phi)
        return x, self.apply_D(U, x)

    def solve_split(self, U, phi):
        """su3_wilson_nf2.py:1443-1445."""
        return self.solve_split_with_intermediate(U, phi)[0]

    def solve_pseudofermion(self, U, phi):
        """solver_form dispatch of su3_wilson_nf2.py:1447-1454."""
        if self.solver_form == "normal":
            return self.solve_normal(U, phi)
        return self.solve_split(U, phi)

    def solve_normal_multi(self, U, rhs, masses: Optional[Sequence[float]] = None, **overrides):
        """Solve k right-hand sides per configuration at once.

        rhs: (B, k, *L, Ns, Nc).  masses: optional per-(b, k) masses, flattened
        config-major.  Each column converges independently (per-column
        scalars and masking); the links are shared by all k columns."""
        shp = tuple(rhs.shape)
        if len(shp) != self.Nd + 4 or shp[0] != self.Bs or shp[2:] != tuple(self.fermion_shape()[1:]):
            raise ValueError(f"rhs must have shape (B, k, *L, Ns, Nc); got {shp}")
        k = int(shp[1])
        flat = rhs.reshape((self.Bs * k, *shp[2:]))
        x = self._solve(U, flat, None, nrhs=k, masses=masses, **overrides)
        return x.reshape(shp)

    # ------------------------------------------------------ fields and actions
    def hot_start(self, scale: float = 0.1):
        raise NotImplementedError

    def random_fermion(self):
        return _fields.random_fermion(self._rng, self.fermion_shape()).astype(self.dtype)

    def sample_pseudofermion(self, U):
        """phi = D^dag eta (su3_wilson_nf2.py:1177-1181)."""
        chi = self.random_fermion()
        return self.apply_Ddag(U, chi)

    def pseudofermion_action(self, U, phi):
        """S = Re <phi, (D^dag D)^-1 phi> per batch element (su3_wilson_nf2.py:1183-1196);
        with a split solver form S = |chi|^2 with D^dag chi = phi, the same value."""
        if self.solver_form in ("split", "eo_split"):
            chi = self.solve_dagger_eo(U, phi) if self.solver_form == "eo_split" else self.solve_dagger(U, phi)
            return _vdot_batch(chi, chi)
        x = self._solve_with_history(U, phi)
        if self.use_solver_guess:
            self.solver_guess.update(x)
        return _vdot_batch(phi, x)

    def eop_pseudofermion_action(self, U, phi):
        """S = Re <phi_e, (S^dag S)^-1 phi_e> with phi_e = P_e phi, or |chi|^2 with
        S^dag chi = phi_e for a split solver form (su3_wilson_nf2.py:1198-1208)."""
        self._check_fermion(phi, "phi")
        phi_e = self.project_even(phi)
        if self.solver_form in ("split", "eo_split"):
            chi = self.solve_eop_even_dagger(U, phi_e)
            return _vdot_batch(chi, chi)
        return _vdot_batch(phi_e, self.solve_eop_even_normal(U, phi_e))

    def apply_gamma5(self, psi):
        """(gamma5 psi) on the spin axis; host-side helper for tests."""
        g5 = gamma5(self.gamma)
        return np.einsum("st,...tc->...sc", g5, np.asarray(psi))

    def normal_residual_stats(self, U, phi, x) -> Dict[str, float]:
        """su3_wilson_nf2.py:1456-1466, with D^dag D evaluated on the GPU in FP64."""
        Ax = self.apply_normal(U, x)
        r = np.asarray(Ax.cpu() if isinstance(Ax, torch.Tensor) else Ax) - np.asarray(
            phi.cpu() if isinstance(phi, torch.Tensor) else phi)
        ph = np.asarray(phi.cpu() if isinstance(phi, torch.Tensor) else phi)
        rel = [float(np.linalg.norm(r[b]) / (np.linalg.norm(ph[b]) + 1e-30)) for b in range(r.shape[0])]
        absn = [float(np.linalg.norm(r[b])) for b in range(r.shape[0])]
        return {"mean_rel_residual": float(np.mean(rel)), "max_rel_residual": float(np.max(rel)),
                "mean_abs_residual": float(np.mean(absn)), "max_abs_residual": float(np.max(absn))}

    def solver_info_stats(self) -> Dict[str, float]:
        """Iteration statistics of the last solve (the reference can only report
        NaN here because jax cg returns info=None, su3_wilson_nf2.py:1468-1554)."""
        it = [s["iters"] for s in self.last_solve_stats]
        if not it:
            return {"mean_iters": float("nan"), "max_iters": float("nan")}
        return {"mean_iters": float(np.mean(it)), "max_iters": float(np.max(it)), "min_iters": float(np.min(it))}
