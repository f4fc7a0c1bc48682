"""CPU tests: the drop-in theory classes accept every constructor field of the
reference dataclasses (SU3WilsonNf2, su3_wilson_nf2.py:594-623 with the
SU3YangMills base fields su3_ym.py:181-188; U1WilsonNf2,
u1_wilson_nf2.py:296-320 with U1YangMills u1_ym.py:124-130), validate them
with the reference's messages, and keep the solver-guess history the way the
reference monomial does (su3_wilson_nf2.py:254-294)."""

import numpy as np
import pytest

from jaxqft_amd.chrono import SolverGuess
from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

# every init field of the reference SU3WilsonNf2 (base class first), with a non-default value
# where the drop-in supports it
SU3_FIELDS = dict(
    lattice_shape=(4, 4, 4, 8), beta=5.6, batch_size=2, layout="BXYMIJ", seed=3, exp_method="su3",
    dtype=np.complex128,
    mass=0.05, wilson_r=1.0, cg_tol=1e-10, cg_maxiter=700, solver_kind="cg", solver_form="normal",
    preconditioner_kind="jacobi", gmres_restart=16, gmres_solve_method="incremental", fermion_bc=(1, 1, 1, -1),
    dirac_kernel="reference", eo_use_half_spinor=False, include_gauge_monomial=False,
    include_fermion_monomial=True, fermion_monomial_kind="eo_preconditioned", gauge_timescale=1,
    fermion_timescale=2, pseudofermion_refresh="ou", pseudofermion_gamma=None, pseudofermion_force_mode="analytic",
    smd_gamma=0.4, auto_refresh_pseudofermions=False, use_solver_guess=True, solver_guess_mode="MRE",
    solver_guess_history=4, jit_dirac_kernels=False, jit_solvers=False,
)

U1_FIELDS = dict(
    lattice_shape=(4, 4, 4, 4), beta=1.0, batch_size=1, layout="BMXYIJ", seed=0, dtype=np.complex64,
    mass=0.1, wilson_r=1.0, cg_tol=1e-8, cg_maxiter=500, solver_kind="cg", solver_form="normal",
    preconditioner_kind="none", gmres_restart=32, gmres_solve_method="batched", fermion_bc=None,
    dirac_kernel="optimized", include_gauge_monomial=True, include_fermion_monomial=True,
    fermion_monomial_kind="unpreconditioned", gauge_timescale=0, fermion_timescale=1, pseudofermion_refresh="heatbath",
    pseudofermion_gamma=None, pseudofermion_force_mode="autodiff", smd_gamma=0.3, auto_refresh_pseudofermions=True,
    jit_dirac_kernels=True, jit_solvers=True,
)


def test_su3_accepts_every_reference_field():
    th = SU3WilsonNf2(**SU3_FIELDS)
    assert th.exp_method == "su3"
    assert th.solver_guess_mode == "mre" and th.solver_guess_history == 4 and th.use_solver_guess
    assert th.fermion_monomial_kind == "eo_preconditioned"
    # the reference derives the OU refresh rate from smd_gamma (su3_wilson_nf2.py:659-661)
    assert th.pseudofermion_gamma == pytest.approx(0.4)
    assert th.fermion_shape() == (2, 4, 4, 4, 8, 4, 3)
    assert th.field_shape() == (2, 4, 4, 4, 8, 4, 3, 3)
    assert SU3WilsonNf2(lattice_shape=(4, 4, 4, 4)).pseudofermion_gamma == pytest.approx(0.3)


def test_u1_accepts_every_reference_field():
    th = U1WilsonNf2(**U1_FIELDS)
    assert th.fermion_bc[-1] == -1  # U(1) default: antiperiodic in the last axis (u1_wilson_nf2.py:461-471)
    assert th.fermion_shape() == (1, 4, 4, 4, 4, 4, 1)


@pytest.mark.parametrize("field,value,msg", [
    ("solver_guess_mode", "cubic", "solver_guess_mode must be one of: last, poly, mre"),
    ("solver_guess_history", 0, "solver_guess_history must be positive"),
    ("gmres_restart", 0, "gmres_restart must be positive"),
    ("fermion_monomial_kind", "twisted", "fermion_monomial_kind must be one of"),
    ("exp_method", "pade", "exp_method must be one of: expm, su3"),
    ("dirac_kernel", "fast", "dirac_kernel must be one of: optimized, reference"),
    ("preconditioner_kind", "ilu", "preconditioner_kind must be one of: none, jacobi"),
])
def test_reference_validation_messages(field, value, msg):
    with pytest.raises(ValueError, match=msg):
        SU3WilsonNf2(lattice_shape=(4, 4, 4, 4), **{field: value})


def test_solver_guess_last_and_poly():
    g = SolverGuess("last", 1)
    phi = np.zeros((1, 2, 2, 2, 2, 4, 3), dtype=np.complex128)
    assert g.predict(phi) is None
    x1, x2, x3 = (np.full(phi.shape, v, dtype=np.complex128) for v in (1.0, 2.0, 4.0))
    g.update(x1)
    g.update(x2)
    assert np.array_equal(g.predict(phi), x2) and len(g.xs) == 1
    # 'poly' over three entries: 3 x_k - 3 x_{k-1} + x_{k-2} (exact on quadratic histories)
    p = SolverGuess("poly", 3)
    for v in (1.0, 4.0, 9.0):
        p.update(np.full(phi.shape, v, dtype=np.complex128))
    assert np.allclose(p.predict(phi), 16.0)
    # entries of another shape are dropped
    other = np.zeros((1, 2, 2, 2, 4, 4, 3), dtype=np.complex128)
    assert p.predict(other) is None and not p.xs
    p.update(x3)
    p.clear()
    assert p.predict(phi) is None


def test_solver_guess_mre_fits_target_and_falls_back():
    rng = np.random.default_rng(4)
    n = 48
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 8 * np.eye(n)
    xs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3)]
    target = A @ (0.7 * xs[2] + 0.3 * xs[1])
    g = SolverGuess("mre", 3)
    for x in xs:
        g.update(x, A @ x)
    got = g.predict(xs[0], target)
    # the target lies in the span of the images: the blended guess reproduces it
    assert np.linalg.norm(A @ got - target) / np.linalg.norm(target) < 1e-8
    # without the target the history extrapolates as 'poly'
    poly = g.predict(xs[0])
    assert np.allclose(poly, 3 * xs[2] - 3 * xs[1] + xs[0])


@pytest.mark.parametrize("cls_name", ["SU3WilsonNf2", "U1WilsonNf2"])
def test_model_surface_used_by_the_reference_callers(cls_name):
    """Every model attribute that the reference's pseudofermion monomials
    (su3_wilson_nf2.py:218-590, u1_wilson_nf2.py:96-293) and measurements
    (core/measurements.py:277-371, :1600-1625) read on the fermion path exists
    on the drop-in, private names included."""
    import importlib

    mod = importlib.import_module("jaxqft_amd." + ("su3_wilson_nf2" if cls_name == "SU3WilsonNf2" else "u1_wilson_nf2"))
    th = getattr(mod, cls_name)(lattice_shape=(4, 4, 4, 4))
    used = ["solver_form", "random_fermion", "dtype", "_project_even", "_project_odd", "solve_normal",
            "solve_normal_with_guess", "solve_eop_even_normal", "solve_eop_even_normal_with_guess",
            "apply_eo_schur_even", "apply_eo_schur_even_dagger", "apply_D", "apply_Ddag", "solve_dagger",
            "solve_dagger_eo", "solve_direct", "solve_split_with_intermediate", "solve_eop_even_dagger",
            "solve_eop_even_split_with_intermediate", "pseudofermion_action", "eop_pseudofermion_action",
            "pseudofermion_force_from_solution", "eop_pseudofermion_force_from_solution", "fermion_shape",
            "gamma", "wilson_r", "dirac", "layout", "lattice_shape", "_links_with_fermion_bc",
            "use_solver_guess", "solver_guess_mode", "solver_guess_history"]
    missing = [n for n in used if not hasattr(th, n)]
    assert not missing, missing
    assert hasattr(th.dirac, "spin_eye")
