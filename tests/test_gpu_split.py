"""GPU tests of the split solver forms and the direct / dagger solves
(su3_wilson_nf2.py:1271-1454), which the drop-in serves with CG on the normal
equations (D^dag D, or S^dag S for the normalized even Schur complement S):

* every solution is checked against the FP64 oracle's operator: D x = b,
  D^dag x = b (directly and through the even-odd reduction), S x_e = P_e b
  and S^dag x_e = P_e b, with x_e zero on the odd sites;
* a split pair (x, y) satisfies D^dag y = phi and D x = y, for 'split' and
  'eo_split';
* the split-form pseudofermion action |chi|^2 (D^dag chi = phi), the even-odd
  one (S^dag chi = P_e phi) and the split-form force equal the normal-form
  values, which use no split solve at all.

Tolerances: the solves stop at the FP64 true residual 1e-12 of the normal
system; the direct residual is then bounded by about cond(D) 1e-12, so 1e-9
catches a wrong operator, sign or parity while leaving room for it."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = (4, 4, 4, 8)
BC = (1, 1, 1, -1)
MASS = 0.1


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def theory(nc, form="normal", kind="cg"):
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    return cls(lattice_shape=L, mass=MASS, fermion_bc=BC, dtype=np.complex128, cg_tol=1e-12, cg_maxiter=4000,
               solver_kind=kind, solver_form=form)


def inputs(nc, seed):
    from jaxqft_amd import fields

    rng = fields.rng_for(seed)
    U = fields.su3_hot_start(rng, 1, L, 0.4) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.4)
    b = fields.random_fermion(rng, (1, *L, 4, nc))
    return U, b


@pytest.mark.parametrize("nc", [3, 1])
def test_direct_and_dagger_solves(nc):
    print(f"[gpu] direct/dagger solves nc={nc}", flush=True)
    from oracle.wilson_oracle import WilsonOracle

    th = theory(nc)
    ora = WilsonOracle(L, nc, MASS, fermion_bc=BC)
    U, b = inputs(nc, 4100 + nc)
    xs = {}
    for name, op in (("solve_direct", ora.apply_D), ("solve_dagger", ora.apply_Ddag),
                     ("solve_direct_eo", ora.apply_D), ("solve_dagger_eo", ora.apply_Ddag)):
        x = getattr(th, name)(U, b)
        assert rel(op(U, x), b) < 1e-9, name
        xs[name] = x
    # the full and the even-odd routes reach the same solution
    assert rel(xs["solve_direct_eo"], xs["solve_direct"]) < 1e-8
    assert rel(xs["solve_dagger_eo"], xs["solve_dagger"]) < 1e-8
    # D and D^dag differ: the two solutions must too
    assert rel(xs["solve_direct"], xs["solve_dagger"]) > 1e-3


@pytest.mark.parametrize("nc", [3, 1])
def test_schur_direct_and_dagger_solves(nc):
    print(f"[gpu] Schur direct/dagger solves nc={nc}", flush=True)
    from oracle.wilson_oracle import WilsonOracle

    th = theory(nc)
    ora = WilsonOracle(L, nc, MASS, fermion_bc=BC)
    U, b = inputs(nc, 4200 + nc)
    b_e = ora.project_even(b)
    x = th.solve_eop_even_direct(U, b)
    assert np.linalg.norm(ora.project_odd(x)) == 0.0
    assert rel(ora.apply_eo_schur_even(U, x, normalized=True), b_e) < 1e-9
    y = th.solve_eop_even_dagger(U, b)
    assert np.linalg.norm(ora.project_odd(y)) == 0.0
    assert rel(ora.apply_eo_schur_even_dagger(U, y, normalized=True), b_e) < 1e-9
    assert rel(x, y) > 1e-3
    # the split pair of the Schur system: S^dag y = P_e phi, S x = y
    x2, y2 = th.solve_eop_even_split_with_intermediate(U, b)
    assert rel(ora.apply_eo_schur_even_dagger(U, y2, normalized=True), b_e) < 1e-9
    assert rel(ora.apply_eo_schur_even(U, x2, normalized=True), y2) < 1e-9


@pytest.mark.parametrize("nc", [3, 1])
@pytest.mark.parametrize("form,kind", [("split", "bicgstab"), ("eo_split", "gmres")])
def test_split_forms_match_the_normal_form(nc, form, kind):
    print(f"[gpu] split form {form} nc={nc}", flush=True)
    from oracle.wilson_oracle import WilsonOracle

    sp = theory(nc, form, kind)
    nm = theory(nc)
    ora = WilsonOracle(L, nc, MASS, fermion_bc=BC)
    U, eta = inputs(nc, 4300 + nc)
    phi = ora.apply_Ddag(U, eta)
    x, y = sp.solve_split_with_intermediate(U, phi)
    assert rel(ora.apply_Ddag(U, y), phi) < 1e-9
    assert rel(ora.apply_D(U, x), y) < 1e-9
    assert rel(sp.solve_split(U, phi), x) < 1e-12
    assert rel(sp.solve_pseudofermion(U, phi), x) < 1e-12
    # phi = D^dag eta: the split pair is (D^-1 eta, eta) and the action |eta|^2
    assert rel(y, eta) < 1e-9
    s_split = sp.pseudofermion_action(U, phi)
    s_norm = nm.pseudofermion_action(U, phi)
    assert abs(s_split[0] - s_norm[0]) / abs(s_norm[0]) < 1e-9
    assert abs(s_split[0] - np.vdot(eta, eta).real) / np.vdot(eta, eta).real < 1e-9
    e_split = sp.eop_pseudofermion_action(U, phi)
    e_norm = nm.eop_pseudofermion_action(U, phi)
    assert abs(e_split[0] - e_norm[0]) / abs(e_norm[0]) < 1e-9
    assert rel(sp.solve_eop_even(U, phi), nm.solve_eop_even(U, phi)) < 1e-8
    f_split = sp.pseudofermion_force(U, phi)
    f_norm = nm.pseudofermion_force(U, phi)
    assert rel(f_split, f_norm) < 1e-8
