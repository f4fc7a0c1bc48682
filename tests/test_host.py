"""CPU-side checks: the C ABI library loads and exports every declared entry
point, its argument validation (no device needed), the Python drop-in's
validation and shapes, the field generators, and the golden fixtures."""

import ctypes
from pathlib import Path

import numpy as np
import pytest

from jaxqft_amd import _native as nat
from jaxqft_amd import fields
from oracle import wilson_oracle as wo

GOLD = Path(__file__).resolve().parent / "golden"


def test_library_exports_declared_symbols():
    lib = nat.load()
    names = nat.declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.wl_version() >= 1


def _geom(L=(4, 4, 4, 4), nc=3, batch=1, nd=4):
    return nat.WlGeom(nd, (ctypes.c_int * 4)(*L), nc, batch)


def _params(r=1.0):
    return nat.WlParams(0.1, r, (ctypes.c_double * 4)(1, 1, 1, -1), (ctypes.c_double * 4)(0, 0, 0, 0))


@pytest.mark.parametrize("geom,params,code,msg", [
    (dict(nd=3), {}, nat.WL_ERR_UNSUPPORTED, "Nd=4"),
    (dict(nc=2), {}, nat.WL_ERR_UNSUPPORTED, "Nc"),
    (dict(batch=0), {}, nat.WL_ERR_BAD_ARG, "batch"),
    (dict(L=(4, 0, 4, 4)), {}, nat.WL_ERR_BAD_ARG, "extent"),
    ({}, dict(r=0.5), nat.WL_ERR_UNSUPPORTED, "wilson_r"),
])
def test_ctx_create_validation(geom, params, code, msg):
    lib = nat.load()
    h = ctypes.c_void_p()
    rc = lib.wl_ctx_create(ctypes.byref(_geom(**geom)), ctypes.byref(_params(**params)), ctypes.byref(h))
    assert rc == code
    assert msg in lib.wl_last_error().decode()


def test_null_handles_rejected():
    lib = nat.load()
    assert lib.wl_apply(None, None, 0, None, None, 1, 0, None) == nat.WL_ERR_BAD_ARG
    assert lib.wl_field_import(None, None, 0, None) == nat.WL_ERR_BAD_ARG
    sp = nat.WlSolveParams()
    assert lib.wl_cg_normal(None, None, None, None, None, 1, 0, ctypes.byref(sp), None, None) == nat.WL_ERR_BAD_ARG
    lib.wl_ctx_destroy(None)
    lib.wl_gauge_release(None)
    lib.wl_field_destroy(None)


def test_every_entry_point_rejects_null_handles():
    """Each entry point that takes a context, gauge, field or batch handle
    returns WL_ERR_BAD_ARG with a message for a null one (checked before any
    device call, so this runs without a GPU); the size queries return 0 and
    the destroy calls accept null."""
    lib = nat.load()
    B = nat.WL_ERR_BAD_ARG
    sp = nat.WlSolveParams()
    t = nat.WlTuning()
    dims = (ctypes.c_int * 4)(4, 4, 4, 4)
    out = ctypes.c_void_p()
    calls = {
        "wl_ctx_set_mass": lambda: lib.wl_ctx_set_mass(None, 0.1),
        "wl_ctx_set_tuning": lambda: lib.wl_ctx_set_tuning(None, 0, 1, ctypes.byref(t)),
        "wl_ctx_get_tuning": lambda: lib.wl_ctx_get_tuning(None, 0, 1, ctypes.byref(t)),
        "wl_ctx_set_eo_slab": lambda: lib.wl_ctx_set_eo_slab(None, 0, 1, 0),
        "wl_gauge_upload": lambda: lib.wl_gauge_upload(None, None, 0, 0, 0, ctypes.byref(out)),
        "wl_gauge_prepare": lambda: lib.wl_gauge_prepare(None, 0, 18, None),
        "wl_field_create": lambda: lib.wl_field_create(None, 0, 1, ctypes.byref(out)),
        "wl_field_export": lambda: lib.wl_field_export(None, None, 0, None),
        "wl_dslash": lambda: lib.wl_dslash(None, None, 0, 18, None, None, 1, None),
        "wl_cg_normal_field": lambda: lib.wl_cg_normal_field(None, None, None, None, 0, 1, ctypes.byref(sp), None, None),
        "wl_eo_apply": lambda: lib.wl_eo_apply(None, None, 0, 1, None, None, 1, 0, None),
        "wl_pf_force": lambda: lib.wl_pf_force(None, None, None, None, None, 0, 0, None),
        "wl_pf_force_field": lambda: lib.wl_pf_force_field(None, None, None, None, None, 0, 0, None),
        "wl_cg_eo_normal": lambda: lib.wl_cg_eo_normal(None, None, None, None, None, 1, 0, ctypes.byref(sp), None,
                                                       None),
        "wl_lime_decode": lambda: lib.wl_lime_decode(None, 0, 8, dims, 4, 1, 12, 0, None, 0, None),
        "wl_batch_create": lambda: lib.wl_batch_create(None, None, 1, ctypes.byref(sp), None, ctypes.byref(out)),
        "wl_batch_admit": lambda: lib.wl_batch_admit(None, 0, None, 0, 0.1),
        "wl_batch_step": lambda: lib.wl_batch_step(None),
        "wl_batch_query": lambda: lib.wl_batch_query(None, None, None),
        "wl_batch_retire": lambda: lib.wl_batch_retire(None, 0, None, 0),
        "wl_ctx_set_profiling": lambda: lib.wl_ctx_set_profiling(None, 1),
        "wl_ctx_get_profile": lambda: lib.wl_ctx_get_profile(None, None),
    }
    for name, call in calls.items():
        assert call() == B, name
        assert lib.wl_last_error(), name
    assert lib.wl_field_bytes(None) == 0
    assert lib.wl_batch_bytes(None, 4, 0) == 0
    lib.wl_batch_destroy(None)
    # every declared entry point that takes a handle is covered above or in the test before
    assert lib.wl_apply_noalloc(None, None, 0, None, None, 1, 0, None) == B
    assert "null context" in lib.wl_last_error().decode()
    covered = set(calls) | {"wl_apply", "wl_apply_noalloc", "wl_field_import", "wl_cg_normal", "wl_ctx_destroy", "wl_gauge_release",
                            "wl_field_destroy", "wl_field_bytes", "wl_batch_bytes", "wl_batch_destroy"}
    assert set(nat.declared_symbols()) - covered == {"wl_version", "wl_last_error", "wl_ctx_create",
                                                     "wl_stream_probe"}


def test_stream_probe_validation():
    """The HBM streaming probe checks its arguments before any launch."""
    lib = nat.load()
    B = nat.WL_ERR_BAD_ARG
    buf = ctypes.c_void_p(1 << 20)  # never dereferenced: every call below fails validation
    assert lib.wl_stream_probe(3, 0, 1, buf, buf, 4096, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_COPY, 0, 3, buf, buf, 4096, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_COPY, 0, 1, buf, buf, 4095, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_COPY, 0, 1, buf, buf, 0, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_READ, 0, 1, None, buf, 4096, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_WRITE, 0, 1, None, None, 4096, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_COPY, 1, 4, ctypes.c_void_p((1 << 20) + 8), buf, 4096, 0, None) == B
    assert lib.wl_stream_probe(nat.WL_STREAM_COPY, 1, 4, buf, buf, 4096, -1, None) == B
    assert "stream probe" in lib.wl_last_error().decode()


def test_theory_validation_and_shapes():
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    th = SU3WilsonNf2(lattice_shape=(4, 4, 4, 8), batch_size=2)
    assert th.fermion_shape() == (2, 4, 4, 4, 8, 4, 3)
    assert th.field_shape() == (2, 4, 4, 4, 4, 8, 3, 3)
    assert th.fermion_bc == (1, 1, 1, 1)
    assert SU3WilsonNf2(lattice_shape=(4,) * 4, layout="B...MIJ").field_shape() == (1, 4, 4, 4, 4, 4, 3, 3)
    u1 = U1WilsonNf2(lattice_shape=(4, 4, 4, 8))
    assert u1.fermion_shape() == (1, 4, 4, 4, 8, 4, 1)
    assert u1.field_shape() == (1, 4, 4, 4, 4, 8)
    assert u1.fermion_bc == (1, 1, 1, -1)
    for bad in (dict(layout="XYZ"), dict(solver_kind="minres"), dict(solver_form="split"),
                dict(solver_form="eo_split"), dict(solver_form="direct", solver_kind="gmres"),
                dict(preconditioner_kind="ilu"), dict(dirac_kernel="fast"), dict(fermion_bc=(1, 1)),
                dict(solver_precision="f8"), dict(solver_recon=8), dict(solver_search="pipelined")):
        with pytest.raises(ValueError):
            SU3WilsonNf2(lattice_shape=(4,) * 4, **bad)
    # the search-direction forms map onto the C ABI's values, recurrence by default
    assert nat.solve_params(1e-10, 10)[0].search_update == nat.WL_SEARCH_RECURRENCE
    assert nat.solve_params(1e-10, 10, search="xpay")[0].search_update == nat.WL_SEARCH_XPAY
    assert nat.WlSolveParams.search_update.offset == nat.WlSolveParams.accum_prec.offset + 4
    # the reference's solver kinds and forms (su3_wilson_nf2.py:634-639): a split form needs a
    # non-Hermitian solver kind there; here every one runs CG on the normal equations
    for kind, form in (("cg", "normal"), ("bicgstab", "normal"), ("gmres", "normal"), ("bicgstab", "split"),
                       ("gmres", "eo_split"), ("BiCGStab", "EO_Split")):
        t = SU3WilsonNf2(lattice_shape=(4,) * 4, solver_kind=kind, solver_form=form)
        assert (t.solver_kind, t.solver_form) == (kind.lower(), form.lower())
    with pytest.raises(ValueError, match="requires a non-Hermitian solver"):
        U1WilsonNf2(lattice_shape=(4,) * 4, solver_form="split")
    assert th.dirac.flops_per_site_matvec(nc=3) == 3600  # the reference's own estimator at Nc=3
    assert th.dirac.flops_per_site_dslash(nc=3) == 3552


def test_parity_projections_under_the_reference_names():
    """_project_even / _project_odd (the names the reference's even-odd monomial
    calls) are the projections: complementary, idempotent, and zero on the
    sites of the other parity (sum of coordinates mod 2)."""
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (2, 3, 4, 2)
    th = SU3WilsonNf2(lattice_shape=L)
    psi = fields.random_fermion(fields.rng_for(7), (1, *L, 4, 3))
    e, o = th._project_even(psi), th._project_odd(psi)
    np.testing.assert_array_equal(e + o, psi)
    np.testing.assert_array_equal(th._project_even(e), e)
    for x in np.ndindex(*L):
        zero = o if sum(x) % 2 == 0 else e
        assert not np.any(zero[(0, *x)])
    np.testing.assert_array_equal(th.project_even(psi), e)


@pytest.mark.parametrize("nc,layout", [(3, "BMXYIJ"), (3, "BXYMIJ"), (1, "BMXYIJ"), (1, "BXYMIJ")])
def test_links_with_fermion_bc(nc, layout):
    """_links_with_fermion_bc multiplies U_mu(x) on the slice x_mu = L_mu - 1 by
    the direction's phase and leaves every other link alone (checked link by
    link), batched and for one configuration; periodic phases return U itself;
    a link array of another rank is refused with the reference's message."""
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from jaxqft_amd.u1_wilson_nf2 import U1WilsonNf2

    L = (2, 3, 2, 4)
    bc = (np.exp(0.4j), 1, -1, np.exp(-1.1j))
    cls = SU3WilsonNf2 if nc == 3 else U1WilsonNf2
    th = cls(lattice_shape=L, fermion_bc=bc, batch_size=2, layout=layout)
    rng = fields.rng_for(11)
    U = fields.su3_hot_start(rng, 2, L, 0.5, layout=layout) if nc == 3 else fields.u1_hot_start(rng, 2, L, 0.5,
                                                                                                layout=layout)
    got = th._links_with_fermion_bc(U)
    for b in range(2):
        for mu in range(4):
            for x in np.ndindex(*L):
                idx = (b, mu, *x) if layout == "BMXYIJ" else (b, *x, mu)
                want = U[idx] * (th.fermion_bc[mu] if x[mu] == L[mu] - 1 else 1.0)
                # one rounding of a complex product (vectorised and scalar products may differ by an ulp)
                np.testing.assert_allclose(got[idx], want, rtol=1e-15, atol=0)
    np.testing.assert_array_equal(th._links_with_fermion_bc(U[0]), got[0])
    per = cls(lattice_shape=L, fermion_bc=(1, 1, 1, 1), batch_size=2, layout=layout)
    assert per._links_with_fermion_bc(U) is U
    with pytest.raises(ValueError, match="Unexpected link tensor rank"):
        th._links_with_fermion_bc(U[0, 0])


def test_su3_field_generator():
    rng = fields.rng_for(5)
    U = fields.su3_hot_start(rng, 1, (4, 4, 4, 4), 0.5).reshape(-1, 3, 3)
    np.testing.assert_allclose(np.einsum("nij,nkj->nik", U, np.conj(U)), np.broadcast_to(np.eye(3), U.shape), atol=1e-13)
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-13)
    # exp of the same algebra element via scipy
    from scipy.linalg import expm

    rng = fields.rng_for(6)
    a = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    X = fields.project_su3_algebra(0.8 * a[None])
    np.testing.assert_allclose(fields.expm_su3_algebra(X)[0], expm(X[0]), atol=1e-13)


def test_point_sources():
    src = fields.point_sources((2, 2, 2, 4), (1, 0, 1, 3))
    assert src.shape == (12, 2, 2, 2, 4, 4, 3)
    assert np.count_nonzero(src) == 12
    for k in range(12):
        assert src[k, 1, 0, 1, 3, k // 3, k % 3] == 1


def _explicit_wilson(U, psi, L, mass, bc, dag):
    """Textbook Wilson operator by a loop over sites, independent of the
    oracle's rolls and sparse gamma tables:
      (D psi)(x) = (m+4) psi(x) - 1/2 sum_mu [ (1 - s g_mu) U~_mu(x) psi(x+mu)
                                             + (1 + s g_mu) U~_mu(x-mu)^dag psi(x-mu) ]
    with s = +1 for D and -1 for D^dag (= g5 D g5), dense gamma matrices
    g0 = s1 x s1, g1 = s1 x s2, g2 = s2 x 1, g3 = s3 x 1, and U~_mu(x) =
    p_mu U_mu(x) on the slice x_mu = L_mu - 1."""
    s1 = np.array([[0, 1], [1, 0]], complex)
    s2 = np.array([[0, -1j], [1j, 0]], complex)
    s3 = np.array([[1, 0], [0, -1]], complex)
    g = [np.kron(s1, s1), np.kron(s1, s2), np.kron(s2, np.eye(2)), np.kron(s3, np.eye(2))]
    sg = -1.0 if dag else 1.0
    nc = psi.shape[-1]
    out = np.zeros_like(psi)
    for x in np.ndindex(*L):
        acc = (mass + 4.0) * psi[(0, *x)]
        for mu in range(4):
            xf, xb = list(x), list(x)
            xf[mu] = (x[mu] + 1) % L[mu]
            xb[mu] = (x[mu] - 1) % L[mu]
            uf = np.asarray(U[(0, mu, *x)]).reshape(nc, nc)
            ub = np.asarray(U[(0, mu, *xb)]).reshape(nc, nc)
            if x[mu] == L[mu] - 1:
                uf = bc[mu] * uf
            if xb[mu] == L[mu] - 1:
                ub = bc[mu] * ub
            fwd = (np.eye(4) - sg * g[mu]) @ (uf @ psi[(0, *xf)].T).T
            bwd = (np.eye(4) + sg * g[mu]) @ (np.conj(ub.T) @ psi[(0, *xb)].T).T
            acc = acc - 0.5 * (fwd + bwd)
        out[(0, *x)] = acc
    return out


@pytest.mark.parametrize("name", ["op_su3_2x2x2x2", "op_su3_4x4x4x4", "op_su3_2x3x4x5", "op_u1_2x2x2x2",
                                  "op_u1_4x4x4x4"])
def test_golden_operator_fixtures_explicit(name):
    """The committed operator fixtures (which the GPU tests compare against)
    equal the explicit site-loop formula above: D and D^dag directly, g5 D
    with g5 = s1 x s3, D^dag D as the explicit D^dag of the explicit D psi,
    and the hopping terms as D - (m+4) and D^dag - (m+4)."""
    z = np.load(GOLD / f"{name}.npz")
    L = tuple(int(v) for v in z["L"])
    mass = float(z["mass"])
    bc = tuple(complex(b) for b in z["bc"])
    U, psi = z["U"], z["psi"]
    D = _explicit_wilson(U, psi, L, mass, bc, dag=False)
    Dd = _explicit_wilson(U, psi, L, mass, bc, dag=True)
    g5 = np.kron(np.array([[0, 1], [1, 0]]), np.diag([1, -1]))
    expect = {"D": D, "Ddag": Dd, "g5D": np.einsum("st,...tc->...sc", g5, D),
              "DdagD": _explicit_wilson(U, D, L, mass, bc, dag=True),
              "K": D - (mass + 4.0) * psi, "Kdag": Dd - (mass + 4.0) * psi}
    for key, ref in expect.items():
        assert np.linalg.norm(z[key] - ref) / np.linalg.norm(ref) < 1e-14, key
    # the oracle reproduces the fixtures bit for bit up to rounding
    ora = wo.WilsonOracle(L, psi.shape[-1], mass, fermion_bc=bc)
    assert np.linalg.norm(ora.apply_D(U, psi) - z["D"]) / np.linalg.norm(z["D"]) < 1e-14


def test_golden_cg_fixture():
    z = np.load(GOLD / "cg_su3_4x4x4x4.npz")
    ora = wo.WilsonOracle((4, 4, 4, 4), 3, 0.1, fermion_bc=(1, 1, 1, -1))
    r = ora.apply_normal(z["U"], z["x"]) - z["b"]
    assert np.linalg.norm(r) / np.linalg.norm(z["b"]) <= 1e-11


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (8, 8, 8, 8), (4, 4, 4, 8)])
def test_golden_index_tables(L):
    z = np.load(GOLD / f"index_{'x'.join(map(str, L))}.npz")
    t = wo.checkerboard_neighbor_maps(L)
    for k in t:
        np.testing.assert_array_equal(t[k], z[k])
        assert t[k].dtype == np.int32


def test_gamma_algebra_checker_residuals():
    """check_gamma_algebra: zero on the exact bases of every dimension; a
    basis with g1 scaled by (1 + e) reports (1 + e)^2 - 1 (its own
    anticommutator is 2 (1 + e)^2); a non-Hermitian entry of size d reports
    hermiticity error sqrt(2) d."""
    from jaxqft_amd.fermions.gamma import build_euclidean_gamma, check_gamma_algebra

    for d in (2, 3, 4, 5, 6):
        r = check_gamma_algebra(build_euclidean_gamma(d))
        assert r["max_rel_clifford_error"] == 0.0 and r["max_hermiticity_error"] == 0.0
    g = build_euclidean_gamma(4)
    g[1] = g[1] * 1.001
    r = check_gamma_algebra(g)
    assert abs(r["max_rel_clifford_error"] - (1.001 ** 2 - 1)) < 1e-12
    assert r["max_hermiticity_error"] == 0.0
    g = build_euclidean_gamma(4)
    g[2, 0, 1] += 1e-3j
    assert abs(check_gamma_algebra(g)["max_hermiticity_error"] - np.sqrt(2) * 1e-3) < 1e-15
