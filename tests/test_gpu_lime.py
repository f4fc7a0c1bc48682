"""GPU parity of the LIME payload decode (wl_lime_decode) against the oracle
and against closed-form fields: bit-exact for every order, precision and
lattice rank, on tile-ragged extents, at 32^4 through a device round trip,
and end to end through the Wilson operator on a decoded configuration."""

import numpy as np
import pytest
import torch

import lime_files as lf

pytestmark = pytest.mark.gpu

CASES = [
    # datatype, precision, colors, spins, ncomplex, datacount, dims
    ("USQCD_F3_ColorMatrix", "F", 3, 4, 9, 4, (4, 3, 2, 5)),
    ("USQCD_D3_ColorMatrix", "D", 3, 4, 9, 4, (10, 3, 2, 13)),   # 5 x 7 tiles: x3 in runs of 7 and 6
    ("QDP_D3_DiracFermion", "D", 3, 4, 12, 1, (18, 2, 3, 11)),   # 8 x 11 tiles: x0 in runs of 8, 8 and 2
    ("Lattice", "F", 3, 4, 12, 1, (33, 2, 2, 35)),               # 8 x 16 tiles, ragged on both axes
    ("Lattice", "F", 3, 4, 12, 1, (5, 1, 3, 2)),
    ("QDP_F3_ColorVector", "F", 3, 0, 3, 2, (3, 4, 5)),
    ("QDP_D1_ColorMatrix", "D", 1, 0, 1, 2, (6, 5)),
    ("Generic", "D", 0, 0, 5, 3, (7,)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] + "_" + c[1] + "_" + "x".join(map(str, c[6])) for c in CASES])
def test_closed_form_field_decodes_exactly(tmp_path, case):
    print(f"[gpu] lime field {case}", flush=True)
    from jaxqft_amd.io import lime
    from oracle import lime_oracle as lo

    datatype, prec, colors, spins, ncx, dc, dims = case
    raw = lf.payload_by_loops(lf.coord_value, dims, dc, ncx, prec)
    p = lf.field_file(tmp_path / "f.lime", dims, raw, datatype, prec, colors, spins, ncx, dc)
    dec = lime.decode_scidac_field(p)
    wb = 4 if prec == "F" else 8
    ref = lo.decode_payload(raw, prec, 2 * wb * ncx, dc, dims, datatype, colors, spins)
    for key in ("site_data", "lattice_data"):
        assert dec[key].dtype == ref[key].dtype and dec[key].shape == ref[key].shape, key
        np.testing.assert_array_equal(dec[key], ref[key])
    np.testing.assert_array_equal(dec["lattice_data"].reshape(dims + (dc, ncx)), lf.coord_lattice(dims, dc, ncx))
    dd = lime.decode_scidac_field(p, device="cuda")
    assert dd["site_data"].is_cuda
    np.testing.assert_array_equal(dd["lattice_data"].cpu().numpy(), ref["lattice_data"])


@pytest.mark.parametrize("prec,dims,ildg", [("D", (8, 8, 8, 8), True), ("F", (8, 8, 8, 8), False),
                                            ("D", (4, 6, 2, 18), True), ("F", (17, 3, 1, 9), True)])
def test_gauge_matches_oracle(tmp_path, prec, dims, ildg):
    print(f"[gpu] lime gauge {prec} {dims}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.io import lime
    from oracle import lime_oracle as lo

    rng = fields.rng_for(4100)
    U = fields.su3_hot_start(rng, 1, dims, 0.7)[0]           # (4, *dims, 3, 3)
    lat = np.moveaxis(U, 0, 4).reshape(*dims, 4, 9)          # (*dims, 4, 9)
    raw = lf.payload_from_lattice(lat, 4, prec)
    p = lf.gauge_file(tmp_path / "g.lime", dims, raw, prec, ildg=ildg)
    wb = 4 if prec == "F" else 8
    ref = lo.gauge_from_lattice(lo.decode_payload(raw, prec, 18 * wb, 4, dims, "ColorMatrix", 3, 4)["lattice_data"],
                                4, batch_size=2)
    got = lime.decode_scidac_gauge(p, batch_size=2)
    assert got.shape == (2, 4, *dims, 3, 3) and got.dtype == ref.dtype
    np.testing.assert_array_equal(got, ref)
    if prec == "D":
        np.testing.assert_array_equal(got[0], U)
    dev = lime.decode_scidac_gauge(p, device="cuda")
    assert dev.is_cuda and dev.shape == (1, 4, *dims, 3, 3)
    np.testing.assert_array_equal(dev.cpu().numpy(), ref[:1])


@pytest.mark.parametrize("prec,dims", [("F", (4, 4, 4, 8)), ("D", (18, 2, 3, 11))])
def test_pseudofermion_leaf_matches_oracle(tmp_path, prec, dims):
    print(f"[gpu] lime pseudofermion {prec} {dims}", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.io import lime
    from oracle import lime_oracle as lo

    rng = fields.rng_for(4200)
    phi = fields.random_fermion(rng, (*dims, 4, 3))
    raw = lf.payload_from_lattice(phi.reshape(*dims, 1, 12), 4, prec)
    p = lf.pseudofermion_file(tmp_path / "pf.lime", dims, raw, prec)
    wb = 4 if prec == "F" else 8
    ref = lo.pseudofermion_from_lattice(
        lo.decode_payload(raw, prec, 24 * wb, 1, dims, "Lattice", 3, 4)["lattice_data"], 1, batch_size=3)
    got, meta = lime.decode_scidac_pseudofermion(p, batch_size=3)
    assert meta.recordtype == 0 and meta.datatype == "Lattice"
    assert got.shape == (3, *dims, 4, 3)
    np.testing.assert_array_equal(got, ref)
    got1, _ = lime.decode_scidac_pseudofermion(p, field_index=1)
    np.testing.assert_array_equal(got1, ref[:1])
    if prec == "D":
        np.testing.assert_array_equal(got1[0], phi)


def test_momentum_diagnostics(tmp_path):
    print("[gpu] lime momentum", flush=True)
    from jaxqft_amd.io import lime
    from oracle import lime_oracle as lo

    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(11)
    h = rng.normal(size=(4, *dims, 3, 3)) + 1j * rng.normal(size=(4, *dims, 3, 3))
    a = 0.5 * (h - np.swapaxes(h.conj(), -1, -2))
    a = a - np.eye(3) * (np.trace(a, axis1=-2, axis2=-1)[..., None, None] / 3)
    for P, tol in [(a, 1e-14), (a + 0.1 * np.eye(3), None)]:
        raw = lf.payload_from_lattice(np.moveaxis(P, 0, 4).reshape(*dims, 4, 9), 4, "D")
        p = lf.gauge_file(tmp_path / "m.lime", dims, raw, "D")
        mom, info = lime.decode_scidac_momentum(p)
        np.testing.assert_array_equal(mom[0], P)
        want = lo.momentum_diagnostics(mom)
        for k in want:
            np.testing.assert_allclose(info[k], want[k], rtol=1e-12, atol=1e-15)
        if tol is not None:
            assert info["rel_antihermitian_error"] < tol and info["max_abs_trace"] < tol


def test_narrowing_and_widening_are_exact():
    print("[gpu] lime narrowing/widening", flush=True)
    from jaxqft_amd.io import lime

    dims = (6, 5, 4, 3)
    rng = np.random.default_rng(5)
    z = (rng.normal(size=(*dims, 4, 9)) + 1j * rng.normal(size=(*dims, 4, 9))) * 10.0 ** rng.integers(-30, 30, (*dims, 4, 9))
    for prec, out, want in [("D", torch.complex64, z.astype(np.complex64)),
                            ("F", torch.complex128, z.astype(np.complex64).astype(np.complex128))]:
        raw = lf.payload_from_lattice(z, 4, prec)
        pay = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()
        got = lime.decode_payload(pay, 4 if prec == "F" else 8, dims, 4, 9, "site", out).cpu().numpy()
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("prec", ["D", "F"])
def test_full_size_device_round_trip_32x32x32x32(prec):
    """A 32^4 SU(3) link field encoded on the device (lexicographic order,
    bytes reversed per word by torch) decodes back bit for bit in all three
    orders."""
    print(f"[gpu] lime 32^4 round trip {prec}", flush=True)
    from jaxqft_amd.io import lime

    L = (32, 32, 32, 32)
    g = torch.Generator(device="cuda").manual_seed(7)
    rdt = torch.float64 if prec == "D" else torch.float32
    wb = 8 if prec == "D" else 4
    U = torch.randn((4, *L, 3, 3, 2), generator=g, device="cuda", dtype=rdt)        # (mu, x0..x3, i, j, re/im)
    lexi = U.permute(4, 3, 2, 1, 0, 5, 6, 7).contiguous()                           # (x3, x2, x1, x0, mu, i, j, re/im)
    pay = lexi.view(torch.uint8).view(-1, wb).flip(-1).reshape(-1).contiguous()
    del lexi
    cdt = torch.complex128 if prec == "D" else torch.complex64
    Uc = torch.view_as_complex(U).reshape(4, *L, 9)
    got = lime.decode_payload(pay, wb, L, 4, 9, "count", cdt)
    assert torch.equal(got, Uc)
    del got
    got = lime.decode_payload(pay, wb, L, 4, 9, "site", cdt)
    assert torch.equal(got, Uc.permute(1, 2, 3, 4, 0, 5))
    del got
    got = lime.decode_payload(pay, wb, L, 4, 9, "lexi", cdt)
    assert torch.equal(got.view(32, 32, 32, 32, 4, 9), Uc.permute(4, 3, 2, 1, 0, 5))


def test_decoded_configuration_drives_the_operator(tmp_path):
    print("[gpu] lime -> resident gauge -> D", flush=True)
    from jaxqft_amd import fields
    from jaxqft_amd.io import lime
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2
    from oracle import lime_oracle as lo
    from oracle.wilson_oracle import WilsonOracle

    L = (4, 4, 4, 8)
    rng = fields.rng_for(4300)
    U = fields.su3_hot_start(rng, 1, L, 0.6)[0]
    raw = lf.payload_from_lattice(np.moveaxis(U, 0, 4).reshape(*L, 4, 9), 4, "D")
    p = lf.gauge_file(tmp_path / "cfg.lime", L, raw, "D")
    th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=(1, 1, 1, -1), dtype=np.complex128)
    rg = th.resident_gauge(lime.decode_scidac_gauge(p, device="cuda"))
    psi = fields.random_fermion(rng, th.fermion_shape())
    out = th.apply_D(rg, psi)
    Uo = lo.gauge_from_lattice(lo.decode_payload(raw, "D", 144, 4, L, "ColorMatrix", 3, 4)["lattice_data"], 4)
    ref = WilsonOracle(L, 3, 0.05, fermion_bc=(1, 1, 1, -1)).apply_D(Uo, psi)
    assert np.max(np.abs(out - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
