"""CPU checks of the LIME reader: record headers, SciDAC field discovery and
metadata, the host-side validation that runs before any device work, the
argument checks of wl_lime_decode, and the pinning of the payload-decode
oracle against closed-form fields written value by value.

The reference tree holds no LIME files; the files here are written by
``tests/lime_files.py`` from the published record layout."""

import ctypes
import struct

import numpy as np
import pytest
import torch

from jaxqft_amd import _native as nat
from jaxqft_amd.io import lime
from oracle import lime_oracle as lo

import lime_files as lf


# ---------------------------------------------------------------- records
def test_record_headers_offsets_padding_and_flags(tmp_path):
    recs = [("alpha", b"x" * 5, lf.MB), ("beta-xml", b"<a>1</a>", 0), ("gamma", b"", lf.ME), ("delta", b"y" * 17, 0)]
    p = tmp_path / "r.lime"
    p.write_bytes(b"".join(lf.record(t, d, f) for t, d, f in recs))
    heads = lime.list_lime_records(p)
    assert [h.type_name for h in heads] == ["alpha", "beta-xml", "gamma", "delta"]
    assert [h.size for h in heads] == [5, 8, 0, 17]
    # each payload is padded to a multiple of 8 bytes
    assert [h.offset for h in heads] == [0, 144 + 8, 2 * 144 + 16, 3 * 144 + 16]
    assert [h.index for h in heads] == [0, 1, 2, 3]
    assert heads[0].is_begin_message and not heads[0].is_end_message
    assert heads[2].is_end_message and not heads[2].is_begin_message
    assert all(h.version == 1 for h in heads)
    assert heads[1].data_offset == heads[1].offset + 144
    assert lime.read_lime_record_data(p, 3) == b"y" * 17
    assert lime.read_lime_record_data(p, heads[1]) == b"<a>1</a>"
    assert lime.read_lime_record_data(p, 2) == b""
    with pytest.raises(IndexError):
        lime.read_lime_record_data(p, 4)
    with pytest.raises(IndexError):
        lime.read_lime_record_data(p, -1)


def test_record_header_errors(tmp_path):
    good = lf.record("alpha", b"abc")
    p = tmp_path / "bad_magic.lime"
    p.write_bytes(good + struct.pack(">IHHQ", 0x12345678, 1, 0, 0) + b"\x00" * 128)
    with pytest.raises(ValueError, match="Bad LIME magic 0x12345678 at offset 152"):
        lime.list_lime_records(p)
    p2 = tmp_path / "short.lime"
    p2.write_bytes(good + good[:100])
    with pytest.raises(ValueError, match="Truncated LIME header at offset 152"):
        lime.list_lime_records(p2)
    p3 = tmp_path / "empty.lime"
    p3.write_bytes(b"")
    assert lime.list_lime_records(p3) == []


def test_type_name_stops_at_first_nul(tmp_path):
    head = struct.pack(">IHHQ", lf.MAGIC, 1, 0, 0) + b"name\x00junk" + b"\x00" * 119
    p = tmp_path / "n.lime"
    p.write_bytes(head)
    assert lime.list_lime_records(p)[0].type_name == "name"


# -------------------------------------------------------- field discovery
def test_gauge_file_fields_and_dims(tmp_path):
    dims = (4, 3, 2, 5)
    pay = b"\x00" * (int(np.prod(dims)) * 4 * 72)
    p = lf.gauge_file(tmp_path / "g.lime", dims, pay, "F")
    fields = lime.find_scidac_fields(p)
    assert len(fields) == 1
    m = fields[0]
    assert m.datatype == "USQCD_F3_ColorMatrix" and m.precision == "F"
    assert (m.colors, m.spins, m.typesize, m.datacount, m.recordtype) == (3, 4, 72, 4, 0)
    assert m.dims == dims
    heads = lime.list_lime_records(p)
    names = [h.type_name for h in heads]
    assert m.private_record_index == names.index("scidac-private-record-xml")
    assert m.record_xml_index == names.index("scidac-record-xml")
    assert m.binary_record_index == names.index("ildg-binary-data")
    assert m.binary_record_type == "ildg-binary-data"
    assert m.file_xml_index == names.index("scidac-private-file-xml")


def test_dims_fall_back_to_ildg_format(tmp_path):
    dims = (4, 4, 2, 6)
    p = lf.gauge_file(tmp_path / "g.lime", dims, b"", "D", file_dims=False)
    m = lime.find_scidac_fields(p)[0]
    assert m.dims == dims and m.file_xml_index is None


def test_partial_ildg_dims_and_no_dims(tmp_path):
    xml = b"<ildgFormat><lx>4</lx><ly>6</ly><lt>8</lt></ildgFormat>"
    body = [lf.record("ildg-format", xml),
            lf.record("scidac-private-record-xml", lf.private_record_xml("QDP_D3_ColorVector", "D", 3, 0, 48, 1)),
            lf.record("scidac-binary-data", b"")]
    p = tmp_path / "p.lime"
    p.write_bytes(b"".join(body))
    assert lime.find_scidac_fields(p)[0].dims == (4, 6)  # the run stops at the first missing tag
    p.write_bytes(b"".join(body[1:]))
    assert lime.find_scidac_fields(p)[0].dims == ()


def test_metadata_defaults_and_unparsable_values(tmp_path):
    xml = (b"<scidacRecord><datatype>Lattice</datatype><colors>x</colors><datacount>0</datacount>"
           b"<recordtype>one</recordtype></scidacRecord>")
    p = tmp_path / "m.lime"
    p.write_bytes(lf.record("scidac-private-file-xml", lf.private_file_xml((2, 2)))
                  + lf.record("scidac-private-record-xml", xml) + lf.record("scidac-binary-data", b""))
    m = lime.find_scidac_fields(p)[0]
    assert m.precision == "F"            # default precision
    assert m.colors == 0 and m.spins == 0 and m.typesize == 0
    assert m.datacount == 1              # 0 is taken as 1
    assert m.recordtype is None          # unparsable
    assert m.record_xml_index is None


def test_descriptor_without_binary_is_skipped_and_binary_search_runs_on(tmp_path):
    a = lf.private_record_xml("A_ColorVector", "F", 3, 0, 24, 1)
    b = lf.private_record_xml("B_ColorVector", "F", 3, 0, 24, 1)
    p = tmp_path / "s.lime"
    p.write_bytes(lf.record("scidac-private-record-xml", a) + lf.record("scidac-private-record-xml", b)
                  + lf.record("scidac-binary-data", b"\x00" * 24))
    fields = lime.find_scidac_fields(p)
    # the first descriptor takes the first binary record after it, as does the second
    assert [f.datatype for f in fields] == ["A_ColorVector", "B_ColorVector"]
    assert fields[0].binary_record_index == fields[1].binary_record_index == 2
    p.write_bytes(lf.record("scidac-private-record-xml", a))
    assert lime.find_scidac_fields(p) == []


def test_pseudofermion_file_has_metadata_field_first(tmp_path):
    dims = (2, 2, 2, 4)
    p = lf.pseudofermion_file(tmp_path / "pf.lime", dims, b"\x00" * (32 * 96), "F")
    fields = lime.find_scidac_fields(p)
    assert [(f.datatype, f.precision, f.recordtype) for f in fields] == [("char", "U", 1), ("Lattice", "F", 0)]
    assert fields[1].typesize == 96 and fields[1].spins == 4 and fields[1].colors == 3


# ------------------------------------------------ validation before the device
@pytest.mark.parametrize("kw,msg", [
    (dict(precision="U"), "Unsupported SciDAC precision"),
    (dict(typesize=70), "Incompatible typesize"),
    (dict(typesize=12, precision="F"), "Expected even re/im"),
    (dict(cut=8), "Binary length mismatch"),
    (dict(cut=3), "multiple of element size"),
    (dict(colors=2), "ColorMatrix payload mismatch"),
    (dict(colors=0), "Invalid colors"),
])
def test_decode_validation_errors(tmp_path, kw, msg):
    dims = (2, 2, 2, 2)
    prec = kw.get("precision", "D")
    wb = 4 if prec == "F" else 8
    typesize = kw.get("typesize", 2 * wb * 9)
    colors = kw.get("colors", 3)
    payload = b"\x00" * (16 * 4 * 144 - kw.get("cut", 0))
    p = tmp_path / "v.lime"
    p.write_bytes(lf.record("scidac-private-file-xml", lf.private_file_xml(dims))
                  + lf.record("scidac-private-record-xml",
                              lf.private_record_xml("QDP_D3_ColorMatrix", prec, colors, 4, typesize, 4))
                  + lf.record("scidac-binary-data", payload))
    with pytest.raises(ValueError, match=msg):
        lime.decode_scidac_field(p)
    with pytest.raises(ValueError, match=msg):
        lime.decode_scidac_gauge(p)


def test_decode_needs_a_colormatrix_for_gauge_and_a_leaf_for_pseudofermion(tmp_path):
    dims = (2, 2, 2, 2)
    p = lf.field_file(tmp_path / "f.lime", dims, b"\x00" * (16 * 96), "QDP_F3_DiracFermion", "F", 3, 4, 12, 1)
    with pytest.raises(ValueError, match="Expected ColorMatrix"):
        lime.decode_scidac_gauge(p)
    q = lf.field_file(tmp_path / "c.lime", dims, b"\x00" * 16, "char", "U", 0, 0, 1, 16, recordtype=1)
    with pytest.raises(ValueError, match="auto-locate"):
        lime.decode_scidac_pseudofermion(q)
    with pytest.raises(IndexError):
        lime.decode_scidac_field(q, field_index=1)


def test_no_dims_volume_inference_errors(tmp_path):
    p = tmp_path / "n.lime"
    p.write_bytes(lf.record("scidac-private-record-xml", lf.private_record_xml("X", "F", 0, 0, 0, 1))
                  + lf.record("scidac-binary-data", b"\x00" * 8))
    with pytest.raises(ValueError, match="Missing dimensions"):
        lime.decode_scidac_field(p)
    p.write_bytes(lf.record("scidac-private-record-xml", lf.private_record_xml("X", "F", 0, 0, 24, 1))
                  + lf.record("scidac-binary-data", b"\x00" * 28))
    with pytest.raises(ValueError, match="Cannot infer volume"):
        lime.decode_scidac_field(p)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device error")
def test_valid_decode_fails_loudly_without_a_device(tmp_path):
    dims = (2, 2, 2, 2)
    p = lf.gauge_file(tmp_path / "g.lime", dims, b"\x00" * (16 * 4 * 144), "D")
    with pytest.raises(RuntimeError, match="HIP device"):
        lime.decode_scidac_gauge(p)


# --------------------------------------------------- C ABI argument checks
def _decode_rc(nbytes=16 * 4 * 9 * 16, wb=8, dims=(2, 2, 2, 2), dc=4, ncx=9, order=nat.WL_LIME_COUNT,
               dtype=nat.WL_C128, payload=4096, out=4096):
    lib = nat.load()
    cd = (ctypes.c_int * len(dims))(*dims)
    return lib.wl_lime_decode(ctypes.c_void_p(payload), nbytes, wb, cd, len(dims), dc, ncx, order,
                              ctypes.c_void_p(out), dtype, None)


@pytest.mark.parametrize("kw,code,msg", [
    (dict(payload=0), nat.WL_ERR_BAD_ARG, "null"),
    (dict(wb=2), nat.WL_ERR_BAD_ARG, "word_bytes"),
    (dict(dtype=7), nat.WL_ERR_BAD_ARG, "dtype"),
    (dict(order=3), nat.WL_ERR_BAD_ARG, "order"),
    (dict(dims=(2, 2, 2, 2, 2)), nat.WL_ERR_UNSUPPORTED, "dimensions"),
    (dict(dims=(2, 0, 2, 2)), nat.WL_ERR_BAD_ARG, "dims"),
    (dict(dc=0), nat.WL_ERR_BAD_ARG, "datacount"),
    (dict(nbytes=100), nat.WL_ERR_BAD_ARG, "payload holds 100 bytes"),
    (dict(payload=4104), nat.WL_ERR_BAD_ARG, "payload not aligned"),
    (dict(out=4100), nat.WL_ERR_BAD_ARG, "output not aligned"),
    (dict(ncx=40000, nbytes=16 * 4 * 40000 * 16), nat.WL_ERR_UNSUPPORTED, "does not fit in LDS"),
])
def test_wl_lime_decode_argument_checks(kw, code, msg):
    rc = _decode_rc(**kw)
    assert rc == code
    assert msg in nat.load().wl_last_error().decode()


# ------------------------------------------------------- oracle pinning
CASES = [
    # datatype, precision, colors, spins, ncomplex, datacount, dims
    ("USQCD_F3_ColorMatrix", "F", 3, 4, 9, 4, (4, 3, 2, 5)),
    ("USQCD_D3_ColorMatrix", "D", 3, 4, 9, 4, (3, 2, 5, 2)),
    ("QDP_D3_DiracFermion", "D", 3, 4, 12, 1, (2, 3, 4, 5)),
    ("Lattice", "F", 3, 4, 12, 1, (5, 1, 3, 2)),
    ("QDP_F3_ColorVector", "F", 3, 0, 3, 2, (3, 4, 5)),
    ("QDP_D1_ColorMatrix", "D", 1, 0, 1, 2, (6, 5)),
    ("Generic", "D", 0, 0, 5, 3, (7,)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] + "_" + c[1] + "_" + "x".join(map(str, c[6])) for c in CASES])
def test_oracle_decodes_closed_form_fields(case):
    datatype, prec, colors, spins, ncx, dc, dims = case
    raw = lf.payload_by_loops(lf.coord_value, dims, dc, ncx, prec)
    wb = 4 if prec == "F" else 8
    dec = lo.decode_payload(raw, prec, 2 * wb * ncx, dc, dims, datatype, colors, spins)
    want = lf.coord_lattice(dims, dc, ncx)
    lat = dec["lattice_data"]
    assert lat.dtype == (np.complex64 if prec == "F" else np.complex128)
    assert lat.shape[:len(dims) + 1] == dims + (dc,)
    np.testing.assert_array_equal(lat.reshape(dims + (dc, ncx)), want)
    # site_data is lexicographic (x0 fastest)
    site = dec["site_data"].reshape(-1, dc, ncx)
    for s, x in enumerate(lf.lexicographic_sites(dims)):
        np.testing.assert_array_equal(site[s], want[x])


def test_oracle_gauge_and_pseudofermion_axes():
    dims = (3, 2, 4, 2)
    raw = lf.payload_by_loops(lf.coord_value, dims, 4, 9, "D")
    lat = lo.decode_payload(raw, "D", 144, 4, dims, "QDP_D3_ColorMatrix", 3, 4)["lattice_data"]
    U = lo.gauge_from_lattice(lat, 4, batch_size=2)
    assert U.shape == (2, 4, *dims, 3, 3)
    for mu in range(4):
        for x in [(0, 0, 0, 0), (2, 1, 3, 1), (1, 0, 2, 1)]:
            for i in range(3):
                for j in range(3):
                    assert U[1, mu][x][i, j] == lf.coord_value(x, mu, 3 * i + j)
    raw = lf.payload_by_loops(lf.coord_value, dims, 1, 12, "F")
    lat = lo.decode_payload(raw, "F", 96, 1, dims, "Lattice", 3, 4)["lattice_data"]
    pf = lo.pseudofermion_from_lattice(lat, 1)
    assert pf.shape == (1, *dims, 4, 3)
    assert pf[0, 2, 1, 3, 1, 2, 1] == lf.coord_value((2, 1, 3, 1), 0, 7)


def test_numpy_writer_matches_loop_writer():
    for dims, dc, ncx, prec in [((3, 2, 4, 5), 4, 9, "D"), ((2, 5, 3), 1, 12, "F"), ((6, 4), 2, 1, "D")]:
        a = lf.coord_lattice(dims, dc, ncx)
        assert lf.payload_from_lattice(a, len(dims), prec) == lf.payload_by_loops(lf.coord_value, dims, dc, ncx, prec)


def test_oracle_momentum_diagnostics():
    rng = np.random.default_rng(3)
    h = rng.normal(size=(1, 4, 2, 2, 2, 2, 3, 3)) + 1j * rng.normal(size=(1, 4, 2, 2, 2, 2, 3, 3))
    a = 0.5 * (h - np.swapaxes(h.conj(), -1, -2))
    a = a - np.eye(3) * (np.trace(a, axis1=-2, axis2=-1)[..., None, None] / 3)
    d = lo.momentum_diagnostics(a)
    assert d["rel_antihermitian_error"] < 1e-15 and d["max_abs_trace"] < 1e-14
    # add a known Hermitian part c*1: ||m + m^dag|| = 2|c| sqrt(3 n), tr = tr a + 3c
    c = 0.25
    m = a + c * np.eye(3)
    d = lo.momentum_diagnostics(m)
    n = a.size // 9
    np.testing.assert_allclose(d["rel_antihermitian_error"], 2 * c * np.sqrt(3 * n) / np.linalg.norm(m), rtol=1e-12)
    np.testing.assert_allclose([d["mean_abs_trace"], d["max_abs_trace"]], [3 * c, 3 * c], rtol=1e-12)
