"""Writers of small SciDAC / ILDG LIME files for the LIME tests.

Byte layout of the published LIME format: every record is a 144-byte
big-endian header (u32 magic 0x456789ab, u16 version 1, u16 flags with
message-begin 0x8000 and message-end 0x4000, u64 payload length, 128-byte
NUL-padded type name) followed by the payload, zero-padded to a multiple of 8
bytes.  A SciDAC lattice field is stored site by site in lexicographic order
(x0 fastest), each site holding ``datacount`` items of big-endian (re, im)
reals.

``payload_by_loops`` writes a field value by value with ``struct.pack`` from
a closed form of the coordinates; ``payload_from_lattice`` writes a NumPy
array of shape (*dims, datacount, ncomplex) and is itself checked against the
loop writer."""

from __future__ import annotations

import itertools
import struct
from pathlib import Path
from typing import Callable, Iterable, List, Optional, Sequence, Tuple

import numpy as np

MAGIC = 0x456789AB
MB, ME = 0x8000, 0x4000


def record(type_name: str, payload: bytes, flags: int = 0, version: int = 1) -> bytes:
    name = type_name.encode("ascii")
    assert len(name) < 128
    head = struct.pack(">IHHQ", MAGIC, version, flags, len(payload)) + name + b"\x00" * (128 - len(name))
    assert len(head) == 144
    return head + payload + b"\x00" * ((-len(payload)) % 8)


def coord_value(x: Sequence[int], d: int, k: int) -> complex:
    """Closed-form test value: unique per (site, item, component) for extents
    up to 16, fewer than 16 items and 64 components; exact in float32."""
    x = tuple(x) + (0,) * (4 - len(x))
    re = x[0] + 16 * x[1] + 256 * x[2] + 4096 * x[3] + 65536 * d
    im = k + 64 * d + 0.5
    return complex(re, im)


def coord_lattice(dims: Sequence[int], datacount: int, ncomplex: int) -> np.ndarray:
    """coord_value on the whole lattice as an array (*dims, datacount, ncomplex)."""
    dims = tuple(dims)
    out = np.zeros(dims + (datacount, ncomplex), dtype=np.complex128)
    for x in itertools.product(*[range(n) for n in dims]):
        for d in range(datacount):
            for k in range(ncomplex):
                out[x + (d, k)] = coord_value(x, d, k)
    return out


def lexicographic_sites(dims: Sequence[int]) -> Iterable[Tuple[int, ...]]:
    """Site coordinates in lexicographic order, x0 fastest."""
    for rev in itertools.product(*[range(n) for n in reversed(tuple(dims))]):
        yield tuple(reversed(rev))


def payload_by_loops(value: Callable[[Tuple[int, ...], int, int], complex], dims: Sequence[int], datacount: int,
                     ncomplex: int, precision: str) -> bytes:
    fmt = ">ff" if precision == "F" else ">dd"
    parts: List[bytes] = []
    for x in lexicographic_sites(dims):
        for d in range(datacount):
            for k in range(ncomplex):
                z = value(x, d, k)
                parts.append(struct.pack(fmt, z.real, z.imag))
    return b"".join(parts)


def payload_from_lattice(a: np.ndarray, nd: int, precision: str) -> bytes:
    """a: (*dims, datacount, ncomplex) complex -> lexicographic big-endian bytes."""
    rev = np.transpose(a, tuple(range(nd - 1, -1, -1)) + tuple(range(nd, a.ndim)))
    re_im = np.ascontiguousarray(rev).astype(np.complex128).view(np.float64)
    return re_im.astype(">f4" if precision == "F" else ">f8").tobytes()


def private_file_xml(dims: Sequence[int]) -> bytes:
    d = " ".join(str(v) for v in dims)
    return (f'<?xml version="1.0" encoding="UTF-8"?><scidacFile><version>1.1</version>'
            f'<spacetime>{len(dims)}</spacetime><dims>{d} </dims><volfmt>0</volfmt></scidacFile>').encode()


def private_record_xml(datatype: str, precision: str, colors: int, spins: int, typesize: int, datacount: int,
                       recordtype: Optional[int] = 0) -> bytes:
    rt = "" if recordtype is None else f"<recordtype>{recordtype}</recordtype>"
    return (f'<?xml version="1.0" encoding="UTF-8"?><scidacRecord><version>1.1</version>'
            f'<date>Thu Jan  1 00:00:00 1970 UTC</date>{rt}<datatype>{datatype}</datatype>'
            f'<precision>{precision}</precision><colors>{colors}</colors><spins>{spins}</spins>'
            f'<typesize>{typesize}</typesize><datacount>{datacount}</datacount></scidacRecord>').encode()


def ildg_format_xml(dims: Sequence[int], precision: str) -> bytes:
    tags = "".join(f"<{t}>{v}</{t}>" for t, v in zip(("lx", "ly", "lz", "lt"), dims))
    bits = 32 if precision == "F" else 64
    return (f'<?xml version="1.0" encoding="UTF-8"?><ildgFormat><version>1.0</version><field>su3gauge</field>'
            f'<precision>{bits}</precision>{tags}</ildgFormat>').encode()


def gauge_file(path: Path, dims: Sequence[int], payload: bytes, precision: str, nc: int = 3, datacount: int = 4,
               ildg: bool = True, file_dims: bool = True) -> Path:
    """A gauge configuration laid out as Chroma / QIO write one."""
    wb = 4 if precision == "F" else 8
    recs = []
    if file_dims:
        recs.append(record("scidac-private-file-xml", private_file_xml(dims), MB))
        recs.append(record("scidac-file-xml", b"<info>test configuration</info>", ME))
    recs.append(record("scidac-private-record-xml",
                       private_record_xml(f"USQCD_{precision}{nc}_ColorMatrix", precision, nc, 4, 2 * wb * nc * nc,
                                          datacount), MB))
    recs.append(record("scidac-record-xml", b"<info>links</info>"))
    if ildg:
        recs.append(record("ildg-format", ildg_format_xml(dims, precision)))
    recs.append(record("ildg-binary-data" if ildg else "scidac-binary-data", payload))
    recs.append(record("scidac-checksum", b"<scidacChecksum><version>1.0</version></scidacChecksum>", ME))
    Path(path).write_bytes(b"".join(recs))
    return Path(path)


def pseudofermion_file(path: Path, dims: Sequence[int], payload: bytes, precision: str, ns: int = 4,
                       nc: int = 3) -> Path:
    """A Chroma pseudofermion file: character-valued global metadata fields
    (precision U, recordtype 1) before the floating-point leaf."""
    wb = 4 if precision == "F" else 8
    meta_payload = b"<pf>monomial 0</pf>\x00\x00\x00"
    recs = [
        record("scidac-private-file-xml", private_file_xml(dims), MB),
        record("scidac-file-xml", b"<info>pseudofermions</info>", ME),
        record("scidac-private-record-xml", private_record_xml("char", "U", 0, 0, 1, len(meta_payload), 1), MB),
        record("scidac-record-xml", b"<info>global metadata</info>"),
        record("scidac-binary-data", meta_payload),
        record("scidac-checksum", b"<scidacChecksum/>", ME),
        record("scidac-private-record-xml",
               private_record_xml("Lattice", precision, nc, ns, 2 * wb * ns * nc, 1, 0), MB),
        record("scidac-record-xml", b"<info>leaf</info>"),
        record("scidac-binary-data", payload),
        record("scidac-checksum", b"<scidacChecksum/>", ME),
    ]
    Path(path).write_bytes(b"".join(recs))
    return Path(path)


def field_file(path: Path, dims: Sequence[int], payload: bytes, datatype: str, precision: str, colors: int,
               spins: int, ncomplex: int, datacount: int, recordtype: Optional[int] = 0) -> Path:
    """One generic SciDAC field."""
    wb = 4 if precision == "F" else 8
    recs = [
        record("scidac-private-file-xml", private_file_xml(dims), MB),
        record("scidac-private-record-xml",
               private_record_xml(datatype, precision, colors, spins, 2 * wb * ncomplex, datacount, recordtype), MB),
        record("scidac-binary-data", payload, ME),
    ]
    Path(path).write_bytes(b"".join(recs))
    return Path(path)
