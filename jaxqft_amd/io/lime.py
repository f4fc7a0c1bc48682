"""SciDAC / ILDG LIME files (Chroma, QIO) with the payload decoded on the GPU.

Mirrors jaxqft/io/lime.py: the same record and field descriptors
(``LimeRecord`` :18-39, ``ScidacFieldMeta`` :42-58), the same functions,
arguments, return values and errors (``iter_lime_records`` :69-102,
``list_lime_records`` :105-106, ``read_lime_record_data`` :109-123,
``find_scidac_fields`` :163-224, ``decode_scidac_field`` :247-331,
``decode_scidac_gauge`` :334-352, ``decode_scidac_momentum`` :355-381,
``decode_scidac_pseudofermion`` :384-420).

A LIME file is a sequence of records, each a 144-byte big-endian header
(magic 0x456789ab, u16 version, u16 flags with message-begin bit 15 and
message-end bit 14, u64 payload length, 128-byte NUL-padded type name)
followed by the payload padded to a multiple of 8 bytes.  A SciDAC field is a
``scidac-private-record-xml`` record (datatype, precision F/D, colors, spins,
typesize, datacount, recordtype) followed by its binary record
(``scidac-binary-data`` or ``ildg-binary-data``); the lattice extents come
from the ``<dims>`` of the ``scidac-private-file-xml`` record or, failing that,
``<lx><ly><lz><lt>`` of ``ildg-format``.  The binary payload holds, for every
site in lexicographic order (x0 fastest), ``datacount`` items of
``typesize`` bytes of big-endian (re, im) reals.

Headers and XML are parsed on the host (they are a few hundred bytes).  The
payload is read into pinned host memory, copied to the device, and decoded by
``wl_lime_decode`` (byte swap, widening or narrowing, and the reorder from
lexicographic sites to the lattice axes or to direction-major links) in one
pass over HBM.  Every ``decode_*`` function returns NumPy arrays like the
reference; with ``device="cuda"`` it returns torch tensors left on the device
instead, ready for ``resident_gauge`` or a solve.
"""

from __future__ import annotations

import ctypes
import re
import struct
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, Iterator, List, Optional, Tuple, Union

import numpy as np
import torch

from .. import _native as nat

LIME_MAGIC = 0x456789AB
LIME_HEADER_BYTES = 144
_HEADER = struct.Struct(">IHHQ128s")
_MB_FLAG = 0x8000
_ME_FLAG = 0x4000
_BINARY_TYPES = ("ildg-binary-data", "scidac-binary-data")

PathLike = Union[str, Path]


@dataclass(frozen=True)
class LimeRecord:
    """One LIME record header; ``offset`` is the header's byte offset."""

    index: int
    offset: int
    version: int
    flags: int
    size: int
    type_name: str

    @property
    def data_offset(self) -> int:
        return self.offset + LIME_HEADER_BYTES

    @property
    def is_begin_message(self) -> bool:
        return bool(self.flags & _MB_FLAG)

    @property
    def is_end_message(self) -> bool:
        return bool(self.flags & _ME_FLAG)


@dataclass(frozen=True)
class ScidacFieldMeta:
    """A SciDAC field: its private-record XML and the binary record it names."""

    datatype: str
    precision: str
    colors: int
    spins: int
    typesize: int
    datacount: int
    dims: Tuple[int, ...]
    recordtype: Optional[int]
    private_record_index: int
    binary_record_index: int
    binary_record_type: str
    record_xml_index: Optional[int]
    file_xml_index: Optional[int]


# ------------------------------------------------------------------ records
def iter_lime_records(path: PathLike) -> Iterator[LimeRecord]:
    """Yield the record headers of a LIME file in file order (lime.py:69-102).

    Raises ValueError on a header cut short by the end of the file or on a
    wrong magic number."""
    p = Path(path)
    with p.open("rb") as f:
        index = 0
        while True:
            offset = f.tell()
            head = f.read(LIME_HEADER_BYTES)
            if not head:
                return
            if len(head) < LIME_HEADER_BYTES:
                raise ValueError(f"Truncated LIME header at offset {offset} in {p}")
            magic, version, flags, size, raw_type = _HEADER.unpack(head)
            if magic != LIME_MAGIC:
                raise ValueError(f"Bad LIME magic {magic:#x} at offset {offset} in {p} (expected {LIME_MAGIC:#x})")
            type_name = raw_type.split(b"\x00", 1)[0].decode("ascii", "replace")
            yield LimeRecord(index, offset, int(version), int(flags), int(size), type_name)
            # payload, then padding to the next multiple of 8 bytes
            f.seek(offset + LIME_HEADER_BYTES + size + ((-size) % 8))
            index += 1


def list_lime_records(path: PathLike) -> List[LimeRecord]:
    return list(iter_lime_records(path))


def _record(path: PathLike, rec: Union[LimeRecord, int]) -> LimeRecord:
    if isinstance(rec, LimeRecord):
        return rec
    heads = list_lime_records(path)
    if rec < 0 or rec >= len(heads):
        raise IndexError(f"Record index {rec} out of range [0,{len(heads)})")
    return heads[rec]


def read_lime_record_data(path: PathLike, rec: Union[LimeRecord, int]) -> bytes:
    """Payload bytes of a record given by header or index (lime.py:109-123)."""
    r = _record(path, rec)
    with Path(path).open("rb") as f:
        f.seek(r.data_offset)
        return f.read(r.size)


# ---------------------------------------------------------------------- XML
def _tag(xml: Optional[str], tag: str) -> Optional[str]:
    """Stripped text of the first <tag ...>...</tag> element, or None."""
    if not xml:
        return None
    m = re.search(r"<" + re.escape(tag) + r"(?:\s[^>]*)?>(.*?)</" + re.escape(tag) + r">", xml, re.S)
    return None if m is None else m.group(1).strip()


def _int_or(text: Optional[str], default):
    if text is None:
        return default
    try:
        return int(text)
    except (TypeError, ValueError):
        return default


def _lattice_dims(file_xml: Optional[str], ildg_xml: Optional[str]) -> Tuple[int, ...]:
    """<dims> of the SciDAC file XML, else the leading run of <lx><ly><lz><lt>
    of the ILDG format record (lime.py:141-160)."""
    d = _tag(file_xml, "dims")
    if d:
        vals = tuple(int(t) for t in d.split())
        if vals:
            return vals
    out: List[int] = []
    for name in ("lx", "ly", "lz", "lt"):
        t = _tag(ildg_xml, name)
        if t is None:
            break
        out.append(int(t))
    return tuple(out)


def find_scidac_fields(path: PathLike) -> List[ScidacFieldMeta]:
    """Every SciDAC field of the file with the binary record it points to
    (lime.py:163-224): for each ``scidac-private-record-xml`` record, the
    first later binary record; a descriptor with none is skipped."""
    heads = list_lime_records(path)
    xml: Dict[int, str] = {}
    with Path(path).open("rb") as f:
        for r in heads:
            if "xml" in r.type_name or r.type_name == "ildg-format":
                f.seek(r.data_offset)
                xml[r.index] = f.read(r.size).decode("utf-8", "replace")
    file_xml_index = next((r.index for r in heads if r.type_name == "scidac-private-file-xml"), None)
    ildg_index = next((r.index for r in heads if r.type_name == "ildg-format"), None)
    dims = _lattice_dims(xml.get(file_xml_index), xml.get(ildg_index))

    fields: List[ScidacFieldMeta] = []
    for pos, r in enumerate(heads):
        if r.type_name != "scidac-private-record-xml":
            continue
        binary = next((h for h in heads[pos + 1:] if h.type_name in _BINARY_TYPES), None)
        if binary is None:
            continue
        text = xml.get(r.index, "")
        nxt = heads[pos + 1] if pos + 1 < len(heads) else None
        fields.append(ScidacFieldMeta(
            datatype=_tag(text, "datatype") or "",
            precision=_tag(text, "precision") or "F",
            colors=int(_int_or(_tag(text, "colors"), 0) or 0),
            spins=int(_int_or(_tag(text, "spins"), 0) or 0),
            typesize=int(_int_or(_tag(text, "typesize"), 0) or 0),
            datacount=int(_int_or(_tag(text, "datacount"), 1) or 1),
            dims=dims,
            recordtype=_int_or(_tag(text, "recordtype"), None),
            private_record_index=r.index,
            binary_record_index=binary.index,
            binary_record_type=binary.type_name,
            record_xml_index=nxt.index if nxt is not None and nxt.type_name == "scidac-record-xml" else None,
            file_xml_index=file_xml_index,
        ))
    return fields


# ------------------------------------------------------------------ decoding
_ORDERS = {"lexi": nat.WL_LIME_LEXI, "site": nat.WL_LIME_SITE, "count": nat.WL_LIME_COUNT}


def _word_bytes(precision: str) -> int:
    """Bytes per real of a SciDAC precision tag (lime.py:227-233)."""
    p = precision.upper().strip()
    if p == "F":
        return 4
    if p == "D":
        return 8
    raise ValueError(f"Unsupported SciDAC precision {precision!r}")


@dataclass(frozen=True)
class _Plan:
    meta: ScidacFieldMeta
    record: LimeRecord
    word_bytes: int
    volume: int
    ncomplex: int           # complex values per item
    item_shape: Tuple[int, ...]


def _plan(path: PathLike, meta: ScidacFieldMeta) -> _Plan:
    """Validate a field's metadata against its binary record and fix the shape
    of one item (lime.py:259-320); every check runs before any device work."""
    rec = _record(path, meta.binary_record_index)
    wb = _word_bytes(meta.precision)
    if meta.typesize % wb != 0:
        raise ValueError(f"Incompatible typesize={meta.typesize} for precision={meta.precision} ({wb} bytes)")
    nreal = meta.typesize // wb
    if nreal % 2 != 0:
        raise ValueError(f"Expected even re/im scalar count per site item, got {nreal}")
    if rec.size % wb != 0:
        raise ValueError(f"buffer size must be a multiple of element size ({rec.size} bytes, {wb}-byte words)")
    nwords = rec.size // wb
    if meta.dims:
        volume = int(np.prod(meta.dims))
    else:
        per_site = meta.datacount * nreal
        if per_site == 0:
            raise ValueError("Missing dimensions and invalid record shape metadata")
        if nwords % per_site != 0:
            raise ValueError("Cannot infer volume from payload size and metadata")
        volume = nwords // per_site
    expected = volume * meta.datacount * nreal
    if nwords != expected:
        raise ValueError(f"Binary length mismatch: got {nwords} float words, expected {expected} "
                         f"(V={volume}, datacount={meta.datacount}, n_real_per_item={nreal})")
    ncx = nreal // 2
    dt = meta.datatype
    if "ColorMatrix" in dt:
        n = meta.colors
        if n <= 0:
            raise ValueError(f"Invalid colors={meta.colors} in {dt}")
        if ncx != n * n:
            raise ValueError(f"ColorMatrix payload mismatch: got {ncx} complex values, expected {n * n}")
        shape: Tuple[int, ...] = (n, n)
    elif "DiracFermion" in dt:
        ns, nc = meta.spins, meta.colors
        if ns <= 0 or nc <= 0:
            raise ValueError(f"Invalid spins/colors ({meta.spins},{meta.colors}) in {dt}")
        if ncx != ns * nc:
            raise ValueError(f"DiracFermion payload mismatch: got {ncx} complex values, expected {ns * nc}")
        shape = (ns, nc)
    elif "ColorVector" in dt:
        nc = meta.colors
        if nc <= 0:
            raise ValueError(f"Invalid colors={meta.colors} in {dt}")
        if ncx != nc:
            raise ValueError(f"ColorVector payload mismatch: got {ncx} complex values, expected {nc}")
        shape = (nc,)
    elif meta.spins > 0 and meta.colors > 0 and ncx == meta.spins * meta.colors:
        # generic "Lattice" records of QIO with explicit spin/colour counts
        shape = (meta.spins, meta.colors)
    else:
        shape = (ncx,)
    return _Plan(meta, rec, wb, volume, ncx, shape)


def _payload_on_device(path: PathLike, rec: LimeRecord) -> Tuple[torch.Tensor, torch.Tensor]:
    """The record's payload read into pinned host memory and queued for copy
    to HBM; returns (device payload, pinned staging buffer)."""
    n = rec.size
    host = torch.empty(n, dtype=torch.uint8, pin_memory=True)
    with Path(path).open("rb") as f:
        f.seek(rec.data_offset)
        got = f.readinto(memoryview(host.numpy()))
    if got != n:
        raise ValueError(f"record {rec.index} is cut short: {got} of {n} payload bytes")
    dev = torch.empty(n, dtype=torch.uint8, device="cuda")
    dev.copy_(host, non_blocking=True)
    return dev, host


def decode_payload(payload: torch.Tensor, word_bytes: int, dims: Tuple[int, ...], datacount: int, ncomplex: int,
                   order: str, dtype: torch.dtype) -> torch.Tensor:
    """Decode a device-resident SciDAC payload with ``wl_lime_decode``.

    ``order``: "lexi" -> (V, datacount, ncomplex); "site" -> (*dims, datacount,
    ncomplex); "count" -> (datacount, *dims, ncomplex).  ``dtype``:
    torch.complex128 or torch.complex64."""
    nat.require_gpu()
    lib = nat.load()
    if dtype not in (torch.complex128, torch.complex64):
        raise ValueError("dtype must be torch.complex128 or torch.complex64")
    if payload.dtype != torch.uint8 or not payload.is_cuda or not payload.is_contiguous():
        raise ValueError("payload must be a contiguous uint8 CUDA tensor")
    dims = tuple(int(v) for v in dims)
    V = int(np.prod(dims))
    if order == "lexi":
        shape = (V, datacount, ncomplex)
    elif order == "site":
        shape = (*dims, datacount, ncomplex)
    elif order == "count":
        shape = (datacount, *dims, ncomplex)
    else:
        raise ValueError(f"order must be one of {sorted(_ORDERS)}")
    out = torch.empty(shape, dtype=dtype, device=payload.device)
    cdims = (ctypes.c_int * len(dims))(*dims)
    nat.check(lib.wl_lime_decode(nat.c_void_p(payload.data_ptr()), payload.numel(), int(word_bytes), cdims,
                                 len(dims), int(datacount), int(ncomplex), _ORDERS[order],
                                 nat.c_void_p(out.data_ptr()),
                                 nat.WL_C128 if dtype == torch.complex128 else nat.WL_C64,
                                 nat.c_void_p(nat.current_stream())))
    return out


def _decode(path: PathLike, plan: _Plan, order: str) -> torch.Tensor:
    """The field of a plan decoded on the device; element type follows the
    file precision (complex64 for F, complex128 for D) as the reference's
    float32 / float64 decode does."""
    dt = torch.complex64 if plan.word_bytes == 4 else torch.complex128
    dims = plan.meta.dims if plan.meta.dims else (plan.volume,)
    if plan.ncomplex == 0:  # an empty item: nothing to decode
        shape = {"lexi": (plan.volume, plan.meta.datacount, 0), "site": (*dims, plan.meta.datacount, 0),
                 "count": (plan.meta.datacount, *dims, 0)}[order]
        return torch.empty(shape, dtype=dt, device="cuda")
    nat.require_gpu()
    payload, staging = _payload_on_device(path, plan.record)
    out = decode_payload(payload, plan.word_bytes, dims, plan.meta.datacount, plan.ncomplex, order, dt)
    torch.cuda.current_stream().synchronize()  # the copy out of the staging buffer has run
    del staging
    return out


def _finish(t: torch.Tensor, device):
    return t if device is not None else t.cpu().numpy()


def _field_meta(path: PathLike, field_index: int) -> ScidacFieldMeta:
    fields = find_scidac_fields(path)
    if field_index < 0 or field_index >= len(fields):
        raise IndexError(f"field_index={field_index} out of range [0,{len(fields)})")
    return fields[field_index]


def decode_scidac_field(path: PathLike, field_index: int = 0, *, device=None) -> dict:
    """Decode one SciDAC field (lime.py:247-331).

    Returns {"meta": ScidacFieldMeta, "site_data": (V, datacount, *item) in
    lexicographic site order, "lattice_data": (*dims, datacount, *item)}; the
    item shape is (Nc, Nc) for ColorMatrix, (Ns, Nc) for DiracFermion and for
    Lattice records with spins and colours, (Nc,) for ColorVector.  As in the
    reference, lattice_data is a view of site_data with the lattice axes
    reversed (no dims: the site data itself)."""
    meta = _field_meta(path, field_index)
    plan = _plan(path, meta)
    site = _finish(_decode(path, plan, "lexi").reshape(plan.volume, meta.datacount, *plan.item_shape), device)
    lat = site
    if meta.dims:
        nd = len(meta.dims)
        axes = (*range(nd - 1, -1, -1), *range(nd, nd + site.ndim - 1))
        lat = site.reshape(*reversed(meta.dims), *site.shape[1:])
        lat = lat.permute(*axes) if isinstance(lat, torch.Tensor) else np.transpose(lat, axes)
    return {"meta": meta, "site_data": site, "lattice_data": lat}


def decode_scidac_gauge(path: PathLike, field_index: int = 0, batch_size: int = 1, *, device=None):
    """A gauge field in the BM...IJ layout (batch, mu, *dims, Nc, Nc)
    (lime.py:334-352), decoded on the device straight into direction-major
    order."""
    meta = _field_meta(path, field_index)
    plan = _plan(path, meta)
    if "ColorMatrix" not in meta.datatype:
        raise ValueError(f"Expected ColorMatrix field, got {meta.datatype}")
    if meta.dims:
        U = _decode(path, plan, "count").reshape(meta.datacount, *meta.dims, *plan.item_shape)
    else:
        # no lattice extents: the reference moves nothing and keeps the site axis first
        U = _decode(path, plan, "lexi").reshape(plan.volume, meta.datacount, *plan.item_shape)
    U = U.unsqueeze(0)
    if int(batch_size) != 1:
        U = U.repeat(int(batch_size), *([1] * (U.dim() - 1)))
    return _finish(U, device)


def decode_scidac_momentum(path: PathLike, field_index: int = 0, batch_size: int = 1, *,
                           check_traceless_antihermitian: bool = True, device=None):
    """A momentum field stored like a gauge field (ColorMatrix, datacount 4), in
    BM...IJ, with the anti-Hermiticity and trace diagnostics of
    lime.py:355-381 (relative ||m + m^dag|| / ||m||, mean and max |tr m|)."""
    mom = decode_scidac_gauge(path, field_index=field_index, batch_size=batch_size, device="cuda")
    info = {}
    if check_traceless_antihermitian:
        m = mom.to(torch.complex128)
        denom = float(torch.linalg.vector_norm(m))
        herm = float(torch.linalg.vector_norm(m + m.conj().transpose(-1, -2)))
        tr = torch.diagonal(m, dim1=-2, dim2=-1).sum(-1).abs()
        info = {
            "rel_antihermitian_error": herm / denom if denom > 0 else 0.0,
            "mean_abs_trace": float(tr.mean()),
            "max_abs_trace": float(tr.max()),
        }
    return _finish(mom, device), info


def decode_scidac_pseudofermion(path: PathLike, field_index: Optional[int] = None, batch_size: int = 1, *,
                                device=None):
    """A pseudofermion field (batch, *dims, Ns, Nc) and its descriptor
    (lime.py:384-420).  With field_index None the first field with precision
    F or D, recordtype 0 or absent, and positive spins and colours is taken
    (Chroma's pseudofermion files carry character-valued metadata fields
    before the floating-point leaf)."""
    fields = find_scidac_fields(path)
    idx = field_index
    if idx is None:
        picks = [i for i, m in enumerate(fields)
                 if m.precision.upper() in ("F", "D") and m.recordtype in (0, None) and m.spins > 0 and m.colors > 0]
        if not picks:
            raise ValueError("Could not auto-locate pseudofermion leaf field; pass --field-index.")
        idx = picks[0]
    meta = _field_meta(path, int(idx))
    plan = _plan(path, meta)
    if meta.dims:
        lat = _decode(path, plan, "site").reshape(*meta.dims, meta.datacount, *plan.item_shape)
    else:
        lat = _decode(path, plan, "lexi").reshape(plan.volume, meta.datacount, *plan.item_shape)
    pf = lat[..., 0, :, :] if meta.datacount == 1 else lat
    pf = pf.unsqueeze(0)
    if int(batch_size) != 1:
        pf = pf.repeat(int(batch_size), *([1] * (pf.dim() - 1)))
    return _finish(pf.contiguous(), device), meta
