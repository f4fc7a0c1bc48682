"""File formats around the Wilson operator: SciDAC / ILDG LIME
(jaxqft/io/__init__.py, jaxqft/io/lime.py)."""

from .lime import (  # noqa: F401
    LIME_HEADER_BYTES,
    LIME_MAGIC,
    LimeRecord,
    ScidacFieldMeta,
    decode_payload,
    decode_scidac_field,
    decode_scidac_gauge,
    decode_scidac_momentum,
    decode_scidac_pseudofermion,
    find_scidac_fields,
    iter_lime_records,
    list_lime_records,
    read_lime_record_data,
)

__all__ = sorted([
    "LimeRecord", "ScidacFieldMeta", "iter_lime_records", "list_lime_records", "read_lime_record_data",
    "find_scidac_fields", "decode_scidac_field", "decode_scidac_gauge", "decode_scidac_momentum",
    "decode_scidac_pseudofermion", "decode_payload",
])
