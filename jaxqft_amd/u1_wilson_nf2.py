"""U(1) Wilson-fermion operator and CG drop-in (jaxqft/models/u1_wilson_nf2.py:296-1025).

Links (B, 4, L0..L3) for BMXYIJ or (B, L0..L3, 4) for BXYMIJ
(u1_ym.py:189-192); fermions (B, L0..L3, 4, 1) (u1_wilson_nf2.py:446-447);
default boundary phases antiperiodic in the last axis (u1_wilson_nf2.py:461-471).
"""

from __future__ import annotations

from dataclasses import dataclass

from . import fields as _fields
from .theory import WilsonNf2Theory


@dataclass
class U1WilsonNf2(WilsonNf2Theory):
    _default_antiperiodic_t = True

    def __post_init__(self):
        super().__post_init__()
        self.Nc = 1

    def field_shape(self):
        if self.layout == "BMXYIJ":
            return (self.Bs, self.Nd, *self.lattice_shape)
        return (self.Bs, *self.lattice_shape, self.Nd)

    def hot_start(self, scale: float = 0.1):
        """exp(i scale N) per link (u1_ym.py:206-214)."""
        return _fields.u1_hot_start(self._rng, self.Bs, self.lattice_shape, scale, self.layout).astype(self.dtype)
