"""SU(3) Wilson-fermion operator and CG drop-in (jaxqft/models/su3_wilson_nf2.py:594-1635).

Links (B, 4, L0..L3, 3, 3) for BMXYIJ or (B, L0..L3, 4, 3, 3) for BXYMIJ
(su3_ym.py:246-249); fermions (B, L0..L3, 4, 3) (su3_wilson_nf2.py:815-817);
default boundary phases periodic (su3_wilson_nf2.py:712-713).
"""

from __future__ import annotations

from dataclasses import dataclass

from . import fields as _fields
from .theory import WilsonNf2Theory


@dataclass
class SU3WilsonNf2(WilsonNf2Theory):
    _default_antiperiodic_t = False

    def __post_init__(self):
        super().__post_init__()
        self.Nc = 3

    def field_shape(self):
        if self.layout == "BMXYIJ":
            return (self.Bs, self.Nd, *self.lattice_shape, 3, 3)
        return (self.Bs, *self.lattice_shape, self.Nd, 3, 3)

    def hot_start(self, scale: float = 0.1):
        """exp(project_su3_algebra(scale (N + iN))) per link (su3_ym.py:366-367)."""
        return _fields.su3_hot_start(self._rng, self.Bs, self.lattice_shape, scale, self.layout).astype(self.dtype)
