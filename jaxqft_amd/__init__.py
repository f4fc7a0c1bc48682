"""MI355X-native Wilson-fermion operator and normal-equation solvers.

Drop-in for the Wilson-fermion path of jaxQFT (``jaxqft.fermions`` and the
operator/solver part of ``SU3WilsonNf2`` / ``U1WilsonNf2``), running on
hand-written HIP kernels for gfx950 through the C ABI of libwilson_amd.
"""

from __future__ import annotations

__all__ = ["SU3WilsonNf2", "U1WilsonNf2", "WilsonDiracOperator", "build_euclidean_gamma", "gamma5",
           "check_gamma_algebra"]


def __getattr__(name):
    if name == "SU3WilsonNf2":
        from .su3_wilson_nf2 import SU3WilsonNf2
        return SU3WilsonNf2
    if name == "U1WilsonNf2":
        from .u1_wilson_nf2 import U1WilsonNf2
        return U1WilsonNf2
    if name in ("WilsonDiracOperator", "build_euclidean_gamma", "gamma5", "check_gamma_algebra"):
        from . import fermions
        return getattr(fermions, name)
    raise AttributeError(name)
