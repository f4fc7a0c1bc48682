"""Euclidean gamma matrices (host constants of the operator).

Same API as the reference module jaxqft/fermions/gamma.py:14-75
(``build_euclidean_gamma``, ``gamma5``, ``check_gamma_algebra``), on NumPy.
The 4D basis g0 = s1 x s1, g1 = s1 x s2, g2 = s2 x 1, g3 = s3 x 1 is the one
the HIP kernels hard-code (their spin projectors are derived from it).
"""

from __future__ import annotations

from typing import Dict

import numpy as np

_S1 = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=np.complex128)
_S2 = np.array([[0.0, -1.0j], [1.0j, 0.0]], dtype=np.complex128)
_S3 = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=np.complex128)


def build_euclidean_gamma(ndim: int, dtype=np.complex128) -> np.ndarray:
    """gamma[mu] with {gamma_mu, gamma_nu} = 2 delta_{mu nu} (gamma.py:14-46)."""
    d = int(ndim)
    if d < 1:
        raise ValueError("ndim must be >= 1")
    if d == 1:
        return np.ones((1, 1, 1), dtype=dtype)
    if d == 2:
        return np.stack([_S1, _S2]).astype(dtype)
    gam = [_S1, _S2]
    cur = 2
    while cur + 1 < d:
        eye_m = np.eye(gam[0].shape[0], dtype=np.complex128)
        gam = [np.kron(_S1, g) for g in gam] + [np.kron(_S2, eye_m), np.kron(_S3, eye_m)]
        cur += 2
    if cur != d:
        g_last = (1j) ** ((d - 1) // 2) * gam[0]
        for mu in range(1, d - 1):
            g_last = g_last @ gam[mu]
        gam = gam + [g_last]
    return np.stack(gam).astype(dtype)


def gamma5(gam: np.ndarray) -> np.ndarray:
    """gamma5 = i^(d/2) g0 g1 ... g_{d-1} for even d (gamma.py:49-57)."""
    d = int(gam.shape[0])
    if d % 2 != 0:
        raise ValueError("gamma5 is defined here only for even ndim")
    out = (1j) ** (d // 2) * gam[0]
    for mu in range(1, d):
        out = out @ gam[mu]
    return out.astype(gam.dtype)


def check_gamma_algebra(gam: np.ndarray) -> Dict[str, float]:
    """Clifford and hermiticity residuals (gamma.py:60-75)."""
    d = int(gam.shape[0])
    n = int(gam.shape[-1])
    eye = np.eye(n, dtype=gam.dtype)
    max_rel = 0.0
    max_herm = 0.0
    for mu in range(d):
        gm = gam[mu]
        max_herm = max(max_herm, float(np.linalg.norm(gm - np.conj(gm.T))))
        for nu in range(d):
            anti = gm @ gam[nu] + gam[nu] @ gm
            target = 2.0 * eye if mu == nu else np.zeros_like(eye)
            rel = float(np.linalg.norm(anti - target)) / (float(np.linalg.norm(target)) + 1e-12)
            max_rel = max(max_rel, rel)
    return {"max_rel_clifford_error": max_rel, "max_hermiticity_error": max_herm}
