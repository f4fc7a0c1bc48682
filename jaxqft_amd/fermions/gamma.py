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
    """Residuals of the Euclidean Clifford algebra and of hermiticity, with
    the keys of the reference's checker (gamma.py:60-75).

    All anticommutators are formed at once: A[mu, nu] = g_mu g_nu + g_nu g_mu
    should equal 2 delta_{mu nu} times the identity.  The Clifford error is
    the largest Frobenius norm of A[mu, nu] - target divided by that of the
    target (plus 1e-12, so an off-diagonal pair, whose target is zero, reports
    its absolute error over 1e-12); the hermiticity error is the largest
    Frobenius norm of g_mu - g_mu^dag."""
    g = np.asarray(gam)
    d, n = int(g.shape[0]), int(g.shape[-1])
    gg = np.einsum("aij,bjk->abik", g, g)
    anti = gg + np.swapaxes(gg, 0, 1)
    target = 2.0 * np.einsum("ab,ij->abij", np.eye(d), np.eye(n))
    err = np.linalg.norm((anti - target).reshape(d, d, n * n), axis=-1)
    scale = np.linalg.norm(target.reshape(d, d, n * n), axis=-1) + 1e-12
    herm = np.linalg.norm((g - np.conj(np.swapaxes(g, -1, -2))).reshape(d, n * n), axis=-1)
    return {"max_rel_clifford_error": float(np.max(err / scale)), "max_hermiticity_error": float(np.max(herm))}
