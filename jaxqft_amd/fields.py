"""Seeded synthetic gauge and fermion fields.

The reference draws its fields with JAX threefry keys, which cannot be
reproduced bit for bit without JAX.  These generators draw from the same
distributions with NumPy's PCG64:

* SU(3) links ``U = exp(X)`` with ``X = project_su3_algebra(eps (N + iN))``
  (``su3_ym.py:238-244``, ``:265-272``, ``:366-367``); ``X = iH`` with ``H``
  traceless Hermitian, so ``U = V diag(exp(i lambda)) V^dag`` from the
  eigen-decomposition ``H = V diag(lambda) V^dag`` is exact to rounding;
* U(1) links ``exp(i eps N)`` (``u1_ym.py:206-214``);
* fermions ``(N + iN)/sqrt(2)`` (``su3_wilson_nf2.py:1169-1175``);
* unit point sources over (spin, colour) (``core/measurements.py:333-337``).
"""

from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

SEED_BASE = 20261203


def rng_for(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed)))


def project_su3_algebra(x: np.ndarray) -> np.ndarray:
    """Traceless anti-Hermitian part (su3_ym.py:63-67)."""
    y = 0.5 * (x - np.conj(np.swapaxes(x, -1, -2)))
    tr = np.trace(y, axis1=-2, axis2=-1)[..., None, None] / 3.0
    return y - tr * np.eye(3, dtype=x.dtype)


def expm_su3_algebra(X: np.ndarray, chunk: int = 1 << 19) -> np.ndarray:
    """exp(X) for traceless anti-Hermitian 3x3 X via the Hermitian eigenproblem."""
    out = np.empty_like(X, dtype=np.complex128)
    flat_in = X.reshape(-1, 3, 3)
    flat_out = out.reshape(-1, 3, 3)
    for s in range(0, flat_in.shape[0], chunk):
        H = -1j * flat_in[s:s + chunk]
        H = 0.5 * (H + np.conj(np.swapaxes(H, -1, -2)))
        lam, V = np.linalg.eigh(H)
        flat_out[s:s + chunk] = np.einsum("nij,nj,nkj->nik", V, np.exp(1j * lam), np.conj(V))
    return out


def su3_hot_start(rng: np.random.Generator, batch: int, lattice_shape: Sequence[int], eps: float,
                  layout: str = "BMXYIJ") -> np.ndarray:
    """Random SU(3) links, complex128, shape (B, Nd, *L, 3, 3) for BMXYIJ."""
    L = tuple(int(v) for v in lattice_shape)
    nd = len(L)
    shape = (int(batch), nd, *L, 3, 3) if layout.upper() == "BMXYIJ" else (int(batch), *L, nd, 3, 3)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return expm_su3_algebra(project_su3_algebra(float(eps) * a))


def u1_hot_start(rng: np.random.Generator, batch: int, lattice_shape: Sequence[int], eps: float,
                 layout: str = "BMXYIJ") -> np.ndarray:
    """Random U(1) links exp(i eps N), complex128, shape (B, Nd, *L)."""
    L = tuple(int(v) for v in lattice_shape)
    nd = len(L)
    shape = (int(batch), nd, *L) if layout.upper() == "BMXYIJ" else (int(batch), *L, nd)
    return np.exp(1j * float(eps) * rng.standard_normal(shape))


def random_fermion(rng: np.random.Generator, shape: Tuple[int, ...]) -> np.ndarray:
    """Complex unit-variance Gaussian, E|z|^2 = 1."""
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)


def point_sources(lattice_shape: Sequence[int], site: Sequence[int], ns: int = 4, nc: int = 3) -> np.ndarray:
    """The ns*nc unit sources at one site, shape (ns*nc, *L, ns, nc), ordered
    spin-major like the (src_spin, src_color) loop of measurements.py:333-337."""
    L = tuple(int(v) for v in lattice_shape)
    out = np.zeros((ns * nc, *L, ns, nc), dtype=np.complex128)
    for s in range(ns):
        for c in range(nc):
            out[(s * nc + c, *tuple(int(v) for v in site), s, c)] = 1.0
    return out


def su3_hot_start_device(seed: int, batch: int, lattice_shape: Sequence[int], eps: float, device="cuda",
                         chunk: int = 1 << 20):
    """Same distribution as su3_hot_start, drawn on the device with a seeded
    torch generator (a different random stream from the NumPy one): the
    exponential of the projected algebra element via torch.linalg.matrix_exp.
    Returns a complex128 tensor (B, 4, *L, 3, 3) on `device`."""
    import torch

    L = tuple(int(v) for v in lattice_shape)
    shape = (int(batch), len(L), *L, 3, 3)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    out = torch.empty(shape, dtype=torch.complex128, device=device)
    flat = out.view(-1, 3, 3)
    eye = torch.eye(3, dtype=torch.complex128, device=device)
    for s in range(0, flat.shape[0], chunk):
        n = min(chunk, flat.shape[0] - s)
        a = torch.complex(torch.randn((n, 3, 3), generator=gen, device=device, dtype=torch.float64),
                          torch.randn((n, 3, 3), generator=gen, device=device, dtype=torch.float64)) * float(eps)
        y = 0.5 * (a - a.conj().transpose(-1, -2))
        y = y - (torch.diagonal(y, dim1=-2, dim2=-1).sum(-1) / 3.0)[:, None, None] * eye
        flat[s:s + n] = torch.linalg.matrix_exp(y)
    return out
