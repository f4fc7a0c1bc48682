"""CPU oracle: a NumPy FP64 restatement of the reference Wilson-fermion path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (``jaxqft_amd``)
never imports it and has no CPU fallback.

Every function cites the reference file:line it restates.  Reference paths are
relative to the jaxQFT source tree.

Conventions (all restated from the reference):

* spinor fields: ``(B, L0, L1, L2, L3, Ns, Nc)`` complex, C-order, last
  lattice axis fastest (``su3_wilson_nf2.py:815-817``, ``u1_wilson_nf2.py:446-447``);
* link fields ``BMXYIJ``: ``(B, Nd, L0, .., L3, Nc, Nc)`` (SU(3)) or
  ``(B, Nd, L0, .., L3)`` (U(1)); ``BXYMIJ`` moves the direction axis after the
  lattice axes (``su3_ym.py:246-249``, ``u1_ym.py:189-192``);
* D psi = (m + r Nd) psi - 1/2 sum_mu [(r - g_mu) U~_mu(x) psi(x+mu)
  + (r + g_mu) U~_mu(x-mu)^dag psi(x-mu)]   (``wilson.py:228-256``, sign=-1),
  D^dag swaps the two projectors (sign=+1);
* U~ carries the fermion boundary phase p_mu on the slice x_mu = L_mu - 1
  (``su3_wilson_nf2.py:726-759``).

Parity pinning: see ``tests/test_oracle.py``.  Every function here is checked
there against something the mathematics fixes independently of this file
(Clifford relations, free-field momentum-space symbol with antiperiodic
momenta, dense brute-force adjoint/LU, gauge covariance, invariants), and the
two reference stencil variants (``_apply_dslash_reference`` and
``_apply_dslash_optimized``) are restated separately and checked against each
other.  The CG iterate sequence of ``jax.scipy.sparse.linalg.cg`` (third
party, not in the reference tree) is restated from its published algorithm;
only its converged solution is pinned (dense LU), not its per-iteration
sequence.
"""

from __future__ import annotations

from typing import Callable, Dict, Optional, Sequence, Tuple

import numpy as np

# ----------------------------------------------------------------------------
# gamma matrices  (gamma.py:14-75)
# ----------------------------------------------------------------------------

_S1 = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=np.complex128)
_S2 = np.array([[0.0, -1.0j], [1.0j, 0.0]], dtype=np.complex128)
_S3 = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=np.complex128)


def build_euclidean_gamma(ndim: int) -> np.ndarray:
    """Recursive Kronecker construction (gamma.py:14-46)."""
    d = int(ndim)
    if d < 1:
        raise ValueError("ndim must be >= 1")
    if d == 1:
        return np.ones((1, 1, 1), dtype=np.complex128)
    if d == 2:
        return np.stack([_S1, _S2])
    gam = [_S1, _S2]
    cur = 2
    while cur + 1 < d:
        m = gam[0].shape[0]
        eye_m = np.eye(m, dtype=np.complex128)
        gam = [np.kron(_S1, g) for g in gam] + [np.kron(_S2, eye_m), np.kron(_S3, eye_m)]
        cur += 2
    if cur == d:
        return np.stack(gam)
    n = (d - 1) // 2
    g_last = (1j) ** n * gam[0]
    for mu in range(1, d - 1):
        g_last = g_last @ gam[mu]
    return np.stack(gam + [g_last])


def gamma5(gam: np.ndarray) -> np.ndarray:
    """gamma5 = i^(d/2) g0 g1 ... (gamma.py:49-57)."""
    d = int(gam.shape[0])
    if d % 2:
        raise ValueError("gamma5 is defined only for even ndim")
    out = (1j) ** (d // 2) * gam[0]
    for mu in range(1, d):
        out = out @ gam[mu]
    return out


def sparse_gamma_tables(gam: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Signed-permutation form of each gamma_mu (wilson.py:50-67)."""
    perm, phase = [], []
    for g in gam:
        nnz = np.sum(np.abs(g) > 1e-7, axis=1)
        if not np.all(nnz == 1):
            raise ValueError("gamma is not a signed permutation")
        idx = np.argmax(np.abs(g), axis=1)
        perm.append(idx.astype(np.int32))
        phase.append(g[np.arange(g.shape[0]), idx])
    return np.stack(perm), np.stack(phase)


# ----------------------------------------------------------------------------
# layout helpers  (su3_ym.py:246-252, u1_ym.py:189-196)
# ----------------------------------------------------------------------------

def links_bmxyij(U: np.ndarray, nd: int, layout: str) -> np.ndarray:
    """Return links with the direction axis at position 1 (BMXYIJ)."""
    lay = layout.upper()
    if lay in ("BMXYIJ", "BM...IJ"):
        return U
    if lay in ("BXYMIJ", "B...MIJ"):
        return np.moveaxis(U, 1 + nd, 1)
    raise ValueError(f"unknown layout {layout}")


def normalize_fermion_bc(bc, nd: int, default_antiperiodic_t: bool = False) -> Tuple[complex, ...]:
    """su3_wilson_nf2.py:711-724 (SU(3), periodic default) and
    u1_wilson_nf2.py:461-471 (U(1), antiperiodic-in-last-axis default)."""
    if bc is None:
        vals = [1.0 + 0.0j] * nd
        if default_antiperiodic_t:
            vals[-1] = -1.0 + 0.0j
    elif isinstance(bc, str):
        txt = bc.strip().lower()
        if txt in ("", "periodic", "p"):
            vals = [1.0 + 0.0j] * nd
        elif txt in ("antiperiodic-t", "anti-t", "ap-t", "apt", "chroma"):
            vals = [1.0 + 0.0j] * nd
            vals[-1] = -1.0 + 0.0j
        else:
            vals = [complex(t.strip().replace("I", "j").replace("i", "j")) for t in bc.split(",") if t.strip()]
    else:
        vals = [complex(v) for v in bc]
    if len(vals) != nd:
        raise ValueError(f"fermion_bc must have Nd={nd} entries; got {len(vals)}")
    return tuple(vals)


def links_with_fermion_bc(U: np.ndarray, bc: Sequence[complex]) -> np.ndarray:
    """Multiply U_mu(x) by p_mu on the slice x_mu = L_mu - 1
    (su3_wilson_nf2.py:726-759; U(1): u1_wilson_nf2.py:473-505).  U is BMXYIJ."""
    nd = len(bc)
    out = np.array(U, dtype=np.complex128, copy=True)
    for mu in range(nd):
        p = complex(bc[mu])
        if abs(p - 1.0) <= 1e-12:
            continue
        sl = [slice(None)] * out.ndim
        sl[1] = mu
        sl[2 + mu] = out.shape[2 + mu] - 1
        out[tuple(sl)] *= p
    return out


# ----------------------------------------------------------------------------
# Wilson stencil  (wilson.py:107-336)
# ----------------------------------------------------------------------------

def _roll_site(x: np.ndarray, shift: int, mu: int) -> np.ndarray:
    """su3_wilson_nf2.py:990-991: jnp.roll on lattice axis 1+mu."""
    return np.roll(x, shift, axis=1 + mu)


def _is_scalar_links(U_mu: np.ndarray, psi: np.ndarray) -> bool:
    return U_mu.ndim == psi.ndim - 2


def color_mul_left(U_mu: np.ndarray, psi: np.ndarray) -> np.ndarray:
    """wilson.py:107-127: out[...,s,a] = sum_b U[...,a,b] psi[...,s,b]."""
    if _is_scalar_links(U_mu, psi):
        return U_mu[..., None, None] * psi
    return np.einsum("...ab,...sb->...sa", U_mu, psi)


def dagger_link(U_mu: np.ndarray, scalar: bool) -> np.ndarray:
    """su3_ym.py:21-22 (matrix adjoint) / u1_wilson_nf2.py:522-524 (conj)."""
    if scalar:
        return np.conj(U_mu)
    return np.conj(np.swapaxes(U_mu, -1, -2))


def gamma_apply(gam: np.ndarray, psi: np.ndarray, mu: int) -> np.ndarray:
    """Dense gamma_mu on the spin axis (wilson.py:145-153, use_sparse=False)."""
    return np.einsum("st,...tc->...sc", gam[mu], psi)


def gamma_apply_sparse(perm: np.ndarray, phase: np.ndarray, psi: np.ndarray, mu: int) -> np.ndarray:
    """Gather + phase form (wilson.py:145-152)."""
    g = np.take(psi, perm[mu], axis=-2)
    return phase[mu].reshape((1,) * (psi.ndim - 2) + (-1, 1)) * g


def spin_project(gam, psi, mu, coeff, r=1.0):
    """(r I + coeff gamma_mu) psi (wilson.py:155-157)."""
    return r * psi + coeff * gamma_apply(gam, psi, mu)


def dslash_reference(U_links: np.ndarray, psi: np.ndarray, sign: int, r: float = 1.0) -> np.ndarray:
    """Literal loop of wilson.py:228-256.  U_links is BMXYIJ with the boundary
    phase already applied; sign=-1 gives K for D, sign=+1 for D^dag."""
    if sign not in (-1, 1):
        raise ValueError("sign must be -1 (D) or +1 (Ddag)")
    nd = U_links.shape[1]
    scalar = U_links.ndim == psi.ndim - 1
    gam = build_euclidean_gamma(nd)
    out = np.zeros_like(psi, dtype=np.complex128)
    for mu in range(nd):
        U_mu = U_links[:, mu]
        psi_xpmu = _roll_site(psi, -1, mu)
        psi_xmmu = _roll_site(psi, +1, mu)
        fwd = color_mul_left(U_mu, psi_xpmu)
        U_mu_xmmu = _roll_site(U_mu, +1, mu)
        bwd = color_mul_left(dagger_link(U_mu_xmmu, scalar), psi_xmmu)
        fwd = spin_project(gam, fwd, mu, +sign, r)
        bwd = spin_project(gam, bwd, mu, -sign, r)
        out = out - 0.5 * (fwd + bwd)
    return out


def dslash_optimized(U_links: np.ndarray, psi: np.ndarray, sign: int, r: float = 1.0) -> np.ndarray:
    """wilson.py:258-284: backward hop as roll(U^dag(x) psi(x), +mu), with the
    sparse (signed permutation) gamma apply."""
    if sign not in (-1, 1):
        raise ValueError("sign must be -1 (D) or +1 (Ddag)")
    nd = U_links.shape[1]
    scalar = U_links.ndim == psi.ndim - 1
    perm, phase = sparse_gamma_tables(build_euclidean_gamma(nd))
    out = np.zeros_like(psi, dtype=np.complex128)
    for mu in range(nd):
        U_mu = U_links[:, mu]
        up = color_mul_left(U_mu, _roll_site(psi, -1, mu))
        fwd = r * up + sign * gamma_apply_sparse(perm, phase, up, mu)
        ub = color_mul_left(dagger_link(U_mu, scalar), psi)
        bwd_local = r * ub - sign * gamma_apply_sparse(perm, phase, ub, mu)
        out = out - 0.5 * (fwd + _roll_site(bwd_local, +1, mu))
    return out


class WilsonOracle:
    """Theory-level operator API of SU3WilsonNf2 / U1WilsonNf2 restated
    (su3_wilson_nf2.py:1036-1110, u1_wilson_nf2.py:526-660).

    ``U`` is given in ``layout``; the boundary phase is applied here, as
    ``_links_with_fermion_bc`` does in the reference."""

    def __init__(self, lattice_shape, nc: int, mass: float, wilson_r: float = 1.0,
                 fermion_bc=None, layout: str = "BMXYIJ", kernel: str = "reference"):
        self.lattice_shape = tuple(int(v) for v in lattice_shape)
        self.Nd = len(self.lattice_shape)
        self.nc = int(nc)
        self.mass = float(mass)
        self.wilson_r = float(wilson_r)
        self.layout = layout
        self.fermion_bc = normalize_fermion_bc(fermion_bc, self.Nd, default_antiperiodic_t=(self.nc == 1))
        self.kernel = kernel
        self.gamma = build_euclidean_gamma(self.Nd)
        self.Ns = self.gamma.shape[-1]
        self.m0 = self.mass + self.wilson_r * self.Nd

    def _links(self, U):
        return links_with_fermion_bc(links_bmxyij(np.asarray(U), self.Nd, self.layout), self.fermion_bc)

    def _dslash(self, Ul, psi, sign):
        f = dslash_reference if self.kernel == "reference" else dslash_optimized
        return f(Ul, np.asarray(psi, dtype=np.complex128), sign, self.wilson_r)

    def apply_D(self, U, psi):
        Ul = self._links(U)
        return self.m0 * psi + self._dslash(Ul, psi, -1)

    def apply_Ddag(self, U, psi):
        Ul = self._links(U)
        return self.m0 * psi + self._dslash(Ul, psi, +1)

    def apply_offdiag(self, U, psi):
        return self._dslash(self._links(U), psi, -1)

    def apply_offdiag_dagger(self, U, psi):
        return self._dslash(self._links(U), psi, +1)

    def apply_normal(self, U, psi):
        return self.apply_Ddag(U, self.apply_D(U, psi))

    def apply_gamma5(self, psi):
        """su3_wilson_nf2.py:1688-1690."""
        return np.einsum("st,...tc->...sc", gamma5(self.gamma), psi)

    def apply_g5D(self, U, psi):
        return self.apply_gamma5(self.apply_D(U, psi))

    # ----------------------------------------------------------- even-odd blocks
    def parity_masks(self):
        """Even/odd masks of shape (1, *L, 1, 1) (su3_wilson_nf2.py:819-829)."""
        even = (parity_field(self.lattice_shape) % 2 == 0).astype(np.float64)[None, ..., None, None]
        return even, 1.0 - even

    def project_even(self, psi):
        """su3_wilson_nf2.py:884-887."""
        return np.asarray(psi) * self.parity_masks()[0]

    def project_odd(self, psi):
        """su3_wilson_nf2.py:889-892."""
        return np.asarray(psi) * self.parity_masks()[1]

    def apply_K_oe(self, U, psi_even):
        """P_o K P_e (su3_wilson_nf2.py:1057-1061)."""
        return self.project_odd(self.apply_offdiag(U, self.project_even(psi_even)))

    def apply_K_eo(self, U, psi_odd):
        """P_e K P_o (su3_wilson_nf2.py:1063-1067)."""
        return self.project_even(self.apply_offdiag(U, self.project_odd(psi_odd)))

    def apply_Kdag_oe(self, U, psi_even):
        """su3_wilson_nf2.py:1069-1073."""
        return self.project_odd(self.apply_offdiag_dagger(U, self.project_even(psi_even)))

    def apply_Kdag_eo(self, U, psi_odd):
        """su3_wilson_nf2.py:1075-1079."""
        return self.project_even(self.apply_offdiag_dagger(U, self.project_odd(psi_odd)))

    def apply_eo_schur_even(self, U, psi_even, normalized: bool = False):
        """S_e = m0 - K_eo K_oe / m0, or 1 - K_eo K_oe / m0^2 (su3_wilson_nf2.py:1112-1121)."""
        xe = self.project_even(psi_even)
        z = self.apply_K_eo(U, self.apply_K_oe(U, xe))
        return xe - z / (self.m0 * self.m0) if normalized else self.m0 * xe - z / self.m0

    def apply_eo_schur_even_dagger(self, U, psi_even, normalized: bool = False):
        """S_e^dag = m0 - K^dag_eo K^dag_oe / m0 (su3_wilson_nf2.py:1138-1146)."""
        xe = self.project_even(psi_even)
        z = self.apply_Kdag_eo(U, self.apply_Kdag_oe(U, xe))
        return xe - z / (self.m0 * self.m0) if normalized else self.m0 * xe - z / self.m0

    def apply_eo_schur_odd(self, U, psi_odd, normalized: bool = False):
        """S_o = m0 - K_oe K_eo / m0 (su3_wilson_nf2.py:1148-1156)."""
        xo = self.project_odd(psi_odd)
        z = self.apply_K_oe(U, self.apply_K_eo(U, xo))
        return xo - z / (self.m0 * self.m0) if normalized else self.m0 * xo - z / self.m0

    def pack_parity(self, psi, parity: int):
        """(B, *L, Ns, Nc) -> (B, V/2, Ns, Nc), sites of one parity in ascending
        flat order (su3_wilson_nf2.py:894-910)."""
        idx = checkerboard_indices(self.lattice_shape)[parity]
        psi = np.asarray(psi)
        return psi.reshape((psi.shape[0], -1) + psi.shape[-2:])[:, idx]

    def unpack_parity(self, psi_c, parity: int):
        """Inverse of pack_parity, zero on the other parity (su3_wilson_nf2.py:912-932)."""
        idx = checkerboard_indices(self.lattice_shape)[parity]
        V = int(np.prod(self.lattice_shape))
        out = np.zeros((psi_c.shape[0], V) + psi_c.shape[-2:], dtype=np.complex128)
        out[:, idx] = psi_c
        return out.reshape((psi_c.shape[0],) + self.lattice_shape + psi_c.shape[-2:])

    def _parity_compact_hop(self, Ul_b, src_c, sign: int, target_block: str):
        """Literal restatement of _apply_dslash_parity_compact_one
        (su3_wilson_nf2.py:934-965), with the exact projectors (r + coeff g_mu)
        of projected_hop (wilson.py:163-169).  Ul_b: links of one configuration
        with the boundary phase, (Nd, *L[, Nc, Nc])."""
        maps = checkerboard_neighbor_maps(self.lattice_shape)
        blk = "oe" if target_block == "oe" else "eo"
        target, plus, minus, bwd = maps["target_" + blk], maps["plus_" + blk], maps["minus_" + blk], maps["bwd_" + blk]
        V = int(np.prod(self.lattice_shape))
        scalar = self.nc == 1
        out = np.zeros_like(src_c, dtype=np.complex128)
        for mu in range(self.Nd):
            U_mu = Ul_b[mu].reshape((V,) if scalar else (V, self.nc, self.nc))
            U_f = U_mu[target]
            U_b = np.conj(U_mu[bwd[mu]]) if scalar else np.conj(np.swapaxes(U_mu[bwd[mu]], -1, -2))
            for U_l, psi_n, coeff in ((U_f, src_c[plus[mu]], float(+sign)), (U_b, src_c[minus[mu]], float(-sign))):
                upsi = U_l[:, None, None] * psi_n if scalar else np.einsum("nij,nsj->nsi", U_l, psi_n)
                out = out - 0.5 * (self.wilson_r * upsi + coeff * np.einsum("st,ntc->nsc", self.gamma[mu], upsi))
        return out

    def apply_eo_schur_even_compact_one(self, U_b, xe_c, daggered: bool = False, normalized: bool = True):
        """Compact even Schur matvec for one configuration (su3_wilson_nf2.py:967-988)."""
        Ul = links_with_fermion_bc(links_bmxyij(np.asarray(U_b)[None], self.Nd, self.layout), self.fermion_bc)[0]
        sign = +1 if daggered else -1
        y = self._parity_compact_hop(Ul, np.asarray(xe_c, dtype=np.complex128), sign, "oe")
        z = self._parity_compact_hop(Ul, y, sign, "eo")
        return xe_c - z / (self.m0 * self.m0) if normalized else self.m0 * xe_c - z / self.m0

    def solve_eop_even_normal(self, U, rhs, x0=None, tol: float = 1e-10, maxiter: int = 5000):
        """CG on S_e^dag S_e (normalized) in the compact even basis, per batch
        element (su3_wilson_nf2.py:1373-1389), with the jax cg rule."""
        out = []
        pb = self.pack_parity(self.project_even(rhs), 0)
        xs = None if x0 is None else self.pack_parity(self.project_even(x0), 0)
        for b in range(pb.shape[0]):
            def A(v, b=b):
                y = self.apply_eo_schur_even_compact_one(np.asarray(U)[b], v, daggered=False, normalized=True)
                return self.apply_eo_schur_even_compact_one(np.asarray(U)[b], y, daggered=True, normalized=True)
            xb, _ = cg(A, pb[b], None if xs is None else xs[b], tol=tol, maxiter=maxiter)
            out.append(xb)
        return self.unpack_parity(np.stack(out), 0)

    # ------------------------------------------------------ pseudofermion force
    def pseudofermion_force_from_solution(self, U, X, Y):
        """Analytic force of S = phi^dag (D^dag D)^-1 phi from X = (D^dag D)^-1 phi
        and Y = D X (su3_wilson_nf2.py:1556-1576; U(1): u1_wilson_nf2.py:1004-1025):
        per link F_mu(x) = TA(k^T), k = c1 - c2,
        c1[a,c] = sum_st conj(Y(x)[s,a]) (r - g_mu)[s,t] (U~_mu(x) X(x+mu))[t,c],
        c2[d,c] = sum_sta conj(Y(x+mu)[s,a]) (r + g_mu)[s,t] U~_mu(x)^dag[a,d] X(x)[t,c],
        TA the traceless anti-Hermitian part (su3_ym.py:63-67) or 1/2 (k - conj k)
        for U(1) (u1_ym.py:21-23).  Returned in the layout of U."""
        Ul = self._links(U)  # BMXYIJ, with the boundary phase
        X = np.asarray(X, dtype=np.complex128)
        Y = np.asarray(Y, dtype=np.complex128)
        eye = np.eye(self.Ns, dtype=np.complex128)
        out = []
        for mu in range(self.Nd):
            U_mu = Ul[:, mu]
            Xp = _roll_site(X, -1, mu)
            Yp = _roll_site(Y, -1, mu)
            p_fwd = self.wilson_r * eye - self.gamma[mu]
            p_bwd = self.wilson_r * eye + self.gamma[mu]
            if self.nc == 1:
                UX = U_mu[..., None, None] * Xp
                Udag = np.conj(U_mu)[..., None, None]
            else:
                UX = np.einsum("...ij,...sj->...si", U_mu, Xp)
                Udag = np.conj(np.swapaxes(U_mu, -1, -2))
            c1 = np.einsum("...sa,st,...tc->...ac", np.conj(Y), p_fwd, UX)
            c2 = np.einsum("...sa,st,...ad,...tc->...dc", np.conj(Yp), p_bwd, Udag, X)
            if self.nc == 1:
                k = (c1 - c2)[..., 0, 0]
                out.append(0.5 * (k - np.conj(k)))
            else:
                k = np.swapaxes(c1 - c2, -1, -2)
                y = 0.5 * (k - np.conj(np.swapaxes(k, -1, -2)))
                tr = np.trace(y, axis1=-2, axis2=-1)[..., None, None] / 3.0
                out.append(y - tr * np.eye(3, dtype=np.complex128))
        F = np.stack(out, axis=1)
        if self.layout.upper() == "BXYMIJ":
            F = np.moveaxis(F, 1, 1 + self.Nd)
        return F

    def eop_pseudofermion_force_from_solution(self, U, X_e, Y_e):
        """Force of S = phi_e^dag (S^dag S)^-1 phi_e with the normalized even
        Schur complement, from X_e and Y_e = S X_e (su3_wilson_nf2.py:1578-1597)."""
        alpha = 1.0 / (self.m0 * self.m0)
        f1 = self.pseudofermion_force_from_solution(U, self.apply_K_oe(U, X_e), -alpha * np.asarray(Y_e))
        f2 = self.pseudofermion_force_from_solution(U, X_e, -alpha * self.apply_Kdag_oe(U, Y_e))
        return f1 + f2

    def normal_residual_stats(self, U, phi, x) -> Dict[str, float]:
        """su3_wilson_nf2.py:1456-1466."""
        r = self.apply_normal(U, x) - phi
        rel = [np.linalg.norm(r[b]) / (np.linalg.norm(phi[b]) + 1e-30) for b in range(r.shape[0])]
        absn = [np.linalg.norm(r[b]) for b in range(r.shape[0])]
        return {"mean_rel_residual": float(np.mean(rel)), "max_rel_residual": float(np.max(rel)),
                "mean_abs_residual": float(np.mean(absn)), "max_abs_residual": float(np.max(absn))}


# ----------------------------------------------------------------------------
# CG  (jax.scipy.sparse.linalg.cg, as called at su3_wilson_nf2.py:1243-1269)
# ----------------------------------------------------------------------------

def cg(A: Callable[[np.ndarray], np.ndarray], b: np.ndarray, x0: Optional[np.ndarray] = None,
       tol: float = 1e-5, atol: float = 0.0, maxiter: Optional[int] = None) -> Tuple[np.ndarray, int]:
    """Restatement of the published jax.scipy.sparse.linalg.cg iteration
    (unpreconditioned): stop when |r|^2 <= max(tol^2 |b|^2, atol^2) on the
    recursively updated residual, or at maxiter.  Returns (x, iterations);
    the reference itself returns (x, None)."""
    b = np.asarray(b, dtype=np.complex128)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.complex128)
    if maxiter is None:
        maxiter = 10 * b.size
    bs = float(np.vdot(b, b).real)
    atol2 = max(tol * tol * bs, atol * atol)
    r = b - A(x)
    p = r.copy()
    gamma = float(np.vdot(r, r).real)
    k = 0
    while gamma > atol2 and k < maxiter:
        Ap = A(p)
        alpha = gamma / float(np.vdot(p, Ap).real)
        x = x + alpha * p
        r = r - alpha * Ap
        gamma_new = float(np.vdot(r, r).real)
        beta = gamma_new / gamma
        gamma = gamma_new
        p = r + beta * p
        k += 1
    return x, k


# ----------------------------------------------------------------------------
# index tables  (su3_wilson_nf2.py:819-882)
# ----------------------------------------------------------------------------

def parity_field(lattice_shape) -> np.ndarray:
    parity = np.zeros(lattice_shape, dtype=np.int32)
    nd = len(lattice_shape)
    for ax, n in enumerate(lattice_shape):
        shp = [1] * nd
        shp[ax] = n
        parity = parity + np.arange(n, dtype=np.int32).reshape(shp)
    return parity


def checkerboard_indices(lattice_shape) -> Tuple[np.ndarray, np.ndarray]:
    """su3_wilson_nf2.py:831-840."""
    flat = parity_field(lattice_shape).reshape(-1)
    return (np.flatnonzero(flat % 2 == 0).astype(np.int32), np.flatnonzero(flat % 2 == 1).astype(np.int32))


def shift_flat_indices(flat_idx: np.ndarray, lattice_shape, mu: int, step: int) -> np.ndarray:
    """su3_wilson_nf2.py:842-845."""
    coords = np.array(np.unravel_index(flat_idx, lattice_shape), dtype=np.int64).T
    coords[:, mu] = (coords[:, mu] + int(step)) % int(lattice_shape[mu])
    return np.ravel_multi_index(coords.T, lattice_shape).astype(np.int32)


def checkerboard_neighbor_maps(lattice_shape) -> Dict[str, np.ndarray]:
    """su3_wilson_nf2.py:847-882 (names follow the reference attributes)."""
    even_idx, odd_idx = checkerboard_indices(lattice_shape)
    vol = int(np.prod(lattice_shape))
    even_inv = np.full((vol,), -1, dtype=np.int32)
    odd_inv = np.full((vol,), -1, dtype=np.int32)
    even_inv[even_idx] = np.arange(even_idx.shape[0], dtype=np.int32)
    odd_inv[odd_idx] = np.arange(odd_idx.shape[0], dtype=np.int32)
    out = {k: [] for k in ("plus_oe", "minus_oe", "bwd_oe", "plus_eo", "minus_eo", "bwd_eo")}
    for mu in range(len(lattice_shape)):
        op = shift_flat_indices(odd_idx, lattice_shape, mu, +1)
        om = shift_flat_indices(odd_idx, lattice_shape, mu, -1)
        out["plus_oe"].append(even_inv[op])
        out["minus_oe"].append(even_inv[om])
        out["bwd_oe"].append(om)
        ep = shift_flat_indices(even_idx, lattice_shape, mu, +1)
        em = shift_flat_indices(even_idx, lattice_shape, mu, -1)
        out["plus_eo"].append(odd_inv[ep])
        out["minus_eo"].append(odd_inv[em])
        out["bwd_eo"].append(em)
    res = {k: np.stack(v, axis=0).astype(np.int32) for k, v in out.items()}
    res["target_oe"] = odd_idx
    res["target_eo"] = even_idx
    res["even_idx"] = even_idx
    res["odd_idx"] = odd_idx
    return res


# ----------------------------------------------------------------------------
# brute force and closed forms used to pin the restatement
# ----------------------------------------------------------------------------

def dense_matrix(apply: Callable[[np.ndarray], np.ndarray], shape) -> np.ndarray:
    """Column-by-column build (pattern of scripts/u1_wilson_nf2/check_exact_det.py:31-57)."""
    n = int(np.prod(shape))
    M = np.zeros((n, n), dtype=np.complex128)
    e = np.zeros(n, dtype=np.complex128)
    for j in range(n):
        e[:] = 0
        e[j] = 1.0
        M[:, j] = np.asarray(apply(e.reshape(shape))).reshape(-1)
    return M


def free_field_symbol(p: Sequence[float], mass: float, r: float = 1.0) -> np.ndarray:
    """Momentum-space free Wilson operator (textbook, not from the reference).

    For a plane wave psi(x) = exp(i p.x) u and U = 1, the hopping sign of
    wilson.py:251-255 (forward projector r - g for D) gives
    -1/2 [(r - g) e^{ip} + (r + g) e^{-ip}] = -r cos p + i g sin p, so
    D(p) = m + r sum_mu (1 - cos p_mu) + i sum_mu g_mu sin p_mu.
    A boundary phase exp(i theta) in direction mu is exact for momenta
    p_mu = (2 pi n + theta) / L_mu."""
    nd = len(p)
    gam = build_euclidean_gamma(nd)
    ns = gam.shape[-1]
    out = (mass + r * sum(1.0 - np.cos(pm) for pm in p)) * np.eye(ns, dtype=np.complex128)
    for mu in range(nd):
        out = out + 1j * np.sin(p[mu]) * gam[mu]
    return out


def plane_wave(lattice_shape, p: Sequence[float], u: np.ndarray) -> np.ndarray:
    """psi(x)_{s,c} = exp(i p.x) u_{s,c}, shape (1, *L, Ns, Nc)."""
    grids = np.meshgrid(*[np.arange(n) for n in lattice_shape], indexing="ij")
    phase = np.exp(1j * sum(pm * g for pm, g in zip(p, grids)))
    return (phase[..., None, None] * u[(None,) * len(lattice_shape)])[None]


def gauge_transform_links(U: np.ndarray, Omega: np.ndarray) -> np.ndarray:
    """su3_ym.py:770-779: U'_mu(x) = Omega^dag(x) U_mu(x) Omega(x+mu), BMXYIJ."""
    nd = U.shape[1]
    out = []
    for mu in range(nd):
        Om_xpmu = np.roll(Omega, -1, axis=1 + mu)
        out.append(np.conj(np.swapaxes(Omega, -1, -2)) @ U[:, mu] @ Om_xpmu)
    return np.stack(out, axis=1)


def gauge_transform_fermion(psi: np.ndarray, Omega: np.ndarray) -> np.ndarray:
    """su3_wilson_nf2.py:1674-1676: psi' = Omega^dag psi."""
    return np.einsum("...ab,...sb->...sa", np.conj(np.swapaxes(Omega, -1, -2)), psi)
