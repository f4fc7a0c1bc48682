"""Chronological initial guesses for the normal-equation solves.

The same predictors as ``_chrono_predict_from_history`` and
``_chrono_predict_mre`` (su3_wilson_nf2.py:85-194), on device tensors (or
NumPy arrays): the Gram system of the minimal-residual extrapolation is a few
complex inner products (torch reductions on the device) and an n x n solve on
the host.  Pass the result to ``solve_normal_with_guess``.
"""

from __future__ import annotations

import math
from typing import List, Optional

import numpy as np
import torch


def _vdot(a, b) -> complex:
    if isinstance(a, torch.Tensor):
        return complex(torch.vdot(a.reshape(-1), b.reshape(-1)).item())
    return complex(np.vdot(np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)))


def _norm(x) -> float:
    return math.sqrt(max(_vdot(x, x).real, 0.0))


def _vdots(pairs) -> List[complex]:
    """Many inner products; on the device they are reduced first and copied
    to the host together (one synchronization instead of one per product)."""
    if pairs and isinstance(pairs[0][0], torch.Tensor):
        vals = torch.stack([torch.vdot(a.reshape(-1), b.reshape(-1)) for a, b in pairs]).cpu().numpy()
        return [complex(v) for v in vals]
    return [_vdot(a, b) for a, b in pairs]


def chrono_predict_from_history(history: List, mode: str, max_hist: int):
    """'last' (the newest solution) or 'poly': x_{k+1} ~ sum_{j<p} (-1)^j C(p, j+1) x_{k-j}
    over the newest p = min(max_hist, len(history)) solutions (su3_wilson_nf2.py:85-111)."""
    if not history:
        return None
    h = history[-max(1, int(max_hist)):]
    if len(h) == 1 or str(mode).lower() != "poly":
        return h[-1]
    p = len(h)
    x0 = h[-1] * 0
    for j in range(p):
        x0 = x0 + float(((-1) ** j) * math.comb(p, j + 1)) * h[-1 - j]
    return x0


def chrono_predict_mre(history_x: List, history_y: List, target, max_hist: int, reg: float = 1e-12):
    """Minimal-residual extrapolation (su3_wilson_nf2.py:114-194): with pairs
    y_i ~ A x_i, coefficients c minimize |target - sum c_i y_i| (older entries
    damped by exp(-0.35 age), Tikhonov term reg * mean diagonal); the guess
    sum c_i x_i is blended toward the newest x by 1 / max(1, |c|_1), and the
    newest x is returned instead whenever the coefficients or the guess look
    unstable or the residual proxy improves by less than 3%."""
    if not history_x or not history_y:
        return None
    pairs = list(zip(history_x, history_y))[-max(1, int(max_hist)):]
    if len(pairs) == 1:
        return pairs[-1][0]
    xs = [p[0] for p in pairs]
    ys = [p[1] for p in pairs]
    n = len(xs)
    w = np.array([math.exp(-0.35 * float((n - 1) - i)) for i in range(n)])
    G = np.zeros((n, n), dtype=np.complex128)
    g = np.zeros(n, dtype=np.complex128)
    idx = [(i, j) for i in range(n) for j in range(i, n)]
    prods = _vdots([(ys[i], target) for i in range(n)] + [(ys[i], ys[j]) for i, j in idx])
    for i in range(n):
        g[i] = w[i] * prods[i]
    for (i, j), v in zip(idx, prods[n:]):
        G[i, j] = w[i] * w[j] * v
        G[j, i] = np.conj(G[i, j])
    lam = float(reg) * max(1.0, abs(float(np.trace(G).real)) / max(1, n))
    Gr = G + lam * np.eye(n)
    try:
        c = np.linalg.solve(Gr, g)
    except np.linalg.LinAlgError:
        c = np.linalg.lstsq(Gr, g, rcond=None)[0]
    c = c * w
    if not np.all(np.isfinite(c)):
        return xs[-1]
    cl1 = float(np.sum(np.abs(c)))
    if float(np.max(np.abs(c))) > 2.0 or cl1 > 3.0:
        return xs[-1]
    x0 = xs[-1] * 0
    for ci, xi in zip(c, xs):
        x0 = x0 + complex(ci) * xi
    beta = 1.0 / max(1.0, cl1)
    x0 = beta * x0 + (1.0 - beta) * xs[-1]
    yc = ys[-1] * 0
    for ci, yi in zip(c, ys):
        yc = yc + complex(ci) * yi
    dm, dl = target - yc, target - ys[-1]
    sq = _vdots([(xs[-1], xs[-1]), (x0, x0), (dm, dm), (dl, dl)])
    n_last, n_x0, r_mre, r_last = (math.sqrt(max(v.real, 0.0)) for v in sq)
    if not math.isfinite(n_x0) or (n_last > 0.0 and n_x0 > 5.0 * n_last):
        return xs[-1]
    if not (math.isfinite(r_mre) and math.isfinite(r_last)) or r_mre >= 0.97 * r_last:
        return xs[-1]
    return x0


class SolverGuess:
    """History of normal-equation solutions and their images y = D x for
    chronological initial guesses, kept per pseudofermion as the reference
    monomial keeps it (WilsonNf2PseudofermionMonomial._predict_solver_guess /
    _update_solver_guess, su3_wilson_nf2.py:254-294): at most `history`
    entries; mode 'last' or 'poly' extrapolates the solutions, 'mre' combines
    them to fit the target eta (phi = D^dag eta) and falls back to 'poly'
    without eta or images."""

    def __init__(self, mode: str = "last", history: int = 1):
        mode = str(mode).lower()
        if mode not in ("last", "poly", "mre"):
            raise ValueError("solver_guess_mode must be one of: last, poly, mre")
        if int(history) <= 0:
            raise ValueError("solver_guess_history must be positive")
        self.mode = mode
        self.history = int(history)
        self.xs: List = []
        self.ys: List = []

    def clear(self) -> None:
        self.xs = []
        self.ys = []

    def predict(self, phi, eta=None):
        """A guess of the same shape as phi, or None with no usable history."""
        shape = tuple(phi.shape)
        keep = [(x, y) for x, y in zip(self.xs, self.ys) if tuple(x.shape) == shape][-self.history:]
        self.xs = [x for x, _ in keep]
        self.ys = [y for _, y in keep]
        if not self.xs:
            return None
        if self.mode == "mre":
            if eta is not None:
                pairs = [(x, y) for x, y in zip(self.xs, self.ys) if y is not None and tuple(y.shape) == tuple(eta.shape)]
                if pairs:
                    x0 = chrono_predict_mre([p[0] for p in pairs], [p[1] for p in pairs], eta, self.history)
                    if x0 is not None:
                        return x0
            return chrono_predict_from_history(self.xs, "poly", self.history)
        return chrono_predict_from_history(self.xs, self.mode, self.history)

    def update(self, x, y=None) -> None:
        self.xs = (self.xs + [x])[-self.history:]
        self.ys = (self.ys + [y])[-self.history:]
