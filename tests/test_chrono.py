"""Chronological initial guesses (jaxqft_amd/chrono.py; su3_wilson_nf2.py:85-194),
CPU: pinned by what the predictors must reproduce exactly -- polynomial
extrapolation is exact on histories that are polynomials of degree < p in the
step (so the binomial weights are the right ones), the minimal-residual
extrapolation returns the exact preimage when the target lies in the span of
the stored images with |c|_1 <= 1, and the documented fallbacks to the newest
solution."""

import numpy as np
import pytest

from jaxqft_amd.chrono import chrono_predict_from_history, chrono_predict_mre


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_poly_extrapolation_exact_on_polynomials(p):
    rng = np.random.default_rng(p)
    coef = [rng.standard_normal(7) + 1j * rng.standard_normal(7) for _ in range(p)]
    hist = [sum(c * (k ** d) for d, c in enumerate(coef)) for k in range(6)]
    nxt = sum(c * (6 ** d) for d, c in enumerate(coef))
    np.testing.assert_allclose(chrono_predict_from_history(hist, "poly", p), nxt, rtol=1e-10, atol=1e-9)


def test_history_modes_and_empty():
    h = [np.ones(3) * k for k in range(4)]
    assert chrono_predict_from_history([], "poly", 3) is None
    np.testing.assert_array_equal(chrono_predict_from_history(h, "last", 3), h[-1])
    np.testing.assert_array_equal(chrono_predict_from_history(h, "unknown", 3), h[-1])
    np.testing.assert_array_equal(chrono_predict_from_history(h[:1], "poly", 3), h[0])


def _spd(n, rng):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a @ a.conj().T + n * np.eye(n)


def test_mre_exact_in_span():
    rng = np.random.default_rng(7)
    A = _spd(24, rng)
    xs = [rng.standard_normal(24) + 1j * rng.standard_normal(24) for _ in range(4)]
    ys = [A @ x for x in xs]
    c = np.array([0.1, -0.2, 0.3j, 0.35])  # |c|_1 = 0.95 <= 1: no blending
    target = sum(ci * yi for ci, yi in zip(c, ys))
    x0 = chrono_predict_mre(xs, ys, target, 4)
    np.testing.assert_allclose(x0, sum(ci * xi for ci, xi in zip(c, xs)), rtol=1e-8, atol=1e-8)


def test_mre_fallbacks():
    rng = np.random.default_rng(8)
    A = _spd(16, rng)
    xs = [rng.standard_normal(16) + 0j for _ in range(3)]
    ys = [A @ x for x in xs]
    assert chrono_predict_mre([], [], ys[0], 3) is None
    np.testing.assert_array_equal(chrono_predict_mre(xs[:1], ys[:1], ys[0], 3), xs[0])
    # target = the newest image: no improvement over it, so the newest solution comes back
    np.testing.assert_array_equal(chrono_predict_mre(xs, ys, ys[-1], 3), xs[-1])
    # large coefficients are refused
    np.testing.assert_array_equal(chrono_predict_mre(xs, ys, 10.0 * ys[0] - 9.0 * ys[1], 3), xs[-1])
