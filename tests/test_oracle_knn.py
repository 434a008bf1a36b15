"""The grid-surrogate oracle against the reference's neighbour-count rules,
exact reproduction of linear and multilinear functions, an independent
weighted least-squares fit and library interpolation (CPU).

The oracle restates the kNN local-linear fit of buildSurrogateMTECUDA /
buildSurrogateROVCUDA (SimulateGPU.cu:5533-5917, multiLocLinReg :2934-3053)
and the multilinear evaluation of Utility::interpolateSurrogate
(Utility.cpp:253-363).  A wrong sign on the centring, a dropped slope term,
weights from the wrong distance or interpolation weights applied to the
wrong dimension each fail one of these.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from transportoptimiser_amd import native as nat


@pytest.mark.parametrize("n,rule,k", [
    (3, 0, 3), (10, 0, 5), (60, 0, 7), (100, 0, 7), (300, 0, 11), (499, 0, 15),
    (500, 0, 15), (4096, 0, 17),
    (3, 1, 3), (10, 1, 5), (60, 1, 5), (99, 1, 5), (100, 1, 10), (499, 1, 10),
    (500, 1, 25), (4096, 1, 97),
])
def test_neighbour_count_rules(oracle, n, rule, k):
    # MTE: 300 -> ceil(0.02*200 + 7) = 11; 499 -> ceil(14.98) = 15;
    # 500 -> ceil(ln 500 + 8) = ceil(14.21) = 15; 4096 -> ceil(16.32) = 17.
    # ROV: integer quotients make [50, 100) -> 5 and [100, 500) -> 10;
    # 4096 -> 25 + ceil(3596 * 0.02) = 25 + 72 = 97.
    assert oracle.knn_k(n, rule) == k


@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_fit_reproduces_linear_functions(oracle, d, normalize):
    rng = np.random.default_rng(d)
    x = rng.uniform(-2.0, 3.0, size=(300, d)) * np.arange(1, d + 1)
    a, b = 0.7, rng.normal(size=d)
    y = a + x @ b
    R = 7
    s = oracle.knn_fit(x, y, R, 12, nat.KNN_NORMALIZE if normalize else 0)
    grids = [s[i * R:(i + 1) * R] for i in range(d)]
    vals = s[d * R:].reshape((R,) * d)
    mesh = np.stack(np.meshgrid(*grids, indexing="ij"), -1)
    np.testing.assert_allclose(vals, a + mesh @ b, rtol=1e-9, atol=1e-9)
    # grid coordinates span the sample range
    for i in range(d):
        assert grids[i][0] == x[:, i].min() and grids[i][-1] == pytest.approx(x[:, i].max())


def _wls_reference(x, y, q, k):
    """Independent statement: stable nearest-k by Euclidean distance, Gaussian
    weights with bandwidth the k-th distance, numpy least squares on the
    sqrt-weighted design."""
    d2 = ((x - q) ** 2).sum(1)
    nb = np.argsort(d2, kind="stable")[:k]
    dist = np.sqrt(d2[nb])
    w = np.exp(-0.5 * (dist / dist[-1]) ** 2)
    X = np.column_stack([np.ones(k), x[nb] - q])
    coef, *_ = np.linalg.lstsq(X * np.sqrt(w)[:, None], y[nb] * np.sqrt(w), rcond=None)
    return coef[0]


@pytest.mark.parametrize("d", [1, 2, 3])
def test_fit_matches_weighted_least_squares(oracle, d):
    rng = np.random.default_rng(10 + d)
    x = rng.uniform(0.0, 1.0, size=(250, d))
    y = np.sin(3 * x[:, 0]) + (x[:, 1] ** 2 if d > 1 else 0) + 0.05 * rng.normal(size=250)
    R, k = 6, 15
    s = oracle.knn_fit(x, y, R, k, 0)
    grids = [s[i * R:(i + 1) * R] for i in range(d)]
    vals = s[d * R:].reshape((R,) * d)
    for idx in [(0,) * d, (R - 1,) * d, tuple(rng.integers(0, R, d)), tuple(rng.integers(0, R, d))]:
        q = np.array([grids[i][idx[i]] for i in range(d)])
        assert vals[idx] == pytest.approx(_wls_reference(x, y, q, k), rel=1e-9, abs=1e-12)


def test_normalized_fit_is_scale_invariant(oracle):
    # In [0,1]-scaled coordinates the fit cannot depend on each dimension's
    # units; in raw coordinates it does (neighbours change).
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, size=(200, 2))
    y = np.cos(4 * x[:, 0]) * x[:, 1]
    scale = np.array([1000.0, 0.01])
    a = oracle.knn_fit(x, y, 5, 10, nat.KNN_NORMALIZE)
    b = oracle.knn_fit(x * scale, y, 5, 10, nat.KNN_NORMALIZE)
    np.testing.assert_allclose(a[10:], b[10:], rtol=1e-9, atol=1e-12)
    c = oracle.knn_fit(x * scale, y, 5, 10, 0)
    assert np.abs(c[10:] - a[10:]).max() > 1e-3


def test_nonpositive_clamp_small_samples_and_duplicates(oracle):
    x = np.linspace(0, 1, 40)[:, None]
    s = oracle.knn_fit(x, 1.0 + x[:, 0], 5, 5, nat.KNN_NONPOSITIVE)
    np.testing.assert_array_equal(s[5:], 0.0)
    s = oracle.knn_fit(x, -2.0 + 0 * x[:, 0], 5, 5, nat.KNN_NONPOSITIVE)
    np.testing.assert_allclose(s[5:], -2.0, rtol=1e-12)
    # fewer than 3 samples: zeros (multiLocLinReg's dataPoints < 3 rule)
    s = oracle.knn_fit(np.array([[0.0], [1.0]]), np.array([3.0, 4.0]), 4, 5, 0)
    np.testing.assert_array_equal(s[4:], 0.0)
    # all samples at one point: the slope is not identifiable; the value is
    # the weighted (here plain) mean of the k nearest, the lowest indices
    xs = np.zeros((10, 1))
    ys = np.arange(10.0)
    s = oracle.knn_fit(xs, ys, 3, 4, 0)
    np.testing.assert_allclose(s[3:], np.mean(ys[:4]), rtol=1e-12)


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_eval_reproduces_multilinear_functions(oracle, d):
    rng = np.random.default_rng(20 + d)
    R = 5
    # uniform grids, as the fit lays them out (the cell search assumes it)
    grids = [np.linspace(*np.sort(rng.uniform(-1, 1, 2)), R) for _ in range(d)]
    coef = rng.normal(size=2 ** d)

    def f(p):  # sum over subsets S of coef_S prod_{i in S} p_i
        out = np.zeros(len(p))
        for c in range(2 ** d):
            term = np.full(len(p), coef[c])
            for i in range(d):
                if (c >> i) & 1:
                    term = term * p[:, i]
            out += term
        return out

    mesh = np.stack(np.meshgrid(*grids, indexing="ij"), -1).reshape(-1, d)
    s = np.concatenate([*grids, f(mesh)])
    pts = rng.uniform(-1.3, 1.3, size=(400, d))  # inside and beyond the grid
    np.testing.assert_allclose(oracle.knn_eval(s, d, R, pts), f(pts), rtol=1e-9, atol=1e-9)
    # against scipy inside the grid, for arbitrary node values
    vals = rng.normal(size=(R,) * d)
    s = np.concatenate([*grids, vals.ravel()])
    lo = np.array([g[0] for g in grids])
    hi = np.array([g[-1] for g in grids])
    inside = rng.uniform(lo, hi, size=(300, d))
    ref = RegularGridInterpolator(grids, vals)(inside)
    np.testing.assert_allclose(oracle.knn_eval(s, d, R, inside), ref, rtol=1e-10, atol=1e-12)
    # nodes return their values
    np.testing.assert_allclose(oracle.knn_eval(s, d, R, mesh), vals.ravel(), rtol=1e-12,
                               atol=1e-14)
