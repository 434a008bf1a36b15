"""GPU grid surrogate against the oracle (roadmc_knn_surrogate_fit / _eval).

Both sides rank neighbours by the same FP64 squared distances (no
contraction, ties to the lower index) and sum them in that order, so the
neighbour sets agree exactly; the fitted values then differ only through exp
(device vs libm, about an ulp) and agree to 1e-9 relative.  Evaluation agrees
to 1e-12.
"""
import numpy as np
import pytest

from transportoptimiser_amd import knn
from transportoptimiser_amd import native as nat

pytestmark = pytest.mark.gpu


def _data(rng, n, d, integer=False):
    if integer:
        x = rng.integers(0, 5, size=(n, d)).astype(np.float64)
    else:
        x = rng.uniform(-1.0, 2.0, size=(n, d)) * (1 + np.arange(d))
    y = np.sin(2 * x[:, 0]) + x[:, -1] ** 2 - 0.3 * x.sum(1) + 0.05 * rng.normal(size=n)
    return x, y


@pytest.mark.parametrize("d,n,R,k,flags", [
    (1, 60, 20, 7, 0),
    (1, 4096, 50, 17, 0),
    (2, 500, 15, 25, nat.KNN_NORMALIZE | nat.KNN_NONPOSITIVE),
    (3, 2000, 12, 65, nat.KNN_NORMALIZE),
    (3, 300, 10, 128, 0),
    (4, 800, 6, 20, nat.KNN_NORMALIZE),
    (8, 400, 3, 30, nat.KNN_NORMALIZE),
    # more than 2048 samples: the radix-select kernel
    (2, 20000, 10, 114, nat.KNN_NORMALIZE),
    (3, 9000, 6, 128, nat.KNN_NORMALIZE | nat.KNN_NONPOSITIVE),
])
def test_fit_matches_oracle(ctx, oracle, d, n, R, k, flags):
    print(f"knn fit: d={d} n={n} R={R} k={k} flags={flags}", flush=True)
    rng = np.random.default_rng(d * 1000 + n)
    x, y = _data(rng, n, d)
    g = knn.fit(ctx, x, y, R, k, normalize=bool(flags & nat.KNN_NORMALIZE),
                nonpositive=bool(flags & nat.KNN_NONPOSITIVE))
    ref = oracle.knn_fit(x, y, R, k, flags)
    np.testing.assert_array_equal(g.surrogate[:d * R], ref[:d * R])
    np.testing.assert_allclose(g.surrogate[d * R:], ref[d * R:], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("n", [300, 5000])
def test_fit_with_tied_distances_matches_oracle(ctx, oracle, n):
    # Integer sample points: many equal distances, broken by sample index
    # (300: lane lists and merge; 5000: radix select, equal keys in index order).
    rng = np.random.default_rng(5)
    x, y = _data(rng, n, 2, integer=True)
    g = knn.fit(ctx, x, y, 9, 12)
    np.testing.assert_allclose(g.surrogate, oracle.knn_fit(x, y, 9, 12, 0), rtol=1e-9,
                               atol=1e-12)


@pytest.mark.parametrize("offset", [0.0, 1e-3])
def test_fit_with_crowded_distances_matches_oracle(ctx, oracle, offset):
    # Coincident samples: two clusters of 2500 (offset 0), so that more
    # samples share the k-th distance than the select kernel sorts in LDS and
    # the equal keys are taken in index order over all samples; or four of
    # 1250 (every other sample moved by the offset), so that 1250 equal keys
    # are sorted in LDS.  The neighbours are the lowest indices of the nearest
    # cluster.
    rng = np.random.default_rng(8)
    x = np.repeat(np.array([[0.0, 0.0], [1.0, 1.0]]), 2500, axis=0)
    x += np.array([0.0, offset]) * (np.arange(5000) % 2)[:, None]
    y = rng.normal(size=5000)
    g = knn.fit(ctx, x, y, 5, 12)
    np.testing.assert_allclose(g.surrogate, oracle.knn_fit(x, y, 5, 12, 0), rtol=1e-9,
                               atol=1e-12)


def test_fit_when_the_subsample_bound_gathers_too_few(ctx, oracle):
    # Every 4th sample (the select kernel's subsample) sits in a tight cluster
    # at a grid point; the others are spread over the square.  The bound from
    # the subsample then holds fewer than k samples, and the full radix select
    # must give the neighbours instead.
    rng = np.random.default_rng(21)
    n, R, k = 6000, 6, 60
    x = rng.uniform(0.0, 1.0, size=(n, 2))
    x[:4] = [[0.0, 0.0], [1.0, 1.0], [0.0, 1.0], [1.0, 0.0]]  # the grid spans [0, 1]^2
    members = np.arange(4, n, 4)
    g1 = np.linspace(0.0, 1.0, R)
    cell = rng.integers(0, R * R, size=members.size)
    centres = np.stack([g1[cell // R], g1[cell % R]], axis=1)
    x[members] = np.clip(centres + rng.uniform(-1e-7, 1e-7, size=centres.shape), 0.0, 1.0)
    y = rng.normal(size=n)
    print(f"knn fit: clustered subsample, n={n} R={R} k={k}", flush=True)
    g = knn.fit(ctx, x, y, R, k)
    np.testing.assert_allclose(g.surrogate, oracle.knn_fit(x, y, R, k, 0), rtol=1e-9,
                               atol=1e-12)


def test_degenerate_inputs(ctx, oracle):
    rng = np.random.default_rng(6)
    # fewer than 3 samples
    g = knn.fit(ctx, [[0.0], [1.0]], [1.0, 2.0], 4, 5)
    np.testing.assert_array_equal(g.surrogate, oracle.knn_fit([[0.0], [1.0]], [1.0, 2.0], 4, 5))
    # a constant dimension, normalised and not
    x = np.column_stack([rng.uniform(size=100), np.full(100, 3.0)])
    y = x[:, 0] ** 2
    for flags in (0, nat.KNN_NORMALIZE):
        g = knn.fit(ctx, x, y, 6, 10, normalize=bool(flags))
        np.testing.assert_allclose(g.surrogate, oracle.knn_fit(x, y, 6, 10, flags), rtol=1e-9,
                                   atol=1e-12)
    # all samples at one point: weighted mean of the k lowest indices
    g = knn.fit(ctx, np.zeros((10, 1)), np.arange(10.0), 3, 4)
    np.testing.assert_allclose(g.values, 1.5, rtol=1e-12)


@pytest.mark.parametrize("d", [1, 2, 3, 5])
def test_eval_matches_oracle(ctx, oracle, d):
    rng = np.random.default_rng(40 + d)
    x, y = _data(rng, 400, d)
    g = knn.fit(ctx, x, y, 6, 15)
    pts = rng.uniform(-2.0, 3.0, size=(5000, d)) * (1 + np.arange(d))
    out = g.predict(ctx, pts)
    np.testing.assert_allclose(out, oracle.knn_eval(g.surrogate, d, 6, pts), rtol=1e-12,
                               atol=1e-13)


def test_neighbour_rules_match_oracle(oracle):
    # The library caps the reference's rule at ROADMC_KNN_MAX_K (128); the ROV
    # rule passes it above 5 650 samples (5 651 -> 25 + ceil(103.02) = 129).
    for n in list(range(0, 700, 7)) + [4096, 5650, 5651, 10000]:
        for rule in (nat.KNN_RULE_MTE, nat.KNN_RULE_ROV):
            assert knn.neighbour_count(n, rule) == min(oracle.knn_k(n, rule), nat.KNN_MAX_K)
    assert oracle.knn_k(5651, nat.KNN_RULE_ROV) == 129
    assert knn.neighbour_count(5651, nat.KNN_RULE_ROV) == 128


def test_rov_surrogate_beyond_the_neighbour_cap(ctx, oracle):
    # 6 000 samples: the reference's ROV rule asks for 135 neighbours, the
    # library fits with 128 (the radix-select kernel, n > 2 048).
    rng = np.random.default_rng(81)
    n = 6000
    iars = rng.uniform(0.0, 0.3, size=(n, 2))
    cost = rng.uniform(0, 0.4, n)
    vals = -1000 * cost * (1 + iars.sum(1)) + 5 * rng.normal(size=n)
    assert oracle.knn_k(n, nat.KNN_RULE_ROV) == 135
    mean, _ = knn.build_surrogate_rov(ctx, iars, cost, vals, np.abs(vals) * 0.1, dim_res=8)
    x = np.column_stack([iars, cost])
    ref = oracle.knn_fit(x, vals, 8, nat.KNN_MAX_K, nat.KNN_NORMALIZE | nat.KNN_NONPOSITIVE)
    np.testing.assert_allclose(mean.surrogate, ref, rtol=1e-9, atol=1e-12)


def test_invalid_arguments(ctx):
    x = np.random.default_rng(7).uniform(size=(50, 2))
    y = x[:, 0]
    for kw in (dict(dim_res=1, k=5), dict(dim_res=5, k=0), dict(dim_res=5, k=129)):
        with pytest.raises(nat.RoadmcError) as e:
            knn.fit(ctx, x, y, **kw)
        assert e.value.code == nat.ERR_INVALID
    with pytest.raises(nat.RoadmcError):
        knn.fit(ctx, np.zeros((20, 9)), np.zeros(20), 2, 5)
    bad = x.copy()
    bad[3, 1] = np.nan
    with pytest.raises(nat.RoadmcError) as e:
        knn.fit(ctx, bad, y, 5, 5)
    assert e.value.code == nat.ERR_NUMERIC


def test_reference_surrogate_builders(ctx, oracle):
    """build_surrogate_mte / build_surrogate_rov with the reference's
    neighbour rules and flags, against the oracle."""
    rng = np.random.default_rng(8)
    n, S = 700, 2
    iars = rng.uniform(0.0, 0.3, size=(n, S))
    pops = 1000 * (1 - iars) + rng.normal(size=(n, S))
    sds = 50 + 100 * iars
    cost = rng.uniform(0, 0.4, n)
    vals = -1000 * cost * (1 + iars.sum(1)) + 5 * rng.normal(size=n)
    grids = knn.build_surrogate_mte(ctx, iars, pops, sds, dim_res=20)
    k = oracle.knn_k(n, nat.KNN_RULE_MTE)
    for s in range(S):
        np.testing.assert_allclose(grids[s][0].surrogate,
                                   oracle.knn_fit(iars[:, s], pops[:, s], 20, k, 0), rtol=1e-9)
        np.testing.assert_allclose(grids[s][1].surrogate,
                                   oracle.knn_fit(iars[:, s], sds[:, s], 20, k, 0), rtol=1e-9)
    mean, sd = knn.build_surrogate_rov(ctx, iars, cost, vals, np.abs(vals) * 0.1, dim_res=10)
    x = np.column_stack([iars, cost])
    k = oracle.knn_k(n, nat.KNN_RULE_ROV)
    ref = oracle.knn_fit(x, vals, 10, k, nat.KNN_NORMALIZE | nat.KNN_NONPOSITIVE)
    np.testing.assert_allclose(mean.surrogate, ref, rtol=1e-9, atol=1e-12)
    assert (mean.values <= 0).all()
