"""The road-matrix oracle against hand geometry, an independent segment test,
closed forms and invariants (CPU).

The oracle restates SpeciesRoadPatches::roadCrossings,
computeTransitionProbabilities and computeSurvivalProbabilities
(include/road/SpeciesRoadPatches.cpp:296-451) with the CPU segment test of
Utility::lineSegmentIntersect (include/utilities/Utility.cpp:3-91).  Each test
is chosen so that a plausible slip fails it: a swapped segment end, a
half-open parameter interval, counting bridges, the wrong cut-off multiplier,
exp(-lambda d) in place of the reference's division by it, a transposed
matrix, a missing 3600.
"""
import math

import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd.landscape import crossed_links
from transportoptimiser_amd.road import RoadSpeciesInput, random_road_species


def _inp(xy, road, **kw):
    xy = np.asarray(xy, dtype=np.float64)
    P = len(xy)
    base = dict(x=xy[:, 0].copy(), y=xy[:, 1].copy(), area=np.ones(P), pref_mean=np.zeros(P),
                pref_sd=np.zeros(P), road=np.asarray(road, dtype=np.float64),
                lambda_mean=0.5, lambda_sd=0.0, lambda_mult=0.0, drop_last_link=False)
    base.update(kw)
    return RoadSpeciesInput(**base)


def _orient_count(xy, segs):
    """Proper-crossing count of every centroid link by orientation signs
    (an algorithm independent of the parametric test)."""
    P = len(xy)
    out = np.zeros((P, P), dtype=np.int64)

    def orient(a, b, c):
        return (b[..., 0] - a[..., 0]) * (c[..., 1] - a[..., 1]) - \
            (b[..., 1] - a[..., 1]) * (c[..., 0] - a[..., 0])

    A = np.broadcast_to(xy[:, None, :], (P, P, 2))
    B = np.broadcast_to(xy[None, :, :], (P, P, 2))
    for q1, q2 in segs:
        q1 = np.asarray(q1)
        q2 = np.asarray(q2)
        d1 = orient(q1, q2, A)
        d2 = orient(q1, q2, B)
        d3 = orient(A, B, q1)
        d4 = orient(A, B, q2)
        out += ((d1 * d2 < 0) & (d3 * d4 < 0)).astype(np.int64)
    np.fill_diagonal(out, 0)
    return out


def test_single_crossing_and_dropped_last_link(oracle):
    inp = _inp([[0, 0], [2, 0]], [[1, -1], [1, 1]])
    cr, _, _ = oracle.road_matrices(inp)
    assert cr.tolist() == [[0, 1], [1, 0]]
    # With one in-range pair, the reference's validCrossings is 0: no link is
    # passed to the segment test and the count stays zero.
    inp.drop_last_link = True
    cr, _, _ = oracle.road_matrices(inp)
    assert cr.tolist() == [[0, 0], [0, 0]]


def test_zigzag_counts_every_segment(oracle):
    road = [[0.2, -1], [0.6, 1], [1.0, -1], [1.4, 1], [1.8, -1]]
    inp = _inp([[0, 0], [2, 0], [0, 5]], road)
    cr, _, _ = oracle.road_matrices(inp)
    assert cr[0, 1] == 4 and cr[1, 0] == 4
    assert cr[0, 2] == 0


def test_bridges_and_tunnels_not_counted(oracle):
    road = [[0.2, -1], [0.6, 1], [1.0, -1], [1.4, 1], [1.8, -1]]
    rt = np.array([nat.SEG_ROAD, nat.SEG_BRIDGE, nat.SEG_TUNNEL, nat.SEG_ROAD], dtype=np.int32)
    cr, _, _ = oracle.road_matrices(_inp([[0, 0], [2, 0]], road, road_type=rt))
    assert cr[0, 1] == 2


def test_touching_counts_parallel_and_collinear_do_not(oracle):
    # Closed parameter interval [0, 1]: a road ending on the link counts.
    cr, _, _ = oracle.road_matrices(_inp([[0, 0], [2, 0]], [[1, 0], [1, 1]]))
    assert cr[0, 1] == 1
    # Road through a centroid: u_b = 0 on the link's end, counted.
    cr, _, _ = oracle.road_matrices(_inp([[0, 0], [2, 0]], [[0, -1], [0, 1]]))
    assert cr[0, 1] == 1
    # Parallel (denominator 0, numerator non-zero) and collinear (0 / 0).
    cr, _, _ = oracle.road_matrices(_inp([[0, 0], [2, 0]], [[0, 1], [2, 1]]))
    assert cr[0, 1] == 0
    cr, _, _ = oracle.road_matrices(_inp([[0, 0], [2, 0]], [[0.5, 0], [1.5, 0]]))
    assert cr[0, 1] == 0
    # A segment just short of the link.
    cr, _, _ = oracle.road_matrices(_inp([[0, 0], [2, 0]], [[1, 0.001], [1, 1]]))
    assert cr[0, 1] == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_counts_match_orientation_test(oracle, seed):
    rng = np.random.default_rng(seed)
    inp = random_road_species(rng, 150, n_road_points=25, bridge_fraction=0.0)
    inp.lambda_mean, inp.lambda_sd = 0.01, 0.0  # cut-off 300 km: every pair in range
    inp.drop_last_link = False
    cr, _, _ = oracle.road_matrices(inp)
    xy = np.stack([inp.x, inp.y], 1)
    segs = list(zip(inp.road[:-1], inp.road[1:]))
    ref = _orient_count(xy, segs)
    np.testing.assert_array_equal(cr, ref)
    np.testing.assert_array_equal(cr > 0, crossed_links(xy, inp.road))
    np.testing.assert_array_equal(cr, cr.T)


def test_drop_last_link_is_the_last_in_range_pair(oracle):
    rng = np.random.default_rng(4)
    inp = random_road_species(rng, 120, bridge_fraction=0.0)
    inp.drop_last_link = False
    full, _, _ = oracle.road_matrices(inp)
    inp.drop_last_link = True
    dropped, _, _ = oracle.road_matrices(inp)
    xy = np.stack([inp.x, inp.y], 1)
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    cut = math.log(0.05) / -(inp.lambda_mean + inp.lambda_sd * inp.lambda_mult)
    i, j = np.nonzero(np.triu(d <= cut, 1))
    last = np.argmax(i * 120 + j)
    li, lj = i[last], j[last]
    expect = full.copy()
    expect[li, lj] = expect[lj, li] = 0
    np.testing.assert_array_equal(dropped, expect)


def test_crossing_and_transition_cutoffs_use_different_multipliers(oracle):
    # lambda_mean 1, sd 0.5, mult 0: transition lambda = 1.  Crossing cut-off
    # ln 20 / 1 = 3.0; transition cut-off ln 20 / 1.5 = 2.0.  A pair 2.5 apart
    # has a crossing count but no transition weight.
    inp = _inp([[0, 0], [2.5, 0]], [[1, -1], [1, 1]], lambda_mean=1.0, lambda_sd=0.5,
               lambda_mult=0.0)
    cr, tr, _ = oracle.road_matrices(inp)
    assert cr[0, 1] == 1
    assert tr[0, 1] == 0.0 and tr[0, 0] == 1.0
    # mult = -1: lambda 0.5; crossing cut-off ln 20 / 0.5 = 6.0, transition
    # cut-off ln 20 / 0.75 = 4.0.  At 5 apart, again count without weight.
    inp = _inp([[0, 0], [5, 0]], [[1, -1], [1, 1]], lambda_mean=1.0, lambda_sd=0.5,
               lambda_mult=-1.0)
    cr, tr, _ = oracle.road_matrices(inp)
    assert cr[0, 1] == 1 and tr[0, 1] == 0.0


def test_transition_rows_and_forms(oracle):
    # Patches on a line at 0, 1, 2 (all within the cut-offs), equal weights.
    xy = [[0, 0], [1, 0], [2, 0]]
    lam = 0.7
    ref = _inp(xy, [[5, 5], [6, 6]], lambda_mean=lam, transition_form=nat.TRANS_REFERENCE)
    dec = _inp(xy, [[5, 5], [6, 6]], lambda_mean=lam, transition_form=nat.TRANS_DECAY)
    _, tr, _ = oracle.road_matrices(ref)
    _, td, _ = oracle.road_matrices(dec)
    np.testing.assert_allclose(tr.sum(1), 1.0, rtol=1e-15)
    np.testing.assert_allclose(td.sum(1), 1.0, rtol=1e-15)
    # Reference form: weight / exp(-lambda d) grows with distance.
    assert tr[0, 2] / tr[0, 1] == pytest.approx(math.exp(lam), rel=1e-14)
    assert tr[0, 1] / tr[0, 0] == pytest.approx(math.exp(lam), rel=1e-14)
    # Decay form: weight x exp(-lambda d).
    assert td[0, 2] / td[0, 1] == pytest.approx(math.exp(-lam), rel=1e-14)
    # Row = origin: the middle patch's row is symmetric, the end rows are not.
    assert tr[1, 0] == pytest.approx(tr[1, 2], rel=1e-15)
    assert tr[0, 2] != pytest.approx(tr[2, 0] * 0.5)


def test_transition_weights_follow_area_and_preference(oracle):
    # Two destinations equidistant from patch 0.
    xy = [[0, 0], [1, 0], [-1, 0]]
    pm = np.array([0.0, 0.3, -0.2])
    ps = np.array([0.0, 0.4, 0.1])
    area = np.array([1.0, 2.0, 0.5])
    h = 0.8
    inp = _inp(xy, [[5, 5], [6, 6]], area=area, pref_mean=pm, pref_sd=ps, pref_mult=h,
               lambda_mean=0.3)
    _, tr, _ = oracle.road_matrices(inp)
    expect = (area[1] / area[2]) * math.exp((pm[1] - pm[2]) + (ps[1] - ps[2]) * h)
    assert tr[0, 1] / tr[0, 2] == pytest.approx(expect, rel=1e-14)
    # Scaling every area leaves the normalised rows unchanged.
    inp.area = area * 7.0
    _, tr7, _ = oracle.road_matrices(inp)
    np.testing.assert_allclose(tr7, tr, rtol=1e-14)


def test_survival_closed_form(oracle):
    road = [[0.2, -1], [0.6, 1], [1.0, -1]]
    inp = _inp([[0, 0], [2, 0], [0, 3]], road, flows=np.array([0.0, 360.0, 720.0]),
               vehicle_width=6.0, animal_length=4.0, animal_speed=1.0)
    cr, _, sv = oracle.road_matrices(inp)
    assert cr[0, 1] == 2
    assert sv.shape == (3, 3, 3)
    np.testing.assert_array_equal(sv[0], 1.0)                 # no traffic
    assert sv[1, 0, 0] == 1.0                                  # no crossing
    # 2 crossings x 360 veh/h x 10 m / (1 x 3600) = 2.
    assert sv[1, 0, 1] == pytest.approx(math.exp(-2.0), rel=1e-15)
    assert sv[2, 0, 1] == pytest.approx(math.exp(-4.0), rel=1e-15)
    np.testing.assert_array_equal(sv[1], sv[1].T)


def test_single_patch(oracle):
    inp = _inp([[0, 0]], [[1, -1], [1, 1]])
    cr, tr, sv = oracle.road_matrices(inp)
    assert cr.tolist() == [[0]] and tr.tolist() == [[1.0]] and sv[0].tolist() == [[1.0]]


def _ratio_class(n, de):
    """numpy statement of the device's division-free verdict on RN(n / de)
    against [0, 1] (road_kernels.hip, seg_hit): 0 outside, 1 inside,
    2 undecided (the device divides)."""
    ad = np.abs(de)
    s = np.where(de < 0, -n, n)
    hi = ad * (1.0 + 2.0 ** -51)
    ok = (ad >= 2.0 ** -900) & (ad <= 2.0 ** 100)
    inside = (s >= 0) & (s <= ad)
    near = ((s < 0) & (s > -2.0 ** -900)) | ((s > ad) & (s <= hi))
    cls = np.where(inside, 1, np.where(near, 2, 0)).astype(np.int8)
    cls[~ok] = 2
    return cls


def test_division_free_classification_agrees_with_division():
    """Every decided class equals the division's verdict, on quotients at and
    within a few ulps of 0 and 1, subnormal and huge operands, both signs."""
    rng = np.random.default_rng(12)
    m = 400000
    de = rng.uniform(0.5, 2.0, m) * 2.0 ** rng.integers(-1074, 1000, m).astype(np.float64)
    de[: m // 8] = rng.uniform(-3, 3, m // 8)
    de = np.where(rng.uniform(size=m) < 0.5, -de, de)
    de = de[(de != 0) & np.isfinite(de)]
    m = len(de)
    kind = rng.integers(0, 5, m)
    steps = rng.integers(-4, 5, m)
    n = np.empty(m)
    one = de.copy()                                        # q = 1
    n[kind == 0] = one[kind == 0]
    n[kind == 1] = (de * rng.uniform(-3, 3, m))[kind == 1]  # generic
    n[kind == 2] = (de * 2.0 ** -rng.integers(1000, 1100, m))[kind == 2]  # q near 0
    n[kind == 3] = -(de * 2.0 ** -rng.integers(1000, 1100, m))[kind == 3]
    n[kind == 4] = 0.0
    # walk some numerators a few ulps either way
    for _ in range(4):
        up = steps > 0
        dn = steps < 0
        n[up] = np.nextafter(n[up], np.inf)
        n[dn] = np.nextafter(n[dn], -np.inf)
        steps = steps - np.sign(steps)
    with np.errstate(all="ignore"):
        q = n / de
    truth = ((q >= 0) & (q <= 1)).astype(np.int8)
    cls = _ratio_class(n, de)
    decided = cls != 2
    np.testing.assert_array_equal(cls[decided], truth[decided])
    # Generic quotients with moderate operands are all but always decided.
    moderate = (kind == 1) & (np.abs(de) > 2.0 ** -800) & (np.abs(de) < 2.0 ** 90)
    assert moderate.sum() > 1000 and decided[moderate].mean() > 0.999
