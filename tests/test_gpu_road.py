"""GPU road -> matrix builder against the oracle (roadmc_build_species_road).

Crossing counts are compared bit-exactly.  Both sides evaluate the same FP64
operations in the same order without contractions and differ only in the exp
implementation (device vs libm, each within about 1 ulp) and in the order of
the transition row sum (a fixed-order tree on the device, a sequential sum in
the oracle).  Survival therefore agrees to 1e-13 relative, and transitions to
1e-12: a sum of P positive terms in two orders differs by at most (P-1) eps,
1.1e-13 at P = 1000.
"""
import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd import road as rd

pytestmark = pytest.mark.gpu

RTOL = 1e-13
RTOL_TRANS = 1e-12


def _check(ctx, oracle, inp):
    print(f"road builder: P={inp.n_patches} points={len(inp.road)} "
          f"form={inp.transition_form} drop={inp.drop_last_link}", flush=True)
    m = rd.build_species_road(ctx, inp)
    cr, tr, sv = oracle.road_matrices(inp)
    np.testing.assert_array_equal(m.crossings, cr)
    np.testing.assert_allclose(m.trans, tr, rtol=RTOL_TRANS, atol=0)
    np.testing.assert_allclose(m.surv, sv, rtol=RTOL, atol=0)
    return m


@pytest.mark.parametrize("P,seed", [(2, 0), (7, 1), (64, 2), (300, 3), (1000, 4)])
@pytest.mark.parametrize("form", [nat.TRANS_REFERENCE, nat.TRANS_DECAY])
def test_matches_oracle(ctx, oracle, P, seed, form):
    inp = rd.random_road_species(np.random.default_rng(seed), P)
    inp.transition_form = form
    m = _check(ctx, oracle, inp)
    assert m.crossings.sum() > 0 or P < 10
    np.testing.assert_allclose(m.trans.sum(1), 1.0, rtol=1e-13)


def test_without_drop_and_without_types(ctx, oracle):
    inp = rd.random_road_species(np.random.default_rng(5), 500, n_road_points=300)
    inp.drop_last_link = False
    inp.road_type = None
    _check(ctx, oracle, inp)


def test_degenerate_roads_and_controls(ctx, oracle):
    rng = np.random.default_rng(6)
    inp = rd.random_road_species(rng, 50)
    inp.road = inp.road[:1]
    inp.road_type = np.zeros(0, dtype=np.int32)
    m = _check(ctx, oracle, inp)
    assert m.crossings.sum() == 0
    inp = rd.random_road_species(rng, 1)
    _check(ctx, oracle, inp)
    inp = rd.random_road_species(rng, 40, n_controls=0)
    m = rd.build_species_road(ctx, inp)
    assert m.surv.shape == (0, 40, 40)


def test_more_segments_than_one_lds_chunk(ctx, oracle):
    # 1300 segments span three LDS chunks of 512.
    inp = rd.random_road_species(np.random.default_rng(7), 400, n_road_points=1301)
    _check(ctx, oracle, inp)


def _boundary_input(xy, road):
    P = len(xy)
    return rd.RoadSpeciesInput(
        x=xy[:, 0].copy(), y=xy[:, 1].copy(), area=np.ones(P), pref_mean=np.zeros(P),
        pref_sd=np.zeros(P), road=road, lambda_mean=0.05, lambda_sd=0.0, lambda_mult=0.0,
        drop_last_link=False)


def test_boundary_geometry_matches_oracle(ctx, oracle):
    """Road vertices on centroids, on links and an ulp off centroids, so that
    segment parameters fall at and within ulps of 0 and 1: the device's
    division-free decision must reproduce the division test exactly."""
    rng = np.random.default_rng(13)
    P = 60
    xy = rng.uniform(0, 4, (P, 2))
    pts = []
    for _ in range(400):
        i, j = rng.integers(0, P, 2)
        kind = rng.integers(0, 4)
        if kind == 0:
            p = xy[i].copy()
        elif kind == 1:
            p = xy[i] + rng.choice([0.25, 0.5, 0.75]) * (xy[j] - xy[i])
        elif kind == 2:
            p = np.nextafter(xy[i], xy[i] + rng.choice([-1.0, 1.0], 2))
        else:
            p = rng.uniform(0, 4, 2)
        pts.append(p)
    m = _check(ctx, oracle, _boundary_input(xy, np.array(pts)))
    assert m.crossings.sum() > 0


def test_integer_grid_geometry_matches_oracle(ctx, oracle):
    """Centroids and road on an integer grid: parallel, collinear and
    touching segments (zero denominators and numerators) throughout."""
    rng = np.random.default_rng(14)
    xy = rng.integers(0, 6, (50, 2)).astype(np.float64)
    xy = np.unique(xy, axis=0)
    road = rng.integers(0, 6, (120, 2)).astype(np.float64)
    m = _check(ctx, oracle, _boundary_input(xy, road))
    assert m.crossings.sum() > 0


def test_partial_outputs(ctx, oracle):
    inp = rd.random_road_species(np.random.default_rng(8), 100)
    cr, tr, sv = oracle.road_matrices(inp)
    m = rd.build_species_road(ctx, inp, crossings=False, trans=False, surv=True)
    assert m.crossings is None and m.trans is None
    np.testing.assert_allclose(m.surv, sv, rtol=RTOL, atol=0)
    m = rd.build_species_road(ctx, inp, crossings=True, trans=False, surv=False)
    np.testing.assert_array_equal(m.crossings, cr)


def test_large_landscape_properties(ctx, oracle):
    # 4096 patches: symmetric non-negative counts, survival a function of the
    # counts, unit row sums, and the oracle's counts.
    inp = rd.random_road_species(np.random.default_rng(9), 4096, n_road_points=120)
    print("road builder: P=4096 properties", flush=True)
    m = rd.build_species_road(ctx, inp)
    cr = m.crossings
    assert (cr == cr.T).all() and (cr >= 0).all() and (np.diag(cr) == 0).all()
    np.testing.assert_allclose(m.trans.sum(1), 1.0, rtol=1e-12)
    c = 0
    k = (inp.vehicle_width + inp.animal_length) * inp.flows[c] / (inp.animal_speed * 3600)
    np.testing.assert_allclose(m.surv[c], np.exp(-cr * k), rtol=1e-13)
    ocr, _, _ = oracle.road_matrices(inp)
    np.testing.assert_array_equal(cr, ocr)


def test_invalid_arguments_raise(ctx):
    inp = rd.random_road_species(np.random.default_rng(10), 20)
    inp.animal_speed = 0.0
    with pytest.raises(nat.RoadmcError) as e:
        rd.build_species_road(ctx, inp)
    assert e.value.code == nat.ERR_INVALID
    inp = rd.random_road_species(np.random.default_rng(10), 20)
    inp.transition_form = 7
    with pytest.raises(nat.RoadmcError):
        rd.build_species_road(ctx, inp)


def test_road_to_mte_end_to_end(ctx, oracle):
    """Matrices from the builder drive the MTE path; the oracle's MTE on the
    oracle's matrices gives the same expected end population."""
    from transportoptimiser_amd.landscape import RoadProblem
    rng = np.random.default_rng(11)
    inp = rd.random_road_species(rng, 96)
    P = inp.n_patches
    cap = rng.uniform(500, 5000, P)
    srp = rd.SpeciesRoadPatches(ctx, inp)
    sd_gpu = srp.to_species(1, 0.8 * cap, cap, np.full(P, 0.05), noise_sd=0.1)
    _, tr, sv = oracle.road_matrices(inp)
    from transportoptimiser_amd.landscape import SpeciesData
    sd_orc = SpeciesData(trans=tr, surv=sv[1], init_pop=0.8 * cap, capacity=cap,
                         growth=np.full(P, 0.05), noise_sd=0.1)
    kw = dict(n_years=20, n_paths=2000, precision=nat.PREC_FP64, seed=5)
    out = ctx.simulate_mte(RoadProblem(species=[sd_gpu], **kw))
    ref = oracle.mte_problem(RoadProblem(species=[sd_orc], **kw), n_threads=8)
    np.testing.assert_allclose(out.end_pops, ref["end_pops"], rtol=1e-9, atol=0)


@pytest.mark.parametrize("precision", [nat.PREC_FP32, nat.PREC_FP64, nat.PREC_BF16X3])
def test_batch_builds_matrices_on_device(ctx, precision):
    """Species that carry their road (trans None) get trans o surv built and
    packed on the device inside the batch; the results equal, bit for bit,
    the batch run on the host copies of the same builder's matrices: the
    packed operand is the same FP64 product, narrowed the same way."""
    from transportoptimiser_amd.landscape import RoadProblem
    rng = np.random.default_rng(21)
    dev, host = [], []
    for r, P in enumerate([40, 129, 300, 17]):
        inp = rd.random_road_species(rng, P, n_controls=3)
        srp = rd.SpeciesRoadPatches(ctx, inp)
        cap = rng.uniform(500, 5000, P)
        control = [0, 2, -1, 1][r]
        kw = dict(noise_sd=0.1, threshold_ratio=0.4)
        args = (0.8 * cap, cap, np.full(P, 0.05))
        d = srp.device_species(control, *args, **kw)
        h = srp.to_species(control if control >= 0 else 2, *args, **kw)
        # Road 2 mixes a road species with a host-matrix one.
        other = None
        if r == 2:
            sp = rd.random_road_species(rng, 50, n_controls=2)
            o = rd.SpeciesRoadPatches(ctx, sp)
            c2 = rng.uniform(500, 5000, 50)
            other = o.to_species(1, 0.5 * c2, c2, np.full(50, 0.02), noise_sd=0.1)
        pk = dict(n_years=15, n_paths=1500, precision=precision, seed=7, road_id=r,
                  drop_tol=1e-6 if r % 2 else 0.0)
        dev.append(RoadProblem(species=[d] + ([other] if other else []), **pk))
        host.append(RoadProblem(species=[h] + ([other] if other else []), **pk))
    print(f"batch road->device pack precision={precision}", flush=True)
    a = ctx.simulate_mte_batch(dev, end_pops=True)
    b = ctx.simulate_mte_batch(host, end_pops=True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.end_pops, y.end_pops)
        np.testing.assert_array_equal(x.mte_prob, y.mte_prob)
        for pm, pn in zip(x.patch_mean, y.patch_mean):
            np.testing.assert_array_equal(pm, pn)
    # A prepared batch (the per-generation replay path) builds them once.
    with ctx.prepare_batch(dev) as bt:
        bt.launch()
        bt.sync()
        c = bt.fetch(end_pops=True)
    for x, y in zip(c, b):
        np.testing.assert_array_equal(x.end_pops, y.end_pops)


def test_device_road_species_rejected_where_unsupported(ctx):
    from transportoptimiser_amd.landscape import RoadProblem
    rng = np.random.default_rng(22)
    inp = rd.random_road_species(rng, 30, n_controls=2)
    srp = rd.SpeciesRoadPatches(ctx, inp)
    cap = rng.uniform(500, 5000, 30)
    d = srp.device_species(5, 0.8 * cap, cap, np.full(30, 0.05))
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch([RoadProblem(species=[d], n_years=5, n_paths=100)])
    assert e.value.code == nat.ERR_INVALID
    d.road_control = 0
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte(RoadProblem(species=[d], n_years=5, n_paths=100), plain=True)
    assert e.value.code == nat.ERR_UNSUPPORTED


def test_batch_road_species_edge_cases(ctx):
    """No flow controls (transitions only, no mortality), a road with no
    vertices, and a one-patch species, each against the host matrices of the
    same builder; a road whose patch count differs from the species' is an
    error."""
    from transportoptimiser_amd.landscape import RoadProblem
    rng = np.random.default_rng(23)
    cases = []
    a = rd.random_road_species(rng, 33, n_controls=0)
    b = rd.random_road_species(rng, 20)
    b.road = np.zeros((0, 2))
    b.road_type = None
    c = rd.random_road_species(rng, 1)
    for inp, control in ((a, -1), (b, 1), (c, 0)):
        srp = rd.SpeciesRoadPatches(ctx, inp)
        P = inp.n_patches
        cap = rng.uniform(500, 5000, P)
        args = (0.8 * cap, cap, np.full(P, 0.05))
        d = srp.device_species(control, *args, noise_sd=0.1)
        m = rd.build_species_road(ctx, inp)
        surv = None if len(inp.flows) == 0 else m.surv[control]
        from transportoptimiser_amd.landscape import SpeciesData
        h = SpeciesData(trans=m.trans, surv=surv, init_pop=args[0], capacity=args[1],
                        growth=args[2], noise_sd=0.1)
        cases.append((d, h))
    kw = dict(n_years=8, n_paths=700, seed=3)
    print("batch road species edge cases", flush=True)
    out_d = ctx.simulate_mte_batch([RoadProblem(species=[d], road_id=i, **kw)
                                    for i, (d, _) in enumerate(cases)], end_pops=True)
    out_h = ctx.simulate_mte_batch([RoadProblem(species=[h], road_id=i, **kw)
                                    for i, (_, h) in enumerate(cases)], end_pops=True)
    for x, y in zip(out_d, out_h):
        np.testing.assert_array_equal(x.end_pops, y.end_pops)
        np.testing.assert_array_equal(x.mte_prob, y.mte_prob)
    bad = cases[0][0]
    bad.init_pop, bad.capacity, bad.growth = bad.init_pop[:-1], bad.capacity[:-1], bad.growth[:-1]
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch([RoadProblem(species=[bad], **kw)])
    assert e.value.code == nat.ERR_INVALID
