"""MTE on the GPU against the CPU oracle (runs on an MI355X).

Bars: FP32 kernels give per-path end populations within 1e-4 relative of the
FP64 oracle on the same Philox streams (split-bf16 products, PREC_BF16X3,
too); FP64 kernels within 1e-10; the MTE
probability may differ only through paths that sit within that tolerance of
the threshold.  Batched and single-road launches are bit-identical.
"""
import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd.landscape import RoadProblem, SpeciesData, make_species

pytestmark = pytest.mark.gpu


def _problem(P, N, T, S=1, law=nat.LAW_LOGNORMAL, endpoint=nat.END_POST_GROWTH, seed=11,
             drop_tol=0.0, prec=nat.PREC_FP32, road_id=0):
    rng = np.random.default_rng(seed + 1000 * P + S)
    sps = [make_species(rng, P) for _ in range(S)]
    for sd in sps:
        sd.params = np.array([0.03, 0.04, 0.2, 0.5, 0.4, 0.3, 0.1, 0.5])
    return RoadProblem(sps, n_years=T, n_paths=N, growth_law=law, endpoint=endpoint,
                       precision=prec, drop_tol=drop_tol, seed=seed, road_id=road_id)


def _check_against_oracle(out, ref, pr, rtol):
    assert out.end_pops.shape == ref["end_pops"].shape
    np.testing.assert_allclose(out.end_pops, ref["end_pops"], rtol=rtol, atol=0)
    for s, sd in enumerate(pr.species):
        np.testing.assert_allclose(out.patch_mean[s], ref["patch_mean"][s], rtol=rtol,
                                   atol=1e-9 * sd.capacity.max())
        thr = sd.threshold_ratio * sd.init_pop.sum()
        near = np.abs(ref["end_pops"][:, s] - thr) <= 2 * rtol * thr
        assert abs(out.mte_prob[s] - ref["mte_prob"][s]) <= near.mean() + 1e-12
    np.testing.assert_allclose(out.end_mean, ref["end_mean"], rtol=rtol)
    np.testing.assert_allclose(out.end_sd, ref["end_sd"], rtol=50 * rtol, atol=1e-9)


CASES = [
    # P, N, T, S, law, endpoint, drop_tol
    (32, 300, 20, 1, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
    (1, 5, 7, 1, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
    (33, 257, 10, 2, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
    (100, 130, 12, 1, nat.LAW_REF_CPU, nat.END_POST_GROWTH, 0.0),
    (70, 97, 15, 2, nat.LAW_REF_GPU, nat.END_PRE_GROWTH, 1e-5),
    (200, 64, 8, 1, nat.LAW_REF_GPU, nat.END_POST_GROWTH, 0.0),
    (300, 40, 6, 1, nat.LAW_LOGNORMAL, nat.END_PRE_GROWTH, 1e-4),
    (640, 33, 4, 1, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
    (1024, 40, 3, 1, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
]


@pytest.mark.parametrize("P,N,T,S,law,endpoint,tol", CASES)
def test_tiled_fp32_matches_oracle(ctx, oracle, P, N, T, S, law, endpoint, tol):
    pr = _problem(P, N, T, S, law, endpoint, drop_tol=tol)
    print(f"case P={P} N={N} T={T} S={S} law={law} end={endpoint}", flush=True)
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, rtol=1e-4)


@pytest.mark.parametrize("P,N,T,S,law,endpoint,tol", CASES)
def test_tiled_fp64_matches_oracle(ctx, oracle, P, N, T, S, law, endpoint, tol):
    pr = _problem(P, N, T, S, law, endpoint, drop_tol=tol, prec=nat.PREC_FP64)
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, rtol=1e-10)


# Split-bf16 products: the state carries about 17 significant bits and a
# product drops only lo.m_lo; per-path end populations stay within 1e-4
# relative of the FP64 oracle, the FP32 kernels' bar.
B3_RTOL = 1e-4


@pytest.mark.parametrize("P,N,T,S,law,endpoint,tol", CASES)
def test_tiled_bf16x3_matches_oracle(ctx, oracle, P, N, T, S, law, endpoint, tol):
    pr = _problem(P, N, T, S, law, endpoint, drop_tol=tol, prec=nat.PREC_BF16X3)
    print(f"bf16x3 case P={P} N={N} T={T} S={S} law={law} end={endpoint}", flush=True)
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    err = np.max(np.abs(out.end_pops - ref["end_pops"]) / np.abs(ref["end_pops"]))
    print(f"  max rel err of end populations {err:.2e}", flush=True)
    _check_against_oracle(out, ref, pr, rtol=B3_RTOL)


def test_batch_bf16x3_generation(ctx):
    # Against the FP64 kernels on a generation's shapes, and batched launches
    # bit-identical to single-road runs.
    Ps = [64, 96, 128, 256, 512, 700, 1000, 1024]
    r3 = [_problem(P, 2000, 20, 2, seed=8, road_id=r, prec=nat.PREC_BF16X3)
          for r, P in enumerate(Ps)]
    r64 = [_problem(P, 2000, 20, 2, seed=8, road_id=r, prec=nat.PREC_FP64)
           for r, P in enumerate(Ps)]
    a = ctx.simulate_mte_batch(r3, end_pops=True)
    b = ctx.simulate_mte_batch(r64, end_pops=True)
    for x, y in zip(a, b):
        print(f"  dMTE {np.max(np.abs(x.mte_prob - y.mte_prob)):.1e} "
              f"rel {np.max(np.abs(x.end_pops - y.end_pops) / y.end_pops):.1e}", flush=True)
        assert np.max(np.abs(x.mte_prob - y.mte_prob)) <= 2e-3
        np.testing.assert_allclose(x.end_pops, y.end_pops, rtol=B3_RTOL)
    for r, x in zip(r3[:3], a[:3]):
        s = ctx.simulate_mte(r, end_pops=True)
        assert np.array_equal(s.end_pops, x.end_pops)


def test_bf16x3_is_for_the_tiled_kernels_only(ctx):
    from transportoptimiser_amd import adjoint
    pr = _problem(40, 100, 5, 1, prec=nat.PREC_BF16X3)
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte(pr, plain=True)
    assert e.value.code == nat.ERR_UNSUPPORTED
    with pytest.raises(nat.RoadmcError) as e:
        adjoint.Adjoint(ctx, pr, checkpoint_every=2)
    assert e.value.code == nat.ERR_UNSUPPORTED


@pytest.mark.parametrize("prec,rtol", [(nat.PREC_FP32, 1e-4), (nat.PREC_FP64, 1e-10)])
@pytest.mark.parametrize("law", [nat.LAW_LOGNORMAL, nat.LAW_REF_CPU, nat.LAW_REF_GPU])
def test_plain_reference_design_matches_oracle(ctx, oracle, prec, rtol, law):
    pr = _problem(45, 150, 10, 2, law, nat.END_POST_GROWTH, prec=prec, drop_tol=1e-5)
    out = ctx.simulate_mte(pr, plain=True)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, rtol=rtol)


def test_patch_noise_stream_matches_oracle(ctx, oracle):
    seed, road, sp = 0xDEADBEEF12, 9, 3
    z = ctx.patch_normals(seed, road, sp, 9, 4, 13)
    tag = oracle.tag(road, sp, 0)
    ref = np.array([[[oracle.patch_normal(seed, tag, p, y, j) for j in range(13)]
                     for y in range(4)] for p in range(9)])
    np.testing.assert_allclose(z, ref, atol=3e-5, rtol=0)


def test_batch_is_bit_identical_to_single_road_runs(ctx):
    roads = [_problem(P, 300, 6, 2, seed=5, road_id=r)
             for r, P in enumerate([64, 150, 300, 700, 40])]
    singles = [ctx.simulate_mte(r, end_pops=True) for r in roads]
    with ctx.prepare_batch(roads) as b:
        for rep in range(3):  # graph replays, with host work between them
            b.launch()
            b.sync()
            outs = b.fetch(end_pops=True)
            for o, s in zip(outs, singles):
                assert np.array_equal(o.end_pops, s.end_pops)
                assert np.array_equal(o.mte_prob, s.mte_prob)
                for a, c in zip(o.patch_mean, s.patch_mean):
                    assert np.array_equal(a, c)
        assert b.kernel_ms() > 0


def test_batch_fp64_generation_vs_fp32(ctx):
    roads32 = [_problem(P, 2000, 20, 2, seed=8, road_id=r)
               for r, P in enumerate([64, 96, 128, 256, 512, 1000])]
    roads64 = [_problem(P, 2000, 20, 2, seed=8, road_id=r, prec=nat.PREC_FP64)
               for r, P in enumerate([64, 96, 128, 256, 512, 1000])]
    a = ctx.simulate_mte_batch(roads32, end_pops=True)
    b = ctx.simulate_mte_batch(roads64, end_pops=True)
    for x, y in zip(a, b):
        assert np.max(np.abs(x.mte_prob - y.mte_prob)) <= 2e-3
        np.testing.assert_allclose(x.end_pops, y.end_pops, rtol=1e-4)


def test_config1_full_size(ctx, oracle):
    from transportoptimiser_amd.landscape import config1
    pr = config1()
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, rtol=1e-4)
    assert abs(out.mte_prob[0] - ref["mte_prob"][0]) <= 2e-3


def test_invalid_inputs_write_nothing(ctx):
    pr = _problem(16, 10, 3)
    pr.species[0].capacity = pr.species[0].capacity.copy()
    pr.species[0].capacity[3] = 0.0
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte(pr)
    assert e.value.code == nat.ERR_NUMERIC
    # Above the LDS tile only roadmc_mte_run takes a road (test_gpu_mte_wide.py);
    # the batched entry points keep ROADMC_MAX_PATCHES.
    big = _problem(8, 4, 2)
    big.species[0] = SpeciesData(trans=np.eye(1025), surv=None, init_pop=np.ones(1025),
                                 capacity=np.ones(1025), growth=None)
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch([big])
    assert e.value.code == nat.ERR_INVALID
    nan = _problem(16, 10, 3)
    nan.species[0].trans = nan.species[0].trans.copy()
    nan.species[0].trans[2, 2] = np.nan
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte(nan)
    assert e.value.code == nat.ERR_NUMERIC


def test_batch_validation_finds_non_finite_matrices_in_road_order(ctx):
    # The batch checks its host matrices over threads, then every road in
    # order: a non-finite entry in any road's or species' matrix is found, and
    # an earlier road's other error still comes first.
    def road(i, S=2):
        return _problem(24 + 8 * i, 10, 2, S=S, seed=40 + i, road_id=i)

    good = [road(i) for i in range(6)]
    print("batch validation: 6 roads, NaN in species 1's surv of road 4", flush=True)
    bad_surv = [road(i) for i in range(6)]
    sp = bad_surv[4].species[1]
    sp.surv = np.ones_like(sp.trans) if sp.surv is None else sp.surv.copy()
    sp.surv[1, 3] = np.inf
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch(bad_surv)
    assert e.value.code == nat.ERR_NUMERIC
    # an out-of-range patch count in road 1 before the NaN of road 3
    order = [road(i) for i in range(5)]
    order[1].species[0] = SpeciesData(trans=np.eye(1025), surv=None, init_pop=np.ones(1025),
                                      capacity=np.ones(1025), growth=None)
    order[3].species[0].trans = order[3].species[0].trans.copy()
    order[3].species[0].trans[0, 0] = np.nan
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch(order)
    assert e.value.code == nat.ERR_INVALID
    # the NaN road before it: the NaN first
    order[1], order[3] = order[3], order[1]
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch(order)
    assert e.value.code == nat.ERR_NUMERIC
    res = ctx.simulate_mte_batch(good)
    assert len(res) == 6 and all(np.isfinite(r.mte_prob).all() for r in res)


def test_injected_allocation_failure(ctx):
    from transportoptimiser_amd import api
    pr = _problem(20, 50, 3)
    ctx.fail_next_allocs(1)
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte(pr, plain=True)
    assert e.value.code == nat.ERR_OOM
    with api.Context(0) as fresh:
        fresh.fail_next_allocs(1)
        with pytest.raises(nat.RoadmcError) as e:
            fresh.simulate_mte(pr)
        assert e.value.code == nat.ERR_OOM
        # The context stays usable after the failure.
        ok = fresh.simulate_mte(pr)
        assert np.all(np.isfinite(ok.end_pops))


# ---- thresholds that give the MTE probability something to decide ----------

def _informative_thresholds(pr, ref, q=0.5):
    """Sets each species' threshold ratio to the q-quantile of the oracle's
    end totals over the initial total, so that the MTE probability sits
    near q instead of at 0 or 1."""
    for s, sd in enumerate(pr.species):
        sd.threshold_ratio = float(np.quantile(ref["end_pops"][:, s], q) / sd.init_pop.sum())
        ref["mte_prob"][s] = np.mean(ref["end_pops"][:, s] >=
                                     sd.threshold_ratio * sd.init_pop.sum())


@pytest.mark.parametrize("prec,rtol", [(nat.PREC_FP32, 1e-4), (nat.PREC_BF16X3, 1e-4),
                                       (nat.PREC_FP64, 1e-10)])
def test_batch_of_a_generation_matches_oracle(ctx, oracle, prec, rtol):
    # Eight roads spanning the generation's patch counts in one persistent
    # launch, against the oracle directly; thresholds at the median end total
    # of each road and species, so that every MTE probability is near 0.5.
    Ps = [64, 100, 160, 256, 400, 640, 800, 1024]
    roads = [_problem(P, 600, 10, 2, seed=31, road_id=r, prec=prec) for r, P in enumerate(Ps)]
    refs = [oracle.mte_problem(r) for r in roads]
    for r, ref in zip(roads, refs):
        _informative_thresholds(r, ref, 0.5)
    print(f"batch of {len(roads)} roads P={Ps} prec={prec}", flush=True)
    outs = ctx.simulate_mte_batch(roads, end_pops=True, patch_mean=True)
    for pr, out, ref in zip(roads, outs, refs):
        assert (0.2 < ref["mte_prob"]).all() and (ref["mte_prob"] < 0.8).all()
        _check_against_oracle(out, ref, pr, rtol=rtol)
        assert np.max(np.abs(out.mte_prob - ref["mte_prob"])) <= 2e-3


# ---- the large single road: 3 species x 256 patches, 100 years -------------

@pytest.mark.parametrize("prec", [nat.PREC_FP32, nat.PREC_BF16X3])
def test_config2_shape_matches_oracle(ctx, oracle, prec):
    # The per-species loop of simulateMTECUDA (SimulateGPU.cu:3680-3843) at the
    # large road's shape (l = 1.5 km, mortality U(0, 0.4), theta = 0.3) on
    # 1 500 paths; then again with thresholds at the median end total.
    from transportoptimiser_amd.landscape import config2
    pr = config2(n_paths=1500)
    pr.precision = prec
    print(f"config2 shape, 1500 paths, prec={prec}", flush=True)
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, rtol=1e-4)
    assert np.max(np.abs(out.mte_prob - ref["mte_prob"])) <= 2e-3
    _informative_thresholds(pr, ref, 0.5)
    out = ctx.simulate_mte(pr)
    assert (0.2 < ref["mte_prob"]).all() and (ref["mte_prob"] < 0.8).all()
    _check_against_oracle(out, ref, pr, rtol=1e-4)
    assert np.max(np.abs(out.mte_prob - ref["mte_prob"])) <= 2e-3


def test_config2_full_size(ctx):
    # 100 000 paths: FP32 and split-bf16 products against the FP64 kernels
    # (themselves oracle-checked above), at theta = 0.3 and at the median.
    from transportoptimiser_amd.landscape import config2
    pr = config2()
    res = {}
    for prec in (nat.PREC_FP64, nat.PREC_FP32, nat.PREC_BF16X3):
        pr.precision = prec
        print(f"config2 full size prec={prec}", flush=True)
        res[prec] = ctx.simulate_mte(pr, end_pops=True)
    ref = res[nat.PREC_FP64]
    assert np.isfinite(ref.end_pops).all()
    for prec in (nat.PREC_FP32, nat.PREC_BF16X3):
        o = res[prec]
        assert np.isfinite(o.end_pops).all() and ((0 <= o.mte_prob) & (o.mte_prob <= 1)).all()
        err = np.max(np.abs(o.end_pops - ref.end_pops) / ref.end_pops)
        print(f"  prec={prec}: max rel err {err:.2e}, MTE {o.mte_prob}", flush=True)
        assert err <= 1e-4
        assert np.max(np.abs(o.mte_prob - ref.mte_prob)) <= 2e-3
        for s, sd in enumerate(pr.species):
            thr = np.median(ref.end_pops[:, s])
            d = abs(np.mean(o.end_pops[:, s] >= thr) - np.mean(ref.end_pops[:, s] >= thr))
            assert d <= 2e-3
