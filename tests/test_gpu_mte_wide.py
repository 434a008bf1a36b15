"""MTE beyond one workgroup's LDS tile (P > 1024 patches) on the GPU.

roadmc_mte_run keeps the state of every path in HBM and runs one blocked
FP32 product per year above ROADMC_MAX_PATCHES.  Same bars as the tiled FP32
kernels against the FP64 oracle on the same Philox streams: per-path end
populations within 1e-4 relative, patch means, MTE only through paths at the
threshold."""
import numpy as np
import pytest

from tests.test_gpu_mte import _check_against_oracle, _problem
from transportoptimiser_amd import native as nat
from transportoptimiser_amd import road as rd
from transportoptimiser_amd.landscape import RoadProblem, SpeciesData, make_species

pytestmark = pytest.mark.gpu

CASES = [
    # P, N, T, S, law, endpoint, drop_tol
    (1500, 96, 4, 1, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
    (2048, 64, 3, 1, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0),
    (1100, 130, 5, 1, nat.LAW_REF_CPU, nat.END_PRE_GROWTH, 1e-5),
    (1030, 67, 6, 2, nat.LAW_REF_GPU, nat.END_POST_GROWTH, 0.0),
    (1200, 50, 4, 1, nat.LAW_LOGNORMAL, nat.END_PRE_GROWTH, 1e-4),
]


@pytest.mark.parametrize("P,N,T,S,law,endpoint,tol", CASES)
def test_wide_fp32_matches_oracle(ctx, oracle, P, N, T, S, law, endpoint, tol):
    pr = _problem(P, N, T, S, law, endpoint, drop_tol=tol)
    print(f"wide case P={P} N={N} T={T} S={S} law={law} end={endpoint}", flush=True)
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, 1e-4)


def test_wide_road_with_a_small_species(ctx, oracle):
    # One species above the tile and one below: the road runs wide.
    rng = np.random.default_rng(5)
    pr = RoadProblem([make_species(rng, 1300), make_species(rng, 200)], n_years=4, n_paths=80,
                     seed=4, road_id=3)
    print("wide road with species P=1300 and P=200", flush=True)
    out = ctx.simulate_mte(pr)
    ref = oracle.mte_problem(pr)
    _check_against_oracle(out, ref, pr, 1e-4)


def test_wide_species_from_its_road_matches_host_matrices(ctx):
    # The movement matrix built on the device from the road equals the one
    # packed on the host from the same builder's matrices: identical results.
    rng = np.random.default_rng(8)
    P = 1200
    inp = rd.random_road_species(rng, P, n_controls=2)
    srp = rd.SpeciesRoadPatches(ctx, inp)
    cap = rng.uniform(500, 5000, P)
    args = (0.8 * cap, cap, np.full(P, 0.05))
    dev = srp.device_species(1, *args, noise_sd=0.1, threshold_ratio=0.4)
    host = srp.to_species(1, *args, noise_sd=0.1, threshold_ratio=0.4)
    print(f"wide species from its road, P={P}", flush=True)
    a = ctx.simulate_mte(RoadProblem([host], n_years=3, n_paths=70, seed=2, drop_tol=1e-6))
    b = ctx.simulate_mte(RoadProblem([dev], n_years=3, n_paths=70, seed=2, drop_tol=1e-6))
    np.testing.assert_array_equal(a.end_pops, b.end_pops)
    np.testing.assert_array_equal(a.mte_prob, b.mte_prob)


def test_wide_refuses_other_precisions_and_batches(ctx):
    pr = _problem(1100, 16, 2, prec=nat.PREC_FP64)
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte(pr)
    assert e.value.code == nat.ERR_UNSUPPORTED
    pr = _problem(1100, 16, 2)
    with pytest.raises(nat.RoadmcError) as e:
        ctx.simulate_mte_batch([pr])
    assert e.value.code == nat.ERR_INVALID
