"""roadmc_rovs_smooth_grad on the GPU: the smoothed-validity value of the
species-state valuation under a fixed policy and its gradient with respect to
the survival matrices, against the FP64 oracle (orc_rovs_smooth_grad, itself
pinned by central differences and by the hard-validity limit) on the same
streams.  The GPU runs FP32 products and an FP32 reverse sweep: the value
within 2e-5 relative, each gradient matrix within 2e-5 of its largest entry
(measured: 3e-6 and 2e-6)."""
import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd import rov_species as rs

pytestmark = pytest.mark.gpu

CASES = [
    # P, S, C, N, T
    (3, 1, 2, 5, 3),      # fewer paths than a tile, one column tile of padding
    (24, 1, 3, 600, 8),
    (60, 2, 2, 400, 6),
    (100, 1, 3, 300, 5),
]


def _problem(P, S, C, N, T, seed=3):
    return rs.synthetic_problem(seed=seed, n_patches=P, n_species=S, n_controls=C, n_paths=N,
                                n_years=T, threshold_ratio=0.97)


def _check(val, g, ref_val, ref_g):
    assert val == pytest.approx(ref_val, rel=2e-5)
    scale = np.abs(ref_g).reshape(ref_g.shape[0] * ref_g.shape[1], -1).max(1)
    err = np.abs(g - ref_g).reshape(scale.shape[0], -1).max(1)
    print(f"value {val:.6f} (oracle {ref_val:.6f}); gradient error / scale {err / scale}",
          flush=True)
    assert np.all(scale > 0)
    assert np.all(err <= 2e-5 * scale)


@pytest.mark.parametrize("P,S,C,N,T", CASES)
def test_smooth_gradient_matches_oracle(ctx, oracle, P, S, C, N, T):
    pr = _problem(P, S, C, N, T)
    pol = np.random.default_rng(P).integers(0, C, size=(T, N)).astype(np.int32)
    tau = 0.02 * pr.species[0].params[3]
    print(f"smooth gradient P={P} S={S} C={C} N={N} T={T}", flush=True)
    val, g = rs.smooth_value_gradient(ctx, pr, pol, tau)
    ref_val, ref_g = oracle.rovs_smooth_grad(pr, pol, tau)
    _check(val, g, ref_val, ref_g)
    # the value alone: the same number; a second call: the same bits (every
    # sum in a fixed order)
    v2, none = rs.smooth_value_gradient(ctx, pr, pol, tau, grad=False)
    assert none is None and v2 == val
    v3, g3 = rs.smooth_value_gradient(ctx, pr, pol, tau)
    assert v3 == val and np.array_equal(g3, g)


def test_smooth_gradient_with_an_unused_control(ctx, oracle):
    # Control 1 never drawn: its survival matrix gets an exactly zero
    # gradient (no rows to contract), the others match the oracle.
    pr = _problem(40, 1, 3, 500, 6)
    pol = 2 * np.random.default_rng(7).integers(0, 2, size=(6, 500)).astype(np.int32)
    tau = 0.02 * pr.species[0].params[3]
    print("smooth gradient with control 1 unused", flush=True)
    val, g = rs.smooth_value_gradient(ctx, pr, pol, tau)
    ref_val, ref_g = oracle.rovs_smooth_grad(pr, pol, tau)
    assert np.all(g[:, 1] == 0.0) and np.all(ref_g[:, 1] == 0.0)
    _check(val, g[:, [0, 2]], ref_val, ref_g[:, [0, 2]])


def test_smooth_gradient_under_the_fitted_policy(ctx, oracle):
    # The natural policy: ALGO4's decisions on its own forward states.
    pr = _problem(32, 1, 3, 2000, 10)
    pr.dim_res = 8
    print("smooth gradient under ALGO4's opt_control", flush=True)
    pol = rs.simulate_rov_species(ctx, pr, paths=True)["opt_control"]
    tau = 0.05 * pr.species[0].params[3]
    val, g = rs.smooth_value_gradient(ctx, pr, pol, tau)
    ref_val, ref_g = oracle.rovs_smooth_grad(pr, pol, tau)
    _check(val, g, ref_val, ref_g)


def test_smooth_gradient_against_its_central_differences(ctx):
    # Beyond the oracle's reach (P = 128, 4000 paths, 20 years): the GPU
    # gradient's largest entries against central differences of the GPU
    # value (h = 1e-2 on the survival entry; FP32 populations leave about
    # 1e-6 relative noise in the value).
    import copy
    pr = _problem(128, 1, 3, 4000, 20)
    pol = np.random.default_rng(9).integers(0, 3, size=(20, 4000)).astype(np.int32)
    tau = 0.05 * pr.species[0].params[3]
    print("smooth gradient P=128 N=4000 T=20 against central differences", flush=True)
    _, g = rs.smooth_value_gradient(ctx, pr, pol, tau)
    h = 1e-2
    for flat in np.argsort(np.abs(g).ravel())[-3:]:
        s_, c, k, j = np.unravel_index(flat, g.shape)
        p1, p2 = copy.deepcopy(pr), copy.deepcopy(pr)
        p1.species[s_].surv[c, k, j] += h
        p2.species[s_].surv[c, k, j] -= h
        v1, _ = rs.smooth_value_gradient(ctx, p1, pol, tau, grad=False)
        v2, _ = rs.smooth_value_gradient(ctx, p2, pol, tau, grad=False)
        fd = (v1 - v2) / (2 * h)
        print(f"  entry {(s_, c, k, j)}: gradient {g[s_, c, k, j]:.6f}, central difference {fd:.6f}",
              flush=True)
        assert g[s_, c, k, j] == pytest.approx(fd, rel=2e-2)


def test_config4_objective_has_a_rov_gradient(ctx):
    # Config 4's objective (mean end population + the smoothed species-state
    # value) on one landscape: the ROV term's gradient with respect to the
    # crossing mortality m (surv_c = 1 - m q_c / q_max) is non-zero and
    # matches central differences of the GPU value under m -> m +- h.
    import copy
    from transportoptimiser_amd import adjoint as adj
    from transportoptimiser_amd import landscape
    rng = np.random.default_rng(21)
    sd = landscape.make_species(rng, 64, mort_max=0.4)
    mte = landscape.RoadProblem([sd], n_years=20, n_paths=2000, seed=4)
    rp = rs.synthetic_problem(seed=6, n_patches=64, n_species=1, n_controls=3, n_paths=2000,
                              n_years=10)
    params = rp.species[0].params.copy()
    params[3] = 0.97 * sd.init_pop.sum()
    rp.species = [adj.rov_species_on(sd, rp.flow, params)]
    pol = np.random.default_rng(2).integers(0, 3, size=(10, 2000)).astype(np.int32)
    tau = 0.05 * params[3]
    print("config 4 objective with the smoothed ROV term", flush=True)
    J, gm, gr = adj.objective_gradient(ctx, mte, rp, pol, tau, rov_weight=0.5)
    J0, gm0, _ = adj.mte_gradient(ctx, mte, 10)
    g_rov = (gm - gm0) / 0.5
    v, _ = rs.smooth_value_gradient(ctx, rp, pol, tau, grad=False)
    assert J == pytest.approx(J0 + 0.5 * v, rel=1e-12)
    assert np.abs(g_rov).max() > 0
    h = 1e-3
    for flat in np.argsort(np.abs(g_rov).ravel())[-3:]:
        k, j = np.unravel_index(flat, g_rov.shape)
        vals = []
        for dm in (h, -h):
            sd2 = copy.deepcopy(sd)
            sd2.surv[k, j] -= dm  # m + dm
            rp2 = copy.deepcopy(rp)
            rp2.species = [adj.rov_species_on(sd2, rp.flow, params)]
            vals.append(rs.smooth_value_gradient(ctx, rp2, pol, tau, grad=False)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"  m[{k}, {j}]: dV/dm {g_rov[k, j]:.6f}, central difference {fd:.6f}", flush=True)
        assert g_rov[k, j] == pytest.approx(fd, rel=2e-2)


def test_smooth_gradient_refusals(ctx):
    pr = _problem(24, 1, 3, 100, 5)
    pol = np.zeros((5, 100), dtype=np.int32)
    for tau in (0.0, -1.0, float("nan")):
        with pytest.raises(nat.RoadmcError) as e:
            rs.smooth_value_gradient(ctx, pr, pol, tau)
        assert e.value.code == nat.ERR_INVALID
    bad = pol.copy()
    bad[2, 7] = 3
    with pytest.raises(nat.RoadmcError) as e:
        rs.smooth_value_gradient(ctx, pr, bad, 1.0)
    assert e.value.code == nat.ERR_INVALID
    pr.precision = nat.PREC_BF16X3
    with pytest.raises(nat.RoadmcError) as e:
        rs.smooth_value_gradient(ctx, pr, pol, 1.0)
    assert e.value.code == nat.ERR_UNSUPPORTED
