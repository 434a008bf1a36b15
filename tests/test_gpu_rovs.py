"""GPU species-state ROV (roadmc_rovs_run) against the oracle.

The forward pass runs in FP32 (or split-bf16) on MFMA on the GPU and in FP64
in the oracle:
totals and surviving fractions agree to 1e-4 relative, and the drawn
controls exactly (the same Philox stream).  The backward induction is FP64 on
both sides; it is compared on identical inputs (the oracle's backward runs on
the GPU's forward arrays): values to 1e-8 relative, decisions path by path.
"""
import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd import rov_species as rs
from tests.test_oracle_rovs import _flat_problem

pytestmark = pytest.mark.gpu


def _run(ctx, pr):
    print(f"rovs: P={pr.n_patches} S={len(pr.species)} C={len(pr.flow)} N={pr.n_paths} "
          f"T={pr.n_years} R={pr.dim_res}", flush=True)
    return rs.simulate_rov_species(ctx, pr, paths=True, grids=True)


PRECISIONS = [nat.PREC_FP32, nat.PREC_BF16X3]


@pytest.mark.parametrize("P,S,C,N,T,R", [
    (24, 1, 3, 2000, 12, 8),
    (40, 2, 2, 1500, 10, 6),
    (64, 1, 4, 3001, 15, 10),
])
def test_matches_oracle(ctx, oracle, P, S, C, N, T, R):
    # ALGO4 runs FP32 forward products (split-bf16 is refused below).
    pr = rs.synthetic_problem(seed=P + S, n_patches=P, n_species=S, n_controls=C, n_paths=N,
                              n_years=T, dim_res=R)
    g = _run(ctx, pr)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    np.testing.assert_array_equal(g["controls"], ctl)
    np.testing.assert_allclose(g["unit_profit"], up, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g["total_pops"], tp, rtol=1e-4)
    np.testing.assert_allclose(g["aars"], aa, rtol=1e-4)
    # backward on the GPU's own forward arrays
    V, oc, out = oracle.rovs_backward(pr, g["controls"], g["total_pops"], g["aars"],
                                      g["unit_profit"])
    scale = np.abs(V).max() + 1e-300
    np.testing.assert_allclose(g["cond_exp"], V, rtol=1e-8, atol=1e-10 * scale)
    assert (g["opt_control"] == oc).mean() > 0.9999
    assert g["value"] == pytest.approx(out["value"], rel=1e-8)
    assert g["first_control"] == out["control"]
    assert g["value_sd"] == pytest.approx(out["sd"], rel=1e-6, abs=1e-9)
    # the state the decisions used
    adj = g["total_pops"][:, :T] * g["aars"][:, :, -1]          # (S, T, N)
    np.testing.assert_allclose(g["adj_pops"].transpose(1, 0, 2), adj, rtol=1e-12)


def test_closed_form_with_constant_profit(ctx):
    T, u = 6, -0.5
    pr = _flat_problem(T=T, u=u, N=777)
    g = _run(ctx, pr)
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    Q = pr.flow[-1]
    assert g["value"] == pytest.approx(Q * u * sum(d ** i for i in range(T)), rel=1e-9)
    assert g["first_control"] == len(pr.flow) - 1
    for t in range(1, T):
        np.testing.assert_allclose(g["cond_exp"][t], Q * u * sum(d ** i for i in range(T - t)),
                                   rtol=1e-9)


def test_unreachable_thresholds_keep_zero_flow(ctx):
    g = _run(ctx, _flat_problem(thr=1e12))
    assert g["value"] == 0.0 and g["first_control"] == 0
    assert (g["opt_control"] == 0).all()


def test_algo4_refuses_split_bf16(ctx):
    pr = rs.synthetic_problem(seed=3, n_patches=24, n_paths=200, n_years=5, dim_res=4)
    pr.precision = nat.PREC_BF16X3
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_UNSUPPORTED


@pytest.mark.parametrize("prec", PRECISIONS)
def test_largest_landscape(ctx, oracle, prec):
    # P = 512 (16 column tiles, 8 waves per workgroup), N not a multiple of 32;
    # split-bf16 products through ALGO6, the method that takes them.
    pr = rs.synthetic_problem(seed=11, n_patches=512, n_controls=2, n_paths=333, n_years=6,
                              dim_res=6)
    pr.precision = prec
    if prec == nat.PREC_BF16X3:
        pr.method = 6
    g = _run(ctx, pr)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    np.testing.assert_allclose(g["total_pops"], tp, rtol=1e-4)
    np.testing.assert_allclose(g["aars"], aa, rtol=1e-4)


def test_invalid_arguments(ctx):
    # (More than 512 patches: test_wide_refuses_algo6_and_split_bf16.)
    from transportoptimiser_amd import native as nat
    pr = rs.synthetic_problem(n_patches=16, n_paths=100, n_years=2)
    with pytest.raises(nat.RoadmcError):
        rs.simulate_rov_species(ctx, pr)


# ---- ALGO6: the recomputation ------------------------------------------------
#
# The GPU recomputes each path in FP32 (the forward kernel's products and
# growth, states as n . r_c); the oracle in FP64.  Realized values are sums of
# Q_c x unit profit over the decisions taken, so they agree to rounding
# wherever the decisions agree; decisions can differ only where two payoffs or
# a state and a threshold are within FP32 rounding of each other.

@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("P,S,C,N,T,R,ratio,ck", [
    (24, 1, 3, 2000, 12, 8, 0.98, 0),
    (40, 2, 2, 1500, 10, 6, 0.9, 3),
    (64, 1, 4, 3001, 15, 10, 0.93, 4),
])
def test_algo6_matches_oracle(ctx, oracle, P, S, C, N, T, R, ratio, ck, prec):
    pr = rs.synthetic_problem(seed=P + S, n_patches=P, n_species=S, n_controls=C, n_paths=N,
                              n_years=T, dim_res=R, threshold_ratio=ratio)
    pr.method, pr.checkpoint_years, pr.precision = 6, ck, prec
    g = _run(ctx, pr)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    np.testing.assert_array_equal(g["controls"], ctl)
    np.testing.assert_allclose(g["total_pops"], tp, rtol=1e-4)
    V, oc, out = oracle.rovs_backward(pr, g["controls"], g["total_pops"], g["aars"],
                                      g["unit_profit"], method=6)
    assert len(np.unique(oc[1:T - 1])) > 1
    scale = np.abs(V).max()
    close = np.abs(g["cond_exp"] - V) <= 1e-9 * scale
    print(f"algo6: decisions equal {np.mean(g['opt_control'] == oc):.6f}, "
          f"values equal {close.mean():.6f}", flush=True)
    assert (g["opt_control"] == oc).mean() > 0.999
    assert close.mean() > 0.999
    assert g["value"] == pytest.approx(out["value"], rel=1e-4)
    assert g["first_control"] == out["control"]


def test_algo6_realized_recursion_without_mortality(ctx):
    from tests.test_oracle_rovs import _stochastic_problem
    T = 7
    pr = _stochastic_problem(T=T, N=1000)
    pr.method = 6
    g = _run(ctx, pr)
    V, oc, up = g["cond_exp"], g["opt_control"], g["unit_profit"]
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    Q = np.asarray(pr.flow)
    assert len(np.unique(oc[1:T - 1])) > 1
    for t in range(1, T - 1):
        np.testing.assert_allclose(V[t], Q[oc[t]] * up[t] + d * V[t + 1], rtol=1e-10, atol=1e-12)


def test_algo6_closed_forms(ctx):
    T, u = 6, -0.5
    pr = _flat_problem(T=T, u=u, N=777)
    pr.method = 6
    g = _run(ctx, pr)
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    Q = pr.flow[-1]
    assert g["value"] == pytest.approx(Q * u * sum(d ** i for i in range(T)), rel=1e-9)
    for t in range(1, T):
        np.testing.assert_allclose(g["cond_exp"][t], Q * u * sum(d ** i for i in range(T - t)),
                                   rtol=1e-9)
    pr = _flat_problem(thr=1e12)
    pr.method = 6
    g = _run(ctx, pr)
    assert g["value"] == 0.0 and (g["opt_control"] == 0).all()


@pytest.mark.parametrize("prec", PRECISIONS)
def test_algo6_checkpoint_interval_does_not_change_results(ctx, prec):
    # Replaying from a checkpoint reproduces the forward populations, so the
    # result is the same for every checkpoint interval (split-bf16 checkpoints
    # keep each entry's hi and lo halves).
    pr = rs.synthetic_problem(seed=5, n_patches=48, n_species=1, n_controls=3, n_paths=1200,
                              n_years=11, dim_res=6, threshold_ratio=0.98)
    pr.method, pr.precision = 6, prec
    ref = _run(ctx, pr)
    for ck in (2, 5, 11):
        pr.checkpoint_years = ck
        g = _run(ctx, pr)
        np.testing.assert_array_equal(g["cond_exp"], ref["cond_exp"])
        np.testing.assert_array_equal(g["opt_control"], ref["opt_control"])


def test_algo6_limits(ctx):
    from transportoptimiser_amd import native as nat
    # Three species of 512 patches need 3 x 66 KiB of tiles: more than the LDS.
    pr = rs.synthetic_problem(seed=3, n_patches=512, n_species=3, n_paths=64, n_years=4,
                              dim_res=3)
    pr.method = 6
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_UNSUPPORTED
    pr = rs.synthetic_problem(n_patches=16, n_paths=100, n_years=5)
    pr.method = 5
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_INVALID
    pr = rs.synthetic_problem(n_patches=16, n_paths=100, n_years=5)
    pr.precision = nat.PREC_FP64
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_UNSUPPORTED


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("P,S", [(512, 2), (256, 4), (33, 7)])
def test_algo6_realized_recursion_at_the_tile_limits(ctx, P, S, prec):
    # The largest species tiles one workgroup holds (2 x 512, 4 x 256 patches)
    # and the most species (7, at a ragged patch count): without mortality
    # the realized recursion holds path by path.
    from tests.test_oracle_rovs import _stochastic_problem
    T = 6
    pr = _stochastic_problem(T=T, N=333, P=P)
    pr.species = [pr.species[0]] * S
    pr.method, pr.precision = 6, prec
    pr.dim_res = 3
    g = _run(ctx, pr)
    V, oc, up = g["cond_exp"], g["opt_control"], g["unit_profit"]
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    Q = np.asarray(pr.flow)
    assert len(np.unique(oc[1:T - 1])) > 1
    for t in range(1, T - 1):
        np.testing.assert_allclose(V[t], Q[oc[t]] * up[t] + d * V[t + 1], rtol=1e-10, atol=1e-12)


# ---- beyond one workgroup's tile: the HBM-state forward pass (ALGO4, FP32) ----

@pytest.mark.parametrize("P,S,C,N,T,R", [
    (600, 1, 3, 500, 6, 6),
    (1100, 2, 2, 300, 5, 5),
])
def test_wide_forward_matches_oracle(ctx, oracle, P, S, C, N, T, R):
    pr = rs.synthetic_problem(seed=P + S, n_patches=P, n_species=S, n_controls=C, n_paths=N,
                              n_years=T, dim_res=R)
    g = _run(ctx, pr)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    np.testing.assert_array_equal(g["controls"], ctl)
    np.testing.assert_allclose(g["unit_profit"], up, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g["total_pops"], tp, rtol=1e-4)
    np.testing.assert_allclose(g["aars"], aa, rtol=1e-4)
    V, oc, out = oracle.rovs_backward(pr, g["controls"], g["total_pops"], g["aars"],
                                      g["unit_profit"])
    scale = np.abs(V).max() + 1e-300
    np.testing.assert_allclose(g["cond_exp"], V, rtol=1e-8, atol=1e-10 * scale)
    assert (g["opt_control"] == oc).mean() > 0.9999
    assert g["value"] == pytest.approx(out["value"], rel=1e-8)
    assert g["first_control"] == out["control"]


def test_wide_refuses_algo6_and_split_bf16(ctx):
    pr = rs.synthetic_problem(seed=4, n_patches=600, n_paths=100, n_years=4, dim_res=4)
    pr.method = 6
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_UNSUPPORTED
    pr.method, pr.precision = 6, nat.PREC_BF16X3
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_UNSUPPORTED
