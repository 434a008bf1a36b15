"""GPU species-state ROV (roadmc_rovs_run) against the oracle.

The forward pass runs in FP32 on MFMA on the GPU and in FP64 in the oracle:
totals and surviving fractions agree to 1e-4 relative, and the drawn
controls exactly (the same Philox stream).  The backward induction is FP64 on
both sides; it is compared on identical inputs (the oracle's backward runs on
the GPU's forward arrays): values to 1e-8 relative, decisions path by path.
"""
import numpy as np
import pytest

from transportoptimiser_amd import rov_species as rs
from tests.test_oracle_rovs import _flat_problem

pytestmark = pytest.mark.gpu


def _run(ctx, pr):
    print(f"rovs: P={pr.n_patches} S={len(pr.species)} C={len(pr.flow)} N={pr.n_paths} "
          f"T={pr.n_years} R={pr.dim_res}", flush=True)
    return rs.simulate_rov_species(ctx, pr, paths=True, grids=True)


@pytest.mark.parametrize("P,S,C,N,T,R", [
    (24, 1, 3, 2000, 12, 8),
    (40, 2, 2, 1500, 10, 6),
    (64, 1, 4, 3001, 15, 10),
])
def test_matches_oracle(ctx, oracle, P, S, C, N, T, R):
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


def test_largest_landscape(ctx, oracle):
    # P = 512 (16 column tiles, 8 waves per workgroup), N not a multiple of 32.
    pr = rs.synthetic_problem(seed=11, n_patches=512, n_controls=2, n_paths=333, n_years=6,
                              dim_res=6)
    g = _run(ctx, pr)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    np.testing.assert_allclose(g["total_pops"], tp, rtol=1e-4)
    np.testing.assert_allclose(g["aars"], aa, rtol=1e-4)


def test_invalid_arguments(ctx):
    from transportoptimiser_amd import native as nat
    pr = rs.synthetic_problem(n_patches=600, n_paths=100, n_years=5)
    with pytest.raises(nat.RoadmcError) as e:
        rs.simulate_rov_species(ctx, pr)
    assert e.value.code == nat.ERR_INVALID
    pr = rs.synthetic_problem(n_patches=16, n_paths=100, n_years=2)
    with pytest.raises(nat.RoadmcError):
        rs.simulate_rov_species(ctx, pr)
