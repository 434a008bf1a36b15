"""ROV on the GPU against the CPU oracle (runs on an MI355X).

The forward price paths are FP64 on both sides and must agree to 1e-10; the
backward induction is FP64 regression on both sides, so values agree to
1e-8 relative.  The Longstaff-Schwartz example pins the regression machinery
end to end.
"""
import json
import pathlib
import time

import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd.rov import RovProblem, Uncertainty, config5, exp_pv, simulate_rov

pytestmark = pytest.mark.gpu
GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def test_forward_paths_match_oracle(ctx, oracle):
    pr = config5(n_paths=3000, n_steps=12)
    pr.uncertainties.append(Uncertainty(nat.SDE_MRJD, 1.0, 1.5, 0.05, 0.3, 0.4, 0.2))
    corr = np.eye(4)
    corr[:3, :3] = pr.correlation
    pr.correlation = corr
    pr.profit_weight = [1.0, -1.5, 0.0, 0.1]
    pr.basis_degree = 2
    r = simulate_rov(ctx, pr, want_paths=True)
    st, _ = oracle.rov_forward(pr)
    np.testing.assert_allclose(r["states"], st, rtol=1e-10, atol=1e-12)


def test_longstaff_schwartz_example_on_gpu(ctx):
    d = json.loads((GOLD / "longstaff_schwartz_2001.json").read_text())
    paths = np.array(d["paths_t0_t1_t2_t3"])
    pr = RovProblem([Uncertainty(nat.SDE_GBM, 1.0, 0.0, 0.0)], n_steps=3, n_paths=8,
                    basis_degree=2, standardize=0, itm_only=1, mode=nat.ROV_AMERICAN,
                    value_mode=nat.VALUE_REALIZED, strike=d["strike"],
                    discount=d["discount_per_period"], flow=[0.0, 0.0],
                    states=paths.T[:, None, :].copy(), controls=np.zeros((3, 8), np.int32))
    r = simulate_rov(ctx, pr, want_paths=True)
    assert r["value"] == pytest.approx(d["value"], abs=6e-5)
    disc = d["discount_per_period"]
    np.testing.assert_allclose(r["coeffs"][2, 0] * disc, d["regression_t2"], atol=2e-3)
    np.testing.assert_allclose(r["coeffs"][1, 0] * disc, d["regression_t1"], atol=2e-3)
    # Exercise decisions of the paper: t=1 paths 4,6,7,8; t=2 none (stopped);
    # t=3 path 3 (1-based).
    ex = r["opt_control"]
    assert list(np.nonzero(ex[0])[0]) == [3, 5, 6, 7]


VARIANTS = [
    dict(),
    dict(pooled=1),
    dict(value_mode=nat.VALUE_REALIZED),
    dict(itm_only=0, hysteresis=0.975),
    dict(standardize=0, basis_degree=2),
]


@pytest.mark.parametrize("kw", VARIANTS)
def test_backward_induction_matches_oracle(ctx, oracle, kw):
    pr = config5(n_paths=20000, n_steps=20)
    for k, v in kw.items():
        setattr(pr, k, v)
    r = simulate_rov(ctx, pr, want_paths=True)
    st, ctrl = oracle.rov_forward(pr)
    ref = oracle.rov_backward(pr, st, ctrl)
    assert r["value"] == pytest.approx(ref["value"], rel=1e-8)
    assert r["first_control"] == ref["first_control"]
    assert r["value_sd"] == pytest.approx(ref["value_sd"], rel=1e-6)
    G = ref["beta"].shape[1]
    np.testing.assert_allclose(r["coeffs"][1:, :G], ref["beta"][1:], rtol=1e-6, atol=1e-8)
    same = np.isclose(r["cond_exp"][1:], ref["V"][1:], rtol=1e-8, atol=1e-10)
    assert same.mean() > 0.999


def test_config5_full_size(ctx, oracle):
    # The full configuration (100 000 paths, 50 steps, cubic basis in 3
    # variables) against the oracle on the same streams.
    pr = config5()
    t = time.perf_counter()
    r = simulate_rov(ctx, pr, want_paths=True)
    dt = time.perf_counter() - t
    print(f"config5: value {r['value']:.3f} sd {r['value_sd']:.3f} control {r['first_control']}"
          f" in {dt * 1e3:.1f} ms", flush=True)
    assert np.isfinite(r["value"]) and r["value"] <= 0.0
    assert r["coeffs"].shape == (50, 2, 20)
    st, ctrl = oracle.rov_forward(pr)
    np.testing.assert_allclose(r["states"], st, rtol=1e-10, atol=1e-12)
    ref = oracle.rov_backward(pr, st, ctrl)
    assert r["value"] == pytest.approx(ref["value"], rel=1e-8)
    assert r["first_control"] == ref["first_control"]
    assert r["value_sd"] == pytest.approx(ref["value_sd"], rel=1e-6)
    np.testing.assert_allclose(r["coeffs"][1:], ref["beta"][1:], rtol=1e-6, atol=1e-8)
    assert np.isclose(r["cond_exp"][1:], ref["V"][1:], rtol=1e-8, atol=1e-10).mean() > 0.999


def test_exp_pv_matches_moment_recursion(ctx):
    # E[x'] = E[x] + k(m - E[x])dt + E[x] p (E[e^J] - 1), J ~ N(-J0^2/2, J0^4)
    # (the Brownian term has mean zero); E[PV] = sum_t (1+g)^t E[x_t]/(1+r)^t.
    u = Uncertainty(nat.SDE_MRJD, 1.0, 1.4, 0.2, 0.25, 0.3, 0.1)
    T, dt, r_, g = 30, 1.0, 0.07, 0.01
    mean, sd = exp_pv(ctx, u, T, dt, r_, g, 200000, 9)
    e, pv = 1.0, 0.0
    ej = np.exp(-0.09 / 2 + 0.3 ** 4 / 2)
    for t in range(T):
        e = e + 0.25 * (1.4 - e) * dt + e * 0.1 * (ej - 1)
        pv += (1 + g) ** t * e / (1 + r_) ** t
    assert mean == pytest.approx(pv, abs=4 * sd / np.sqrt(200000))


@pytest.mark.parametrize("u,T,dt,r_,g", [
    (Uncertainty(nat.SDE_MRJD, 1.0, 1.4, 0.2, 0.25, 0.3, 0.1), 30, 1.0, 0.07, 0.01),
    (Uncertainty(nat.SDE_MRJD, 2.5, 1.0, 0.1, 0.6, 0.5, 0.3), 12, 0.5, 0.04, 0.0),
])
def test_exp_pv_matches_oracle(ctx, oracle, u, T, dt, r_, g):
    # The same Philox draws in FP64 on both sides: mean and SD (1/N) to 1e-10.
    N = 100000
    mean, sd = exp_pv(ctx, u, T, dt, r_, g, N, 21)
    _, m_ref, sd_ref = oracle.exp_pv(u, T, dt, r_, g, N, 21)
    assert mean == pytest.approx(m_ref, rel=1e-10)
    assert sd == pytest.approx(sd_ref, rel=1e-10)


def test_american_refuses_regressed_values(ctx):
    pr = RovProblem([Uncertainty(nat.SDE_GBM, 1.0, 0.0, 0.2)], n_steps=3, n_paths=100,
                    mode=nat.ROV_AMERICAN, value_mode=nat.VALUE_REGRESSED, strike=1.1,
                    basis_degree=2)
    with pytest.raises(nat.RoadmcError) as e:
        simulate_rov(ctx, pr)
    assert e.value.code == nat.ERR_UNSUPPORTED


def test_rov_invalid_inputs(ctx):
    pr = config5(n_paths=100, n_steps=3)
    pr.correlation = np.array([[1, 2, 0], [2, 1, 0], [0, 0, 1]], dtype=float)
    with pytest.raises(nat.RoadmcError) as e:
        simulate_rov(ctx, pr)
    assert e.value.code == nat.ERR_NUMERIC
    pr = config5(n_paths=100, n_steps=3)
    pr.uncertainties = pr.uncertainties * 2  # 6 variables, cubic: 84 > 35
    pr.correlation = None
    pr.profit_weight = [0.0] * 6
    with pytest.raises(nat.RoadmcError) as e:
        simulate_rov(ctx, pr)
    assert e.value.code == nat.ERR_INVALID
