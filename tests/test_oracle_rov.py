"""The ROV oracle against published answers, library routines and closed
forms (CPU)."""
import json
import pathlib

import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd.rov import RovProblem, Uncertainty, config5

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def ls_problem():
    d = json.loads((GOLD / "longstaff_schwartz_2001.json").read_text())
    paths = np.array(d["paths_t0_t1_t2_t3"])  # (N, T+1)
    states = paths.T[:, None, :].copy()       # (T+1, 1, N)
    pr = RovProblem([Uncertainty(nat.SDE_GBM, 1.0, 0.0, 0.0)], n_steps=3, n_paths=8,
                    basis_degree=2, standardize=0, itm_only=1, mode=nat.ROV_AMERICAN,
                    value_mode=nat.VALUE_REALIZED, strike=d["strike"],
                    discount=d["discount_per_period"], flow=[0.0, 0.0])
    return pr, states, d


def test_longstaff_schwartz_example(oracle):
    pr, states, d = ls_problem()
    ctrl = np.zeros((3, 8), dtype=np.int32)
    r = oracle.rov_backward(pr, states, ctrl)
    # The paper rounds to 3-4 digits.
    assert r["value"] == pytest.approx(d["value"], abs=6e-5)
    disc = d["discount_per_period"]
    np.testing.assert_allclose(r["beta"][2, 0] * disc, d["regression_t2"], atol=2e-3)
    np.testing.assert_allclose(r["beta"][1, 0] * disc, d["regression_t1"], atol=2e-3)


def test_normal_equations_match_lstsq(oracle):
    rng = np.random.default_rng(0)
    X = rng.normal(size=(500, 3))
    Phi = np.column_stack([np.ones(500), X, X[:, 0] ** 2, X[:, 0] * X[:, 1]])
    y = Phi @ np.array([1.0, -2.0, 0.5, 3.0, 0.25, -1.0]) + 0.1 * rng.normal(size=500)
    beta = oracle.solve(Phi.T @ Phi, Phi.T @ y)
    ref, *_ = np.linalg.lstsq(Phi, y, rcond=None)
    np.testing.assert_allclose(beta, ref, rtol=1e-9, atol=1e-12)


def test_solver_pivots_on_magnitude(oracle):
    # A zero leading pivot with a large sub-diagonal entry: the reference's
    # test fabsf(A[..] > maxElem) compares before taking the magnitude and
    # would keep the zero pivot (SimulateGPU.cu:6312).
    A = np.array([[0.0, 1.0, 2.0], [-5.0, 1.0, 0.0], [1.0, -3.0, 4.0]])
    b = np.array([1.0, 2.0, 3.0])
    np.testing.assert_allclose(oracle.solve(A, b), np.linalg.solve(A, b), rtol=1e-12)
    H = np.array([[1.0 / (i + j + 1) for j in range(6)] for i in range(6)])
    xb = np.arange(1.0, 7.0)
    np.testing.assert_allclose(oracle.solve(H, H @ xb), xb, rtol=1e-6)


def _problem(kind, **kw):
    u = Uncertainty(kind, **kw)
    return RovProblem([u], n_steps=20, n_paths=20000, seed=3, flow=[0.0, 1.0])


def test_ou_euler_moments(oracle):
    k, mu, s, x0, dt = 0.4, 1.0, 0.2, 1.6, 1.0
    pr = _problem(nat.SDE_OU, x0=x0, mean=mu, sd=s, reversion=k)
    st, _ = oracle.rov_forward(pr)
    m, v = x0, 0.0
    for t in range(1, 21):
        m = mu + (m - mu) * (1 - k * dt)
        v = (1 - k * dt) ** 2 * v + s * s * dt
        x = st[t, 0]
        assert x.mean() == pytest.approx(m, abs=4 * np.sqrt(v / len(x)) + 1e-12)
        assert x.var() == pytest.approx(v, rel=0.05)


def test_gbm_mean(oracle):
    mu, s = 0.02, 0.15
    pr = _problem(nat.SDE_GBM, x0=1.0, mean=mu, sd=s)
    st, _ = oracle.rov_forward(pr)
    for t in (1, 10, 20):
        x = st[t, 0]
        se = np.sqrt((np.exp(s * s * t) - 1) * np.exp(2 * mu * t) / len(x))
        assert x.mean() == pytest.approx(np.exp(mu * t), abs=4 * se)


def test_mrjd_mean_recursion(oracle):
    # E[x'] = E[x](1 - k dt) + k m dt + E[x] p (E[e^J'] - 1), with
    # J' = J^2 z' - J^2/2 ~ N(-J^2/2, J^4) (SimulateGPU.cu:940-952).
    k, m, s, J, p = 0.3, 2.0, 0.05, 0.4, 0.2
    pr = _problem(nat.SDE_MRJD, x0=1.0, mean=m, sd=s, reversion=k, jump_size=J, jump_prob=p)
    st, _ = oracle.rov_forward(pr)
    ej = np.exp(-J * J / 2 + J ** 4 / 2)
    e = 1.0
    for t in range(1, 21):
        e = e * (1 - k) + k * m + e * p * (ej - 1)
        x = st[t, 0]
        assert x.mean() == pytest.approx(e, abs=4 * x.std() / np.sqrt(len(x)))


def test_correlation_and_controls(oracle):
    pr = config5(n_paths=20000, n_steps=5)
    st, ctrl = oracle.rov_forward(pr)
    u0, u1 = pr.uncertainties[0], pr.uncertainties[1]
    z0 = (st[1, 0] - st[0, 0] - u0.reversion * (u0.mean - st[0, 0])) / u0.sd
    z1 = (st[1, 1] - st[0, 1] - u1.reversion * (u1.mean - st[0, 1])) / u1.sd
    z2 = np.log(st[1, 2] / st[0, 2]) - (0.02 - 0.15 ** 2 / 2)
    c = np.corrcoef([z0, z1, z2])
    assert c[0, 1] == pytest.approx(0.3, abs=0.03)
    assert c[0, 2] == pytest.approx(0.3, abs=0.03)
    assert abs(np.mean(ctrl == 1) - 0.5) < 0.02


def test_switching_decisions_are_consistent(oracle):
    # With zero continuation (one step left), the best control is argmin of
    # the clamped current payoffs: the high-flow control wins exactly where
    # the unit profit is negative.
    pr = config5(n_paths=4000, n_steps=2)
    st, ctrl = oracle.rov_forward(pr)
    r = oracle.rov_backward(pr, st, ctrl)
    up = pr.unit_cost + np.tensordot(pr.profit_weight, st[1], axes=(0, 0))
    want = np.minimum(150.0 * st[1, 2] * up, 0.0)
    np.testing.assert_allclose(r["V"][1], want, rtol=1e-12, atol=1e-12)
