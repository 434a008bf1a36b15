"""The species-state ROV oracle (the reference's ALGO4) against invariants
and closed forms (CPU).

The oracle restates forwardPathKernel (SimulateGPU.cu:789-956),
computePathStates (:959-1003), the ALGO4 loop (:4401-4998),
backwardInduction (:2333-2784) and firstPeriodInduction (:2210-2330).  A
transposed movement matrix, a missing /100 in the growth, a wrong discount or
an off-by-one in the backward recursion fails one of these.
"""
import numpy as np
import pytest

from transportoptimiser_amd import rov_species as rs


def _flat_problem(P=6, C=3, N=400, T=6, thr=0.0, u=-0.5, lv=0.0, identity=False, surv=None,
                  params=None):
    rng = np.random.default_rng(P * 100 + C)
    trans = np.eye(P) if identity else rng.uniform(0.1, 1.0, (P, P))
    trans /= trans.sum(1, keepdims=True)
    s = np.ones((C, P, P)) if surv is None else surv
    cap = rng.uniform(500, 5000, P)
    prm = np.array([5.0, 5.0, 0.0, thr, 0.0, 0.0, 0.0, lv]) if params is None else params
    sp = rs.RovsSpecies(0.8 * cap, cap, trans, s, prm)
    # deterministic prices (no diffusion, no jumps, at their means) and
    # composition: unit profit = 1 + 0.2 * 1.5 - 2 * 0.9 * 1.0 = u
    ore = (1.0 + 0.2 * 1.5 - u) / 2.0
    return rs.RovsProblem(
        species=[sp], flow=np.linspace(0, 100, C),
        fuels=[rs.Price(1.5, 1.5, 0.0, 0.3)], commodities=[rs.Price(1.0, 1.0, 0.0, 0.3)],
        ore_mean=np.array([ore]), ore_sd=np.array([0.0]), fuel_cost=np.array([0.2]),
        fuel_idx=np.array([0], dtype=np.int32), unit_cost=1.0, unit_revenue=2.0,
        n_years=T, n_paths=N, dim_res=5, hysteresis=1.0, seed=3)


def test_aar_is_one_without_mortality(oracle):
    pr = _flat_problem()
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    np.testing.assert_allclose(aa, 1.0, rtol=1e-12)


def test_aar_equals_constant_survival(oracle):
    P, C = 5, 3
    s = np.stack([np.full((P, P), v) for v in (1.0, 0.9, 0.7)])
    pr = _flat_problem(P=P, C=C, identity=True, surv=s)
    _, _, aa, _ = oracle.rovs_forward(pr)
    for c, v in enumerate((1.0, 0.9, 0.7)):
        np.testing.assert_allclose(aa[0, :, c, :], v, rtol=1e-12)


def test_growth_closed_form(oracle):
    # Identity movement, no mortality, constant growth-rate mean (no
    # reversion, no noise, no jumps), no local variability:
    # n <- n (1 + g (K - n) / K / 100) per patch, the reference's /100.
    pr = _flat_problem(identity=True, T=5, N=8)
    _, tp, _, _ = oracle.rovs_forward(pr)
    sp = pr.species[0]
    n = sp.init_pop.copy()
    ref = [n.sum()]
    for _ in range(5):
        n = n * (1 + 5.0 * (sp.capacity - n) / sp.capacity / 100.0)
        ref.append(n.sum())
    np.testing.assert_allclose(tp[0, :, 0], ref, rtol=1e-12)
    np.testing.assert_allclose(tp[0, :, 7], ref, rtol=1e-12)


def test_migration_orientation(oracle):
    # One year with a movement matrix whose rows are far from uniform and
    # survival that differs by origin: the total after the move is
    # sum_j (n M)_j, which fails for M transposed.
    rng = np.random.default_rng(9)
    P, C = 5, 2
    trans = rng.uniform(0.1, 1.0, (P, P))
    trans[0] *= 6.0
    trans /= trans.sum(1, keepdims=True)
    surv = np.stack([np.ones((P, P)), np.broadcast_to(rng.uniform(0.5, 1.0, (P, 1)), (P, P))])
    pr = _flat_problem(P=P, C=C, surv=surv.copy())
    pr.species[0].trans = trans
    _, tp, aa, _ = oracle.rovs_forward(pr)
    n0 = pr.species[0].init_pop
    expect = (n0 @ (trans * surv[1])).sum() / n0.sum()
    assert aa[0, 0, 1, 0] == pytest.approx(expect, rel=1e-12)
    wrong = (n0 @ (trans * surv[1]).T).sum() / n0.sum()
    assert abs(wrong - expect) > 1e-3


def test_controls_are_uniform(oracle):
    pr = _flat_problem(N=6000, T=4, C=3)
    ctl, _, _, _ = oracle.rovs_forward(pr)
    freq = np.bincount(ctl.ravel(), minlength=3) / ctl.size
    np.testing.assert_allclose(freq, 1 / 3, atol=0.01)


def test_unit_profit_closed_form(oracle):
    pr = _flat_problem(u=-0.4)
    _, _, _, up = oracle.rovs_forward(pr)
    np.testing.assert_allclose(up, -0.4, rtol=1e-12)


def test_backward_closed_form_with_constant_profit(oracle):
    # Always valid (threshold 0), constant unit profit u < 0: the highest flow
    # wins every step, the fitted grids reproduce the constant continuation,
    # and V_t = Q u (1 + d + ... + d^(T-1-t)) with d = 1 / (1 + rrr dt / 100).
    T, u = 6, -0.5
    pr = _flat_problem(T=T, u=u)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    V, oc, out = oracle.rovs_backward(pr, ctl, tp, aa, up)
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    Q = pr.flow[-1]
    for t in range(1, T):
        expect = Q * u * sum(d ** i for i in range(T - t))
        np.testing.assert_allclose(V[t], expect, rtol=1e-9)
        assert (oc[t] == len(pr.flow) - 1).all()
    assert out["value"] == pytest.approx(Q * u * sum(d ** i for i in range(T)), rel=1e-9)
    assert out["control"] == len(pr.flow) - 1
    assert out["sd"] == pytest.approx(0.0, abs=1e-9)


def test_backward_only_zero_flow_when_thresholds_unreachable(oracle):
    pr = _flat_problem(thr=1e12)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    V, oc, out = oracle.rovs_backward(pr, ctl, tp, aa, up)
    assert (oc == 0).all() and np.all(V == 0.0)
    assert out["value"] == 0.0 and out["control"] == 0


def test_backward_no_profit_means_zero_flow(oracle):
    # Unit profit > 0 is a cost: every payoff clamps to 0 and control 0 (the
    # first) is kept.
    pr = _flat_problem(u=0.3)
    ctl, tp, aa, up = oracle.rovs_forward(pr)
    V, oc, out = oracle.rovs_backward(pr, ctl, tp, aa, up)
    assert np.all(V == 0.0) and (oc == 0).all()
