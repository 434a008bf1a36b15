"""The smoothed-validity value of the species-state valuation under a fixed
policy and its gradient (orc_rovs_smooth_grad), pinned on the CPU:

- the reverse-mode gradient against central differences of the oracle's own
  value (FP64, h = 1e-5 on survival entries of both species and several
  controls);
- with tau -> 0 the value equals the hard-validity realized value computed
  here from orc_rovs_forward's arrays (an independent forward pass under the
  same controls), with the thresholds binding on part of the paths;
- a policy that keeps control 0 everywhere has no gradient."""
import copy

import numpy as np
import pytest

from oracle import pyoracle as orc
from transportoptimiser_amd import rov_species as rs


def _problem(ratio=0.97, N=40, T=6):
    return rs.synthetic_problem(seed=3, n_patches=10, n_species=2, n_controls=3, n_paths=N,
                                n_years=T, threshold_ratio=ratio)


def _hard_value(pr, ctl, tp, aa, up):
    T, S = pr.n_years, len(pr.species)
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    A = np.stack([tp[s, :T] * np.take_along_axis(aa[s], ctl[:, None, :], 1)[:, 0]
                  for s in range(S)])
    thr = np.array([sp.params[3] for sp in pr.species])[:, None, None]
    valid = (A >= thr).all(0)
    Q = np.asarray(pr.flow)
    cash = d ** np.arange(T)[:, None] * up * (Q[0] + (Q[ctl] - Q[0]) * valid)
    return cash.sum(0).mean(), valid.mean()


def test_gradient_matches_central_differences():
    pr = _problem()
    pol = np.random.default_rng(1).integers(0, 3, size=(pr.n_years, pr.n_paths)).astype(np.int32)
    tau = 0.02 * pr.species[0].params[3]
    J, g = orc.rovs_smooth_grad(pr, pol, tau)
    assert np.isfinite(J) and np.abs(g).max() > 0
    h = 1e-5
    for (s, c, k, j) in [(0, 2, 1, 3), (1, 1, 4, 4), (0, 1, 2, 7), (1, 2, 0, 9), (0, 0, 5, 5),
                         (1, 0, 8, 2)]:
        p1, p2 = copy.deepcopy(pr), copy.deepcopy(pr)
        p1.species[s].surv[c, k, j] += h
        p2.species[s].surv[c, k, j] -= h
        Jp, _ = orc.rovs_smooth_grad(p1, pol, tau, grad=False)
        Jm, _ = orc.rovs_smooth_grad(p2, pol, tau, grad=False)
        fd = (Jp - Jm) / (2 * h)
        assert g[s, c, k, j] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(g).max())


def test_sharp_limit_is_the_hard_validity_value():
    pr = _problem(N=200)
    ctl, tp, aa, up = orc.rovs_forward(pr)
    Jh, frac = _hard_value(pr, ctl, tp, aa, up)
    assert 0.3 < frac < 0.95  # the thresholds bind on part of the path-years
    J, _ = orc.rovs_smooth_grad(pr, ctl, 1e-9 * pr.species[0].params[3], grad=False)
    assert J == pytest.approx(Jh, rel=1e-12)


def test_control_zero_policy_has_no_gradient():
    pr = _problem()
    pol = np.zeros((pr.n_years, pr.n_paths), dtype=np.int32)
    J, g = orc.rovs_smooth_grad(pr, pol, 0.02 * pr.species[0].params[3])
    assert np.all(g == 0.0)
    ctl, tp, aa, up = orc.rovs_forward(pr)
    d = 1.0 / (1.0 + pr.rrr * pr.time_step / 100.0)
    ref = (d ** np.arange(pr.n_years)[:, None] * up * pr.flow[0]).sum(0).mean()
    assert J == pytest.approx(ref, rel=1e-12, abs=1e-12)
