"""Adjoint on the GPU (runs on an MI355X): against the FP64 oracle adjoint,
against central finite differences of the FP64 forward kernel at the
256-patch / 100-year / 50,000-path size, and through torch.autograd.gradcheck
of a torch.autograd.Function wrapper (test harness only)."""
import time

import numpy as np
import pytest
import torch

from transportoptimiser_amd import native as nat
from transportoptimiser_amd.adjoint import Adjoint, mte_gradient
from transportoptimiser_amd.landscape import RoadProblem, SpeciesData, make_species

pytestmark = pytest.mark.gpu


def _problem(P, N, T, seed=7, prec=nat.PREC_FP64, strong=False):
    rng = np.random.default_rng(seed + P)
    sd = make_species(rng, P, mort_max=0.4)
    if strong:
        sd.growth = rng.uniform(0.2, 0.6, size=P)
    return RoadProblem([sd], n_years=T, n_paths=N, seed=seed, precision=prec)


@pytest.mark.parametrize("prec,rtol", [(nat.PREC_FP64, 1e-9), (nat.PREC_FP32, 2e-4)])
def test_adjoint_matches_oracle(ctx, oracle, prec, rtol):
    pr = _problem(40, 300, 23, prec=prec, strong=True)
    J, gm, gr = mte_gradient(ctx, pr, checkpoint_every=10)
    Jr, gmr, grr = oracle.mte_adjoint(pr.species[0], 0, 23, 300, pr.seed)
    assert J == pytest.approx(Jr, rel=rtol)
    scale = np.abs(gmr).max()
    np.testing.assert_allclose(gm, gmr, rtol=rtol, atol=rtol * scale)
    np.testing.assert_allclose(gr, grr, rtol=rtol, atol=rtol * np.abs(grr).max())


def test_checkpoint_interval_does_not_change_gradient(ctx):
    pr = _problem(64, 200, 17)
    a = mte_gradient(ctx, pr, checkpoint_every=4)
    b = mte_gradient(ctx, pr, checkpoint_every=17)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(a[2], b[2], rtol=1e-12)


def _objective_fp64(ctx, sd, N, T, seed):
    pr = RoadProblem([sd], n_years=T, n_paths=N, seed=seed, precision=nat.PREC_FP64)
    return ctx.simulate_mte(pr, end_pops=False, patch_mean=False).end_mean[0]


def test_config4_gradient_vs_central_differences(ctx):
    # 256 patches, 1 species, 100 years, 50,000 paths, checkpoints every 10
    # years; FP32 adjoint against central differences (h = 1e-3) of the FP64
    # forward on 64 random mortality entries and 8 growth rates.
    P, N, T = 256, 50000, 100
    pr = _problem(P, N, T, seed=4, prec=nat.PREC_FP32)
    sd = pr.species[0]
    t = time.perf_counter()
    J, gm, gr = mte_gradient(ctx, pr, checkpoint_every=10)
    print(f"config4 adjoint (fwd+bwd, FP32): {time.perf_counter() - t:.3f}s", flush=True)
    pr64 = _problem(P, N, T, seed=4, prec=nat.PREC_FP64)
    J64, gm64, gr64 = mte_gradient(ctx, pr64, checkpoint_every=10)
    rng = np.random.default_rng(1)
    cut = np.argwhere(sd.surv < 1.0)
    picks = cut[rng.choice(len(cut), size=64, replace=False)]
    h = 1e-3
    worst = 0.0
    for k, j in picks:
        sp, sm = sd.surv.copy(), sd.surv.copy()
        sp[k, j] -= h
        sm[k, j] += h
        fp = _objective_fp64(ctx, SpeciesData(sd.trans, sp, sd.init_pop, sd.capacity, sd.growth,
                                              noise_sd=sd.noise_sd), N, T, pr.seed)
        fm = _objective_fp64(ctx, SpeciesData(sd.trans, sm, sd.init_pop, sd.capacity, sd.growth,
                                              noise_sd=sd.noise_sd), N, T, pr.seed)
        fd = (fp - fm) / (2 * h)
        worst = max(worst, abs(gm[k, j] - fd) / max(abs(fd), 1e-3 * np.abs(gm64).max()))
        assert gm64[k, j] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(gm64).max())
    print(f"config4: max rel error of FP32 adjoint vs central differences {worst:.2e}")
    assert worst <= 1e-3
    for j in rng.choice(P, size=8, replace=False):
        rp, rm = sd.growth.copy(), sd.growth.copy()
        rp[j] += h
        rm[j] -= h
        fp = _objective_fp64(ctx, SpeciesData(sd.trans, sd.surv, sd.init_pop, sd.capacity, rp,
                                              noise_sd=sd.noise_sd), N, T, pr.seed)
        fm = _objective_fp64(ctx, SpeciesData(sd.trans, sd.surv, sd.init_pop, sd.capacity, rm,
                                              noise_sd=sd.noise_sd), N, T, pr.seed)
        fd = (fp - fm) / (2 * h)
        assert gr[j] == pytest.approx(fd, rel=1e-3)
    np.testing.assert_allclose(gm, gm64, rtol=1e-3, atol=1e-4 * np.abs(gm64).max())


class MteObjective(torch.autograd.Function):
    """J(m, r) = mean total end population, forward and backward through the
    C ABI (FP64)."""

    @staticmethod
    def forward(fctx, mort, growth, ctx, base):
        sd0 = base.species[0]
        sd = SpeciesData(sd0.trans, 1.0 - mort.detach().cpu().numpy(), sd0.init_pop,
                         sd0.capacity, growth.detach().cpu().numpy(), noise_sd=sd0.noise_sd)
        pr = RoadProblem([sd], n_years=base.n_years, n_paths=base.n_paths, seed=base.seed,
                         precision=nat.PREC_FP64)
        adj = Adjoint(ctx, pr, checkpoint_every=3)
        fctx.adj = adj
        return torch.tensor(adj.objective, dtype=torch.float64)

    @staticmethod
    def backward(fctx, g):
        # gradcheck may run backward more than once per forward; the adjoint
        # state is released with fctx.
        gm, gr = fctx.adj.backward(float(g))
        return (torch.from_numpy(gm), torch.from_numpy(gr), None, None)


def test_autograd_gradcheck(ctx):
    base = _problem(6, 16, 5, strong=True)
    sd = base.species[0]
    mort = torch.tensor(1.0 - sd.surv, dtype=torch.float64, requires_grad=True)
    growth = torch.tensor(sd.growth, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda m, r: MteObjective.apply(m, r, ctx, base),
                                    (mort, growth), eps=1e-6, atol=1e-6, rtol=1e-5)
