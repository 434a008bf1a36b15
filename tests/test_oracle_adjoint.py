"""The adjoint oracle against central finite differences of the FP64 oracle
forward (CPU)."""
import numpy as np
import pytest

from transportoptimiser_amd import native as nat
from transportoptimiser_amd.landscape import SpeciesData, make_species


def _objective(oracle, sd, T, N, seed):
    tot, _ = oracle.mte_species(sd, 0, 0, T, 1.0, nat.LAW_LOGNORMAL, nat.END_POST_GROWTH, 0.0,
                                seed, 0, N, n_threads=2)
    return tot.mean()


def test_adjoint_matches_finite_differences(oracle):
    rng = np.random.default_rng(2)
    sd = make_species(rng, 9, mort_max=0.4)
    sd.growth = rng.uniform(0.2, 0.6, size=9)  # strong growth: non-trivial G'
    T, N, seed = 7, 24, 3
    J, gm, gr = oracle.mte_adjoint(sd, 0, T, N, seed)
    assert J == pytest.approx(_objective(oracle, sd, T, N, seed), rel=1e-13)
    h = 1e-6
    for (k, j) in [(0, 1), (3, 3), (5, 2), (8, 7), (2, 6)]:
        sp, sm = sd.surv.copy(), sd.surv.copy()
        sp[k, j] -= h  # m up = surv down
        sm[k, j] += h
        fp = _objective(oracle, SpeciesData(sd.trans, sp, sd.init_pop, sd.capacity, sd.growth,
                                            noise_sd=sd.noise_sd), T, N, seed)
        fm = _objective(oracle, SpeciesData(sd.trans, sm, sd.init_pop, sd.capacity, sd.growth,
                                            noise_sd=sd.noise_sd), T, N, seed)
        assert gm[k, j] == pytest.approx((fp - fm) / (2 * h), rel=1e-5, abs=1e-8)
    for j in (0, 4, 8):
        rp, rm = sd.growth.copy(), sd.growth.copy()
        rp[j] += h
        rm[j] -= h
        fp = _objective(oracle, SpeciesData(sd.trans, sd.surv, sd.init_pop, sd.capacity, rp,
                                            noise_sd=sd.noise_sd), T, N, seed)
        fm = _objective(oracle, SpeciesData(sd.trans, sd.surv, sd.init_pop, sd.capacity, rm,
                                            noise_sd=sd.noise_sd), T, N, seed)
        assert gr[j] == pytest.approx((fp - fm) / (2 * h), rel=1e-5)


def test_mortality_gradient_sign(oracle):
    # More mortality on any link never adds animals.
    rng = np.random.default_rng(4)
    sd = make_species(rng, 12)
    _, gm, _ = oracle.mte_adjoint(sd, 0, 5, 16, 1)
    assert np.all(gm <= 1e-12)
