"""Learned surrogate of the road evaluation with Monte Carlo fallback.

The GA in the reference evaluates most roads of a generation through a
surrogate fitted to fully simulated sample roads (RoadGA::computeSurrogate,
include/main/RoadGA.cpp:1553-2038; fit in SimulateGPU.cu:5533-5968).  Here the
surrogate is an ensemble of MLPs (8 road/landscape features -> 64 -> 64 ->
(MTE, ROV), tanh) trained on the GPU by libroadmc's kernels on roads that the
simulator evaluated; a road whose ensemble members disagree on MTE by more
than `max_spread` is simulated in full instead.  The ROV target is the
config-5 price-state valuation of the road's unit cost, or (rov_model
"species") the reference's species-state valuation (ALGO4) of the road's own
species, which depends on the road's mortality as well as its unit cost, as
the reference's ROV surrogate over (adjusted populations, unit profit) does.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as nat
from .landscape import RoadProblem, make_species
from .rov import RovProblem, config5, simulate_rov
from .rov_species import Price, RovsProblem, simulate_rov_species, species_on

N_FEATURES = 8
N_OUTPUTS = 2
ROV_SCALE = 1000.0


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MlpEnsemble:
    """libroadmc MLP ensemble with Xavier-uniform initialisation."""

    def __init__(self, ctx, n_in=N_FEATURES, n_out=N_OUTPUTS, members=4, seed=0):
        self.ctx = ctx
        h = C.c_void_p()
        nat.check(nat.lib().roadmc_mlp_create(ctx.handle, n_in, 64, n_out, members, C.byref(h)),
                  ctx.handle)
        self.handle = h
        self.n_in, self.n_out, self.members = n_in, n_out, members
        n = C.c_int32()
        nat.check(nat.lib().roadmc_mlp_n_params(h, C.byref(n)))
        self.n_params = n.value
        rng = np.random.default_rng(seed)
        params = np.concatenate([self._init(rng) for _ in range(members)]).astype(np.float32)
        self.set_params(params)

    def _init(self, rng):
        H = 64
        parts = []
        for fan_in, fan_out in ((self.n_in, H), (H, H), (H, self.n_out)):
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-lim, lim, size=fan_in * fan_out))
            parts.append(np.zeros(fan_out))
        return np.concatenate(parts)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        nat.check(nat.lib().roadmc_mlp_set_params(self.handle, _fp(p)), self.ctx.handle)

    def get_params(self):
        p = np.zeros(self.members * self.n_params, dtype=np.float32)
        nat.check(nat.lib().roadmc_mlp_get_params(self.handle, _fp(p)), self.ctx.handle)
        return p

    def gradient(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.members * self.n_params, dtype=np.float32)
        l = np.zeros(self.members, dtype=np.float32)
        nat.check(nat.lib().roadmc_mlp_gradient(self.handle, _fp(X), _fp(Y), len(X), _fp(g),
                                                _fp(l)), self.ctx.handle)
        return l, g.reshape(self.members, self.n_params)

    def train(self, X, Y, steps=2000, lr=3e-3):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        l = np.zeros(self.members, dtype=np.float32)
        nat.check(nat.lib().roadmc_mlp_train(self.handle, _fp(X), _fp(Y), len(X), steps, lr,
                                             _fp(l)), self.ctx.handle)
        return l

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((self.members, len(X), self.n_out), dtype=np.float32)
        nat.check(nat.lib().roadmc_mlp_predict(self.handle, _fp(X), len(X), _fp(out)),
                  self.ctx.handle)
        return out

    def close(self):
        if self.handle:
            nat.lib().roadmc_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- roads of the surrogate study ---------------------------------------------

@dataclass
class SurrogateRoad:
    problem: RoadProblem   # MTE problem (one species)
    unit_cost: float       # the road's unit operating cost (ROV input)
    features: np.ndarray   # (8,)


def deterministic_projection(sd, n_years: int = 50):
    """Mean-field run of the same model with the year noise switched off:
    total population after n_years/2 and n_years relative to the start."""
    M = sd.trans * (sd.surv if sd.surv is not None else 1.0)
    n = np.asarray(sd.init_pop, dtype=np.float64).copy()
    K, g = sd.capacity, sd.growth
    half = 1.0
    for t in range(n_years):
        n = n @ M
        n = n * (1.0 + g * (1.0 - n / K))
        if t == n_years // 2 - 1:
            half = n.sum() / sd.init_pop.sum()
    return half, n.sum() / sd.init_pop.sum()


def road_features(sd, unit_cost: float, n_years: int = 50) -> np.ndarray:
    """8 features: log P, one-year movement loss fraction L, mean growth rate
    r, L/r, the noise-free projection of the total at T/2 and T (relative to
    the start; about 1/N_paths of the Monte Carlo cost), the year-noise SD and
    the unit cost."""
    w = sd.init_pop / sd.init_pop.sum()
    L = float(np.sum(w[:, None] * sd.trans * (1.0 - sd.surv)))
    r = float(np.sum(w * sd.growth))
    half, end = deterministic_projection(sd, n_years)
    return np.array([np.log(sd.n_patches), L, r, L / r, end, half, sd.noise_sd, unit_cost])


def _make_road(args):
    seed, r, n_paths, n_years = args
    rng = np.random.default_rng([seed, r])
    P = int(round(np.exp(rng.uniform(np.log(32), np.log(128)))))
    sd = make_species(rng, P, length_scale_km=rng.uniform(1.0, 3.0),
                      mort_max=rng.uniform(0.1, 0.5), noise_sd=rng.uniform(0.05, 0.2),
                      threshold_ratio=0.5)
    uc = float(rng.uniform(0.0, 0.4))
    pr = RoadProblem([sd], n_years=n_years, n_paths=n_paths, seed=seed, road_id=r)
    return SurrogateRoad(pr, uc, road_features(sd, uc, n_years))


def make_roads(n: int, seed: int, n_paths: int = 8000, n_years: int = 50,
               workers: int = 8) -> list:
    """Random roads: P log-uniform in [32, 128], length scale U(1, 3) km,
    mortality cap U(0.1, 0.5), year noise U(0.05, 0.2), unit cost U(0, 0.4);
    road r draws from its own generator seeded (seed, r)."""
    args = [(seed, r, n_paths, n_years) for r in range(n)]
    if workers > 1:
        import multiprocessing as mp
        # spawned, not forked: the caller may already hold an initialised
        # HIP runtime, which a forked child must not inherit.  A calling
        # script therefore needs the usual `if __name__ == "__main__":` guard.
        with mp.get_context("spawn").Pool(workers) as pool:
            return pool.map(_make_road, args, chunksize=16)
    return [_make_road(a) for a in args]


def rov_problem(unit_cost: float, n_paths: int, n_steps: int, seed: int) -> RovProblem:
    """The config-5 valuation of a road with the given unit cost.  Targets
    use realized cash flows under the fitted policy (with few paths the
    regressed-value recursion compounds cubic-fit extrapolation errors over
    50 steps; realized values keep the targets comparable at any path
    count)."""
    from . import native
    pr = config5(seed=5, n_paths=n_paths, n_steps=n_steps)
    pr.unit_cost = unit_cost
    pr.seed = seed
    pr.value_mode = native.VALUE_REALIZED
    return pr


# Validity threshold of the species-state valuation, as a fraction of the
# initial total: the highest flow is valid on 15-85% of the forward
# path-years across the study's roads.
ROV_THRESHOLD_RATIO = 0.85


def species_rov_problem(sd, unit_cost: float, n_paths: int, n_years: int, seed: int,
                        road_id: int = 0) -> RovsProblem:
    """The species-state valuation (the reference's ALGO4) of a road: its
    species on the road's landscape (species_on: flows 0, 50, 100, crossing
    mortality scaling with the flow), growth 5% a year with local variability
    1, the validity threshold at ROV_THRESHOLD_RATIO of the initial total; one
    fuel and one ore price (the prices of rov_species.synthetic_problem); unit
    cost 0.6 + 2.5 x the road's unit cost, unit revenue 2."""
    flow = np.array([0.0, 50.0, 100.0])
    params = np.array([5.0, 5.0, 0.1, ROV_THRESHOLD_RATIO * float(np.sum(sd.init_pop)), 0.5, 0.1,
                       0.05, 1.0])
    return RovsProblem(
        species=[species_on(sd, flow, params)], flow=flow,
        fuels=[Price(1.5, 1.5, 0.05, 0.3, 0.1, 0.02)],
        commodities=[Price(2.0, 2.0, 0.08, 0.3, 0.1, 0.02)],
        ore_mean=np.array([0.5]), ore_sd=np.array([0.05]),
        fuel_cost=np.array([0.2]), fuel_idx=np.array([0], dtype=np.int32),
        unit_cost=0.6 + 2.5 * unit_cost, unit_revenue=2.0, n_years=n_years, n_paths=n_paths,
        dim_res=12, seed=seed, road_id=road_id)


def simulate(ctx, roads, rov_paths: int = 100000, rov_steps: int = 50,
             rov_model: str = "price") -> np.ndarray:
    """Full Monte Carlo targets (MTE of species 0, ROV / 1000) for roads.  The
    ROV is the config-5 price-state valuation of the road's unit cost
    (rov_model "price", rov_problem; config 5's 100,000 paths by default) or
    the species-state valuation of the road's species ("species",
    species_rov_problem)."""
    outs = ctx.simulate_mte_batch([r.problem for r in roads], end_pops=False, patch_mean=False)
    y = np.zeros((len(roads), 2))
    for i, (r, o) in enumerate(zip(roads, outs)):
        seed = 1000 + r.problem.road_id
        if rov_model == "species":
            v = simulate_rov_species(ctx, species_rov_problem(
                r.problem.species[0], r.unit_cost, rov_paths, rov_steps, seed,
                r.problem.road_id))
        else:
            v = simulate_rov(ctx, rov_problem(r.unit_cost, rov_paths, rov_steps, seed=seed))
        y[i] = (o.mte_prob[0], v["value"] / ROV_SCALE)
    return y


class Surrogate:
    """Standardised MLP ensemble over road features with MC fallback."""

    def __init__(self, ctx, members=4, seed=0, max_spread=0.01):
        self.ctx = ctx
        self.net = MlpEnsemble(ctx, members=members, seed=seed)
        self.max_spread = max_spread

    def fit(self, X, Y, steps=3000, lr=3e-3):
        """Full-batch Adam: half the steps at lr, a quarter at lr/3, a quarter
        at lr/10 (one Adam state throughout)."""
        self.xm, self.xs = X.mean(0), X.std(0) + 1e-9
        self.ym, self.ys = Y.mean(0), Y.std(0) + 1e-9
        Xs, Ys = (X - self.xm) / self.xs, (Y - self.ym) / self.ys
        self.net.train(Xs, Ys, steps // 2, lr)
        self.net.train(Xs, Ys, steps // 4, lr / 3)
        return self.net.train(Xs, Ys, steps - steps // 2 - steps // 4, lr / 10)

    def predict(self, X):
        """(mean (n, 2), spread (n, 2)) in target units; spread = SD over members."""
        out = self.net.predict((X - self.xm) / self.xs) * self.ys + self.ym
        return out.mean(0), out.std(0)

    def evaluate(self, roads, **sim_kw):
        """Surrogate values, with roads whose MTE spread exceeds max_spread
        simulated in full.  Returns (values (n, 2), simulated mask)."""
        X = np.stack([r.features for r in roads])
        mean, spread = self.predict(X)
        fall = spread[:, 0] > self.max_spread
        vals = mean.astype(np.float64)
        if fall.any():
            idx = np.nonzero(fall)[0]
            vals[idx] = simulate(self.ctx, [roads[i] for i in idx], **sim_kw)
        return vals, fall
