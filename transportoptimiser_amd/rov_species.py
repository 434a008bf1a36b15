"""Species-state real-options valuation: the reference's ALGO4.

Host mirror of SimulateGPU::simulateROVCUDA with ROVMethod ALGO4
(include/gpuCode/SimulateGPU.cu:3870-4998): the state of a path is the
adjusted population of each species under the highest flow (total population
x surviving fraction) and the unit profit; controls are the traffic flow
levels; the continuation value per control and step is a kNN local-linear
grid over the paths that drew that control.  ``roadmc_rovs_run`` runs all of
it on the GPU (rovs_kernels.hip, knn_kernels.hip).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native as nat
from .landscape import crossed_links, random_polyline


@dataclass
class RovsSpecies:
    init_pop: np.ndarray   # (P,)
    capacity: np.ndarray   # (P,)
    trans: np.ndarray      # (P, P) row = origin
    surv: np.ndarray       # (C, P, P) survival under each control
    params: np.ndarray     # (8,) growth current, mean, noise SD, threshold, reversion,
                           # jump size, jump probability, local variability


@dataclass
class Price:
    x0: float
    mean: float
    sd: float
    reversion: float
    jump_size: float = 0.0
    jump_prob: float = 0.0
    kind: int = 0          # ROADMC_SDE_* (0: the reference's MRJD law)


@dataclass
class RovsProblem:
    species: list
    flow: np.ndarray                   # (C,) control C-1 is the highest flow
    fuels: list = field(default_factory=list)        # [Price]
    commodities: list = field(default_factory=list)  # [Price]
    ore_mean: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ore_sd: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fuel_cost: np.ndarray = field(default_factory=lambda: np.zeros(0))  # per vehicle class
    fuel_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    unit_cost: float = 0.0
    unit_revenue: float = 0.0
    n_years: int = 50
    time_step: float = 1.0
    rrr: float = 7.0
    n_paths: int = 10000
    dim_res: int = 20
    profit_scaling: float = 2.0        # SimulateGPU.cu:3900
    hysteresis: float = 0.975          # SimulateGPU.cu:2763
    seed: int = 0
    road_id: int = 0
    method: int = 4                    # 4: ALGO4; 6: ALGO6 (forward recomputation)
    checkpoint_years: int = 0          # ALGO6: forward populations kept every k years (0: auto)
    precision: int = 0                 # products: PREC_FP32 (0) or PREC_BF16X3 (2)

    @property
    def n_patches(self) -> int:
        return int(np.asarray(self.species[0].init_pop).shape[0])


class _Marshal:
    def __init__(self, pr: RovsProblem):
        self.keep = []
        P, C_ = pr.n_patches, len(pr.flow)
        sp = (nat.RovsSpecies * len(pr.species))()
        for i, s in enumerate(pr.species):
            surv = np.asarray(s.surv, dtype=np.float64)
            if surv.shape != (C_, P, P):
                raise ValueError("surv must be (n_controls, P, P)")
            sp[i].init_pop = self._arr(s.init_pop)
            sp[i].capacity = self._arr(s.capacity)
            sp[i].trans = self._arr(np.asarray(s.trans).T)   # column-major
            sp[i].surv = self._arr(surv.transpose(0, 2, 1))  # [C] column-major
            for k in range(8):
                sp[i].params[k] = float(s.params[k])
        self.keep.append(sp)
        prices = list(pr.fuels) + list(pr.commodities)
        unc = (nat.Uncertainty * max(len(prices), 1))()
        for i, q in enumerate(prices):
            unc[i].kind, unc[i].x0, unc[i].mean, unc[i].sd = q.kind, q.x0, q.mean, q.sd
            unc[i].reversion, unc[i].jump_size, unc[i].jump_prob = q.reversion, q.jump_size, q.jump_prob
        self.keep.append(unc)
        p = nat.RovsProblem()
        p.n_patches = P
        p.n_species = len(pr.species)
        p.species = sp
        p.n_controls = C_
        p.flow = self._arr(pr.flow)
        p.n_fuels = len(pr.fuels)
        p.n_commodities = len(pr.commodities)
        p.uncertainties = unc
        p.ore_mean = self._arr(pr.ore_mean)
        p.ore_sd = self._arr(pr.ore_sd)
        p.n_vehicles = len(pr.fuel_cost)
        p.fuel_cost = self._arr(pr.fuel_cost)
        fi = np.ascontiguousarray(pr.fuel_idx, dtype=np.int32)
        self.keep.append(fi)
        p.fuel_idx = fi.ctypes.data_as(nat._ip)
        p.unit_cost, p.unit_revenue = pr.unit_cost, pr.unit_revenue
        p.n_years, p.time_step, p.rrr = pr.n_years, pr.time_step, pr.rrr
        p.n_paths, p.dim_res = pr.n_paths, pr.dim_res
        p.profit_scaling, p.hysteresis = pr.profit_scaling, pr.hysteresis
        p.seed, p.road_id, p.method = pr.seed, pr.road_id, pr.method
        p.checkpoint_years = pr.checkpoint_years
        p.precision = pr.precision
        self.problem = p

    def _arr(self, a):
        b = np.ascontiguousarray(a, dtype=np.float64)
        self.keep.append(b)
        return b.ctypes.data_as(nat._dp)


def simulate_rov_species(ctx, pr: RovsProblem, *, paths: bool = False,
                         grids: bool = False) -> dict:
    """``roadmc_rovs_run``.  Returns value, value_sd, first_control and, with
    ``paths``, the per-path arrays (cond_exp, opt_control, adj_pops,
    unit_profit, controls, total_pops, aars); with ``grids`` the fitted
    grids."""
    m = _Marshal(pr)
    T, N, S, C_ = pr.n_years, pr.n_paths, len(pr.species), len(pr.flow)
    D, R = S + 1, pr.dim_res
    out = {"value": np.zeros(1), "value_sd": np.zeros(1),
           "first_control": np.zeros(1, dtype=np.int32)}
    if paths:
        out.update(cond_exp=np.zeros((T, N)), opt_control=np.zeros((T, N), dtype=np.int32),
                   adj_pops=np.zeros((T, S, N)), unit_profit=np.zeros((T, N)),
                   controls=np.zeros((T, N), dtype=np.int32),
                   total_pops=np.zeros((S, T + 1, N)), aars=np.zeros((S, T, C_, N)))
    if grids:
        out["grids"] = np.zeros((T, C_, D * R + R ** D))
    r = nat.RovsResult()
    for name, arr in out.items():
        ptr = arr.ctypes.data_as(nat._ip if arr.dtype == np.int32 else nat._dp)
        setattr(r, name, ptr)
    nat.check(nat.lib().roadmc_rovs_run(ctx.handle, C.byref(m.problem), C.byref(r)), ctx.handle)
    out["value"] = float(out["value"][0])
    out["value_sd"] = float(out["value_sd"][0])
    out["first_control"] = int(out["first_control"][0])
    return out


def species_on(sd, flow, params) -> RovsSpecies:
    """The valuation's species on an MTE species' landscape (a
    landscape.SpeciesData): the same transitions, capacities and initial
    populations; under flow level q_c the crossing mortality m = 1 - sd.surv
    scales with q_c / max q (surv_c = 1 - m q_c / q_max, the rule of
    synthetic_problem).  ``params``: the 8 species parameters (growth current,
    mean, noise SD, threshold, reversion, jump size, jump probability, local
    variability)."""
    q = np.asarray(flow, dtype=np.float64)
    q = q / q.max()
    m = 1.0 - np.asarray(sd.surv, dtype=np.float64)
    surv = np.stack([1.0 - m * qc for qc in q])
    return RovsSpecies(np.asarray(sd.init_pop, dtype=np.float64),
                       np.asarray(sd.capacity, dtype=np.float64),
                       np.asarray(sd.trans, dtype=np.float64), surv,
                       np.asarray(params, dtype=np.float64))


def smooth_value_gradient(ctx, pr: RovsProblem, policy, tau: float, *, grad: bool = True):
    """``roadmc_rovs_smooth_grad``: the smoothed-validity value under the
    policy [T][N] (validity prod_s sigmoid((A_s - threshold_s) / tau)) and,
    with ``grad``, its gradient with respect to each species' survival
    matrices, (S, C, P, P) with rows the origin patch like ``surv``."""
    m = _Marshal(pr)
    T, N, S, C_, P = pr.n_years, pr.n_paths, len(pr.species), len(pr.flow), pr.n_patches
    pol = np.ascontiguousarray(policy, dtype=np.int32)
    if pol.shape != (T, N):
        raise ValueError("policy must be (n_years, n_paths)")
    value = np.zeros(1)
    g = np.zeros((S, C_, P, P)) if grad else None
    nat.check(nat.lib().roadmc_rovs_smooth_grad(
        ctx.handle, C.byref(m.problem), pol.ctypes.data_as(nat._ip), float(tau),
        value.ctypes.data_as(nat._dp), g.ctypes.data_as(nat._dp) if grad else None), ctx.handle)
    # column-major [j][k] per matrix -> rows = origin k
    return float(value[0]), (g.transpose(0, 1, 3, 2).copy() if grad else None)


def synthetic_problem(seed: int = 5, n_patches: int = 64, n_species: int = 1,
                      n_controls: int = 3, n_paths: int = 10000, n_years: int = 30,
                      dim_res: int = 12, threshold_ratio: float = 0.965) -> RovsProblem:
    """A seeded landscape and road: exp(-d / 2 km) transitions, crossing
    mortality on the links a random road cuts, rising with the flow level;
    growth rates in percent per year as the reference's species law takes
    them; thresholds threshold_ratio x the initial total; one MRJD fuel price
    and one ore price; flows 0, ..., 100."""
    rng = np.random.default_rng(seed)
    P, C_ = n_patches, n_controls
    flow = np.linspace(0.0, 100.0, C_)
    species = []
    for _ in range(n_species):
        xy = rng.uniform(0.0, 10.0, size=(P, 2))
        d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
        trans = np.exp(-d / 2.0)
        trans /= trans.sum(1, keepdims=True)
        cut = crossed_links(xy, random_polyline(rng, 10.0)).astype(np.float64)
        mort = rng.uniform(0.0, 0.3, size=(P, P)) * cut
        surv = np.stack([1.0 - mort * (q / flow[-1]) for q in flow])
        cap = rng.uniform(500.0, 5000.0, size=P)
        init = 0.8 * cap
        params = np.array([5.0, 5.0, 0.1, threshold_ratio * init.sum(), 0.5, 0.1, 0.05, 1.0])
        # Each matrix column-major, the C ABI's layout (Eigen's default, in
        # which the reference keeps them): passed to the library without a copy.
        trans = np.asfortranarray(trans)
        surv = np.ascontiguousarray(surv.transpose(0, 2, 1)).transpose(0, 2, 1)
        species.append(RovsSpecies(init, cap, trans, surv, params))
    return RovsProblem(
        species=species, flow=flow,
        fuels=[Price(1.5, 1.5, 0.05, 0.3, 0.1, 0.02)],
        commodities=[Price(2.0, 2.0, 0.08, 0.3, 0.1, 0.02)],
        ore_mean=np.array([0.5]), ore_sd=np.array([0.05]),
        fuel_cost=np.array([0.2]), fuel_idx=np.array([0], dtype=np.int32),
        unit_cost=1.0, unit_revenue=2.0, n_years=n_years, n_paths=n_paths,
        dim_res=dim_res, seed=seed, road_id=7)
