"""ROV problems: exogenous price uncertainties valued by regression Monte
Carlo (SimulateGPU::simulateROVCUDA, include/gpuCode/SimulateGPU.h:187-191;
expected PV, SimulateGPU::expPV, :206)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native as nat


@dataclass
class Uncertainty:
    kind: int = nat.SDE_MRJD
    x0: float = 1.0
    mean: float = 1.0
    sd: float = 0.1
    reversion: float = 0.0
    jump_size: float = 0.0
    jump_prob: float = 0.0


@dataclass
class RovProblem:
    uncertainties: list
    correlation: np.ndarray | None = None
    flow: list = field(default_factory=lambda: [0.0, 150.0])
    unit_cost: float = 0.0
    profit_weight: list | None = None
    traffic_index: int = -1
    n_steps: int = 50
    time_step: float = 1.0
    rrr: float = 7.0
    discount: float = 0.0
    n_paths: int = 100000
    basis_degree: int = 3
    standardize: int = 1
    itm_only: int = 1
    pooled: int = 0
    hysteresis: float = 1.0
    mode: int = nat.ROV_SWITCHING
    value_mode: int = nat.VALUE_REGRESSED
    strike: float = 0.0
    seed: int = 0
    states: np.ndarray | None = None      # (T+1, U, N)
    controls: np.ndarray | None = None    # (T, N)

    @property
    def n_controls(self) -> int:
        return 2 if self.mode == nat.ROV_AMERICAN else len(self.flow)

    def basis_size(self) -> int:
        n = C.c_int32()
        nat.check(nat.lib().roadmc_rov_basis_size(len(self.uncertainties), self.basis_degree,
                                                  C.byref(n)))
        return n.value


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def simulate_rov(ctx, pr: RovProblem, *, want_paths: bool = False) -> dict:
    """Runs roadmc_rov_run; returns value, value_sd, first_control, coeffs
    (T, C, K) and, with want_paths, cond_exp (T+1, N), opt_control (T, N),
    states (T+1, U, N)."""
    keep = []
    U, T, N = len(pr.uncertainties), pr.n_steps, pr.n_paths
    uncs = (nat.Uncertainty * U)()
    for i, u in enumerate(pr.uncertainties):
        uncs[i] = nat.Uncertainty(u.kind, u.x0, u.mean, u.sd, u.reversion, u.jump_size,
                                  u.jump_prob)

    def arr(a, dtype=np.float64):
        if a is None:
            return None
        b = np.ascontiguousarray(a, dtype=dtype)
        keep.append(b)
        return b.ctypes.data_as(C.POINTER(C.c_double if dtype == np.float64 else C.c_int32))

    p = nat.RovProblem()
    p.n_uncertainties = U
    p.unc = uncs
    p.correlation = arr(pr.correlation)
    p.n_controls = len(pr.flow)
    p.flow = arr(pr.flow)
    p.unit_cost = pr.unit_cost
    p.profit_weight = arr(pr.profit_weight if pr.profit_weight is not None else np.zeros(U))
    p.traffic_index = pr.traffic_index
    p.n_steps = T
    p.time_step = pr.time_step
    p.rrr = pr.rrr
    p.discount = pr.discount
    p.n_paths = N
    p.basis_degree = pr.basis_degree
    p.standardize = pr.standardize
    p.itm_only = pr.itm_only
    p.pooled = pr.pooled
    p.hysteresis = pr.hysteresis
    p.mode = pr.mode
    p.value_mode = pr.value_mode
    p.strike = pr.strike
    p.seed = pr.seed
    p.states = arr(pr.states)
    p.controls = arr(pr.controls, np.int32)
    K = pr.basis_size()
    Cn = pr.n_controls
    value = np.zeros(1)
    vsd = np.zeros(1)
    fc = np.zeros(1, dtype=np.int32)
    coeffs = np.zeros((T, Cn, K))
    r = nat.RovResult()
    r.value = _dp(value)
    r.value_sd = _dp(vsd)
    r.first_control = fc.ctypes.data_as(C.POINTER(C.c_int32))
    r.coeffs = _dp(coeffs)
    out = {}
    if want_paths:
        cond = np.zeros((T + 1, N))
        optc = np.zeros((T, N), dtype=np.int32)
        states = np.zeros((T + 1, U, N))
        r.cond_exp = _dp(cond)
        r.opt_control = optc.ctypes.data_as(C.POINTER(C.c_int32))
        r.states = _dp(states)
        out.update(cond_exp=cond, opt_control=optc, states=states)
    nat.check(nat.lib().roadmc_rov_run(ctx.handle, C.byref(p), C.byref(r)), ctx.handle)
    out.update(value=float(value[0]), value_sd=float(vsd[0]), first_control=int(fc[0]),
               coeffs=coeffs)
    return out


def exp_pv(ctx, u: Uncertainty, n_steps: int, time_step: float, rrr: float, growth: float,
           n_paths: int, seed: int) -> tuple[float, float]:
    """SimulateGPU::expPV: mean and SD (1/N) of sum_t (1+g)^t x_t / (1+r)^t."""
    unc = nat.Uncertainty(u.kind, u.x0, u.mean, u.sd, u.reversion, u.jump_size, u.jump_prob)
    m = C.c_double()
    s = C.c_double()
    nat.check(nat.lib().roadmc_exp_pv(ctx.handle, C.byref(unc), n_steps, time_step, rrr, growth,
                                      n_paths, seed, C.byref(m), C.byref(s)), ctx.handle)
    return m.value, s.value


def config5(seed: int = 5, n_paths: int = 100000, n_steps: int = 50) -> RovProblem:
    """Two mean-reverting OU prices (fuel, commodity; kappa ~ U(0.1, 1),
    sigma ~ U(0.1, 0.3)), one GBM traffic index (mu 0.02, sigma 0.15), pairwise
    correlation 0.3, controls = flows {0, 150} (the reference's test programme
    uses {0, 75, 150}, tests/test_simple.cpp:213-216), cubic basis."""
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.1, 1.0, size=2)
    s = rng.uniform(0.1, 0.3, size=2)
    uncs = [Uncertainty(nat.SDE_OU, 1.0, 1.0, s[0], k[0]),
            Uncertainty(nat.SDE_OU, 1.0, 1.0, s[1], k[1]),
            Uncertainty(nat.SDE_GBM, 1.0, 0.02, 0.15)]
    corr = np.full((3, 3), 0.3)
    np.fill_diagonal(corr, 1.0)
    return RovProblem(uncs, correlation=corr, flow=[0.0, 150.0], unit_cost=0.2,
                      profit_weight=[1.0, -1.5, 0.0], traffic_index=2, n_steps=n_steps,
                      time_step=1.0, rrr=7.0, n_paths=n_paths, basis_degree=3, standardize=1,
                      itm_only=1, pooled=0, hysteresis=1.0, seed=seed)
