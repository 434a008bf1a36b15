"""The reference's grid surrogate: kNN local-linear regression on a grid.

Host mirror of SimulateGPU::buildSurrogateMTECUDA / buildSurrogateROVCUDA
(include/gpuCode/SimulateGPU.cu:5533-5917) and of their evaluation,
Optimiser::evaluateSurrogateModelMTEML / evaluateSurrogateModelROVCR
(include/main/Optimiser.cpp:205-272) with Utility::interpolateSurrogate
(include/utilities/Utility.cpp:253-363).  Fit and evaluation run on the GPU
(``roadmc_knn_surrogate_fit`` / ``roadmc_knn_surrogate_eval``).

* MTE: per species, a 1-D grid over the road's initial animal-at-risk
  fraction (AAR, SpeciesRoadPatches::computeAAR, SpeciesRoadPatches.cpp:
  452-470), one grid for the mean end population and one for its SD.
* ROV: one (S+1)-D grid over the species' initial AARs and the road's unit
  operating cost, neighbours searched in [0,1]-scaled coordinates, values
  clamped to <= 0 (rovCorrection), again mean and SD.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as nat


def neighbour_count(n_samples: int, rule: int) -> int:
    """The reference's k for n_samples (``roadmc_knn_k``)."""
    k = C.c_int32()
    nat.check(nat.lib().roadmc_knn_k(n_samples, rule, C.byref(k)))
    return k.value


@dataclass
class GridSurrogate:
    """A fitted grid: surrogate vector [n_dims*dim_res + dim_res^n_dims]."""

    n_dims: int
    dim_res: int
    surrogate: np.ndarray

    @property
    def grid(self) -> list:
        R = self.dim_res
        return [self.surrogate[d * R:(d + 1) * R] for d in range(self.n_dims)]

    @property
    def values(self) -> np.ndarray:
        R = self.dim_res
        return self.surrogate[self.n_dims * R:].reshape((R,) * self.n_dims)

    def predict(self, ctx, points) -> np.ndarray:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, self.n_dims)
        out = np.empty(len(p))
        s = np.ascontiguousarray(self.surrogate)
        nat.check(nat.lib().roadmc_knn_surrogate_eval(
            ctx.handle, self.n_dims, self.dim_res, s.ctypes.data_as(nat._dp), len(p),
            p.ctypes.data_as(nat._dp), out.ctypes.data_as(nat._dp)), ctx.handle)
        return out


def fit(ctx, x, y, dim_res: int, k: int, *, normalize: bool = False,
        nonpositive: bool = False) -> GridSurrogate:
    """Fits y over the sample points x (n, d) on a dim_res^d grid."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    y = np.ascontiguousarray(y, dtype=np.float64)
    n, d = x.shape
    if y.shape != (n,):
        raise ValueError("y must have one value per sample")
    out = np.empty(d * dim_res + dim_res ** d)
    flags = (nat.KNN_NORMALIZE if normalize else 0) | (nat.KNN_NONPOSITIVE if nonpositive else 0)
    nat.check(nat.lib().roadmc_knn_surrogate_fit(
        ctx.handle, n, d, x.ctypes.data_as(nat._dp), y.ctypes.data_as(nat._dp), dim_res, k,
        flags, out.ctypes.data_as(nat._dp)), ctx.handle)
    return GridSurrogate(d, dim_res, out)


def build_surrogate_mte(ctx, iars, pops, pops_sd, dim_res: int = 20):
    """buildSurrogateMTECUDA for every species: iars, pops, pops_sd are
    (n_samples, n_species).  Returns [(mean grid, SD grid)] per species."""
    iars = np.asarray(iars, dtype=np.float64)
    n = iars.shape[0]
    k = neighbour_count(n, nat.KNN_RULE_MTE)
    return [(fit(ctx, iars[:, s], np.asarray(pops)[:, s], dim_res, k),
             fit(ctx, iars[:, s], np.asarray(pops_sd)[:, s], dim_res, k))
            for s in range(iars.shape[1])]


def build_surrogate_rov(ctx, iars, unit_cost, values, values_sd, dim_res: int = 20):
    """buildSurrogateROVCUDA: predictors (AAR of each species, unit cost),
    neighbours in [0,1]-scaled coordinates, values clamped to <= 0.
    Returns (mean grid, SD grid)."""
    iars = np.asarray(iars, dtype=np.float64).reshape(len(unit_cost), -1)
    x = np.column_stack([iars, np.asarray(unit_cost, dtype=np.float64)])
    k = neighbour_count(len(x), nat.KNN_RULE_ROV)
    mean = fit(ctx, x, values, dim_res, k, normalize=True, nonpositive=True)
    sd = fit(ctx, x, values_sd, dim_res, k, normalize=True)
    return mean, sd


def initial_aar(trans: np.ndarray, surv: np.ndarray, init_pop: np.ndarray) -> float:
    """SpeciesRoadPatches::computeAAR for one control: 1 - sum(R^T n) / sum(n),
    R = trans * surv."""
    n = np.asarray(init_pop, dtype=np.float64)
    return float(1.0 - (n @ (trans * surv)).sum() / n.sum())
