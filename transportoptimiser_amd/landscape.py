"""Synthetic landscapes and roads for the benchmark configurations.

The reference builds these arrays from habitat maps and a designed road
(SpeciesRoadPatches::createSpeciesModel, include/road/SpeciesRoadPatches.cpp:
12-60, 383-451); that builder is upstream of the Monte Carlo path.  Here the
same arrays come from a seeded generator:

* patch centroids uniform in a square of ``side_km``;
* transition matrix T_ij = exp(-d_ij / length_scale), row-normalised (rows
  sum to 1 like ``transProbs``, SpeciesRoadPatches.h:302);
* a random polyline road; a link (i, j) is crossed when the segment between
  the two centroids intersects the road (the segment test of
  SimulateGPU.cu:465-476); crossing mortality m_ij ~ U(0, mort_max) on crossed
  links, so survival S = 1 - m;
* growth r ~ U(0.02, 0.10), capacity K ~ U(500, 5000), n0 = 0.8 K.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import native as nat


@dataclass
class SpeciesData:
    """One species on one road: the data SpeciesRoadPatches exposes."""

    trans: np.ndarray | None     # (P, P) row = origin, rows sum to 1
    surv: np.ndarray | None      # (P, P) survival for the control in force
    init_pop: np.ndarray         # (P,)
    capacity: np.ndarray         # (P,)
    growth: np.ndarray | None    # (P,)
    params: np.ndarray = field(default_factory=lambda: np.zeros(8))
    noise_sd: float = 0.1
    threshold_ratio: float = 0.5
    # With trans None: the road (a road.RoadSpeciesInput) from which the
    # library builds trans and the survival of road_control on the device
    # (-1: the last control).
    road: object = None
    road_control: int = -1

    @property
    def n_patches(self) -> int:
        return int(self.init_pop.shape[0])

    def movement(self, drop_tol: float = 0.0) -> np.ndarray:
        m = self.trans * (self.surv if self.surv is not None else 1.0)
        if drop_tol > 0:
            m = np.where(m.astype(np.float32) > np.float32(drop_tol), m, 0.0)
        return m


@dataclass
class RoadProblem:
    species: list[SpeciesData]
    n_years: int = 50
    time_step: float = 1.0
    n_paths: int = 10000
    growth_law: int = nat.LAW_LOGNORMAL
    endpoint: int = nat.END_POST_GROWTH
    precision: int = nat.PREC_FP32
    drop_tol: float = 0.0
    seed: int = 0
    road_id: int = 0


def _segments_cross(p1, p2, q1, q2) -> np.ndarray:
    """Proper intersection of segments p1p2 (arrays [..., 2]) with q1q2."""

    def orient(a, b, c):
        return (b[..., 0] - a[..., 0]) * (c[..., 1] - a[..., 1]) - \
               (b[..., 1] - a[..., 1]) * (c[..., 0] - a[..., 0])

    d1 = orient(q1, q2, p1)
    d2 = orient(q1, q2, p2)
    d3 = orient(p1, p2, q1)
    d4 = orient(p1, p2, q2)
    return (d1 * d2 < 0) & (d3 * d4 < 0)


def _edge_point(rng: np.random.Generator, side_km: float, edge: int) -> np.ndarray:
    t = rng.uniform(0.0, side_km)
    return np.array([[t, 0.0], [side_km, t], [t, side_km], [0.0, t]][edge])


def random_polyline(rng: np.random.Generator, side_km: float, n_vertices: int = 5) -> np.ndarray:
    """A road entering through one edge of the square and leaving through
    another, with n_vertices - 2 interior bends."""
    e0 = int(rng.integers(0, 4))
    e1 = (e0 + int(rng.integers(1, 4))) % 4
    a, b = _edge_point(rng, side_km, e0), _edge_point(rng, side_km, e1)
    t = np.sort(rng.uniform(0.0, 1.0, size=n_vertices - 2))
    mid = a[None, :] + t[:, None] * (b - a)[None, :]
    mid += rng.normal(0.0, 0.1 * side_km, size=mid.shape)
    mid = np.clip(mid, 0.0, side_km)
    return np.concatenate([a[None, :], mid, b[None, :]], axis=0)


def _cross2(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    return u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]


def crossed_links(xy: np.ndarray, road: np.ndarray) -> np.ndarray:
    """Boolean (P, P) matrix: the centroid link i-j properly crosses the road.

    Uses orient(a, b, c) = a x b + b x c + c x a (2-D cross products), so the
    tests against a fixed road vertex are an outer sum of per-patch terms and
    one (P, P) matrix of centroid cross products."""
    cxy = np.outer(xy[:, 0], xy[:, 1])
    C = cxy - cxy.T                         # xy_i x xy_j
    cut = np.zeros(C.shape, dtype=bool)
    for s in range(road.shape[0] - 1):
        q1, q2 = road[s], road[s + 1]
        side = _cross2(q2 - q1, xy - q1)    # orient(q1, q2, p_i)
        w1 = _cross2(xy, q1[None, :])       # xy_i x q1
        w2 = _cross2(xy, q2[None, :])
        d3 = C + w1[None, :] - w1[:, None]  # orient(p_i, p_j, q1)
        d4 = C + w2[None, :] - w2[:, None]
        cut |= (np.multiply.outer(side, side) < 0) & (d3 * d4 < 0)
    return cut


def _sample_cut_fraction(xy, road, idx_i, idx_j) -> float:
    a, b = xy[idx_i], xy[idx_j]
    hit = np.zeros(idx_i.shape[0], dtype=bool)
    for s in range(road.shape[0] - 1):
        hit |= _segments_cross(a, b, road[s], road[s + 1])
    return float(hit.mean())


def make_species(rng: np.random.Generator, n_patches: int, *, side_km: float = 10.0,
                 length_scale_km: float = 2.0, mort_max: float = 0.3,
                 cut_target: float = 0.3, tries: int = 16, noise_sd: float = 0.1,
                 threshold_ratio: float = 0.5) -> SpeciesData:
    P = int(n_patches)
    xy = rng.uniform(0.0, side_km, size=(P, 2))
    d = np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1))
    trans = np.exp(-d / length_scale_km)
    trans /= trans.sum(axis=1, keepdims=True)
    # Pick, among a few random roads, the one whose crossed-link fraction
    # (estimated on a sample of links) is closest to the target.
    idx_i = rng.integers(0, P, size=4096)
    idx_j = rng.integers(0, P, size=4096)
    best_road, best_err = None, np.inf
    for _ in range(tries):
        road = random_polyline(rng, side_km)
        err = abs(_sample_cut_fraction(xy, road, idx_i, idx_j) - cut_target)
        if err < best_err:
            best_road, best_err = road, err
    cut = crossed_links(xy, best_road)
    mort = np.where(cut, rng.uniform(0.0, mort_max, size=(P, P)), 0.0)
    surv = 1.0 - mort
    growth = rng.uniform(0.02, 0.10, size=P)
    cap = rng.uniform(500.0, 5000.0, size=P)
    init = 0.8 * cap
    params = np.array([0.05, 0.05, 0.1, threshold_ratio, 0.5, 0.05, 0.1, 0.2])
    # Column-major, the layout of the C ABI's matrices (Eigen's default, in
    # which the reference keeps them): passed to the library without a copy.
    return SpeciesData(trans=np.asfortranarray(trans), surv=np.asfortranarray(surv),
                       init_pop=init, capacity=cap, growth=growth, params=params,
                       noise_sd=noise_sd, threshold_ratio=threshold_ratio)


def config1(seed: int = 1, n_paths: int = 10000, n_years: int = 50) -> RoadProblem:
    """1 species, 32 patches, l = 2 km, mortality U(0, 0.3), theta = 0.5."""
    rng = np.random.default_rng(seed)
    sp = make_species(rng, 32, length_scale_km=2.0, mort_max=0.3, threshold_ratio=0.5)
    return RoadProblem([sp], n_years=n_years, n_paths=n_paths, seed=seed, road_id=0)


def config2(seed: int = 2, n_paths: int = 100000, n_years: int = 100) -> RoadProblem:
    """3 species x 256 patches, l = 1.5 km, mortality U(0, 0.4), theta = 0.3."""
    rng = np.random.default_rng(seed)
    sps = [make_species(rng, 256, length_scale_km=1.5, mort_max=0.4, threshold_ratio=0.3)
           for _ in range(3)]
    return RoadProblem(sps, n_years=n_years, n_paths=n_paths, seed=seed, road_id=0)


def generation_sizes(rng: np.random.Generator, n_roads: int, p_lo: int = 64,
                     p_hi: int = 1024) -> np.ndarray:
    """Patch counts log-uniform in [p_lo, p_hi]."""
    return np.rint(np.exp(rng.uniform(np.log(p_lo), np.log(p_hi), size=n_roads))).astype(int)


def _road(args) -> RoadProblem:
    seed, r, P, n_paths, n_years, n_species = args
    rng = np.random.default_rng([seed, r])
    sps = [make_species(rng, int(P), length_scale_km=2.0, mort_max=0.3, threshold_ratio=0.5)
           for _ in range(n_species)]
    return RoadProblem(sps, n_years=n_years, n_paths=n_paths, seed=seed, road_id=r)


def config3(seed: int = 3, n_roads: int = 512, n_paths: int = 20000, n_years: int = 50,
            n_species: int = 2, workers: int = 1) -> list[RoadProblem]:
    """A GA generation: patch counts log-uniform in [64, 1024], random polyline
    roads, 2 species, 50 years, 20,000 paths.  Road r draws from its own
    generator seeded (seed, r), so roads can be built in parallel."""
    sizes = generation_sizes(np.random.default_rng(seed), n_roads)
    args = [(seed, r, int(P), n_paths, n_years, n_species) for r, P in enumerate(sizes)]
    if workers > 1:
        import multiprocessing as mp
        # spawned, not forked: the caller may already hold an initialised
        # HIP runtime, which a forked child must not inherit.  A calling
        # script therefore needs the usual `if __name__ == "__main__":` guard.
        with mp.get_context("spawn").Pool(workers) as pool:
            return pool.map(_road, args, chunksize=4)
    return [_road(a) for a in args]
