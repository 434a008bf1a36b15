"""Road -> crossing counts, transition and survival matrices on the GPU.

Host mirror of the reference's per-species road model
(SpeciesRoadPatches::createSpeciesModel, include/road/SpeciesRoadPatches.cpp:
12-60): ``habitatPatchDistances``, ``roadCrossings``,
``computeTransitionProbabilities`` and ``computeSurvivalProbabilities``
(:270-451) run as one call of ``roadmc_build_species_road``, whose kernels
(csrc/road_kernels.hip) evaluate the segment test of
Utility::lineSegmentIntersect (include/utilities/Utility.cpp:3-91) in FP64.

The results feed the Monte Carlo path: ``SpeciesRoadPatches.to_species``
returns the ``SpeciesData`` that ``Context.simulate_mte`` takes, with the
survival matrix of the chosen flow control.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native as nat
from .landscape import SpeciesData


@dataclass
class RoadSpeciesInput:
    """What SpeciesRoadPatches reads for one species and one road."""

    x: np.ndarray                  # (P,) patch centroids
    y: np.ndarray                  # (P,)
    area: np.ndarray               # (P,)
    pref_mean: np.ndarray          # (P,) habitat preference mean of each patch's type
    pref_sd: np.ndarray            # (P,)
    road: np.ndarray               # (n_points, 2) road vertices
    road_type: np.ndarray | None = None   # (n_points - 1,) SEG_*; None: all open road
    flows: np.ndarray = field(default_factory=lambda: np.array([100.0]))  # vehicles / hour
    pref_mult: float = 0.0         # scenario habitat-preference multiplier
    lambda_mean: float = 1.0
    lambda_sd: float = 0.0
    lambda_mult: float = 0.0
    vehicle_width: float = 2.0     # sum of vehicle width x proportion
    animal_length: float = 0.5
    animal_speed: float = 1.0
    transition_form: int = nat.TRANS_REFERENCE
    drop_last_link: bool = True

    @property
    def n_patches(self) -> int:
        return int(np.asarray(self.x).shape[0])


@dataclass
class RoadMatrices:
    crossings: np.ndarray | None   # (P, P) int32, symmetric
    trans: np.ndarray | None       # (P, P) row = origin, rows sum to 1
    surv: np.ndarray | None        # (C, P, P)


def _d(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def road_struct(inp: RoadSpeciesInput, keep: list) -> "nat.RoadSpecies":
    """The ``roadmc_road_species`` struct for ``inp``; the arrays it points
    into are appended to ``keep``, which must outlive every call using it."""
    P = inp.n_patches
    pa = [_d(inp.x), _d(inp.y), _d(inp.area), _d(inp.pref_mean), _d(inp.pref_sd)]
    for a in pa:
        if a.shape != (P,):
            raise ValueError("patch arrays must all have shape (P,)")
    road = _d(inp.road).reshape(-1, 2)
    rx, ry = _d(road[:, 0]), _d(road[:, 1])
    flows = _d(inp.flows).reshape(-1)
    rt = None
    if inp.road_type is not None:
        rt = np.ascontiguousarray(inp.road_type, dtype=np.int32)
        if rt.shape != (max(len(road) - 1, 0),):
            raise ValueError("road_type needs one entry per segment")
    keep.extend(pa + [rx, ry, flows] + ([rt] if rt is not None else []))
    rs = nat.RoadSpecies()
    rs.n_patches = P
    rs.x, rs.y, rs.area, rs.pref_mean, rs.pref_sd = (a.ctypes.data_as(nat._dp) for a in pa)
    rs.pref_mult = float(inp.pref_mult)
    rs.lambda_mean = float(inp.lambda_mean)
    rs.lambda_sd = float(inp.lambda_sd)
    rs.lambda_mult = float(inp.lambda_mult)
    rs.n_road_points = len(road)
    rs.road_x = rx.ctypes.data_as(nat._dp)
    rs.road_y = ry.ctypes.data_as(nat._dp)
    rs.road_type = rt.ctypes.data_as(nat._ip) if rt is not None else None
    rs.n_controls = len(flows)
    rs.flow = flows.ctypes.data_as(nat._dp) if len(flows) else None
    rs.vehicle_width = float(inp.vehicle_width)
    rs.animal_length = float(inp.animal_length)
    rs.animal_speed = float(inp.animal_speed)
    rs.transition_form = int(inp.transition_form)
    rs.drop_last_link = 1 if inp.drop_last_link else 0
    keep.append(rs)
    return rs


def build_species_road(ctx, inp: RoadSpeciesInput, *, crossings: bool = True,
                       trans: bool = True, surv: bool = True) -> RoadMatrices:
    """One call of ``roadmc_build_species_road`` on ``ctx``'s device."""
    P = inp.n_patches
    keep: list = []
    rs = road_struct(inp, keep)
    flows = keep[7]
    cr = np.empty(P * P, dtype=np.int32) if crossings else None
    tr = np.empty(P * P) if trans else None
    sv = np.empty(len(flows) * P * P) if surv else None
    ptr = lambda a, t: a.ctypes.data_as(t) if a is not None else None  # noqa: E731
    nat.check(nat.lib().roadmc_build_species_road(ctx.handle, C.byref(rs), ptr(cr, nat._ip),
                                                  ptr(tr, nat._dp), ptr(sv, nat._dp)),
              ctx.handle)
    # Column-major views: no copies (surv[c] is Fortran-ordered, as the MTE
    # path marshals it).
    return RoadMatrices(
        crossings=cr.reshape(P, P, order="F") if cr is not None else None,
        trans=tr.reshape(P, P, order="F") if tr is not None else None,
        surv=sv.reshape(len(flows), P, P).transpose(0, 2, 1) if sv is not None else None)


class SpeciesRoadPatches:
    """The reference's per-species road model, built on the GPU.

    ``create_species_model`` mirrors SpeciesRoadPatches::createSpeciesModel
    (SpeciesRoadPatches.cpp:12-60): distances, crossings, transitions and
    survival in one device call.  Attribute names follow the reference's
    getters (getCrossings, getTransProbs, getSurvivalProbs).
    """

    def __init__(self, ctx, inp: RoadSpeciesInput):
        self.ctx = ctx
        self.inp = inp
        self.crossings: np.ndarray | None = None
        self.trans_probs: np.ndarray | None = None
        self.surv_probs: list[np.ndarray] = []

    def create_species_model(self) -> None:
        m = build_species_road(self.ctx, self.inp)
        self.crossings = m.crossings
        self.trans_probs = m.trans
        self.surv_probs = list(m.surv)

    def to_species(self, control: int, init_pop: np.ndarray, capacity: np.ndarray,
                   growth: np.ndarray, **kw) -> SpeciesData:
        """SpeciesData for the Monte Carlo path under flow control ``control``."""
        if self.trans_probs is None:
            self.create_species_model()
        return SpeciesData(trans=self.trans_probs, surv=self.surv_probs[control],
                           init_pop=np.asarray(init_pop, dtype=np.float64),
                           capacity=np.asarray(capacity, dtype=np.float64),
                           growth=np.asarray(growth, dtype=np.float64), **kw)

    def device_species(self, control: int, init_pop: np.ndarray, capacity: np.ndarray,
                       growth: np.ndarray, **kw) -> SpeciesData:
        """SpeciesData that carries the road instead of matrices: the batched
        Monte Carlo path builds trans and the survival of ``control`` on the
        device and packs them straight into its movement operand, so no
        matrix crosses PCIe (``control`` -1: the last)."""
        return SpeciesData(trans=None, surv=None, road=self.inp, road_control=int(control),
                           init_pop=np.asarray(init_pop, dtype=np.float64),
                           capacity=np.asarray(capacity, dtype=np.float64),
                           growth=np.asarray(growth, dtype=np.float64), **kw)


def random_road_species(rng: np.random.Generator, n_patches: int, *, side_km: float = 10.0,
                        n_road_points: int = 40, n_controls: int = 3,
                        bridge_fraction: float = 0.1) -> RoadSpeciesInput:
    """A seeded synthetic species and road: uniform centroids in a square of
    ``side_km``, a random-walk road across it, a share of bridge/tunnel
    segments, lambda chosen so the crossing cut-off spans a few km."""
    xy = rng.uniform(0.0, side_km, size=(n_patches, 2))
    t = np.linspace(0.0, 1.0, n_road_points)
    y0, y1 = rng.uniform(0.1 * side_km, 0.9 * side_km, size=2)
    road = np.stack([t * side_km,
                     y0 + (y1 - y0) * t + rng.normal(0.0, 0.05 * side_km, n_road_points)], 1)
    rt = np.where(rng.uniform(size=n_road_points - 1) < bridge_fraction,
                  rng.integers(1, 3, size=n_road_points - 1), nat.SEG_ROAD).astype(np.int32)
    n_types = 4
    ptype = rng.integers(0, n_types, size=n_patches)
    pm = rng.uniform(-1.0, 1.0, n_types)[ptype]
    ps = rng.uniform(0.0, 0.5, n_types)[ptype]
    return RoadSpeciesInput(
        x=xy[:, 0].copy(), y=xy[:, 1].copy(), area=rng.uniform(0.5, 2.0, n_patches),
        pref_mean=pm, pref_sd=ps, road=road, road_type=rt,
        flows=rng.uniform(10.0, 400.0, n_controls), pref_mult=float(rng.normal()),
        lambda_mean=1.0, lambda_sd=0.1, lambda_mult=float(rng.normal()),
        vehicle_width=1.9, animal_length=0.7, animal_speed=2.5)
