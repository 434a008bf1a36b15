"""Python host API over libroadmc, shaped like the reference's dispatch layer.

``Context.simulate_mte`` plays the role of ``SimulateGPU::simulateMTECUDA``
(include/gpuCode/SimulateGPU.h:152-156): it takes the species-road arrays and
returns the per-path end populations (the reference's ``endPops`` N x S
matrix) together with the statistics ``Simulator::simulateMTE`` derives from
them (Simulator.cpp:184-212).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as nat
from .landscape import RoadProblem, SpeciesData


@dataclass
class MteOutput:
    mte_prob: np.ndarray          # (S,)
    end_mean: np.ndarray          # (S,)
    end_sd: np.ndarray            # (S,)
    patch_mean: list              # S arrays of (P_s,)
    end_pops: np.ndarray | None   # (N, S) per-path end totals


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _Marshal:
    """Keeps the numpy buffers of one or more problems alive for a call."""

    def __init__(self, problems: list[RoadProblem]):
        self.keep = []
        self.species_arrays = []
        self.problems = (nat.MteProblem * len(problems))()
        for i, pr in enumerate(problems):
            sps = (nat.Species * len(pr.species))()
            for s, sd in enumerate(pr.species):
                sps[s] = self._species(sd)
            self.species_arrays.append(sps)
            p = self.problems[i]
            p.n_species = len(pr.species)
            p.species = sps
            p.n_years = pr.n_years
            p.time_step = pr.time_step
            p.n_paths = pr.n_paths
            p.growth_law = pr.growth_law
            p.endpoint = pr.endpoint
            p.precision = pr.precision
            p.drop_tol = pr.drop_tol
            p.seed = pr.seed
            p.road_id = pr.road_id

    def _arr(self, a, order="C"):
        if a is None:
            return None
        b = np.require(np.asarray(a, dtype=np.float64), requirements=[order])
        self.keep.append(b)
        return _dptr(b)

    def _species(self, sd: SpeciesData) -> nat.Species:
        s = nat.Species()
        s.n_patches = sd.n_patches
        s.init_pop = self._arr(sd.init_pop)
        s.capacity = self._arr(sd.capacity)
        s.growth = self._arr(sd.growth)
        # Column-major (Eigen) storage with (i, j) = origin i -> destination j.
        s.trans = self._arr(sd.trans, "F")
        s.surv = self._arr(sd.surv, "F")
        if sd.trans is None and sd.road is not None:
            from .road import road_struct
            rs = road_struct(sd.road, self.keep)
            s.road = C.pointer(rs)
            s.road_control = int(sd.road_control)
        for k in range(8):
            s.params[k] = float(sd.params[k])
        s.noise_sd = float(sd.noise_sd)
        s.threshold_ratio = float(sd.threshold_ratio)
        return s


class _Outputs:
    def __init__(self, problems: list[RoadProblem], end_pops: bool, patch_mean: bool):
        self.results = (nat.MteResult * len(problems))()
        self.host = []
        for i, pr in enumerate(problems):
            S = len(pr.species)
            h = {
                "mte_prob": np.full(S, np.nan),
                "end_mean": np.full(S, np.nan),
                "end_sd": np.full(S, np.nan),
                "patch_mean": np.full(sum(sd.n_patches for sd in pr.species), np.nan)
                if patch_mean else None,
                "end_pops": np.full(pr.n_paths * S, np.nan) if end_pops else None,
            }
            r = self.results[i]
            r.mte_prob = _dptr(h["mte_prob"])
            r.end_mean = _dptr(h["end_mean"])
            r.end_sd = _dptr(h["end_sd"])
            r.patch_mean = _dptr(h["patch_mean"]) if patch_mean else None
            r.end_pops = _dptr(h["end_pops"]) if end_pops else None
            self.host.append(h)
        self.problems = problems

    def outputs(self) -> list[MteOutput]:
        out = []
        for pr, h in zip(self.problems, self.host):
            S = len(pr.species)
            pm = []
            if h["patch_mean"] is not None:
                off = 0
                for sd in pr.species:
                    pm.append(h["patch_mean"][off:off + sd.n_patches].copy())
                    off += sd.n_patches
            ep = None
            if h["end_pops"] is not None:
                ep = h["end_pops"].reshape(S, pr.n_paths).T.copy()
            out.append(MteOutput(h["mte_prob"], h["end_mean"], h["end_sd"], pm, ep))
        return out


class Batch:
    """A prepared generation whose inputs sit in device memory."""

    def __init__(self, ctx: "Context", problems: list[RoadProblem]):
        self.ctx = ctx
        self.problems = problems
        m = _Marshal(problems)
        h = C.c_void_p()
        nat.check(nat.lib().roadmc_batch_prepare(ctx.handle, m.problems, len(problems),
                                                 C.byref(h)), ctx.handle)
        self.handle = h

    def launch(self) -> None:
        nat.check(nat.lib().roadmc_batch_launch(self.handle), self.ctx.handle)

    def sync(self) -> None:
        nat.check(nat.lib().roadmc_batch_sync(self.handle), self.ctx.handle)

    def kernel_ms(self) -> float:
        v = C.c_double()
        nat.check(nat.lib().roadmc_batch_kernel_ms(self.handle, C.byref(v)), self.ctx.handle)
        return v.value

    def fetch(self, end_pops: bool = True, patch_mean: bool = True) -> list[MteOutput]:
        o = _Outputs(self.problems, end_pops, patch_mean)
        nat.check(nat.lib().roadmc_batch_fetch(self.handle, o.results), self.ctx.handle)
        return o.outputs()

    def close(self) -> None:
        if self.handle:
            nat.lib().roadmc_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Context:
    """Owns a libroadmc context (one HIP stream and its device workspace)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = nat.lib().roadmc_ctx_create(device, C.byref(h))
        if rc != nat.OK:
            raise nat.RoadmcError(rc, f"roadmc_ctx_create(device={device}) failed")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            nat.lib().roadmc_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- MTE ------------------------------------------------------------------
    def simulate_mte(self, problem: RoadProblem, *, plain: bool = False, end_pops: bool = True,
                     patch_mean: bool = True) -> MteOutput:
        m = _Marshal([problem])
        o = _Outputs([problem], end_pops, patch_mean)
        fn = nat.lib().roadmc_mte_run_plain if plain else nat.lib().roadmc_mte_run
        nat.check(fn(self.handle, m.problems, o.results), self.handle)
        return o.outputs()[0]

    def simulate_mte_batch(self, problems: list[RoadProblem], *, end_pops: bool = False,
                           patch_mean: bool = True) -> list[MteOutput]:
        m = _Marshal(problems)
        o = _Outputs(problems, end_pops, patch_mean)
        nat.check(nat.lib().roadmc_mte_run_batch(self.handle, m.problems, len(problems),
                                                 o.results), self.handle)
        return o.outputs()

    def prepare_batch(self, problems: list[RoadProblem]) -> Batch:
        return Batch(self, problems)

    def autotune(self, n_paths: int = 20000, n_years: int = 10,
                 precision: int = nat.PREC_FP32) -> list[dict]:
        """Times every valid tile setting of each patch-count bucket
        (``roadmc_autotune_ex``), keeps the fastest in the context and
        returns the trials."""
        cap = 256
        trials = (nat.TileTrial * cap)()
        n = C.c_int32()
        nat.check(nat.lib().roadmc_autotune_ex(self.handle, precision, n_paths, n_years, trials,
                                               cap, C.byref(n)), self.handle)
        return [{"precision": t.precision, "bucket_max_p": t.bucket_max_p, "rows": t.rows,
                 "split": t.split, "per_cu": t.per_cu, "chosen": bool(t.chosen), "ms": t.ms}
                for t in trials[:min(n.value, cap)]]

    def set_tile(self, precision: int, bucket_max_p: int, rows: int, split: int = 1,
                 per_cu: int = 0) -> None:
        nat.check(nat.lib().roadmc_set_tile(self.handle, precision, bucket_max_p, rows, split,
                                            per_cu), self.handle)

    def tile(self, precision: int, bucket_max_p: int) -> dict:
        r, s_, pc = C.c_int32(), C.c_int32(), C.c_int32()
        nat.check(nat.lib().roadmc_get_tile(self.handle, precision, bucket_max_p, C.byref(r),
                                            C.byref(s_), C.byref(pc)), self.handle)
        return {"rows": r.value, "split": s_.value, "per_cu": pc.value}

    def tile_table(self, precision: int = nat.PREC_FP32) -> dict:
        return {b: self.tile(precision, b) for b in (128, 256, 512, 1024)}

    def tiles(self) -> dict:
        mx = (C.c_int32 * 8)()
        rows = (C.c_int32 * 8)()
        n = C.c_int32()
        nat.check(nat.lib().roadmc_get_tiles(self.handle, mx, rows, C.byref(n)), self.handle)
        return {int(mx[i]): int(rows[i]) for i in range(n.value)}

    def patch_normals(self, seed: int, road_id: int, species: int, n_paths: int, n_years: int,
                      n_patches: int) -> np.ndarray:
        out = np.empty(n_paths * n_years * n_patches, dtype=np.float32)
        nat.check(nat.lib().roadmc_patch_normals(
            self.handle, seed, road_id, species, n_paths, n_years, n_patches,
            out.ctypes.data_as(C.POINTER(C.c_float))), self.handle)
        return out.reshape(n_paths, n_years, n_patches)

    def fail_next_allocs(self, n: int) -> None:
        nat.check(nat.lib().roadmc_debug_fail_allocs(self.handle, n), self.handle)


def normal_cdf_inverse(p: float) -> float:
    """Utility::NormalCDFInverse (Utility.cpp:165-193); ValueError outside (0,1)."""
    v = C.c_double()
    rc = nat.lib().roadmc_normal_cdf_inverse(float(p), C.byref(v))
    if rc != nat.OK:
        raise ValueError(f"Invalid input argument ({p}); must be larger than 0 but less than 1.")
    return v.value


def mte_penalty(threshold, end_mean, end_sd, confidence_interval, initial_population,
                cost_per_animal) -> float:
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
            (threshold, end_mean, end_sd, initial_population, cost_per_animal)]
    out = C.c_double()
    nat.check(nat.lib().roadmc_mte_penalty(len(arrs[0]), _dptr(arrs[0]), _dptr(arrs[1]),
                                           _dptr(arrs[2]), float(confidence_interval),
                                           _dptr(arrs[3]), _dptr(arrs[4]), C.byref(out)))
    return out.value


def dense_to_csc(dense: np.ndarray, tol: float = 1e-5):
    """SimulateGPU::dense2Sparse on a (rows, cols) array (column-major walk)."""
    d = np.asfortranarray(dense, dtype=np.float32)
    rows, cols = d.shape
    val = np.zeros(rows * cols, dtype=np.float32)
    ri = np.zeros(rows * cols, dtype=np.int32)
    pc = np.zeros(cols, dtype=np.int32)
    nnz = C.c_int32()
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    nat.check(nat.lib().roadmc_dense_to_csc(d.ctypes.data_as(fp), rows, cols, tol,
                                            val.ctypes.data_as(fp), ri.ctypes.data_as(ip),
                                            pc.ctypes.data_as(ip), C.byref(nnz)))
    n = nnz.value
    return val[:n], ri[:n], pc
