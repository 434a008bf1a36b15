"""ctypes binding of libroadmc (include/roadmc.h).

The library is loaded from ``transportoptimiser_amd/lib/libroadmc.so``; if it
is missing, :func:`lib` raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libroadmc.so"

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_OOM = 3
ERR_NUMERIC = 4
ERR_UNSUPPORTED = 5

ROVS_ALGO4 = 4
ROVS_ALGO6 = 6

LAW_LOGNORMAL = 0
LAW_REF_CPU = 1
LAW_REF_GPU = 2
END_POST_GROWTH = 0
END_PRE_GROWTH = 1
PREC_FP32 = 0
PREC_FP64 = 1
PREC_BF16X3 = 2  # split-bf16 products, tiled MTE kernels only
MAX_PATCHES = 1024

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class Species(C.Structure):
    pass  # fields set below, after RoadSpecies


class MteProblem(C.Structure):
    _fields_ = [
        ("n_species", C.c_int32),
        ("species", C.POINTER(Species)),
        ("n_years", C.c_int32),
        ("time_step", C.c_double),
        ("n_paths", C.c_int64),
        ("growth_law", C.c_int32),
        ("endpoint", C.c_int32),
        ("precision", C.c_int32),
        ("drop_tol", C.c_double),
        ("seed", C.c_uint64),
        ("road_id", C.c_uint32),
    ]


class MteResult(C.Structure):
    _fields_ = [
        ("mte_prob", _dp),
        ("end_mean", _dp),
        ("end_sd", _dp),
        ("patch_mean", _dp),
        ("end_pops", _dp),
    ]


SDE_MRJD = 0
SDE_OU = 1
SDE_GBM = 2
ROV_SWITCHING = 0
ROV_AMERICAN = 1
VALUE_REGRESSED = 0
VALUE_REALIZED = 1
_i32p = C.POINTER(C.c_int32)


class Uncertainty(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("x0", C.c_double),
        ("mean", C.c_double),
        ("sd", C.c_double),
        ("reversion", C.c_double),
        ("jump_size", C.c_double),
        ("jump_prob", C.c_double),
    ]


class RovProblem(C.Structure):
    _fields_ = [
        ("n_uncertainties", C.c_int32),
        ("unc", C.POINTER(Uncertainty)),
        ("correlation", _dp),
        ("n_controls", C.c_int32),
        ("flow", _dp),
        ("unit_cost", C.c_double),
        ("profit_weight", _dp),
        ("traffic_index", C.c_int32),
        ("n_steps", C.c_int32),
        ("time_step", C.c_double),
        ("rrr", C.c_double),
        ("discount", C.c_double),
        ("n_paths", C.c_int64),
        ("basis_degree", C.c_int32),
        ("standardize", C.c_int32),
        ("itm_only", C.c_int32),
        ("pooled", C.c_int32),
        ("hysteresis", C.c_double),
        ("mode", C.c_int32),
        ("value_mode", C.c_int32),
        ("strike", C.c_double),
        ("seed", C.c_uint64),
        ("states", _dp),
        ("controls", _i32p),
    ]


class RovResult(C.Structure):
    _fields_ = [
        ("value", _dp),
        ("value_sd", _dp),
        ("first_control", _i32p),
        ("coeffs", _dp),
        ("cond_exp", _dp),
        ("opt_control", _i32p),
        ("states", _dp),
    ]


SEG_ROAD = 0
SEG_BRIDGE = 1
SEG_TUNNEL = 2
TRANS_REFERENCE = 0
TRANS_DECAY = 1
KNN_NORMALIZE = 1
KNN_NONPOSITIVE = 2
KNN_RULE_MTE = 0
KNN_RULE_ROV = 1
KNN_MAX_K = 128
NOALLOC_MAX_PATHS = 1 << 20


class RoadSpecies(C.Structure):
    _fields_ = [
        ("n_patches", C.c_int32),
        ("x", _dp),
        ("y", _dp),
        ("area", _dp),
        ("pref_mean", _dp),
        ("pref_sd", _dp),
        ("pref_mult", C.c_double),
        ("lambda_mean", C.c_double),
        ("lambda_sd", C.c_double),
        ("lambda_mult", C.c_double),
        ("n_road_points", C.c_int32),
        ("road_x", _dp),
        ("road_y", _dp),
        ("road_type", _ip),
        ("n_controls", C.c_int32),
        ("flow", _dp),
        ("vehicle_width", C.c_double),
        ("animal_length", C.c_double),
        ("animal_speed", C.c_double),
        ("transition_form", C.c_int32),
        ("drop_last_link", C.c_int32),
    ]


class RovsSpecies(C.Structure):
    _fields_ = [
        ("init_pop", _dp),
        ("capacity", _dp),
        ("trans", _dp),
        ("surv", _dp),
        ("params", C.c_double * 8),
    ]


class RovsProblem(C.Structure):
    _fields_ = [
        ("n_patches", C.c_int32),
        ("n_species", C.c_int32),
        ("species", C.POINTER(RovsSpecies)),
        ("n_controls", C.c_int32),
        ("flow", _dp),
        ("n_fuels", C.c_int32),
        ("n_commodities", C.c_int32),
        ("uncertainties", C.POINTER(Uncertainty)),
        ("ore_mean", _dp),
        ("ore_sd", _dp),
        ("n_vehicles", C.c_int32),
        ("fuel_cost", _dp),
        ("fuel_idx", _ip),
        ("unit_cost", C.c_double),
        ("unit_revenue", C.c_double),
        ("n_years", C.c_int32),
        ("time_step", C.c_double),
        ("rrr", C.c_double),
        ("n_paths", C.c_int64),
        ("dim_res", C.c_int32),
        ("profit_scaling", C.c_double),
        ("hysteresis", C.c_double),
        ("seed", C.c_uint64),
        ("road_id", C.c_uint32),
        ("method", C.c_int32),
        ("checkpoint_years", C.c_int32),
        ("precision", C.c_int32),
    ]


class RovsResult(C.Structure):
    _fields_ = [
        ("value", _dp),
        ("value_sd", _dp),
        ("first_control", _ip),
        ("cond_exp", _dp),
        ("opt_control", _ip),
        ("adj_pops", _dp),
        ("unit_profit", _dp),
        ("controls", _ip),
        ("total_pops", _dp),
        ("aars", _dp),
        ("grids", _dp),
    ]


SURR_MTE = 0
SURR_CONTROLLED = 1
NAME_LEN = 64
MAX_HABITAT_TYPES = 8


class TileTrial(C.Structure):
    _fields_ = [("precision", C.c_int32), ("bucket_max_p", C.c_int32), ("rows", C.c_int32),
                ("split", C.c_int32), ("per_cu", C.c_int32), ("chosen", C.c_int32),
                ("ms", C.c_double)]


class CommodityFile(C.Structure):
    _fields_ = [("name", C.c_char * NAME_LEN)] + [
        (k, C.c_double) for k in ("current", "mean", "sd", "reversion", "jump_prob",
                                  "poisson_jump", "ore_mean", "ore_sd")]


class VehicleFile(C.Structure):
    _fields_ = [("name", C.c_char * NAME_LEN), ("width", C.c_double), ("length", C.c_double),
                ("proportion", C.c_double), ("load", C.c_double), ("fuel", C.c_int32),
                ("a", C.c_double), ("agr", C.c_double), ("av", C.c_double),
                ("avsq", C.c_double), ("travel_cost", C.c_double)]


class SpeciesFile(C.Structure):
    _fields_ = [("name", C.c_char * NAME_LEN), ("n_habitat_types", C.c_int32),
                ("habitat_name", (C.c_char * 16) * MAX_HABITAT_TYPES),
                ("max_pop", C.c_double * MAX_HABITAT_TYPES),
                ("cost_per_m2", C.c_double * MAX_HABITAT_TYPES),
                ("pref_mean", C.c_double * MAX_HABITAT_TYPES),
                ("pref_sd", C.c_double * MAX_HABITAT_TYPES)] + [
        (k, C.c_double) for k in ("growth_current", "growth_mean", "growth_sd",
                                  "growth_reversion", "growth_jump_prob", "growth_poisson_jump")]


MAX_LIST, MAX_PROGRAMS, MAX_FILES, PATH_LEN, MAX_CONTROLS = 16, 8, 16, 256, 8


class RoadmcList(C.Structure):
    _fields_ = [("n", C.c_int32), ("v", C.c_double * MAX_LIST)]

    def values(self) -> list:
        return [self.v[i] for i in range(self.n)]


RUN_LISTS = ("animal_bridges", "hab_pref_sd", "movement_sd", "ranging_sd", "growth_mean",
             "growth_sd", "price_mean", "price_sd", "ore_sd", "pop_level", "comparison_cost")


class RunSettings(C.Structure):
    _fields_ = [
        ("confidence_interval", C.c_double), ("use_gpu", C.c_int32),
        ("surrogate_dim_res", C.c_int32), ("interpolation", C.c_int32),
        ("traffic_growth", C.c_double), ("rrr", C.c_double), ("n_years", C.c_int32),
        ("time_step", C.c_double), ("n_paths", C.c_int32), ("rov_dim_res", C.c_int32),
        ("optimisation_method", C.c_int32), ("rov_method", C.c_int32),
    ] + [(k, RoadmcList) for k in RUN_LISTS] + [
        ("n_programs", C.c_int32), ("n_flows", C.c_int32 * MAX_PROGRAMS),
        ("flows", (C.c_double * MAX_CONTROLS) * MAX_PROGRAMS),
        ("n_switching", C.c_int32 * MAX_PROGRAMS),
        ("switching", (C.c_double * (1 << MAX_CONTROLS)) * MAX_PROGRAMS),
        ("bridge", C.c_int32 * MAX_PROGRAMS),
        ("n_fuels", C.c_int32), ("n_transported", C.c_int32),
        ("n_commodities", C.c_int32), ("n_vehicles", C.c_int32), ("n_species", C.c_int32),
        ("commodity", (C.c_char * PATH_LEN) * MAX_FILES),
        ("vehicle", (C.c_char * PATH_LEN) * MAX_FILES),
        ("species", (C.c_char * PATH_LEN) * MAX_FILES),
        ("road_exists", C.c_int32), ("travel_time", C.c_double), ("road_length", C.c_double),
        ("accident_cost", C.c_double), ("fuel_per_journey", RoadmcList),
    ]


Species._fields_ = [
    ("n_patches", C.c_int32),
    ("init_pop", _dp),
    ("capacity", _dp),
    ("growth", _dp),
    ("trans", _dp),
    ("surv", _dp),
    ("params", C.c_double * 8),
    ("noise_sd", C.c_double),
    ("threshold_ratio", C.c_double),
    ("road", C.POINTER(RoadSpecies)),
    ("road_control", C.c_int32),
]


# Every symbol include/roadmc.h declares: name -> (restype, argtypes).
SIGNATURES: dict[str, tuple] = {
    "roadmc_device_count": (C.c_int, [_ip]),
    "roadmc_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "roadmc_ctx_destroy": (None, [C.c_void_p]),
    "roadmc_last_error": (C.c_char_p, [C.c_void_p]),
    "roadmc_debug_fail_allocs": (C.c_int, [C.c_void_p, C.c_int]),
    "roadmc_mte_run": (C.c_int, [C.c_void_p, C.POINTER(MteProblem), C.POINTER(MteResult)]),
    "roadmc_mte_run_plain": (C.c_int, [C.c_void_p, C.POINTER(MteProblem), C.POINTER(MteResult)]),
    "roadmc_mte_run_batch": (C.c_int, [C.c_void_p, C.POINTER(MteProblem), C.c_int32,
                                       C.POINTER(MteResult)]),
    "roadmc_batch_prepare": (C.c_int, [C.c_void_p, C.POINTER(MteProblem), C.c_int32,
                                       C.POINTER(C.c_void_p)]),
    "roadmc_batch_launch": (C.c_int, [C.c_void_p]),
    "roadmc_batch_sync": (C.c_int, [C.c_void_p]),
    "roadmc_batch_fetch": (C.c_int, [C.c_void_p, C.POINTER(MteResult)]),
    "roadmc_batch_kernel_ms": (C.c_int, [C.c_void_p, _dp]),
    "roadmc_batch_destroy": (None, [C.c_void_p]),
    "roadmc_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "roadmc_autotune_ex": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(TileTrial), C.c_int32, _ip]),
    "roadmc_set_tile": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32]),
    "roadmc_get_tile": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _ip, _ip]),
    "roadmc_get_tiles": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "roadmc_dense_to_csc": (C.c_int, [_fp, C.c_int32, C.c_int32, C.c_float, _fp, _ip, _ip, _ip]),
    "roadmc_normal_cdf_inverse": (C.c_int, [C.c_double, _dp]),
    "roadmc_mte_penalty": (C.c_int, [C.c_int32, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]),
    "roadmc_patch_normals": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.c_int32, C.c_int32, C.c_int32, _fp]),
    "roadmc_build_species_road": (C.c_int, [C.c_void_p, C.POINTER(RoadSpecies), _ip, _dp,
                                            _dp]),
    "roadmc_knn_k": (C.c_int, [C.c_int32, C.c_int32, _ip]),
    "roadmc_knn_surrogate_fit": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, C.c_int32,
                                           C.c_int32, C.c_int32, _dp]),
    "roadmc_knn_surrogate_eval": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int32,
                                            _dp, _dp]),
    "roadmc_rovs_run": (C.c_int, [C.c_void_p, C.POINTER(RovsProblem), C.POINTER(RovsResult)]),
    "roadmc_rovs_smooth_grad": (C.c_int, [C.c_void_p, C.POINTER(RovsProblem), _ip, C.c_double,
                                          _dp, _dp]),
    "roadmc_surrogate_data_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _dp,
                                              C.c_double, _ip]),
    "roadmc_surrogate_data_read": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _dp,
                                             _ip]),
    "roadmc_read_commodity": (C.c_int, [C.c_char_p, C.POINTER(CommodityFile)]),
    "roadmc_read_vehicle": (C.c_int, [C.c_char_p, C.POINTER(VehicleFile)]),
    "roadmc_read_species": (C.c_int, [C.c_char_p, C.POINTER(SpeciesFile)]),
    "roadmc_read_run_settings": (C.c_int, [C.c_char_p, C.POINTER(RunSettings)]),
    "roadmc_mte_adjoint_fwd": (C.c_int, [C.c_void_p, C.POINTER(MteProblem), C.c_int32,
                                         C.POINTER(C.c_void_p), _dp]),
    "roadmc_mte_adjoint_bwd": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp]),
    "roadmc_adjoint_destroy": (None, [C.c_void_p]),
    "roadmc_mlp_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    "roadmc_mlp_destroy": (None, [C.c_void_p]),
    "roadmc_mlp_n_params": (C.c_int, [C.c_void_p, _ip]),
    "roadmc_mlp_set_params": (C.c_int, [C.c_void_p, _fp]),
    "roadmc_mlp_get_params": (C.c_int, [C.c_void_p, _fp]),
    "roadmc_mlp_gradient": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, _fp, _fp]),
    "roadmc_mlp_train": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, C.c_int32, C.c_float, _fp]),
    "roadmc_mlp_predict": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "roadmc_rov_basis_size": (C.c_int, [C.c_int32, C.c_int32, _ip]),
    "roadmc_rov_run": (C.c_int, [C.c_void_p, C.POINTER(RovProblem), C.POINTER(RovResult)]),
    "roadmc_exp_pv": (C.c_int, [C.c_void_p, C.POINTER(Uncertainty), C.c_int32, C.c_double,
                                C.c_double, C.c_double, C.c_int64, C.c_uint64, _dp, _dp]),
    "roadmc_exp_pv_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Uncertainty), C.c_int32,
                                        C.c_double, C.c_double, C.c_double, C.c_int64,
                                        C.c_uint64, _dp, _dp]),
}

_LIB = None


def lib() -> C.CDLL:
    """Loads libroadmc.so once; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: run `python -m transportoptimiser_amd.build` first")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


class RoadmcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libroadmc error {code}: {msg}")
        self.code = code


def check(code: int, ctx=None) -> None:
    if code != OK:
        msg = lib().roadmc_last_error(ctx).decode() if ctx else ""
        raise RoadmcError(code, msg)
