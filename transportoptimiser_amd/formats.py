"""On-disk formats of the reference, read and written by libroadmc's host code.

* ``surrogate_data``: the per-scenario file of surrogate training samples
  (RoadGA::writeOutSurrogateData / getExistingSurrogateData,
  include/main/RoadGA.cpp:2977-3070, 4553-4611), for import, export and
  resume of the sample set behind the surrogates.
* the commodity, vehicle and species input files of the reference's data
  directory (read by RoadGA::initialiseFromTextInput, RoadGA.cpp:3464-4550),
  converted to the Monte Carlo inputs of this library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
import numpy as np

from . import native as nat
from .rov_species import Price


def write_surrogate_data(path: str, kind: int, rows: np.ndarray,
                         max_benefit: float | None = None) -> int:
    """Appends samples; returns how many were written.  CONTROLLED samples
    beyond ``max_benefit`` are dropped, as the reference drops them (and all
    of them when it is not finite); None keeps every finite sample."""
    if max_benefit is None:
        max_benefit = np.finfo(np.float64).max
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    width = rows.shape[1]
    n_species = width // 3 if kind == nat.SURR_MTE else width - 3
    n = C.c_int32()
    nat.check(nat.lib().roadmc_surrogate_data_write(str(path).encode(), kind, n_species,
                                                    rows.shape[0], rows.ctypes.data_as(nat._dp),
                                                    max_benefit, C.byref(n)))
    return n.value


def read_surrogate_data(path: str, kind: int, n_species: int) -> np.ndarray:
    width = 3 * n_species if kind == nat.SURR_MTE else n_species + 3
    n = C.c_int32()
    nat.check(nat.lib().roadmc_surrogate_data_read(str(path).encode(), kind, n_species, 0, None,
                                                   C.byref(n)))
    out = np.zeros((n.value, width))
    nat.check(nat.lib().roadmc_surrogate_data_read(str(path).encode(), kind, n_species, n.value,
                                                   out.ctypes.data_as(nat._dp), C.byref(n)))
    return out


def read_commodity(path: str) -> nat.CommodityFile:
    o = nat.CommodityFile()
    nat.check(nat.lib().roadmc_read_commodity(str(path).encode(), C.byref(o)))
    return o


def read_vehicle(path: str) -> nat.VehicleFile:
    o = nat.VehicleFile()
    nat.check(nat.lib().roadmc_read_vehicle(str(path).encode(), C.byref(o)))
    return o


def read_species(path: str) -> nat.SpeciesFile:
    o = nat.SpeciesFile()
    nat.check(nat.lib().roadmc_read_species(str(path).encode(), C.byref(o)))
    return o


# Optimiser::ROVType (Optimiser.h:74-82), counted from 0.
_ROV_TYPES = ("ALGO1", "ALGO2", "ALGO3", "ALGO4", "ALGO5", "ALGO6", "ALGO7")


@dataclass
class RunSettings:
    """The Monte Carlo subset of the main input file (roadmc_read_run_settings).
    Absent numbers are NaN or -1; absent lists are empty."""

    confidence_interval: float
    use_gpu: int
    surrogate_dim_res: int
    interpolation: int
    traffic_growth: float
    rrr: float
    n_years: int
    time_step: float
    n_paths: int
    rov_dim_res: int
    optimisation_method: int
    rov_method: int
    lists: dict
    programs: list          # [(flows, switching costs, bridge flag)]
    n_fuels: int
    n_transported: int
    commodities: list
    vehicles: list
    species: list
    road_exists: int
    travel_time: float
    road_length: float
    accident_cost: float
    fuel_per_journey: list

    def simulation_args(self, program: int = 0) -> dict:
        """Keyword arguments of a RovsProblem (rov_species) that the file fixes:
        the horizon, step, discount rate, path count, grid resolution and the
        flows of one traffic program (SimulateGPU.cu:3870-3936 reads the same
        settings from the optimiser)."""
        flows = self.programs[program][0] if self.programs else []
        return {"n_years": self.n_years, "time_step": self.time_step, "rrr": self.rrr,
                "n_paths": self.n_paths, "dim_res": self.rov_dim_res,
                "flow": np.asarray(flows, dtype=np.float64)}

    def rov_type(self) -> str:
        """The ROV method the reference runs for this file: RoadGA casts the
        written code to Optimiser::ROVType, which counts from 0, whatever the
        file's legend says (RoadGA.cpp:3809-3810)."""
        if 0 <= self.rov_method < len(_ROV_TYPES):
            return _ROV_TYPES[self.rov_method]
        return "invalid"


def read_run_settings(path: str) -> RunSettings:
    o = nat.RunSettings()
    nat.check(nat.lib().roadmc_read_run_settings(str(path).encode(), C.byref(o)))
    progs = [([o.flows[k][i] for i in range(o.n_flows[k])],
              [o.switching[k][i] for i in range(o.n_switching[k])], bool(o.bridge[k]))
             for k in range(o.n_programs)]

    def names(arr, n):
        return [arr[i].value.decode() for i in range(n)]

    return RunSettings(
        confidence_interval=o.confidence_interval, use_gpu=o.use_gpu,
        surrogate_dim_res=o.surrogate_dim_res, interpolation=o.interpolation,
        traffic_growth=o.traffic_growth, rrr=o.rrr, n_years=o.n_years, time_step=o.time_step,
        n_paths=o.n_paths, rov_dim_res=o.rov_dim_res,
        optimisation_method=o.optimisation_method, rov_method=o.rov_method,
        lists={k: getattr(o, k).values() for k in nat.RUN_LISTS}, programs=progs,
        n_fuels=o.n_fuels, n_transported=o.n_transported,
        commodities=names(o.commodity, o.n_commodities),
        vehicles=names(o.vehicle, o.n_vehicles), species=names(o.species, o.n_species),
        road_exists=o.road_exists, travel_time=o.travel_time, road_length=o.road_length,
        accident_cost=o.accident_cost, fuel_per_journey=o.fuel_per_journey.values())


def commodity_price(c: nat.CommodityFile, mean_mult: float = 1.0,
                    sd_mult: float = 1.0) -> Price:
    """The packed uncertainty of SimulateGPU.cu:4020-4065: current, mean x
    multiplier, SD, reversion, jump size and probability x SD multiplier."""
    return Price(x0=c.current, mean=c.mean * mean_mult, sd=c.sd * sd_mult,
                 reversion=c.reversion, jump_size=c.poisson_jump * sd_mult,
                 jump_prob=c.jump_prob * sd_mult)


def species_params(s: nat.SpeciesFile, threshold: float, local_variability: float,
                   gr_mult: float = 1.0, gr_sd_mult: float = 1.0) -> np.ndarray:
    """The eight species parameters of SimulateGPU.cu:3990-4016 from a species
    file's growth block (the threshold and local variability come from the
    species' full parameter set)."""
    return np.array([s.growth_current * gr_mult, s.growth_mean * gr_mult,
                     s.growth_sd * gr_sd_mult, threshold, s.growth_reversion,
                     s.growth_poisson_jump * gr_sd_mult, s.growth_jump_prob * gr_sd_mult,
                     local_variability])
