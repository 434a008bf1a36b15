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
