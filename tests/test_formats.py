"""On-disk formats (host code of libroadmc; runs on the CPU).

The readers are checked against the reference's own input files
(tests/golden/input_data, copied from its tests/Input_Data), the surrogate
data file against the writer's format (RoadGA.cpp:2977-3070) and its reader's
parsing rule (RoadGA.cpp:4553-4611).
"""
import pathlib

import numpy as np
import pytest

from transportoptimiser_amd import formats as fm
from transportoptimiser_amd import native as nat
from transportoptimiser_amd import rov_species as rs

GOLD = pathlib.Path(__file__).resolve().parent / "golden" / "input_data"


def test_commodity_files():
    d = fm.read_commodity(GOLD / "diesel")
    assert d.name == b"diesel"
    assert (d.current, d.mean, d.sd, d.reversion, d.jump_prob, d.poisson_jump) == \
        (1.2, 1.2, 0.05, 0.05, 0.05, 0.05)
    assert (d.ore_mean, d.ore_sd) == (0.0, 0.0)
    o = fm.read_commodity(GOLD / "ore")
    assert o.name == b"ore" and o.current == 30 and o.mean == 30 and o.ore_mean == 1.0
    p = fm.read_commodity(GOLD / "petrol")
    assert p.current == 1.05 and p.mean == 1.05
    q = fm.commodity_price(d, mean_mult=2.0, sd_mult=3.0)
    assert (q.x0, q.mean, q.sd, q.jump_size, q.jump_prob) == \
        pytest.approx((1.2, 2.4, 0.15, 0.15, 0.15))


def test_vehicle_files():
    v = fm.read_vehicle(GOLD / "vehicle_small")
    assert v.name == b"small"
    assert (v.width, v.length, v.proportion, v.load, v.fuel) == (1.7, 2.5, 0.6, 0.5, 2)
    assert (v.a, v.agr, v.av, v.avsq, v.travel_cost) == (171.3978, 9.8565, 2.5314, 0.0219, 13.23)
    m = fm.read_vehicle(GOLD / "vehicle_medium")
    g = fm.read_vehicle(GOLD / "vehicle_large")
    assert (m.width, m.length, m.proportion, m.fuel) == (2.5, 2.0, 0.3, 2)
    assert (g.length, g.proportion, g.fuel, g.travel_cost) == (10.0, 0.1, 3, 35.28)
    # the proportions of the three classes sum to one
    assert v.proportion + m.proportion + g.proportion == pytest.approx(1.0)


def test_species_file():
    s = fm.read_species(GOLD / "species_1")
    assert s.name == b"Koala" and s.n_habitat_types == 5
    names = [s.habitat_name[i].value for i in range(5)]
    assert names == [b"PRIMARY", b"MARGINAL", b"OTHER", b"CLEAR", b"ROAD"]
    np.testing.assert_array_equal(list(s.max_pop)[:5], [0.0005, 0.00025, 0.0001, 0.0, 0.0])
    np.testing.assert_array_equal(list(s.cost_per_m2)[:5], [0, 5e7, 5e6, 0, 0])
    np.testing.assert_array_equal(list(s.pref_mean)[:5], [0, -0.262, -0.396, -0.373, -1000])
    np.testing.assert_array_equal(list(s.pref_sd)[:5], [1e8, 0.073, 0.120, 0.175, 0])
    assert (s.growth_current, s.growth_mean, s.growth_sd, s.growth_reversion,
            s.growth_jump_prob, s.growth_poisson_jump) == (0.014, 0.014, 0.005, 0.1, 0.1, 0.1)
    prm = fm.species_params(s, threshold=100.0, local_variability=0.2, gr_mult=2.0)
    np.testing.assert_allclose(prm, [0.028, 0.028, 0.005, 100.0, 0.1, 0.1, 0.1, 0.2])


def test_missing_files_and_keys(tmp_path):
    with pytest.raises(nat.RoadmcError):
        fm.read_commodity(tmp_path / "absent")
    bad = tmp_path / "bad"
    bad.write_text("Name : x\nCurrent Value : 1\n")
    with pytest.raises(nat.RoadmcError):
        fm.read_commodity(bad)


def test_surrogate_data_mte_round_trip(tmp_path):
    f = tmp_path / "surrogate_data"
    rows = np.array([[0.0123456789, 1234.56789, 12.3456789, 0.5, 2000.0, 30.0],
                     [0.2, 987654321.0, 1e-9, 0.0, 0.0, 0.0]])
    assert fm.write_surrogate_data(f, nat.SURR_MTE, rows) == 2
    lines = f.read_text().splitlines()
    assert lines[0] == "#" * 100 and lines[2] == "#" * 100
    assert lines[1] == ("######################################## SURROGATE RESULTS "
                        "#########################################")
    assert lines[3] == ("IAR_000                  END_POP_00000000000000000END_POP_SD"
                        "000000000000000" "IAR_001                  END_POP_00000000000000001"
                        "END_POP_SD000000000000001")
    # width 20, 6 significant digits, "     " between a species' fields
    assert lines[4].startswith(f"{0.0123457:>20}     {1234.57:>20}     {12.3457:>20}")
    assert len(lines[4]) == 6 * 20 + 4 * 5
    back = fm.read_surrogate_data(f, nat.SURR_MTE, 2)
    np.testing.assert_allclose(back, rows, rtol=5e-6)
    # appending keeps one header
    assert fm.write_surrogate_data(f, nat.SURR_MTE, rows[:1]) == 1
    assert len(f.read_text().splitlines()) == 7
    assert fm.read_surrogate_data(f, nat.SURR_MTE, 2).shape == (3, 6)


def test_surrogate_data_controlled_filter(tmp_path):
    f = tmp_path / "surrogate_data"
    rows = np.array([[0.1, -5.0, -100.0, 10.0],
                     [0.2, -5.0, -1e9, 10.0],     # |value| above the maximum benefit
                     [0.3, -5.0, -200.0, 1e9]])   # |SD| above it
    assert fm.write_surrogate_data(f, nat.SURR_CONTROLLED, rows, max_benefit=-1e6) == 1
    lines = f.read_text().splitlines()
    assert lines[3] == "IAR_  0                  INITIAL_UNIT_PROFIT      AVERAGE_VALUE" \
                       "            VALUE_SD"
    np.testing.assert_allclose(fm.read_surrogate_data(f, nat.SURR_CONTROLLED, 1), rows[:1])
    # a non-finite maximum benefit writes nothing, as the reference
    assert fm.write_surrogate_data(f, nat.SURR_CONTROLLED, rows, max_benefit=np.inf) == 0


def test_surrogate_data_reads_the_reference_zero_fill(tmp_path):
    # The reference's first MTE write leaves the stream's fill character at
    # '0' (setfill in the header), so that file's numbers are zero-padded.
    f = tmp_path / "surrogate_data"
    hdr = "#" * 100 + "\n" + "x\n" + "#" * 100 + "\nIAR_000 ...\n"
    f.write_text(hdr + "0000000000000.123456     00000000000001234.5     0000000000000012.25\n")
    np.testing.assert_allclose(fm.read_surrogate_data(f, nat.SURR_MTE, 1),
                               [[0.123456, 1234.5, 12.25]])
    f.write_text(hdr + "1 2\n")
    with pytest.raises(nat.RoadmcError) as e:
        fm.read_surrogate_data(f, nat.SURR_MTE, 1)
    assert e.value.code == nat.ERR_NUMERIC


def test_valuation_inputs_from_reference_files(oracle):
    """A species-state valuation built from the reference's input files: the
    time-0 unit profit is unit cost + sum_v fuel use x fuel price - unit
    revenue x ore content x ore price at the files' current values."""
    petrol, diesel = fm.read_commodity(GOLD / "petrol"), fm.read_commodity(GOLD / "diesel")
    ore = fm.read_commodity(GOLD / "ore")
    veh = [fm.read_vehicle(GOLD / n) for n in ("vehicle_small", "vehicle_medium",
                                               "vehicle_large")]
    # the vehicles' fuel field indexes the main file's commodity list
    # (1 ore, 2 petrol, 3 diesel); fuels here are [petrol, diesel]
    fuel_idx = np.array([{2: 0, 3: 1}[v.fuel] for v in veh], dtype=np.int32)
    fuel_cost = np.array([0.1 * v.proportion for v in veh])
    pr = rs.synthetic_problem(n_patches=8, n_paths=64, n_years=4)
    pr.fuels = [fm.commodity_price(petrol), fm.commodity_price(diesel)]
    pr.commodities = [fm.commodity_price(ore)]
    pr.ore_mean, pr.ore_sd = np.array([ore.ore_mean]), np.array([ore.ore_sd])
    pr.fuel_cost, pr.fuel_idx = fuel_cost, fuel_idx
    pr.unit_cost, pr.unit_revenue = 3.0, 0.2
    _, _, _, up = oracle.rovs_forward(pr)
    fuel = sum(fc * (petrol.current, diesel.current)[i] for fc, i in zip(fuel_cost, fuel_idx))
    np.testing.assert_allclose(up[0], 3.0 + fuel - 0.2 * ore.ore_mean * ore.current, rtol=1e-12)


def test_run_settings_of_the_reference_test_input_file():
    # The older layout of tests/Input_Data/INPUT_FILE: "Number of Commodities",
    # a legend counting methods from 1, programs without a count or bridge flag.
    r = fm.read_run_settings(GOLD / "INPUT_FILE")
    assert r.confidence_interval == 0.95 and r.use_gpu == 1
    assert r.surrogate_dim_res == 30 and r.interpolation == 1
    assert r.traffic_growth == 0.0 and r.rrr == 7.0
    assert r.n_years == 50 and r.time_step == 1.0
    assert r.n_paths == 50000 and r.rov_dim_res == 40
    assert r.optimisation_method == 3 and r.rov_method == 4
    # The reference casts the written 4 to ROVType ALGO5 (an empty case).
    assert r.rov_type() == "ALGO5"
    assert r.lists["growth_mean"] == [0.5, 1.0, 1.5]
    assert r.lists["growth_sd"] == [1.0, 2.0, 3.0]
    assert r.lists["price_mean"] == [1.0]
    assert r.lists["price_sd"] == [1.0, 2.0, 3.0]
    assert r.lists["ore_sd"] == [1.0, 2.0, 3.0]
    assert r.lists["pop_level"] == [0.8, 0.9, 1.0]
    assert r.lists["animal_bridges"] == [0.0] and r.lists["comparison_cost"] == []
    assert r.programs == [([0.0, 75.0, 150.0], [0.0] * 8, False)]
    assert r.n_fuels == -1 and r.n_transported == 3
    assert r.commodities == ["Input_Data/Commodities/ore", "Input_Data/Commodities/petrol",
                             "Input_Data/Commodities/diesel"]
    assert r.vehicles == ["Input_Data/Vehicles/small", "Input_Data/Vehicles/medium",
                          "Input_Data/Vehicles/large"]
    assert r.species == ["Input_data/Species/species1"]
    assert r.road_exists == -1 and np.isnan(r.travel_time)
    args = r.simulation_args()
    assert (args["n_years"], args["time_step"], args["rrr"], args["n_paths"],
            args["dim_res"]) == (50, 1.0, 7.0, 50000, 40)
    np.testing.assert_array_equal(args["flow"], [0.0, 75.0, 150.0])


def test_run_settings_current_layout(tmp_path):
    # The layout RoadGA::initialiseFromTextInput reads (RoadGA.cpp:3896-4500):
    # a program count, a bridge flag, fuels and transported commodities
    # counted apart, and the existing-road section.
    text = """###
## ECONOMIC PARAMETERS
Required Rate of Return (per cen per annum)     : 5.5
Number of Years                                 : 30
Step Size (Years)                               : 0.5
Number of Simulation Paths                      : 100000
ROV Dimension Resolution                        : 40
Optimisation Method                             : 3
ROV Method                                      : 5
Population Growth Rate Mean Multipliers         : 1
Minimum Population Levels (proportion of min)   : 0.9
Comparison Road Variable Cost Multipliers       : 0.5,1
Programs                                        : 2
Program 1   : 0,200,400 : 0,0,0,0,0,0,0,0 : FALSE
Program 2   : 0,100 : 1,2,3,4 : TRUE
Number of Fuels                                 : 2
Number of Transported Commodities               : 0
Commodity 1                                     : Experiments/Input_Data/Commodities/diesel
Commodity 2                                     : Experiments/Input_Data/Commodities/petrol
Number of Species                               : 1
Species 1                                       : Experiments/Input_Data/Species/Species_3/Species_3
Road Exists                                     : TRUE
Travel Time (hours)                             : 0.5
Length (m)                                      : 15000
Fuel Consumption Per Journey                    : 2,3.3,6.7
Accident Cost per Unit Traffic                  : 1000
"""
    p = tmp_path / "INPUT_FILE"
    p.write_text(text)
    r = fm.read_run_settings(p)
    assert (r.rrr, r.n_years, r.time_step, r.n_paths) == (5.5, 30, 0.5, 100000)
    assert r.rov_method == 5 and r.rov_type() == "ALGO6"
    assert r.programs == [([0.0, 200.0, 400.0], [0.0] * 8, False),
                          ([0.0, 100.0], [1.0, 2.0, 3.0, 4.0], True)]
    assert r.lists["comparison_cost"] == [0.5, 1.0]
    assert (r.n_fuels, r.n_transported) == (2, 0)
    assert r.commodities[1] == "Experiments/Input_Data/Commodities/petrol"
    assert r.road_exists == 1 and r.road_length == 15000.0 and r.accident_cost == 1000.0
    assert r.fuel_per_journey == [2.0, 3.3, 6.7]
    assert np.isnan(r.confidence_interval) and r.use_gpu == -1
    # A program count that disagrees with the program lines is an error.
    p.write_text(text.replace("Programs                                        : 2",
                              "Programs                                        : 3"))
    with pytest.raises(nat.RoadmcError):
        fm.read_run_settings(p)
    with pytest.raises(nat.RoadmcError):
        fm.read_run_settings(tmp_path / "missing")
