"""The ctypes mirrors of the public structs against include/roadmc.h (CPU).

A C program built with gcc from the header prints every struct's size and
each field's offset; the Python mirrors in transportoptimiser_amd.native must
agree, so that arrays of structs passed through the C ABI have the same
stride and every field lands where the library reads it.
"""
import ctypes as C
import pathlib
import shutil
import subprocess

import pytest

from transportoptimiser_amd import native as nat

REPO = pathlib.Path(__file__).resolve().parent.parent

PAIRS = [
    ("roadmc_species", nat.Species),
    ("roadmc_mte_problem", nat.MteProblem),
    ("roadmc_mte_result", nat.MteResult),
    ("roadmc_uncertainty", nat.Uncertainty),
    ("roadmc_rov_problem", nat.RovProblem),
    ("roadmc_rov_result", nat.RovResult),
    ("roadmc_road_species", nat.RoadSpecies),
    ("roadmc_rovs_species", nat.RovsSpecies),
    ("roadmc_rovs_problem", nat.RovsProblem),
    ("roadmc_rovs_result", nat.RovsResult),
    ("roadmc_commodity_file", nat.CommodityFile),
    ("roadmc_vehicle_file", nat.VehicleFile),
    ("roadmc_species_file", nat.SpeciesFile),
    ("roadmc_list", nat.RoadmcList),
    ("roadmc_run_settings", nat.RunSettings),
    ("roadmc_tile_trial", nat.TileTrial),
]


def _oracle_pairs():
    from oracle import pyoracle as po
    return [("orc_species", po.OrcSpecies), ("orc_rov_spec", po.OrcRovSpec),
            ("orc_rovs_spec", po.OrcRovsSpec)]


def _check(tmp_path, header, include_dir, pairs):
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{header}"',
             "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, *_ in py._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{include_dir}", str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.splitlines())
    for cname, py in pairs:
        assert int(got[f"{cname} size"]) == C.sizeof(py), cname
        for fname, *_ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_struct_layouts_match_header(tmp_path):
    _check(tmp_path, "roadmc.h", REPO / "include", PAIRS)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_struct_layouts_match_header(tmp_path):
    _check(tmp_path, "oracle.h", REPO / "oracle", _oracle_pairs())


def test_header_constants_match_python():
    """Every integer #define ROADMC_<NAME> of the header that the Python mirror
    also names (as <NAME>) has the same value there."""
    import re
    text = (REPO / "include" / "roadmc.h").read_text()
    seen = 0
    for name, val in re.findall(r"^#define ROADMC_(\w+)\s+(-?\d+)\b", text, re.M):
        if hasattr(nat, name):
            assert getattr(nat, name) == int(val), name
            seen += 1
    assert seen >= 20


def test_neighbour_count_is_capped_at_the_header_limit():
    """roadmc_knn_k never returns more than ROADMC_KNN_MAX_K, the limit
    roadmc_knn_surrogate_fit enforces (no GPU needed for the count)."""
    import re
    text = (REPO / "include" / "roadmc.h").read_text()
    cap = int(re.search(r"^#define ROADMC_KNN_MAX_K\s+(\d+)", text, re.M).group(1))
    from transportoptimiser_amd import knn
    assert knn.neighbour_count(10000, nat.KNN_RULE_ROV) == cap
    assert knn.neighbour_count(5650, nat.KNN_RULE_ROV) == 128 == cap
    assert knn.neighbour_count(4096, nat.KNN_RULE_ROV) == 97
