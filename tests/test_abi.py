"""The C-ABI library builds, loads and exports what include/roadmc.h declares
(CPU; no compute calls)."""
import ctypes
import pathlib
import re
import subprocess

import pytest

from transportoptimiser_amd import native as nat

REPO = pathlib.Path(__file__).resolve().parent.parent
HEADERS = sorted((REPO / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(roadmc_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "roadmc_mte_run" in names and "roadmc_mte_run_batch" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", str(nat.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(roadmc_\w+)", out))
    for name in declared_functions():
        assert name in exported, name
        assert hasattr(lib, name)


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(nat.SIGNATURES)


def test_null_context_is_rejected():
    lib = nat.lib()
    assert lib.roadmc_mte_run(None, None, None) == nat.ERR_INVALID
    assert lib.roadmc_batch_launch(None) == nat.ERR_INVALID
    assert lib.roadmc_device_count(None) == nat.ERR_INVALID


def test_host_helpers_without_gpu():
    from transportoptimiser_amd import api
    from scipy import stats
    assert api.normal_cdf_inverse(0.975) == pytest.approx(stats.norm.ppf(0.975), abs=4.5e-4)
    with pytest.raises(ValueError):
        api.normal_cdf_inverse(1.0)
    pen = api.mte_penalty([150.0, 10.0], [100.0, 100.0], [10.0, 10.0], 0.95, [2.0, 1.0],
                          [3.0, 1.0])
    lvl = stats.norm.ppf(0.05) * 10 + 100
    assert pen == pytest.approx((150 - lvl) * 6.0, abs=0.05)


def test_host_csc_matches_oracle(oracle):
    import numpy as np
    from transportoptimiser_amd import api
    rng = np.random.default_rng(3)
    A = rng.uniform(0, 3e-5, size=(17, 19)).astype(np.float32)
    for tol in (1e-5, 0.0):
        a = api.dense_to_csc(A, tol)
        b = oracle.dense_to_csc(A, tol)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_cxx_shim_compiles_and_links(tmp_path):
    src = REPO / "tests" / "native" / "shim_driver.cpp"
    exe = tmp_path / "shim_driver"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{REPO / 'include'}",
                    f"-I{REPO / 'transportoptimiser_amd' / 'csrc'}", str(src), "-o", str(exe),
                    f"-L{nat.LIB_PATH.parent}", "-lroadmc",
                    f"-Wl,-rpath,{nat.LIB_PATH.parent}"], check=True)
    # Without a GPU the shim must throw the reference's const char* error.
    out = subprocess.run([str(exe), "--no-gpu-expected"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
