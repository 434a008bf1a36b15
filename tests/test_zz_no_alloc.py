"""roadmc_exp_pv_noalloc (no device allocation, a fixed device array) against
roadmc_exp_pv through the C ABI (runs on an MI355X).

Two calls in a row, the second with fewer paths, so the second reads an array
that still holds the first call's values beyond its own paths; both agree
with the allocating entry point to FP64 rounding (the statistics are reduced
on the device there, on the host in roadmc_exp_pv)."""
import ctypes as C

import pytest

from transportoptimiser_amd import native as nat

pytestmark = pytest.mark.gpu


def _call(fn, ctx, u, T, dt, r, g, n, seed):
    unc = nat.Uncertainty(u[0], u[1], u[2], u[3], u[4], u[5], u[6])
    m, s = C.c_double(), C.c_double()
    rc = fn(ctx.handle, C.byref(unc), T, dt, r, g, n, seed, C.byref(m), C.byref(s))
    return rc, m.value, s.value


def test_noalloc_matches_allocating_entry_point(ctx):
    lib = nat.lib()
    u = (nat.SDE_MRJD, 1.0, 1.4, 0.2, 0.25, 0.3, 0.1)
    for n, T, seed in ((300000, 30, 5), (70001, 12, 6), (1, 4, 7)):
        print(f"no-alloc exp_pv: n_paths={n} T={T} seed={seed}", flush=True)
        rc0, m0, s0 = _call(lib.roadmc_exp_pv, ctx, u, T, 1.0, 0.07, 0.01, n, seed)
        rc1, m1, s1 = _call(lib.roadmc_exp_pv_noalloc, ctx, u, T, 1.0, 0.07, 0.01, n, seed)
        assert rc0 == nat.OK and rc1 == nat.OK
        assert m1 == pytest.approx(m0, rel=1e-12)
        assert s1 == pytest.approx(s0, rel=1e-10, abs=1e-14)


def test_noalloc_rejects_more_paths_than_its_array(ctx):
    lib = nat.lib()
    u = (nat.SDE_MRJD, 1.0, 1.4, 0.2, 0.25, 0.3, 0.1)
    rc, _, _ = _call(lib.roadmc_exp_pv_noalloc, ctx, u, 5, 1.0, 0.07, 0.0,
                     nat.NOALLOC_MAX_PATHS + 1, 1)
    assert rc == nat.ERR_INVALID
    assert b"ROADMC_NOALLOC_MAX_PATHS" in lib.roadmc_last_error(ctx.handle)
    rc, m, _ = _call(lib.roadmc_exp_pv_noalloc, ctx, u, 5, 1.0, 0.07, 0.0,
                     nat.NOALLOC_MAX_PATHS, 1)
    assert rc == nat.OK and m > 0
