"""Calibration helper for rocprofv3 traffic counters.

``python transportoptimiser_amd/profiling.py copy --bytes B`` copies B bytes
device-to-device once (after a warm-up copy) so that FETCH_SIZE and WRITE_SIZE
of that copy kernel can be compared with B: the ratio scales the counters read
for the MTE kernel into HBM bytes.

``python transportoptimiser_amd/profiling.py summarize <stats.csv> [...]``
prints the kernel-trace statistics rows of the named kernels.

``python transportoptimiser_amd/profiling.py uniform --patches P`` runs one
batched launch of 8 roads that all have P patches (2 species, 50 years,
20,000 paths), the per-size workload used for counter profiles of the
persistent MTE kernel; ``--precision bf16x3`` runs the split-bf16 products.

``python transportoptimiser_amd/profiling.py rov [--repeat R]`` runs the ROV
valuation of 3 uncertainties (2 OU + 1 GBM), 2 controls, 50 steps, 100,000
paths and a cubic basis R times, printing the host time of each call.

``python transportoptimiser_amd/profiling.py adjoint [--repeat R]`` runs the
adjoint of one species with 256 patches, 100 years, 50,000 paths and
checkpoints every 10 years (forward and backward) R times.

``python transportoptimiser_amd/profiling.py road [--patches P --points V]``
builds the crossing, transition and survival matrices of a synthetic species
of P patches and a road of V vertices (3 flow controls) R times.

``python transportoptimiser_amd/profiling.py knn [--samples N --dims D --res R]``
fits an R^D grid surrogate from N samples (scaled coordinates, values <= 0,
the reference's neighbour count) and evaluates it at 10^6 points.

``python transportoptimiser_amd/profiling.py rovs [--patches P --species S
--controls C --paths N --years T --res R --method 4|6]`` runs the
species-state ROV (the reference's ALGO4, or ALGO6 with the forward
recomputation) on the synthetic landscape R times.

``python transportoptimiser_amd/profiling.py wide [--patches P --paths N --years T]``
runs one road of one species with P > 1024 patches (the HBM-state path of
roadmc_mte_run) R times, printing the host time and TFLOP/s of each call.

``python transportoptimiser_amd/profiling.py dbstats <results.db>`` prints the
per-kernel call count and duration statistics of a rocprofv3 kernel trace
(the rocpd SQLite output).

``python transportoptimiser_amd/profiling.py traffic --kernel NAME --fetch F.db
--write W.db --copy-fetch CF.db --copy-write CW.db --copy-bytes B`` prints the
HBM bytes per dispatch of kernel NAME: FETCH_SIZE and WRITE_SIZE (KiB per
dispatch), each scaled by B divided by what the same counter reads for the
measured B-byte copy.
"""
from __future__ import annotations

import argparse
import csv
import sys


def copy(nbytes: int) -> None:
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {n * 4} bytes twice (second copy is the measured one)")


def uniform(P: int, n_roads: int = 8, precision: str = "fp32") -> None:
    import os
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd import landscape as ls
    from transportoptimiser_amd import native as nat
    from transportoptimiser_amd.api import Context
    prec = nat.PREC_BF16X3 if precision == "bf16x3" else nat.PREC_FP32
    rng = np.random.default_rng(P)
    roads = [ls.RoadProblem([ls.make_species(rng, P) for _ in range(2)], n_years=50,
                            n_paths=20000, seed=1, road_id=r, precision=prec)
             for r in range(n_roads)]
    flops = sum(2.0 * r.n_paths * P * P * r.n_years * len(r.species) for r in roads)
    ctx = Context(0)
    with ctx.prepare_batch(roads) as b:
        b.launch()
        b.sync()
        ms = b.kernel_ms()
    ctx.close()
    print(f"P={P} {precision}: {n_roads} roads, kernel {ms:.1f} ms, "
          f"{flops / ms / 1e9:.1f} TFLOP/s", flush=True)


def rov(repeat: int) -> None:
    import os
    import time
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd.api import Context
    from transportoptimiser_amd.rov import config5, simulate_rov
    ctx = Context(0)
    pr = config5()
    for i in range(repeat):
        t = time.perf_counter()
        r = simulate_rov(ctx, pr)
        print(f"rov call {i}: {(time.perf_counter() - t) * 1e3:.2f} ms, value {r['value']:.4f}",
              flush=True)
    ctx.close()


def adjoint(repeat: int) -> None:
    import os
    import time
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd import landscape
    from transportoptimiser_amd.adjoint import mte_gradient
    from transportoptimiser_amd.api import Context
    ctx = Context(0)
    pr = landscape.RoadProblem([landscape.make_species(np.random.default_rng(4), 256,
                                                       mort_max=0.4)],
                               n_years=100, n_paths=50000, seed=4)
    for i in range(repeat):
        t = time.perf_counter()
        obj, gm, gr = mte_gradient(ctx, pr, 10)
        print(f"adjoint call {i}: {(time.perf_counter() - t) * 1e3:.2f} ms, objective {obj:.6e}, "
              f"|dJ/dm| {np.abs(gm).sum():.6e}", flush=True)
    ctx.close()


def road(patches: int, points: int, repeat: int) -> None:
    """Road -> matrix builder on a synthetic landscape of ``patches`` patches
    and a road of ``points`` vertices (kernel trace target)."""
    import os
    import time

    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd.api import Context
    from transportoptimiser_amd.road import build_species_road, random_road_species
    ctx = Context(0)
    inp = random_road_species(np.random.default_rng(1), patches, n_road_points=points)
    for i in range(repeat):
        t = time.perf_counter()
        m = build_species_road(ctx, inp)
        print(f"road call {i}: P={patches} segments={points - 1} "
              f"{(time.perf_counter() - t) * 1e3:.2f} ms, crossings {int(m.crossings.sum())}",
              flush=True)
    ctx.close()


def knn(n: int, dims: int, res: int, repeat: int) -> None:
    """Grid-surrogate fit (ROV layout: scaled coordinates, values <= 0, the
    reference's neighbour count for n samples) and an evaluation of 10^6
    points."""
    import os
    import time

    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd import knn as kn
    from transportoptimiser_amd import native as nat
    from transportoptimiser_amd.api import Context
    ctx = Context(0)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 1.0, size=(n, dims))
    y = -np.exp(x.sum(1)) + 0.05 * rng.normal(size=n)
    k = kn.neighbour_count(n, nat.KNN_RULE_ROV)
    pts = rng.uniform(0.0, 1.0, size=(1000000, dims))
    for i in range(repeat):
        t = time.perf_counter()
        g = kn.fit(ctx, x, y, res, k, normalize=True, nonpositive=True)
        t1 = time.perf_counter()
        g.predict(ctx, pts)
        t2 = time.perf_counter()
        print(f"knn call {i}: fit {res}^{dims} grid from {n} samples (k={k}) "
              f"{(t1 - t) * 1e3:.2f} ms, eval 1e6 points {(t2 - t1) * 1e3:.2f} ms", flush=True)
    ctx.close()


def rovs_grad(patches: int, paths: int, years: int, repeat: int, tau_frac: float) -> None:
    """Smoothed-validity gradient of the species-state valuation (1 species,
    3 controls, thresholds at 0.7 of the initial total) under ALGO4's fitted
    policy, tau = tau_frac x the threshold (kernel trace target)."""
    import os
    import time

    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd import rov_species as rs
    from transportoptimiser_amd.api import Context
    ctx = Context(0)
    pr = rs.synthetic_problem(seed=5, n_patches=patches, n_species=1, n_controls=3,
                              n_paths=paths, n_years=years, dim_res=20, threshold_ratio=0.7)
    pol = rs.simulate_rov_species(ctx, pr, paths=True)["opt_control"]
    tau = tau_frac * pr.species[0].params[3]
    for i in range(repeat):
        t = time.perf_counter()
        v, g = rs.smooth_value_gradient(ctx, pr, pol, tau)
        print(f"rovs grad call {i}: P={patches} C=3 N={paths} T={years} tau={tau_frac} x threshold: "
              f"{(time.perf_counter() - t) * 1e3:.2f} ms, value {v:.4f}, "
              f"|dV/dsurv| {np.abs(g).sum():.6e}", flush=True)
    ctx.close()


def rovs(patches: int, species: int, controls: int, paths: int, years: int, res: int,
         repeat: int, method: int = 4, precision: str = "fp32") -> None:
    """Species-state ROV (ALGO4 or ALGO6) on the synthetic landscape, thresholds at 0.7
    of the initial total (every control has in-the-money paths for the first
    half of the horizon at P = 256)."""
    import os
    import time

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd import native as nat
    from transportoptimiser_amd import rov_species as rs
    from transportoptimiser_amd.api import Context
    ctx = Context(0)
    pr = rs.synthetic_problem(seed=5, n_patches=patches, n_species=species, n_controls=controls,
                              n_paths=paths, n_years=years, dim_res=res, threshold_ratio=0.7)
    pr.method = method
    pr.precision = nat.PREC_BF16X3 if precision == "bf16x3" else nat.PREC_FP32
    for i in range(repeat):
        t = time.perf_counter()
        r = rs.simulate_rov_species(ctx, pr)
        print(f"rovs call {i}: ALGO{method} {precision} P={patches} S={species} C={controls} N={paths} T={years} "
              f"R={res}: {(time.perf_counter() - t) * 1e3:.2f} ms, value {r['value']:.4f}, "
              f"control {r['first_control']}", flush=True)
    ctx.close()


def summarize(paths: list[str]) -> None:
    for p in paths:
        with open(p, newline="") as f:
            rows = list(csv.DictReader(f))
        for r in rows:
            print({k: r[k] for k in r if k in ("Name", "Calls", "TotalDurationNs",
                                                "AverageNs", "Percentage", "MinNs", "MaxNs")})


def _db(path: str):
    import sqlite3
    return sqlite3.connect(f"file:{path}?mode=ro", uri=True)


def dbstats(path: str) -> None:
    c = _db(path)
    rows = list(c.execute("select name, count(*), sum(duration), avg(duration), min(duration), "
                          "max(duration) from kernels group by name order by sum(duration) desc"))
    total = sum(r[2] for r in rows) or 1.0
    print(f"{'calls':>6} {'total_ms':>11} {'avg_us':>11} {'min_us':>11} {'max_us':>11} {'pct':>6}  kernel")
    for name, n, tot, avg, mn, mx in rows:
        print(f"{n:>6} {tot / 1e6:>11.3f} {avg / 1e3:>11.1f} {mn / 1e3:>11.1f} {mx / 1e3:>11.1f} "
              f"{100 * tot / total:>6.2f}  {name[:110]}")


def counters(path: str, kernel: str) -> None:
    c = _db(path)
    rows = c.execute("select dispatch_id, kernel_name, counter_name, sum(value) from "
                     "counters_collection group by dispatch_id, counter_name "
                     "order by dispatch_id, counter_name")
    for d, name, cn, v in rows:
        if kernel in name:
            print(f"dispatch {d}: {cn} = {v:.6e}")


def _counter(path: str, kernel: str, counter: str) -> list[float]:
    c = _db(path)
    rows = c.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                     "where counter_name = ? group by dispatch_id order by dispatch_id", (counter,))
    return [float(v) for _, name, v in rows if kernel in name]


def traffic(a) -> None:
    fetch = _counter(a.fetch, a.kernel, "FETCH_SIZE")
    write = _counter(a.write, a.kernel, "WRITE_SIZE")
    cf = max(_counter(a.copy_fetch, "copyBuffer", "FETCH_SIZE"))
    cw = max(_counter(a.copy_write, "copyBuffer", "WRITE_SIZE"))
    if not fetch or not write:
        raise SystemExit(f"no dispatch of {a.kernel} in the counter files")
    kf = a.copy_bytes / (cf * 1024.0)
    kw = a.copy_bytes / (cw * 1024.0)
    f_kib = sum(fetch) / len(fetch)
    w_kib = sum(write) / len(write)
    hbm = f_kib * 1024.0 * kf + w_kib * 1024.0 * kw
    print(f"kernel: {a.kernel}")
    print(f"dispatches_fetch: {len(fetch)}")
    print(f"dispatches_write: {len(write)}")
    print(f"fetch_kib_per_dispatch: {f_kib:.1f}")
    print(f"write_kib_per_dispatch: {w_kib:.1f}")
    print(f"copy_bytes: {a.copy_bytes}")
    print(f"copy_fetch_kib: {cf:.1f}")
    print(f"copy_write_kib: {cw:.1f}")
    print(f"fetch_scale: {kf:.4f}")
    print(f"write_scale: {kw:.4f}")
    print(f"hbm_bytes_per_launch: {hbm:.6e}")


def wide(patches: int, paths: int, years: int, repeat: int) -> None:
    import os
    import time
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from transportoptimiser_amd import landscape
    from transportoptimiser_amd.api import Context
    ctx = Context(0)
    sp = landscape.make_species(np.random.default_rng(6), patches)
    pr = landscape.RoadProblem([sp], n_years=years, n_paths=paths, seed=6)
    Pp = (patches + 127) // 128 * 128
    flops = 2.0 * paths * Pp * Pp * years
    for i in range(repeat):
        t = time.perf_counter()
        out = ctx.simulate_mte(pr, end_pops=False, patch_mean=False)
        ms = (time.perf_counter() - t) * 1e3
        print(f"wide call {i}: P={patches} N={paths} T={years}: {ms:.2f} ms, "
              f"{flops / ms / 1e9:.1f} TFLOP/s (padded to {Pp}), MTE {out.mte_prob[0]:.4f}",
              flush=True)
    ctx.close()


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=1 << 30)
    s = sub.add_parser("summarize")
    s.add_argument("paths", nargs="+")
    ad = sub.add_parser("adjoint")
    ad.add_argument("--repeat", type=int, default=3)
    rv = sub.add_parser("rov")
    rv.add_argument("--repeat", type=int, default=5)
    u = sub.add_parser("uniform")
    u.add_argument("--patches", type=int, required=True)
    u.add_argument("--roads", type=int, default=8)
    u.add_argument("--precision", choices=["fp32", "bf16x3"], default="fp32")
    rd = sub.add_parser("road")
    rd.add_argument("--patches", type=int, default=1024)
    rd.add_argument("--points", type=int, default=200)
    rd.add_argument("--repeat", type=int, default=3)
    kn = sub.add_parser("knn")
    kn.add_argument("--samples", type=int, default=4096)
    kn.add_argument("--dims", type=int, default=3)
    kn.add_argument("--res", type=int, default=20)
    kn.add_argument("--repeat", type=int, default=3)
    rv2 = sub.add_parser("rovs")
    rv2.add_argument("--patches", type=int, default=256)
    rv2.add_argument("--species", type=int, default=1)
    rv2.add_argument("--controls", type=int, default=3)
    rv2.add_argument("--paths", type=int, default=100000)
    rv2.add_argument("--years", type=int, default=50)
    rv2.add_argument("--res", type=int, default=20)
    rv2.add_argument("--repeat", type=int, default=3)
    rv2.add_argument("--method", type=int, default=4, choices=(4, 6))
    rv2.add_argument("--precision", choices=["fp32", "bf16x3"], default="fp32")
    rg = sub.add_parser("rovsgrad")
    rg.add_argument("--patches", type=int, default=256)
    rg.add_argument("--paths", type=int, default=100000)
    rg.add_argument("--years", type=int, default=50)
    rg.add_argument("--repeat", type=int, default=3)
    rg.add_argument("--tau", type=float, default=0.05)
    wd = sub.add_parser("wide")
    wd.add_argument("--patches", type=int, default=2048)
    wd.add_argument("--paths", type=int, default=20000)
    wd.add_argument("--years", type=int, default=10)
    wd.add_argument("--repeat", type=int, default=3)
    d = sub.add_parser("dbstats")
    d.add_argument("path")
    k = sub.add_parser("counters")
    k.add_argument("path")
    k.add_argument("--kernel", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    a = ap.parse_args()
    if a.cmd == "copy":
        copy(a.bytes)
    elif a.cmd == "adjoint":
        adjoint(a.repeat)
    elif a.cmd == "rov":
        rov(a.repeat)
    elif a.cmd == "rovs":
        rovs(a.patches, a.species, a.controls, a.paths, a.years, a.res, a.repeat, a.method,
             a.precision)
    elif a.cmd == "rovsgrad":
        rovs_grad(a.patches, a.paths, a.years, a.repeat, a.tau)
    elif a.cmd == "knn":
        knn(a.samples, a.dims, a.res, a.repeat)
    elif a.cmd == "road":
        road(a.patches, a.points, a.repeat)
    elif a.cmd == "uniform":
        uniform(a.patches, a.roads, a.precision)
    elif a.cmd == "wide":
        wide(a.patches, a.paths, a.years, a.repeat)
    elif a.cmd == "dbstats":
        dbstats(a.path)
    elif a.cmd == "counters":
        counters(a.path, a.kernel)
    elif a.cmd == "traffic":
        traffic(a)
    else:
        summarize(a.paths)
    return 0


if __name__ == "__main__":
    sys.exit(main())
