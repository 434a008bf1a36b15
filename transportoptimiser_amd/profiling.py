"""Calibration helper for rocprofv3 traffic counters.

``python transportoptimiser_amd/profiling.py copy --bytes B`` copies B bytes
device-to-device once (after a warm-up copy) so that FETCH_SIZE and WRITE_SIZE
of that copy kernel can be compared with B: the ratio scales the counters read
for the MTE kernel into HBM bytes.

``python transportoptimiser_amd/profiling.py summarize <stats.csv> [...]``
prints the kernel-trace statistics rows of the named kernels.
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


def summarize(paths: list[str]) -> None:
    for p in paths:
        with open(p, newline="") as f:
            rows = list(csv.DictReader(f))
        for r in rows:
            print({k: r[k] for k in r if k in ("Name", "Calls", "TotalDurationNs",
                                                "AverageNs", "Percentage", "MinNs", "MaxNs")})


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=1 << 30)
    s = sub.add_parser("summarize")
    s.add_argument("paths", nargs="+")
    a = ap.parse_args()
    if a.cmd == "copy":
        copy(a.bytes)
    else:
        summarize(a.paths)
    return 0


if __name__ == "__main__":
    sys.exit(main())
