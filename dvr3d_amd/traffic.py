"""HBM traffic of the LDL^T trailing update from rocprofv3 counter runs.

Reads the CSV output of five rocprofv3 runs (made with `--output-format csv`):

  fetch, write    --pmc FETCH_SIZE / --pmc WRITE_SIZE over `bench.py --ldl-probe N`
                  (one LDL^T factorisation of an N x N matrix), kernel
                  k_gemm_lower_w64 only
  copy_fetch,     the same counters over `bench.py --copy-probe` (three 512 MiB
  copy_write      device copies): FETCH_SIZE / WRITE_SIZE of a copy of known size
                  give the factors that turn each counter into bytes
  trace           --kernel-trace over the same ldl probe (durations)

and writes the per-launch figures of the trailing update (every dispatch of
the persistent lower-tile kernel k_gemm_lower_w64 is one, in factorisation
order; the last, under 64 rows, runs on the 64 x 64 kernel) as JSON:
algorithmic bytes (the two panels read, 8 * 2 * m * K, plus the lower
triangle of A22 read and written, 8 * m^2, with m = N - 2 PW - PW i for
launch i and K = PW = 256), measured bytes, their ratio and the average
duration.

usage: python -m dvr3d_amd.traffic --n 20000 --fetch D1 --write D2 \\
           --copy-fetch D3 --copy-write D4 --trace D5 --out profiles/ldl_update_traffic.json
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os

PW = 256   # LDL^T panel width (trailing-update K)
TILE = 64


def _rows(d, pattern):
    files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not files:
        raise FileNotFoundError(f"no {pattern} under {d}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


UPDATE_KERNEL = "k_gemm_lower_w64"


def _rows_m(n, nl):
    """Rows of launch i's trailing block."""
    return [n - 2 * PW - PW * i for i in range(nl)]


def _counter(d, counter):
    """Per-dispatch values of one counter for the trailing updates, in
    dispatch order."""
    vals = {}
    for r in _rows(d, "*counter_collection.csv"):
        if r["Counter_Name"] != counter or UPDATE_KERNEL not in r["Kernel_Name"]:
            continue
        did = int(r["Dispatch_Id"])
        vals[did] = vals.get(did, 0.0) + float(r["Counter_Value"])
    return [vals[k] for k in sorted(vals)]


def _copy_kib(d, counter):
    vals = {}
    for r in _rows(d, "*counter_collection.csv"):
        if r["Counter_Name"] == counter and "copyBuffer" in r["Kernel_Name"]:
            did = int(r["Dispatch_Id"])
            vals[did] = vals.get(did, 0.0) + float(r["Counter_Value"])
    v = sorted(vals.values())
    return v[len(v) // 2]


def main(argv=None):
    ap = argparse.ArgumentParser()
    for k in ("fetch", "write", "copy-fetch", "copy-write", "trace", "out"):
        ap.add_argument("--" + k, required=True)
    ap.add_argument("--n", type=int, required=True)
    ap.add_argument("--copy-bytes", type=int, default=512 << 20)
    a = ap.parse_args(argv)
    fetch = _counter(a.fetch, "FETCH_SIZE")
    write = _counter(a.write, "WRITE_SIZE")
    # counters are in KiB; a copy of B bytes reads and writes B bytes
    f_scale = a.copy_bytes / 1024.0 / _copy_kib(a.copy_fetch, "FETCH_SIZE")
    w_scale = a.copy_bytes / 1024.0 / _copy_kib(a.copy_write, "WRITE_SIZE")
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            for r in _rows(a.trace, "*kernel_trace.csv") if UPDATE_KERNEL in r["Kernel_Name"]]
    nl = min(len(fetch), len(write))
    ms = _rows_m(a.n, nl)
    alg = [8.0 * 2 * m * PW + 8.0 * m * m for m in ms]
    fb = [v * 1024.0 * f_scale for v in fetch[:nl]]
    wb = [v * 1024.0 * w_scale for v in write[:nl]]
    out = {
        "probe_n": a.n, "panel_width": PW, "launches": nl,
        "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
        "avg_algorithmic_bytes": sum(alg) / nl,
        "avg_fetch_bytes": sum(fb) / nl,
        "avg_write_bytes": sum(wb) / nl,
        "avg_measured_bytes": (sum(fb) + sum(wb)) / nl,
        "measured_over_algorithmic": (sum(fb) + sum(wb)) / sum(alg),
        "avg_duration_us": sum(durs) / len(durs) / 1e3 if durs else None,
        "trace_launches": len(durs),
        "per_launch": [{"m": m, "algorithmic_mb": round(x / 1e6, 1), "fetched_mb": round(f / 1e6, 1),
                        "written_mb": round(w / 1e6, 1)}
                       for m, x, f, w in list(zip(ms, alg, fb, wb))[:: max(1, nl // 5)]],
    }
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "per_launch"}))


if __name__ == "__main__":
    main()
