"""Per-kernel summary of rocprofv3 CSV output (kernel trace and PMC passes).

    python -m wfe_amd.prof_summary --trace T.csv --fetch F.csv --write W.csv \
        --points 576000 [--json out.json]

`--trace` is a `*_kernel_trace.csv` of a `--kernel-trace` run; `--fetch`
and `--write` are the `*_counter_collection.csv` files of one
`--pmc FETCH_SIZE` and one `--pmc WRITE_SIZE` run of the same program. The
counters are scaled by what they read for `k_copy_probe` (each launch reads
and writes `--copy-bytes`), which the benchmark runs in the same process:
FETCH_SIZE and WRITE_SIZE are reported in KiB per dispatch and the scale
turns them into bytes moved to and from HBM.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import statistics
import sys

COPY_KERNEL = "k_copy_probe"


def short_name(kernel: str) -> str:
    n = kernel.split("(")[0]
    for p in ("void ", "wfe::dev::"):
        n = n.replace(p, "")
    return n.strip()


def durations(path: str) -> dict[str, list[float]]:
    out: dict[str, list[float]] = collections.defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            out[short_name(r["Kernel_Name"])].append(
                (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    return out


def counter(path: str, name: str) -> dict[str, list[float]]:
    out: dict[str, list[float]] = collections.defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == name:
                out[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return out


def copy_scale(vals: dict[str, list[float]], copy_bytes: int) -> float | None:
    for k, v in vals.items():
        if COPY_KERNEL in k and v:
            return copy_bytes / (statistics.median(v) * 1024.0)
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--points", type=int, default=0, help="interior points (per-point figures)")
    ap.add_argument("--json")
    a = ap.parse_args(argv)

    dur = durations(a.trace) if a.trace else {}
    fet = counter(a.fetch, "FETCH_SIZE") if a.fetch else {}
    wri = counter(a.write, "WRITE_SIZE") if a.write else {}
    sf = copy_scale(fet, a.copy_bytes)
    sw = copy_scale(wri, a.copy_bytes)
    names = sorted(set(dur) | set(fet) | set(wri),
                   key=lambda n: -sum(dur.get(n, [0.0])))
    print(f"# FETCH_SIZE scale {sf if sf is None else round(sf, 3)}, "
          f"WRITE_SIZE scale {sw if sw is None else round(sw, 3)} "
          f"(copy probe, {a.copy_bytes} B read + written per launch)")
    hdr = f"{'kernel':42s} {'calls':>6s} {'avg_us':>9s} {'fetch_MB':>10s} {'write_MB':>10s}"
    if a.points:
        hdr += f" {'B/pt':>8s} {'GB/s':>8s}"
    print(hdr)
    table = {}
    for n in names:
        d = dur.get(n)
        avg_us = statistics.mean(d) if d else None
        fb = statistics.mean(fet[n]) * 1024.0 * sf if n in fet and sf else None
        wb = statistics.mean(wri[n]) * 1024.0 * sw if n in wri and sw else None
        row = {"calls": len(d) if d else len(fet.get(n, [])), "avg_us": avg_us,
               "fetch_bytes": fb, "write_bytes": wb}
        table[n] = row
        line = (f"{n[:42]:42s} {row['calls']:6d} "
                f"{avg_us if avg_us is not None else float('nan'):9.2f} "
                f"{fb / 1e6 if fb is not None else float('nan'):10.2f} "
                f"{wb / 1e6 if wb is not None else float('nan'):10.2f}")
        if a.points:
            tot = (fb or 0.0) + (wb or 0.0)
            bw = tot / (avg_us * 1e-6) / 1e9 if avg_us and tot else float("nan")
            line += f" {tot / a.points:8.1f} {bw:8.0f}"
        print(line)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"fetch_scale": sf, "write_scale": sw, "copy_bytes": a.copy_bytes,
                       "kernels": table}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
