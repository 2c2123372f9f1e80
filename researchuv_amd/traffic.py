"""HBM traffic per launch of a kernel from two rocprofv3 counter runs (FETCH_SIZE, then WRITE_SIZE).

Usage (from the repository root, after the two profiled runs of the same command):

    python -m researchuv_amd.traffic --fetch DIR_A --write DIR_B --kernel 'kb_spmv<float, 2, 0, 0, 0>' \
        --copy-kernel copyBuffer --copy-bytes 1073741824 --algorithmic BYTES_PER_LAUNCH --out FILE.json

FETCH_SIZE and WRITE_SIZE are read in KiB per dispatch and averaged over the dispatches whose name contains
`--kernel`.  Each is scaled by what the same counter reads for the calibration copy of `--copy-bytes` bytes
(read once, written once) in the same run: scale = copy bytes / (counter KiB * 1024).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os


def _rows(path):
    files = [path] if os.path.isfile(path) else glob.glob(os.path.join(path, "**", "*counter_collection.csv"),
                                                          recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {path}")
    for f in files:
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def mean_counter(path, counter, name_part):
    """Mean over dispatches whose kernel name contains name_part of the per-dispatch sum of `counter`."""
    per = {}
    for r in _rows(path):
        if r.get("Counter_Name") != counter or name_part not in r.get("Kernel_Name", ""):
            continue
        key = r.get("Dispatch_Id") or r.get("Correlation_Id")
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    if not per:
        raise ValueError(f"no {counter} records for kernels matching {name_part!r}")
    return sum(per.values()) / len(per), len(per)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--copy-kernel", required=True)
    ap.add_argument("--copy-bytes", type=float, required=True)
    ap.add_argument("--algorithmic", type=float, default=None, help="algorithmic bytes per launch, for the ratio")
    ap.add_argument("--out", required=True)
    ap.add_argument("--table", nargs="*", default=[],
                    help="also print read / write MB per dispatch of the kernels whose names contain these")
    a = ap.parse_args(argv)
    f_k, n_f = mean_counter(a.fetch, "FETCH_SIZE", a.kernel)
    w_k, n_w = mean_counter(a.write, "WRITE_SIZE", a.kernel)
    f_c, _ = mean_counter(a.fetch, "FETCH_SIZE", a.copy_kernel)
    w_c, _ = mean_counter(a.write, "WRITE_SIZE", a.copy_kernel)
    sf = a.copy_bytes / (f_c * 1024.0)
    sw = a.copy_bytes / (w_c * 1024.0)
    fetch_b = f_k * 1024.0 * sf
    write_b = w_k * 1024.0 * sw
    out = {"kernel": a.kernel, "dispatches": [n_f, n_w], "fetch_kib_raw": f_k, "write_kib_raw": w_k,
           "copy_fetch_kib": f_c, "copy_write_kib": w_c, "copy_bytes": a.copy_bytes, "scale_fetch": sf,
           "scale_write": sw, "hbm_fetch_bytes_per_launch": fetch_b, "hbm_write_bytes_per_launch": write_b,
           "hbm_bytes_per_launch": fetch_b + write_b}
    if a.algorithmic:
        out["algorithmic_bytes_per_launch"] = a.algorithmic
        out["traffic_over_algorithmic"] = (fetch_b + write_b) / a.algorithmic
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")
    print(json.dumps(out))
    if a.table:
        print("| Kernel | Dispatches | Read (MB) | Write (MB) |")
        print("|---|---:|---:|---:|")
        for k in a.table:
            fk, nk = mean_counter(a.fetch, "FETCH_SIZE", k)
            wk, _ = mean_counter(a.write, "WRITE_SIZE", k)
            print(f"| `{k}` | {nk} | {fk * 1024.0 * sf / 1e6:.1f} | {wk * 1024.0 * sw / 1e6:.1f} |")


if __name__ == "__main__":
    main()
