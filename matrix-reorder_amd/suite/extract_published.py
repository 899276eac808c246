"""Write rmat_s24_published.json: the reference's published MERBIT records for
its Graph500 R-MAT s24 sweep, as data.

Runs only next to the reference (it reads two of its data files):
  config/config3.txt      id, file name (rmat_s24_ef{E}_{A}_{B}.mtx: a = A/1e4,
                          b = c = B/1e4), rows, cols, nnz of the symmetric file
  test_rmat_meribit.txt   per id an "origin" and a "gorder1" record:
                          id&method&&reorder_ms&relabel_ms&r1..r10&tile_ms&merbit_ms
The records were measured on an sm_89 GPU with FP32 values 1 and x = 1; the
multiplied matrix holds twice the file's nonzeros (symmetric doubling).

Usage: python matrix-reorder_amd/suite/extract_published.py [reference root]
"""
from __future__ import annotations

import json
import pathlib
import re
import sys

HERE = pathlib.Path(__file__).resolve().parent


def main() -> None:
    ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    cfg = {}
    for line in (ref / "config" / "config3.txt").read_text().splitlines():
        parts = line.split()
        if len(parts) < 5:
            continue
        m = re.search(r"rmat_s(\d+)_ef(\d+)_(\d+)_(\d+)\.mtx", parts[1])
        if not m:
            continue
        cfg[int(parts[0])] = {"scale": int(m.group(1)), "ef": int(m.group(2)), "a": int(m.group(3)) / 1e4,
                              "b": int(m.group(4)) / 1e4, "file_nnz": int(parts[4])}
    rows = {}
    for line in (ref / "test_rmat_meribit.txt").read_text().splitlines():
        f = line.strip().split("&")
        if len(f) != 17:
            continue
        gid, method = int(f[0]), f[1]
        rec = rows.setdefault(gid, dict(cfg[gid], id=gid))
        rec[f"{method}_merbit_ms"] = float(f[16])
        rec[f"{method}_tile_ms"] = float(f[15])
        if method != "origin":
            rec[f"{method}_reorder_ms"] = float(f[3])
            rec[f"{method}_relabel_ms"] = float(f[4])
    out = {"source": "config/config3.txt + test_rmat_meribit.txt (sm_89, FP32, values 1, x 1)",
           "matrices": [rows[k] for k in sorted(rows)]}
    (HERE / "rmat_s24_published.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(rows)} matrices written")


if __name__ == "__main__":
    main()
