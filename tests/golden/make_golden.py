"""Write the golden fixtures under tests/golden/ from the reference.

Runs in the development container only (it needs /root/reference):
  * builds oracle/_ref/ from the reference sources (make -C oracle ref):
    the reference Integer_Heap header, its exported ridge model headers and its
    RCM sources, each behind a small driver of ours;
  * heap_traces.json   scripted push/modify/pop traces and the pop order the
                       reference heap (include/integer_heap.cuh) produces;
  * predict_ref.json   feature vectors and the speedups of the reference's
                       rd_sampled_{merbit,csr}_model.hpp;
  * heap_refresh_traces.json  traces with refresh and PMAX resets, the heap
                       operations of the tile-aware order;
  * rcm_ref.json       small COO graphs and R_Cuthill_Mckee's permutation
                       (rcm/src/rcm.cpp, one OpenMP thread: its reversal loop
                       shares a temporary between threads);
  * model_merbit_sampled.json / model_csr_sampled.json  copies of the fitted
                       model data files of ridge/out/*/sampled/.

The traces are generated with the oracle heap deciding which nodes are still
live, so every modify targets a node that is in the heap with a non-negative
net key, as Gorder guarantees; the expected pops then come from the reference.
"""
from __future__ import annotations

import json
import os
import pathlib
import shutil
import subprocess
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
REF = pathlib.Path(os.environ.get("MR_REFERENCE", "/root/reference"))
sys.path.insert(0, str(ROOT / "oracle"))
import oracle  # noqa: E402


def run(binary: str, stdin: str, env=None) -> str:
    res = subprocess.run([str(ROOT / "oracle" / "_ref" / binary)], input=stdin, capture_output=True, text=True,
                         check=True, env=env)
    return res.stdout


def heap_traces(rng) -> list:
    out = []
    for t in range(40):
        n = int(rng.integers(1, 60))
        buckets = 4096
        ops = []
        net = np.zeros(n, np.int64)
        live = np.ones(n, bool)
        for v in range(n):
            ops.append((0, v, 0))
        # start node bonus as Gorder does
        s = int(rng.integers(0, n))
        ops.append((1, s, 1))
        net[s] += 1
        steps = int(rng.integers(n, 6 * n + 1))
        for _ in range(steps):
            if not live.any():
                break
            r = rng.random()
            if r < 0.2:
                ops.append((2, 0, 0))
                popped = oracle.heap_trace(n, buckets, np.array(ops, np.int32))
                live[:] = True
                live[popped] = False
            else:
                cand = np.flatnonzero(live)
                v = int(rng.choice(cand))
                if r < 0.65 or net[v] == 0:
                    d = int(rng.integers(1, 3)) if r < 0.3 else 1
                else:
                    d = -1
                ops.append((1, v, d))
                net[v] += d
        while live.any():
            ops.append((2, 0, 0))
            popped = oracle.heap_trace(n, buckets, np.array(ops, np.int32))
            live[:] = True
            live[popped] = False
        text = f"{n} {buckets} {len(ops)}\n" + "\n".join(f"{a} {b} {c}" for a, b, c in ops) + "\n"
        pops = [int(x) for x in run("heap_trace", text).split()]
        out.append({"n": n, "buckets": buckets, "ops": ops, "pops": pops})
    return out


def heap_refresh_traces(rng) -> list:
    """Traces in the pattern of the tile-aware order (src/class.cu:3158-3365):
    +1 modifies, refreshes of nodes with a positive key (in any order, some
    twice) followed by PMAX = 0, and pops."""
    out = []
    for t in range(30):
        n = int(rng.integers(2, 50))
        buckets = 4096
        ops = [(0, v, 0) for v in range(n)]
        key = np.zeros(n, np.int64)
        live = np.ones(n, bool)
        s = int(rng.integers(0, n))
        ops.append((1, s, 1))
        key[s] += 1
        while live.any():
            r = rng.random()
            cand = np.flatnonzero(live)
            if r < 0.25:
                ops.append((2, 0, 0))
                popped = oracle.heap_trace(n, buckets, np.array(ops, np.int32))
                live[:] = True
                live[popped] = False
                key[~live] = 0
            elif r < 0.35:
                hot = np.flatnonzero(live & (key > 0))
                for v in rng.permutation(np.concatenate([hot, hot[: len(hot) // 2]])):
                    ops.append((3, int(v), 0))
                key[hot] = 0
                ops.append((4, 0, 0))
            else:
                v = int(rng.choice(cand))
                ops.append((1, v, 1))
                key[v] += 1
        text = f"{n} {buckets} {len(ops)}\n" + "\n".join(f"{a} {b} {c}" for a, b, c in ops) + "\n"
        pops = [int(x) for x in run("heap_trace", text).split()]
        out.append({"n": n, "buckets": buckets, "ops": ops, "pops": pops})
    return out


def predict_ref(rng) -> list:
    rows = []
    feats = [np.zeros(10), np.full(10, 0.1)]
    for _ in range(30):
        f = rng.dirichlet(np.ones(10) * 0.5)
        feats.append(f)
    for _ in range(8):
        feats.append(rng.uniform(-1, 2, 10))
    text = "\n".join(" ".join(repr(float(v)) for v in f) for f in feats) + "\n"
    lines = run("model_eval", text).strip().split("\n")
    for f, line in zip(feats, lines):
        a, b = line.split()
        rows.append({"feat": [float(v) for v in f], "merbit": float(a), "csr": float(b)})
    return rows


def rcm_ref(rng) -> list:
    out = []
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for t in range(25):
        n = int(rng.integers(2, 120))
        m = int(rng.integers(0, 4 * n))
        rows = rng.integers(0, n, m).tolist()
        cols = rng.integers(0, n, m).tolist()
        if t % 5 == 0 and m > 3:   # a few self loops and duplicates
            rows[0] = cols[0]
            rows.append(rows[1]); cols.append(cols[1])
        text = f"{n} {len(rows)}\n" + "\n".join(f"{a} {b}" for a, b in zip(rows, cols)) + "\n"
        perm = [int(x) for x in run("rcm_driver", text, env).split()]
        out.append({"n": n, "rows": rows, "cols": cols, "perm": perm})
    return out


def main() -> None:
    if not REF.exists():
        raise SystemExit(f"{REF} not found: fixtures can only be regenerated next to the reference")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "all", "ref", f"REF={REF}"], check=True)
    rng = np.random.default_rng(20261203)
    (HERE / "heap_traces.json").write_text(json.dumps(heap_traces(rng)))
    (HERE / "predict_ref.json").write_text(json.dumps(predict_ref(rng), indent=0))
    (HERE / "rcm_ref.json").write_text(json.dumps(rcm_ref(rng)))
    (HERE / "heap_refresh_traces.json").write_text(json.dumps(heap_refresh_traces(rng)))
    shutil.copy(REF / "ridge/out/merbit/sampled/merbit_sampled_deploy_train_all_aug_model.json",
                HERE / "model_merbit_sampled.json")
    shutil.copy(REF / "ridge/out/csr/sampled/csr_sampled_deploy_train_all_aug_model.json",
                HERE / "model_csr_sampled.json")
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
