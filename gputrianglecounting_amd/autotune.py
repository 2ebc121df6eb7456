"""Autotuner front end: runs the library's tuner (tc_autotune_i32 /
tc_autotune_batch_i32, include/tc.h) on a workload and keeps the results in
``tuned.json`` beside this file, keyed by workload name, so later runs can
start from the tuned launch configuration without re-timing.

Searched space (library side, in stages around the fastest so far): the
oriented scheme's partner side x partners per wave, then its count
kernel's LDS window x workgroup size, then the piece size (partition
granularity), then lookup x long rows; last the edge-list scheme's kernel
variants (workgroup size, merge steps per lane, LDS ring length) x
small-edge packing threshold x level-1 ranges per resident workgroup.
Every candidate must reproduce the first candidate's per-vertex counts bit
for bit."""
from __future__ import annotations

import json
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
CFG_KEYS = ("wg_size", "items", "ring", "small_max", "ranges_per_wg", "algo", "chunk", "depth", "search_ratio", "probe_ratio", "oriented", "row_groups", "lookup", "long_rows",
            "window", "orient_wg", "piece")


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_entry(key: str, best: dict, rows: list, meta: dict | None = None, path: str = TABLE) -> None:
    table = load_table(path)
    table[key] = {"best": {k: best[k] for k in CFG_KEYS}, "candidates": rows, "meta": meta or {},
                  "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S")}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def best_cfg(key: str, path: str = TABLE):
    """Tuned configuration for ``key`` or None."""
    e = load_table(path).get(key)
    return dict(e["best"]) if e else None


def apply(counter, key: str, path: str = TABLE) -> bool:
    """Make the tuned configuration of ``key`` the counter's default."""
    cfg = best_cfg(key, path)
    if cfg is None:
        return False
    counter.set_default_cfg(**cfg)
    return True


def tune_graph(counter, key, off, ind, out, reps=2, meta=None, path: str = TABLE):
    best, rows = counter.autotune(off, ind, out, reps=reps, set_default=True)
    save_entry(key, best, rows, meta, path)
    return best, rows


def tune_batch(counter, key, tab, off_cat, ind_cat, nv_total, nnz_total, out, reps=2, meta=None, path: str = TABLE):
    best, rows = counter.autotune_batch(tab, off_cat, ind_cat, nv_total, nnz_total, out, reps=reps, set_default=True)
    save_entry(key, best, rows, meta, path)
    return best, rows


def main(argv=None) -> int:
    """``python -m gputrianglecounting_amd.autotune batch4096`` tunes the mixed
    4096-graph batch (n log-uniform in [64, 65536], ef 8..32, seed 7) and
    stores its entry in the table."""
    import argparse

    import torch

    from . import graph as G
    from .tc import TriangleCounter

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("workload", choices=["batch4096"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--table", default=TABLE)
    a = ap.parse_args(argv)
    b = G.gen_batch(G=4096, n_lo=64, n_hi=65536, ef_lo=8, ef_hi=32, seed=7)
    tc = TriangleCounter(0)
    try:
        tab, off, ind = tc.batch_to_device(b)
        out = torch.empty(b.total_vertices, dtype=torch.int32, device="cuda")
        meta = {"graphs": b.count, "vertices": b.total_vertices, "arcs": b.total_arcs, "seed": 7,
                "device": torch.cuda.get_device_name(0)}
        best, rows = tune_batch(tc, a.workload, tab, off, ind, b.total_vertices, b.total_arcs, out, reps=a.reps,
                                meta=meta, path=a.table)
        ok = [r for r in rows if r["identical"]]
        print(json.dumps({"best": {k: best[k] for k in CFG_KEYS}, "candidates": len(rows),
                          "best_ms": min(r["ms"] for r in ok) if ok else None}))
    finally:
        tc.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
