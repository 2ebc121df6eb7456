"""Autotuner front end: runs the library's tuner (tc_autotune_i32 /
tc_autotune_batch_i32, include/tc.h) on a workload and keeps the results in
``tuned.json`` beside this file, keyed by workload name, so later runs can
start from the tuned launch configuration without re-timing.

Searched space (library side): kernel variant (workgroup size, merge steps
per lane, LDS ring length) x small-edge packing threshold x level-1 ranges
per resident workgroup.  Every candidate must reproduce the first
candidate's per-vertex counts bit for bit."""
from __future__ import annotations

import json
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
CFG_KEYS = ("wg_size", "items", "ring", "small_max", "ranges_per_wg", "algo", "chunk", "depth", "search_ratio", "probe_ratio", "oriented", "row_groups", "lookup", "long_rows")


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
