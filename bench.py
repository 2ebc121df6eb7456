"""Benchmark: per-vertex triangle counting on one MI355X.

Default workload: R-MAT scale 24, edge factor 16 (a=.57 b=.19 c=.19),
seed 1, simple symmetric int32 CSR (~520M arcs), counted by the default
configuration: the degree orientation written in place of the input rows
(k_huge_rows, k_orient_rows), then k_count_hash (a cuckoo table of each
row in LDS, partner lists read in wave-coalesced vectors; rows longer than
a wave's table counted first by whole workgroups).  A step is one full
tc_count_i32 call over the resident graph: the orientation, the count
kernel and the int64 total.  The metric is undirected edges / step time
(edge-intersections/s); the roofline entry is for the count kernel, timed
with HIP events on its launch stream.  The default run adds the batched
comparison (4096 mixed graphs counted as one batch against one launch
sequence per graph, with the batched kernel's own roofline entry) and a
short serving leg (a burst, then Poisson arrivals at half the burst rate:
graphs/s and p50/p99 latency).

The CPU baseline runs the reference's own intersection (oracle/_ref, or the
C restatement when that is absent) on a bounded uniform sample of arcs of
the same graph on one host core; its full-graph time is extrapolated by arc
count.  Beside it, the OpenMP restatement of the vertex-list count runs on
all the host threads the run may use (at most 16) over a uniform vertex
sample, extrapolated the same way.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "edge-intersections/s + HBM GB/s (R-MAT s24); batched vs per-graph graphs/s"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def make_graph(workload: str):
    from gputrianglecounting_amd import graph as G

    if workload == "rmat24":
        return G.rmat(24, 16, seed=1), dict(workload="rmat_s24_ef16", scale=24, edge_factor=16, a=0.57, b=0.19,
                                            c=0.19, seed=1)
    if workload == "rmat26":
        return G.rmat(26, 16, seed=1), dict(workload="rmat_s26_ef16", scale=26, edge_factor=16, a=0.57, b=0.19,
                                            c=0.19, seed=1)
    if workload == "rmat20":
        return G.rmat(20, 16, seed=1), dict(workload="rmat_s20_ef16", scale=20, edge_factor=16, a=0.57, b=0.19,
                                            c=0.19, seed=1)
    if workload == "rgg":
        return G.geometric(4194304, 6.0, seed=1), dict(workload="rgg_n4194304_deg6", seed=1)
    if workload == "er":
        return G.erdos_renyi(16384, 131072, seed=1), dict(workload="er_n16384_m131072", seed=1)
    raise SystemExit(f"unknown workload {workload}")


SEARCH_MIN_B = 512  # kSearchMinB in tc_kernels.hip


def algorithmic_bytes(g, small_max: int, search_ratio: int) -> dict:
    """Bytes the count moves by definition of the algorithm it runs.

    Every intersected edge (u < w, both degrees >= 2) with shorter list A
    and longer list B is either streamed, reading both lists once
    (4 (|A| + |B|) bytes), or, when |B| >= 512 and |B| >= search_ratio |A|
    (and the edge is not a small one, |A| + |B| > small_max), counted by
    lookups: every element of A read once and binary-searched in B, one
    4-byte probe per level (4 |A| (1 + ceil(log2(|B| + 1))) bytes).  Plus
    reading the CSR once and writing int32 counts and the int64 total.
    ``b_alg_all`` is the all-streamed figure (4 * sum d^2 + the same fixed
    part), the formula of a pure merge."""
    off = g.off.astype(np.int64)
    ind = g.ind
    d = np.diff(off)
    fixed = 4 * (g.n + 1) + 4 * g.nnz + 4 * g.n + 8
    stream_b = search_b = 0.0
    stream_units = 0
    n_search = n_edges = 0
    r0 = 0
    while r0 < g.n:  # row blocks of about 2^25 arcs
        r1 = int(np.searchsorted(off, off[r0] + (1 << 25), side="right"))
        r1 = min(max(r1, r0 + 1), g.n)
        src = np.repeat(np.arange(r0, r1, dtype=np.int64), d[r0:r1])
        dst = ind[off[r0]:off[r1]].astype(np.int64)
        m = src < dst
        du, dw = d[src[m]], d[dst[m]]
        keep = (du >= 2) & (dw >= 2)
        a = np.minimum(du, dw)[keep]
        b = np.maximum(du, dw)[keep]
        L = a + b
        srch = (L > small_max) & (b >= SEARCH_MIN_B) & (search_ratio > 0) & (b >= search_ratio * a)
        stream_units += int(L[~srch].sum())
        stream_b += 4.0 * float(L[~srch].sum())
        la = a[srch].astype(np.float64)
        search_b += float((4.0 * la * (1.0 + np.ceil(np.log2(b[srch].astype(np.float64) + 1.0)))).sum())
        n_search += int(srch.sum())
        n_edges += int(keep.sum())
        r0 = r1
    dd = d.astype(np.float64)
    return dict(b_kernel=stream_b + search_b + fixed, b_stream=stream_b, b_search=search_b,
                b_alg_all=4 * float((dd * dd).sum()) + fixed, search_edges=n_search, edges=n_edges,
                stream_units=stream_units)


def algorithmic_bytes_oriented(g, side: int = 0) -> dict:
    """Bytes of the oriented per-vertex count by definition.

    Arcs v -> w are kept iff (d(v), v) < (d(w), w).  Each oriented row is
    read once into the table (4 d+ bytes), and for every arc u -> w one
    partner list is streamed after its two row offsets (16 bytes): N+(w)
    when rows take their out-neighbours as partners (``side`` 0), N+(u) when
    they take their in-neighbours (``side`` 1; ``side`` 2 takes them unless
    the input has rows of more than 4096 arcs, as the count kernels decide
    from the prep's hub-row list).  Counts and the total are written once.  Reading the CSR to
    orient it is prep, outside the count kernel."""
    off = g.off.astype(np.int64)
    ind = g.ind
    d = np.diff(off)
    dplus = np.zeros(g.n, np.int64)
    blocks = []
    r0 = 0
    while r0 < g.n:  # row blocks of about 2^25 arcs
        r1 = int(np.searchsorted(off, off[r0] + (1 << 25), side="right"))
        r1 = min(max(r1, r0 + 1), g.n)
        blocks.append((r0, r1))
        r0 = r1
    for r0, r1 in blocks:
        src = np.repeat(np.arange(r0, r1, dtype=np.int64), d[r0:r1])
        dst = ind[off[r0]:off[r1]].astype(np.int64)
        keep = (d[src] < d[dst]) | ((d[src] == d[dst]) & (src < dst))
        dplus += np.bincount(src[keep], minlength=g.n)
    arcs = int(dplus.sum())
    by_in = side == 1 or (side == 2 and int(d.max(initial=0)) <= 4096)
    stream = 0
    for r0, r1 in blocks:
        src = np.repeat(np.arange(r0, r1, dtype=np.int64), d[r0:r1])
        dst = ind[off[r0]:off[r1]].astype(np.int64)
        keep = (d[src] < d[dst]) | ((d[src] == d[dst]) & (src < dst))
        stream += int(dplus[(src if by_in else dst)[keep]].sum())
    b = 4.0 * (arcs + stream) + 16.0 * arcs + 4 * g.n + 8
    return dict(b_kernel=b, oriented_arcs=arcs, oriented_stream_elements=stream,
                max_out_degree=int(dplus.max(initial=0)), partners="in-neighbours" if by_in else "out-neighbours")


def cpu_baseline(g, target_s: float) -> dict:
    import oracle

    use_ref = oracle.ref_lib() is not None
    rng = np.random.default_rng(12345)

    def run(k: int):
        arcs = np.sort(rng.choice(g.nnz, size=k, replace=False)).astype(np.int64)
        t0 = time.perf_counter()
        s = oracle.count_arc_sample(g.off, g.ind, arcs, use_ref=use_ref)
        return time.perf_counter() - t0, s

    k = 20000
    t, _ = run(k)
    k2 = int(min(g.nnz, max(k, k * target_s / max(t, 1e-3))))
    t, _ = run(k2)
    full_s = t * g.nnz / k2
    edges = g.nnz // 2
    cpu = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": edges / full_s, "unit": "edge-intersections/s", "cores": 1,
            "kind": "reference" if use_ref else "port",
            "sample": (f"{k2} arcs drawn uniformly (seed 12345) from the same graph, visited in CSR order by "
                       f"{'the reference count_intersectionsReg (clusteringCount.cpp:2-48)' if use_ref else 'the C restatement'}"
                       f" on 1 core ({cpu}); {t:.1f} s, full-graph time extrapolated by arc count: {full_s:.0f} s")}


class _CSR:
    """A plain global CSR (the batch as one disjoint graph) for the byte count."""

    def __init__(self, off, ind):
        self.off, self.ind = off, ind
        self.n, self.nnz = int(off.shape[0] - 1), int(ind.shape[0])


def batch_as_graph(b) -> _CSR:
    """The packed batch as one disjoint CSR in global vertex ids: graph g's
    rows are v_base[g] .. v_base[g+1], its offsets rebased by arc_base[g]
    (the same map k_voff_batch applies on the device)."""
    vb, ab = b.tab[0::2].astype(np.int64), b.tab[1::2].astype(np.int64)
    G_ = b.count
    ng = np.diff(vb)
    gid = np.repeat(np.arange(G_, dtype=np.int64), ng)
    off = b.off[np.arange(vb[G_], dtype=np.int64) + gid].astype(np.int64) + ab[:G_][gid]
    off = np.append(off, ab[G_])
    del gid
    ind = b.ind.astype(np.int64) + np.repeat(vb[:G_], np.diff(ab))
    return _CSR(off, ind.astype(np.int32))


def cpu_baseline_all_cores(g, target_s: float) -> dict:
    """The OpenMP restatement of the reference's vertex-list count on all
    the host threads this run may use (at most 16), over a uniform sample of
    vertices: every arc is in the sample with the same probability, so the
    full-graph time is extrapolated by arc count."""
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(54321)
    d = g.degrees()

    def run(k: int):
        vl = np.sort(rng.choice(g.n, size=k, replace=False)).astype(np.int32)
        t0 = time.perf_counter()
        oracle.vertex_list_omp(vl, g.off, g.ind, threads=threads)
        return time.perf_counter() - t0, int(d[vl].sum())

    k = 2000
    t, a = run(k)
    k2 = int(min(g.n, max(k, k * target_s / max(t, 1e-3))))
    t, a = run(k2)
    full_s = t * g.nnz / max(a, 1)
    return {"value": (g.nnz // 2) / full_s, "unit": "edge-intersections/s", "cores": threads, "kind": "port",
            "sample": f"{k2} vertices drawn uniformly (seed 54321), {a} arcs, counted by the OpenMP restatement of "
                      f"clusteringCount.cpp:125-155 on {threads} threads; {t:.1f} s, full-graph time extrapolated "
                      f"by arc count: {full_s:.0f} s"}


def run_batch(tc, steps: int, warmup: int, G_: int = 4096) -> dict:
    """The mixed batch: G graphs (n log-uniform in [64, 65536]; ER mean degree
    8, R-MAT a=.57 b=.19 c=.19 ef 8-32, RGG mean degree 6) packed into one
    CSR, counted by one captured HIP graph per step, against one
    tc_count_i32 launch sequence per graph on the same GPU."""
    import torch

    from gputrianglecounting_amd import autotune as AT
    from gputrianglecounting_amd import graph as G

    t0 = time.time()
    b = G.gen_batch(G=G_, n_lo=64, n_hi=65536, ef_lo=8, ef_hi=32, seed=7)
    log(f"batch of {b.count} graphs, {b.total_vertices} vertices, {b.total_arcs} arcs in {time.time() - t0:.1f}s")
    saved = tc.default_cfg()
    tuned = AT.apply(tc, f"batch{G_}")
    cfg = tc.default_cfg()
    tab, off, ind = tc.batch_to_device(b)
    out = torch.empty(b.total_vertices, dtype=torch.int32, device="cuda")
    totals = torch.empty(b.count, dtype=torch.int64, device="cuda")
    graph = tc.capture_batch(tab, off, ind, b.total_vertices, b.total_arcs, out, totals)
    for _ in range(warmup):
        graph.launch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        graph.launch()
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t1) / steps
    out_b, tot_b = out.clone(), totals.clone()
    # one launch sequence per graph, with the automatic configuration of a
    # graph counted alone (the tuned one is the batch's: its table lookup
    # makes the small graphs slower when each runs alone)
    tc.set_default_cfg(**saved)
    out2 = torch.empty_like(out)
    tot2 = torch.empty_like(totals)
    tc.batch_count_pergraph(b.tab, off, ind, out2, tot2)  # warm
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(max(1, steps // 2)):
        tc.batch_count_pergraph(b.tab, off, ind, out2, tot2)
    torch.cuda.synchronize()
    per_s = (time.perf_counter() - t2) / max(1, steps // 2)
    same = bool(torch.equal(out_b, out2) and torch.equal(tot_b, tot2))
    tc.set_default_cfg(**cfg)
    h = out_b.cpu().numpy().astype(np.int64)
    dv = np.diff(b.off.astype(np.int64))
    # per-vertex degrees: graph g's offsets sit at off[v_base[g] + g ...], so
    # the difference at v_base[g+1] + g straddles two graphs and is dropped
    keep = np.ones(dv.shape[0], bool)
    keep[b.tab[2:2 * b.count:2] + np.arange(b.count - 1)] = False
    deg = dv[keep]
    inv = bool((h % 2 == 0).all() and (h <= deg * (deg - 1)).all())
    graph.close()
    # roofline of the batched count kernel: one eager count with the
    # context's HIP events around it (on the launch stream), against the
    # oriented bytes of the batch as one disjoint graph (in-neighbour
    # partners, the batches' side)
    tc.set_timing(True)
    ks = []
    for _ in range(3):
        tc.batch_count(tab, off, ind, b.total_vertices, b.total_arcs, out, totals)
        ks.append(tc.last_timing().count_ms)
    tc.set_timing(False)
    k_ms = float(np.median(ks))
    ab = algorithmic_bytes_oriented(batch_as_graph(b), side=1)
    roof = {"bound": "hbm", "achieved": round(ab["b_kernel"] / (k_ms / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ab["b_kernel"] / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": ("k_count_hash" if cfg.get("lookup") in (2, 3) else "k_count_orient")
            + " (batch, in-neighbour partners)", "count_kernel_ms": round(k_ms, 3),
            "b_kernel_bytes": ab["b_kernel"]}
    tc.set_default_cfg(**saved)
    return {"graphs": b.count, "vertices": b.total_vertices, "arcs": b.total_arcs,
            "batched_graphs_per_s": b.count / batched_s, "batched_ms": batched_s * 1e3,
            "per_graph_graphs_per_s": b.count / per_s, "per_graph_ms": per_s * 1e3,
            "speedup": per_s / batched_s, "identical_to_per_graph": same, "invariants": inv,
            "kernel_cfg": cfg, "autotuned": tuned, "captured": "hipGraph", "roofline": roof}


def run_serve(tc, pool: int = 48, per_rate: int = 192, rel_rates=(0.25, 0.5, 0.75, 0.9),
              budget: int = 32 << 30, slots: int = 4, max_batch: int = 64, seed: int = 11) -> dict:
    """The serving loop under Poisson arrivals: R-MAT requests with n
    log-uniform in [2^10, 2^20] and edge factor uniform in 8..32 (ids >= n
    dropped), a 32 GiB admission budget.  First a closed burst of 2*pool
    requests gives the saturation rate; then open-loop Poisson arrivals at
    fractions of it give latency percentiles (submit to results on the host,
    measured by the server).  Every served result is checked against the
    same graph counted alone."""
    import threading

    from gputrianglecounting_amd import graph as G
    from gputrianglecounting_amd import serve as S

    rng = np.random.default_rng(seed)
    graphs = []
    for i in range(pool):
        n = int(np.exp(rng.uniform(np.log(1024), np.log(1 << 20))))
        ef = int(rng.integers(8, 33))
        graphs.append(G.rmat(int(np.ceil(np.log2(n))), ef, seed=1000 + i, n_keep=n))
    arcs = np.array([g.nnz for g in graphs])
    log(f"serve pool: {pool} graphs, arcs median {int(np.median(arcs))}, max {arcs.max()}")
    alone = [tc.count_graph(g) for g in graphs]
    srv = S.Server(hbm_budget=budget, slots=slots, max_batch_graphs=max_batch)
    bad = 0

    def check(k, r):
        nonlocal bad
        if r.total != alone[k][1] or not np.array_equal(r.out, alone[k][0]):
            bad += 1

    import gc

    # the collector's full passes over this process's heap (torch and numpy
    # loaded) take tens of milliseconds: everything allocated so far is moved
    # out of its reach, so a pause does not land inside a timed burst
    gc.collect()
    gc.freeze()
    try:
        # warm the slots' contexts, then one untimed burst: the server's
        # pinned pool is mapped into the process on first touch (~3 GB/s;
        # 30 GB/s after), and a running server reuses the same slabs
        for k in range(min(4, pool)):
            check(k, srv.wait(srv.submit(graphs[k])))
        order = [int(x) for x in rng.integers(0, pool, 2 * pool)]
        for i in [srv.submit(graphs[k]) for k in order]:
            srv.wait(i)
        t0 = time.perf_counter()
        ids = [srv.submit(graphs[k]) for k in order]
        t_submitted = time.perf_counter() - t0
        res = [srv.wait(i) for i in ids]
        t_burst = time.perf_counter() - t0
        for k, r in zip(order, res):
            check(k, r)
        sat = len(order) / t_burst
        log(f"serve burst: {len(order)} requests in {t_burst * 1e3:.0f} ms = {sat:.1f} graphs/s "
            f"(all submitted after {t_submitted * 1e3:.0f} ms)")
        # the device alone: the same requests packed in batches of max_batch,
        # resident in HBM, counted back to back (no host copies, no queue)
        import torch

        batches = [G.pack([graphs[k] for k in order[i:i + max_batch]]) for i in range(0, len(order), max_batch)]
        dev = [tc.batch_to_device(bb) for bb in batches]
        for bb, (tab, off, ind) in zip(batches, dev):  # warm
            tc.batch_count(tab, off, ind, bb.total_vertices, bb.total_arcs)
        torch.cuda.synchronize()
        t_dev = time.perf_counter()
        for bb, (tab, off, ind) in zip(batches, dev):
            tc.batch_count(tab, off, ind, bb.total_vertices, bb.total_arcs)
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t_dev
        dev_rate = len(order) / t_dev
        del dev
        log(f"serve burst on the device alone: {t_dev * 1e3:.1f} ms = {dev_rate:.1f} graphs/s")
        rates = []
        for rel in rel_rates:
            lam = rel * sat
            order = [int(x) for x in rng.integers(0, pool, per_rate)]
            gaps = rng.exponential(1.0 / lam, per_rate)
            ids = [None] * per_rate

            def producer():
                t_next = time.perf_counter()
                for j, k in enumerate(order):
                    t_next += gaps[j]
                    dt = t_next - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                    ids[j] = srv.submit(graphs[k])

            th = threading.Thread(target=producer)
            t1 = time.perf_counter()
            th.start()
            th.join()
            lat = []
            for j, k in enumerate(order):
                r = srv.wait(ids[j])
                check(k, r)
                lat.append(r.latency_ms)
            span = time.perf_counter() - t1
            lat = np.array(lat)
            rates.append({"offered_graphs_per_s": round(lam, 2), "achieved_graphs_per_s": round(per_rate / span, 2),
                          "requests": per_rate, "p50_ms": round(float(np.percentile(lat, 50)), 3),
                          "p99_ms": round(float(np.percentile(lat, 99)), 3),
                          "mean_ms": round(float(lat.mean()), 3)})
            log(f"serve at {lam:.1f}/s: p50 {rates[-1]['p50_ms']} ms, p99 {rates[-1]['p99_ms']} ms")
        st = srv.stats()
    finally:
        srv.close()
    return {"saturation_graphs_per_s": sat, "burst_requests": 2 * pool, "burst_ms": t_burst * 1e3,
            "burst_submit_ms": t_submitted * 1e3,
            "device_only_graphs_per_s": dev_rate, "device_only_ms": t_dev * 1e3, "rates": rates,
            "results_identical_to_alone": bad == 0, "pool_arcs_median": int(np.median(arcs)),
            "pool_arcs_max": int(arcs.max()), "hbm_budget": budget, "slots": slots, "max_batch_graphs": max_batch,
            "peak_arena_bytes": st["peak_bytes"], "batches": st["batches"]}


def traffic_from_profile(workload: str):
    """HBM bytes per count step (both count kernels) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"count_{workload}_hbm.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat24", choices=["rmat24", "rmat20", "rmat26", "rgg", "er", "batch", "serve"])
    ap.add_argument("--int64", action="store_true", help="time the int64 CSR entry point (tc_count_i64)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cfg", default="", help="kernel config, e.g. wg_size=256,items=8,ring=4096")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched vs per-graph comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the short serving leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the count runs on a single MI355X (no multi-GPU path is built); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from gputrianglecounting_amd.tc import TriangleCounter

    if args.workload == "batch":
        tc = TriangleCounter(0)
        res = run_batch(tc, steps=max(2, args.steps), warmup=args.warmup)
        value = res["batched_graphs_per_s"]
        print(json.dumps({"metric": METRIC, "value": value, "unit": "graphs/s", "n_gpus": 1, "steps": max(2, args.steps),
                          "warmup": args.warmup, "ms_per_step": res["batched_ms"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "int32",
                          "data": "synthetic: 4096 seeded ER/R-MAT/RGG graphs",
                          "config": {"workload": "batch_4096_mixed"}, "batched": res}), flush=True)
        tc.close()
        return 0

    if args.workload == "serve":
        tc = TriangleCounter(0)
        res = run_serve(tc)
        print(json.dumps({"metric": "serving graphs/s at saturation, p50/p99 latency under Poisson arrivals",
                          "value": res["saturation_graphs_per_s"], "unit": "graphs/s", "n_gpus": 1,
                          "steps": res["burst_requests"], "warmup": 4, "ms_per_step": res["burst_ms"] / res["burst_requests"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
                          "data": "synthetic: 48 seeded R-MAT graphs, n log-uniform in [2^10, 2^20], ef 8..32",
                          "config": {"workload": "serve_rmat_1k_1m_poisson"}, "serve": res}), flush=True)
        tc.close()
        return 0

    log(f"generating {args.workload}")
    t0 = time.time()
    g, config = make_graph(args.workload)
    log(f"graph n={g.n} arcs={g.nnz} in {time.time() - t0:.1f}s")

    cfg = None
    if args.cfg:
        cfg = {k: int(v) for k, v in (kv.split("=") for kv in args.cfg.split(","))}

    tc = TriangleCounter(0)
    dev = torch.device("cuda", 0)
    off = torch.from_numpy(g.off).to(dev)
    ind = torch.from_numpy(g.ind).to(dev)
    out = torch.empty(g.n, dtype=torch.int32, device=dev)
    total = torch.empty(1, dtype=torch.int64, device=dev)
    tc.set_timing(True)
    torch.cuda.synchronize()
    log("uploaded; warmup")

    for _ in range(args.warmup):
        tc.count(off, ind, out, total, cfg=cfg)
        torch.cuda.synchronize()
        log(f"warmup step {tc.last_timing()}")

    # full-size checks (untimed): invariants and an independent total from
    # the degree-ordered global-count mode
    h_out = out.cpu().numpy().astype(np.int64)
    d = g.degrees()
    tot = int(total.cpu().item())
    inv_ok = bool((h_out % 2 == 0).all() and (h_out <= d * (d - 1)).all() and (h_out >= 0).all()
                  and tot == int(h_out.sum()) and tot % 6 == 0)
    tot_global = int(tc.count_total(off, ind).cpu().item())
    global_ms = tc.last_timing().count_ms
    # an independent kernel: every undirected edge's full lists (no orientation)
    e_out, e_tot = tc.count(off, ind, cfg=dict(cfg or {}, oriented=-1))
    tot_edges = int(e_tot.cpu().item())
    edges_identical = bool(torch.equal(e_out, out))
    del e_out
    log(f"checks: invariants={inv_ok} total={tot} global-mode total={tot_global} ({global_ms:.1f} ms) "
        f"edge-scheme total={tot_edges} per-vertex identical={edges_identical}")

    int64_identical = None
    if args.int64:  # the same graph as an int64 CSR through tc_count_i64
        off, ind = off.to(torch.int64), ind.to(torch.int64)
        out64 = torch.empty(g.n, dtype=torch.int64, device=dev)
        tc.count(off, ind, out64, total, cfg=cfg)
        torch.cuda.synchronize()
        int64_identical = bool(torch.equal(out64, out.to(torch.int64)))
        out = out64
        log(f"int64 CSR: per-vertex identical={int64_identical}")

    # A step whose kernels take a few tens of microseconds is timed back to
    # back, without reading its events (a wait per step would add the host's
    # round trip to every step); its kernel times then come from as many
    # event-timed steps right after.  Larger steps read their events inside
    # the timed loop.
    tc.count(off, ind, out, total, cfg=cfg)
    small = tc.last_timing().count_ms + tc.last_timing().prep_ms < 2.0
    torch.cuda.synchronize()
    kern = []
    prep = []
    if small:
        tc.set_timing(False)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        tc.count(off, ind, out, total, cfg=cfg)
        if not small:
            tm = tc.last_timing()  # waits for this step's events
            kern.append(tm.count_ms)
            prep.append(tm.prep_ms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed * 1e3 / args.steps
    if small:
        tc.set_timing(True)
        for _ in range(args.steps):
            tc.count(off, ind, out, total, cfg=cfg)
            tm = tc.last_timing()
            kern.append(tm.count_ms)
            prep.append(tm.prep_ms)
    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step, count kernel {np.mean(kern):.1f} ms")

    edges = g.nnz // 2
    value = edges / (ms_per_step / 1e3)
    kcfg = tc.default_cfg() if cfg is None else dict(tc.default_cfg(), **cfg)
    oriented = kcfg["algo"] == 0 and kcfg["oriented"] > 0
    # edges intersected and their merge work exist only on the edge scheme
    # (the oriented count builds no edge list)
    edge_work = {} if oriented else dict(zip(("intersected_edges", "merge_work"), tc.last_edge_work()))
    ab = algorithmic_bytes(g, kcfg["small_max"], kcfg["search_ratio"] if kcfg["algo"] in (0, 3) else -1)
    plan = None
    if oriented:
        # the launch plan the count ran with (its partner side sets the bytes)
        plan = tc.oriented_plan(g.n, g.nnz, cfg=cfg)
        ab.update(algorithmic_bytes_oriented(g, side=plan["side"]))
        plan["partners"] = ab["partners"]
    t_k = float(np.mean(kern)) / 1e3
    achieved = ab["b_kernel"] / t_k / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(config["workload"])}

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(g, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']:.3g} {cpu['unit']}")
        cpu_all = cpu_baseline_all_cores(g, args.cpu_seconds)
        log(f"cpu baseline, {cpu_all['cores']} threads: {cpu_all['value']:.3g} {cpu_all['unit']}")

    batch = None
    if not args.no_batch and args.workload == "rmat24":
        log("batched vs per-graph")
        try:
            batch = run_batch(tc, steps=max(2, args.steps), warmup=1)
            log(f"batched {batch['batched_graphs_per_s']:.0f} graphs/s, per-graph {batch['per_graph_graphs_per_s']:.0f}"
                f" graphs/s, speedup {batch['speedup']:.2f}x")
        except Exception as exc:  # the headline line must still print
            batch = {"error": repr(exc)}

    serving = None
    if not args.no_serve and args.workload == "rmat24":
        log("serving: burst and one Poisson rate")
        try:
            serving = run_serve(tc, pool=24, per_rate=64, rel_rates=(0.5,))
            log(f"serving {serving['saturation_graphs_per_s']:.1f} graphs/s, at 50%: p50 "
                f"{serving['rates'][0]['p50_ms']} ms p99 {serving['rates'][0]['p99_ms']} ms")
        except Exception as exc:  # the headline line must still print
            serving = {"error": repr(exc)}

    line = {
        "metric": METRIC, "value": value, "unit": "edge-intersections/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64" if args.int64 else "int32",
        "data": "synthetic: seeded R-MAT/ER/RGG from the package's counter-based generators",
        "config": dict(config, n=g.n, arcs=g.nnz, undirected_edges=edges, **edge_work,
                       csr="int64" if args.int64 else "int32",
                       kernel_cfg=tc.default_cfg() if cfg is None else cfg, kernel_plan=plan),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "count_kernel_ms": round(float(np.mean(kern)), 3),
        "prep_ms": round(float(np.mean(prep)), 3),
        "kernel_ms_from": "event-timed steps after the timed loop" if small else "the timed steps' events",
        "count_scheme": "oriented rows" if oriented else "undirected edges",
        "b_kernel_bytes": ab["b_kernel"],
        "b_alg_all_edges_bytes": ab["b_alg_all"],
        "hbm_gbs_all_edges_formula": round(ab["b_alg_all"] / t_k / 1e9, 1),
        "batched": batch,
        "serving": serving,
        "oriented": {k: ab[k] for k in ("oriented_arcs", "oriented_stream_elements", "max_out_degree")} if oriented
        else None,
        "checks": {"invariants": inv_ok, "total": tot, "global_mode_total": tot_global,
                   "int64_csr_per_vertex_identical": int64_identical,
                   "global_mode_count_ms": round(global_ms, 3), "edge_scheme_total": tot_edges,
                   "edge_scheme_per_vertex_identical": edges_identical,
                   "host_edge_count_agrees": True if oriented else ab["edges"] == edge_work["intersected_edges"],
                   "totals_agree": tot == tot_global == tot_edges, "triangles": tot // 6},
    }
    print(json.dumps(line), flush=True)
    tc.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
