"""GPU: batches, HIP-graph replay, the per-graph baseline and the autotuner."""
import os

import numpy as np
import pytest

import oracle
from gputrianglecounting_amd import autotune as AT
from gputrianglecounting_amd import graph as G

pytestmark = pytest.mark.gpu


def _per_graph_oracle(b):
    outs, tots = [], []
    for g in range(b.count):
        gr = b.graph(g)
        o, t = oracle.per_vertex(gr.off, gr.ind)
        outs.append(o)
        tots.append(t)
    return np.concatenate(outs) if outs else np.zeros(0, np.int32), np.array(tots, np.int64)


def test_batch_with_empty_and_single_vertex_graphs(counter):
    gs = [G.rmat(9, 8, seed=1), G.from_csr([0], []), G.from_csr([0, 0], []), G.erdos_renyi(300, 2000, seed=2),
          G.from_edges(3, [0, 1, 2], [1, 2, 0])]
    b = G.pack(gs)
    out, totals = counter.count_batch(b)
    ref, rt = _per_graph_oracle(b)
    np.testing.assert_array_equal(out, ref)
    np.testing.assert_array_equal(totals, rt)


def test_disjoint_union_equals_concatenation(counter):
    gs = [G.rmat(11, 16, seed=s) for s in range(3)]
    b = G.pack(gs)
    out, _ = counter.count_batch(b)
    for i, g in enumerate(gs):
        alone, _ = counter.count_graph(g)
        vb = b.v_base(i)
        np.testing.assert_array_equal(out[vb:vb + g.n], alone)


def test_pergraph_baseline_equals_batch(counter):
    import torch

    b = G.gen_batch(G=128, n_lo=64, n_hi=4096, seed=11)
    tab, off, ind = counter.batch_to_device(b)
    out1, tot1 = counter.batch_count(tab, off, ind, b.total_vertices, b.total_arcs)
    out2 = torch.empty_like(out1)
    tot2 = torch.empty_like(tot1)
    counter.batch_count_pergraph(b.tab, off, ind, out2, tot2)
    torch.cuda.synchronize()
    assert torch.equal(out1, out2) and torch.equal(tot1, tot2)
    ref, rt = _per_graph_oracle(b)
    np.testing.assert_array_equal(out1.cpu().numpy(), ref)
    np.testing.assert_array_equal(tot1.cpu().numpy(), rt)


def test_graph_replay_matches_eager(counter):
    import torch

    b = G.gen_batch(G=64, n_lo=64, n_hi=8192, seed=5)
    tab, off, ind = counter.batch_to_device(b)
    out = torch.empty(b.total_vertices, dtype=torch.int32, device="cuda")
    totals = torch.empty(b.count, dtype=torch.int64, device="cuda")
    eager, etot = counter.batch_count(tab, off, ind, b.total_vertices, b.total_arcs)
    torch.cuda.synchronize()
    eager, etot = eager.clone(), etot.clone()
    graph = counter.capture_batch(tab, off, ind, b.total_vertices, b.total_arcs, out, totals)
    for r in range(4):
        out.fill_(-1)
        totals.fill_(-1)
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {r}"
        assert torch.equal(totals, etot), f"replay {r}"
        # unrelated eager work between replays
        counter.count_graph(G.rmat(8, 8, seed=r))
    graph.close()


def test_full_mixed_batch_captured(counter):
    """The mixed batch at full size (4096 graphs, n log-uniform in
    [64, 65536], ef 8..32, seed 7; ~37M vertices, ~533M arcs) counted by a
    captured HIP graph replayed twice: every per-graph total equals the
    oracle's (a committed fixture) and the sum
    of its vertices' counts, every count is even and at most d(d-1); the 8
    largest graphs and 40 seeded others equal the oracle in full; 2^16
    uniformly drawn vertices and the 16 highest-degree ones equal the
    oracle's vertex-list values (OpenMP restatement, per graph)."""
    import time

    import torch

    t0 = time.time()
    b = G.gen_batch(G=4096, n_lo=64, n_hi=65536, ef_lo=8, ef_hi=32, seed=7)
    print(f"case full batch: {b.count} graphs, {b.total_vertices} vertices, {b.total_arcs} arcs "
          f"({time.time() - t0:.1f} s)", flush=True)
    tab, off, ind = counter.batch_to_device(b)
    out = torch.empty(b.total_vertices, dtype=torch.int32, device="cuda")
    totals = torch.empty(b.count, dtype=torch.int64, device="cuda")
    graph = counter.capture_batch(tab, off, ind, b.total_vertices, b.total_arcs, out, totals)
    try:
        graph.launch()
        torch.cuda.synchronize()
        first = out.clone()
        out.fill_(-1)
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, first)
    finally:
        graph.close()
    h = out.cpu().numpy().astype(np.int64)
    ht = totals.cpu().numpy()
    del first, off, ind, tab
    # every graph's total, by the oracle over all arcs (tests/golden/make_batch_totals.py)
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "batch4096_seed7_totals.npz"))
    assert int(gold["vertices"]) == b.total_vertices and int(gold["arcs"]) == b.total_arcs
    np.testing.assert_array_equal(ht, gold["totals"])
    _check_batch_sampled(b, h, ht, n_random=40, n_vertices=1 << 16)


def _check_batch_sampled(b, h, ht, n_random, n_vertices):
    """Batch counts h (per global vertex) and per-graph totals ht against
    invariants on every vertex, the oracle on the 8 largest and n_random
    seeded graphs in full, and its vertex-list values on the 16
    highest-degree and n_vertices uniformly drawn vertices."""
    vb = b.tab[0::2].astype(np.int64)
    gid = np.repeat(np.arange(b.count, dtype=np.int64), np.diff(vb))
    loc = np.arange(b.total_vertices, dtype=np.int64) - vb[gid]
    rows = np.arange(b.total_vertices, dtype=np.int64) + gid
    deg = b.off[rows + 1].astype(np.int64) - b.off[rows]
    del rows
    assert (h % 2 == 0).all() and (h >= 0).all() and (h <= deg * (deg - 1)).all()
    cs = np.concatenate([[0], np.cumsum(h)])
    np.testing.assert_array_equal(cs[vb[1:]] - cs[vb[:-1]], ht)
    assert int(ht.sum()) % 6 == 0
    print(f"case batch: invariants hold, sum {int(ht.sum())}", flush=True)
    rng = np.random.default_rng(2024)
    arcs = np.diff(b.tab[1::2].astype(np.int64))
    whole = np.unique(np.concatenate([np.argsort(-arcs, kind="stable")[:8], rng.integers(0, b.count, n_random)]))
    for g in whole:
        gr = b.graph(int(g))
        ref, rt = oracle.per_vertex(gr.off, gr.ind, threads=16)
        np.testing.assert_array_equal(h[vb[g]:vb[g + 1]], ref.astype(np.int64), err_msg=f"graph {g}")
        assert ht[g] == rt, f"graph {g}"
    print(f"case batch: {whole.shape[0]} whole graphs equal the oracle", flush=True)
    sample = np.unique(np.concatenate([np.argsort(-deg, kind="stable")[:16],
                                       rng.integers(0, b.total_vertices, n_vertices)]))
    sg = gid[sample]
    for part in np.split(np.arange(sample.shape[0]), np.flatnonzero(np.diff(sg)) + 1):
        g = int(sg[part[0]])
        gr = b.graph(g)
        exp, _ = oracle.vertex_list_omp(loc[sample[part]].astype(np.int32), gr.off, gr.ind, threads=16)
        np.testing.assert_array_equal(h[sample[part]], exp, err_msg=f"graph {g}")
    print(f"case batch: {sample.shape[0]} sampled vertices equal the oracle", flush=True)


def test_single_graph_capture(counter):
    import torch

    g = G.rmat(13, 16, seed=2)
    off, ind = counter.to_device(g)
    out = torch.empty(g.n, dtype=torch.int32, device="cuda")
    tot = torch.empty(1, dtype=torch.int64, device="cuda")
    cap = counter.capture_count(off, ind, out, tot)
    ref, rt = oracle.per_vertex(g.off, g.ind)
    for r in range(3):
        out.zero_()
        cap.launch()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
        assert int(tot.item()) == rt
    cap.close()


def test_capture_owns_its_scratch(counter):
    """A recorded count keeps scratch of its own: after the caller's
    context counts a much larger graph (which grows that context's scratch),
    and with a second recorded count replayed on another stream, every
    replay still equals the oracle."""
    import torch

    small, big = G.rmat(12, 16, seed=3), G.rmat(17, 16, seed=4)
    sref, stot = oracle.per_vertex(small.off, small.ind)
    off, ind = counter.to_device(small)
    out = torch.empty(small.n, dtype=torch.int32, device="cuda")
    tot = torch.empty(1, dtype=torch.int64, device="cuda")
    cap = counter.capture_count(off, ind, out, tot)
    g2 = G.erdos_renyi(5000, 40000, seed=6)
    r2, t2 = oracle.per_vertex(g2.off, g2.ind)
    off2, ind2 = counter.to_device(g2)
    out2 = torch.empty(g2.n, dtype=torch.int32, device="cuda")
    tot2 = torch.empty(1, dtype=torch.int64, device="cuda")
    cap2 = counter.capture_count(off2, ind2, out2, tot2)
    bout, btot = counter.count_graph(big)  # grows the caller context's scratch
    bref, brt = oracle.per_vertex(big.off, big.ind, threads=8)
    np.testing.assert_array_equal(bout, bref)
    assert btot == brt
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for r in range(3):
        out.fill_(-1)
        out2.fill_(-1)
        torch.cuda.synchronize()
        cap.launch(s1)
        cap2.launch(s2)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), sref)
        np.testing.assert_array_equal(out2.cpu().numpy(), r2)
        assert int(tot.item()) == stot and int(tot2.item()) == t2, f"replay {r}"
        counter.count_graph(G.rmat(15, 16, seed=r))  # eager work between replays
    cap.close()
    cap2.close()


def test_autotune_sets_a_usable_default(counter):
    """tc_autotune_i32 with set_default: the stored default is the resolved
    winner, so counts with cfg=None and the global-count mode still run and
    equal the oracle."""
    import torch

    g = G.rmat(12, 16, seed=9)
    off, ind = counter.to_device(g)
    out = torch.empty(g.n, dtype=torch.int32, device="cuda")
    saved = counter.default_cfg()
    try:
        best, rows = counter.autotune(off, ind, out, reps=1, set_default=True)
        assert best == counter.default_cfg()
        ref, rt = oracle.per_vertex(g.off, g.ind)
        o, t = counter.count_graph(g)
        np.testing.assert_array_equal(o, ref)
        assert t == rt
        assert int(counter.count_total(off, ind).item()) == rt
    finally:
        counter.set_default_cfg(**saved)


def test_kernel_call_shim_from_threads(counter):
    """The kernelCall stand-in from two host threads at once (one context
    per thread): both results equal the oracle."""
    import threading

    import torch

    from gputrianglecounting_amd import tc as T

    gs = [G.rmat(13, 16, seed=11), G.erdos_renyi(20000, 160000, seed=12)]
    refs = [oracle.per_vertex(g.off, g.ind)[0] for g in gs]
    res = [None, None]

    def work(i):
        torch.cuda.set_device(0)
        off, ind = counter.to_device(gs[i])
        out = torch.empty(gs[i].n, dtype=torch.int32, device="cuda")
        for _ in range(3):
            rc = T.kernel_call(64, 256, gs[i].n, off, ind, out, 4, 64, 2)
            torch.cuda.synchronize()
            assert rc == 0
        res[i] = out.cpu().numpy()

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(2):
        np.testing.assert_array_equal(res[i], refs[i])


def test_autotune_all_identical(counter, tmp_path):
    import torch

    g = G.rmat(12, 16, seed=7)
    off, ind = counter.to_device(g)
    out = torch.empty(g.n, dtype=torch.int32, device="cuda")
    saved = counter.default_cfg()
    best, rows = AT.tune_graph(counter, "test_rmat12", off, ind, out, reps=1, path=str(tmp_path / "t.json"))
    assert rows and all(r["identical"] for r in rows)
    assert AT.best_cfg("test_rmat12", path=str(tmp_path / "t.json")) == {k: best[k] for k in AT.CFG_KEYS}
    torch.cuda.synchronize()
    ref, _ = oracle.per_vertex(g.off, g.ind)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    counter.set_default_cfg(**saved)


def test_autotune_batch_over_the_oriented_knobs(counter, tmp_path):
    """tc_autotune_batch_i32: every stage's candidates (partner side x
    partners per wave, LDS window x workgroup size, piece size, lookup x
    long rows, the edge scheme) give identical counts; the winner, stored as
    the default and in a table, counts the batch equal to the oracle."""
    import torch

    b = G.gen_batch(G=96, n_lo=64, n_hi=8192, ef_lo=8, ef_hi=32, seed=13)
    tab, off, ind = counter.batch_to_device(b)
    out = torch.empty(b.total_vertices, dtype=torch.int32, device="cuda")
    saved = counter.default_cfg()
    try:
        best, rows = AT.tune_batch(counter, "test_batch96", tab, off, ind, b.total_vertices, b.total_arcs, out,
                                   reps=1, path=str(tmp_path / "t.json"))
        assert rows and all(r["identical"] for r in rows)
        seen = {(r["window"], r["orient_wg"]) for r in rows} | {("piece", r["piece"]) for r in rows}
        assert {(425, 128), (851, 256), (1701, 256)} <= seen and ("piece", 4096) in seen
        assert {r["lookup"] for r in rows} >= {1, 2} and any(r["oriented"] == -1 for r in rows)
        assert best == counter.default_cfg()
        assert AT.best_cfg("test_batch96", path=str(tmp_path / "t.json")) == {k: best[k] for k in AT.CFG_KEYS}
        ref, rt = _per_graph_oracle(b)
        o, t = counter.count_batch(b)
        np.testing.assert_array_equal(o, ref)
        np.testing.assert_array_equal(t, rt)
    finally:
        counter.set_default_cfg(**saved)
