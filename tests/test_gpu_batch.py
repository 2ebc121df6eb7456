"""GPU: batches, HIP-graph replay, the per-graph baseline and the autotuner."""
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
