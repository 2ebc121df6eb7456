"""GPU parity: the HIP counter (through the C-ABI) against the CPU oracle.

Bit-exact per-vertex int32 counts and int64 totals on closed-form and seeded
random graphs, every compiled kernel variant, the one-lane-per-edge
cross-check kernel, LCC in FP64, the kernelCall stand-in, validation, the
global-count mode and batches."""
import numpy as np
import pytest

import oracle
from gputrianglecounting_amd import graph as G
from gputrianglecounting_amd import tc as T
from tests import graphs as GR

pytestmark = pytest.mark.gpu

# (algo, wg_size, items, ring, chunk, depth): algo 0 = streamed kernel with A windows, 3 = streamed with
# workgroup/ring phase, 2 = Merge Path rings only
VARIANTS = [(0, 256, 8, 4096, 16, 2), (0, 256, 8, 4096, 16, 4), (0, 128, 16, 4096, 16, 2), (0, 512, 4, 4096, 16, 2),
            (3, 256, 8, 4096, 16, 2), (3, 256, 8, 4096, 16, 4), (3, 512, 4, 4096, 16, 2), (3, 128, 16, 4096, 16, 2),
            (2, 256, 8, 4096, 0, 0), (2, 256, 16, 8192, 0, 0), (2, 512, 8, 8192, 0, 0), (2, 128, 16, 4096, 0, 0),
            (2, 256, 4, 2048, 0, 0), (2, 512, 4, 4096, 0, 0)]


_REF = {}


def _ref(g):
    key = (g.name, g.n, g.nnz)
    if key not in _REF:
        _REF[key] = oracle.per_vertex(g.off, g.ind)
    return _REF[key]


def _check(counter, g, cfg=None):
    print(f"case {g.name} n={g.n} nnz={g.nnz} cfg={cfg}", flush=True)
    out, tot = counter.count_graph(g, cfg=cfg)
    ref, rtot = _ref(g)
    assert out.dtype == np.int32 and out.shape == (g.n,)
    np.testing.assert_array_equal(out, ref)
    assert tot == rtot == int(out.astype(np.int64).sum())


@pytest.mark.parametrize("g", GR.small_cases(), ids=lambda g: g.name)
def test_small_closed_forms(counter, g):
    _check(counter, g)


@pytest.mark.parametrize("g", GR.random_cases(), ids=lambda g: g.name)
def test_random_graphs(counter, g):
    _check(counter, g)


_ORIENTED = [dict(oriented=1), dict(oriented=1, probe_ratio=1), dict(oriented=1, probe_ratio=-1),
             dict(oriented=1, row_groups=1), dict(oriented=1, row_groups=2), dict(oriented=1, row_groups=8),
             dict(oriented=1, row_groups=8, probe_ratio=1),
             dict(oriented=2), dict(oriented=2, row_groups=1), dict(oriented=2, row_groups=8, probe_ratio=1),
             dict(oriented=1, row_groups=4), dict(oriented=1, row_groups=16), dict(oriented=1, row_groups=32),
             dict(oriented=2, row_groups=32),
             dict(lookup=2), dict(lookup=2, oriented=1), dict(lookup=2, oriented=2), dict(lookup=2, row_groups=1),
             dict(lookup=2, oriented=1, row_groups=2), dict(lookup=2, oriented=1, row_groups=8),
             dict(lookup=2, oriented=2, row_groups=32), dict(lookup=1), dict(lookup=3), dict(lookup=3, row_groups=1),
             dict(lookup=3, oriented=2, row_groups=16), dict(lookup=2, long_rows=2),
             dict(long_rows=2, row_groups=1), dict(lookup=3, long_rows=2, oriented=2, row_groups=8)]


@pytest.mark.parametrize("cfg", _ORIENTED, ids=str)
def test_oriented(counter, cfg):
    """Degree-oriented per-vertex counting (k_count_orient): each triangle
    found once and its three vertices credited; same counts as the oracle."""
    for g in GR.small_cases() + GR.random_cases() + _search_graphs():
        _check(counter, g, cfg=cfg)


@pytest.mark.parametrize("cfg", [dict(long_rows=2), dict(long_rows=2, row_groups=1), dict(long_rows=2, oriented=2),
                                 dict(long_rows=2, lookup=2), dict(long_rows=2, lookup=2, oriented=2, row_groups=8),
                                 dict(long_rows=2, lookup=3), dict(long_rows=1, lookup=2), dict(long_rows=1)], ids=str)
def test_long_rows(counter, cfg):
    """Rows whose N+ exceeds a wave's LDS window, counted by whole
    workgroups sharing one table (k_count_long) or in windows: dense
    graphs (K_1800, rows of up to 1799 out-neighbours), hub-heavy R-MAT, and
    two hubs adjacent to everything (in-neighbour partners beyond the
    16-bit counter run of 16384)."""
    import os

    threads = max(1, min(16, os.cpu_count() or 1))
    n = 1800
    g = GR.complete(n)
    out, tot = counter.count_graph(g, cfg=cfg)
    assert (out == (n - 1) * (n - 2)).all() and tot == n * (n - 1) * (n - 2)
    k = 40000
    leaves = np.arange(2, k + 2)
    g2 = G.from_edges(k + 2, np.concatenate([[0], np.zeros(k, np.int64), np.ones(k, np.int64)]),
                      np.concatenate([[1], leaves, leaves]), name=f"k2_join_{k}")
    exp = np.full(k + 2, 2, np.int64)
    exp[:2] = 2 * k
    out, tot = counter.count_graph(g2, cfg=cfg)
    np.testing.assert_array_equal(out, exp)
    assert tot == 6 * k
    for g in [G.rmat(17, 16, seed=8), G.rmat(14, 64, seed=9), G.geometric(600000, 8.0, seed=3)]:
        key = (g.name, g.n, g.nnz)
        if key not in _REF:
            _REF[key] = oracle.per_vertex(g.off, g.ind, threads=threads)
        _check(counter, g, cfg=cfg)


def test_oriented_long_rows(counter):
    """Oriented rows longer than one wave's LDS window (851 elements) are
    walked in windows: in K_1800 the lowest vertices have up to 1799
    out-neighbours (three windows); closed form out[v] = (n-1)(n-2)."""
    n = 1800
    g = GR.complete(n)
    for cfg in _ORIENTED:
        out, tot = counter.count_graph(g, cfg=cfg)
        assert (out == (n - 1) * (n - 2)).all()
        assert tot == n * (n - 1) * (n - 2)


def test_one_way_arcs_stay_in_bounds(counter):
    """A CSR that is not symmetric can keep more oriented arcs than a
    symmetric one of the same size (here m leaves -> hub, hub -> two sinks:
    m of m + 2 arcs point up), more than the orientation's capacity.  The
    rows are cut to the capacity; the count runs and finds no triangle."""
    import torch

    m = 1000
    off = np.concatenate([np.arange(m + 1), [m + 2, m + 2, m + 2]]).astype(np.int32)
    ind = np.concatenate([np.full(m, m), [m + 1, m + 2]]).astype(np.int32)
    d_off = torch.from_numpy(off).cuda()
    d_ind = torch.from_numpy(ind).cuda()
    assert counter.validate(d_off, d_ind) == T.TC_EGRAPH
    for cfg in (dict(oriented=1), dict(oriented=2), dict(oriented=-1), dict(algo=3)):
        out, tot = counter.count(d_off, d_ind, cfg=cfg)
        torch.cuda.synchronize()
        assert int(tot.item()) == 0 and int(out.abs().sum().item()) == 0
    assert int(counter.count_total(d_off, d_ind).item()) == 0


def test_relabelling_invariance(counter):
    """The orientation breaks degree ties by vertex id, so a relabelled
    graph is oriented differently; the counts must follow the relabelling
    exactly (out'[p(v)] = out[v]) on both schemes."""
    rng = np.random.default_rng(11)
    for g in [G.rmat(13, 16, seed=5), GR.complete(60), G.geometric(20000, 6.0, seed=2)]:
        perm = rng.permutation(g.n).astype(np.int64)
        d = np.diff(g.off.astype(np.int64))
        src = np.repeat(np.arange(g.n, dtype=np.int64), d)
        h = G.from_edges(g.n, perm[src], perm[g.ind.astype(np.int64)])
        assert h.nnz == g.nnz
        ref, rtot = _ref(g)
        for cfg in (None, dict(oriented=2), dict(oriented=-1)):
            print(f"case relabelled {g.name} cfg={cfg}", flush=True)
            out, tot = counter.count_graph(h, cfg=cfg)
            np.testing.assert_array_equal(out[perm], ref)
            assert tot == rtot


def test_tied_hubs(counter):
    """Two hubs of equal degree 70001, beyond 16 bits, tie in the degree
    order and are ordered by id.  K_2 joined to 70000 leaves (every leaf in
    one triangle with both hubs): closed form hubs 2*70000, leaves 2, on
    every scheme and in the global-count mode."""
    k = 70000
    leaves = np.arange(2, k + 2)
    src = np.concatenate([[0], np.zeros(k, np.int64), np.ones(k, np.int64)])
    dst = np.concatenate([[1], leaves, leaves])
    g = G.from_edges(k + 2, src, dst, name=f"k2_join_{k}")
    exp = np.full(k + 2, 2, np.int64)
    exp[:2] = 2 * k
    for cfg in (None, dict(oriented=2), dict(row_groups=1), dict(oriented=-1)):
        out, tot = counter.count_graph(g, cfg=cfg)
        np.testing.assert_array_equal(out, exp)
        assert tot == 6 * k
    off, ind = counter.to_device(g)
    assert int(counter.count_total(off, ind).item()) == 6 * k


def test_full_size_configs(counter):
    """The parity configurations at full size: R-MAT s20 ef16 and RGG
    n=4,194,304 (mean degree 6) against the oracle (OpenMP), bit for bit,
    on the oriented and the edge scheme, and their FP64 local clustering
    coefficients bit for bit against the oracle's out/(d(d-1))."""
    import os

    import torch

    threads = max(1, min(16, os.cpu_count() or 1))
    for g in [G.geometric(4194304, 6.0, seed=1), G.rmat(20, 16, seed=1)]:
        print(f"case full size {g.name} n={g.n} nnz={g.nnz}: oracle on {threads} threads", flush=True)
        ref, rtot = oracle.per_vertex(g.off, g.ind, threads=threads)
        for cfg in (None, dict(oriented=-1)):
            print(f"case full size {g.name} cfg={cfg}", flush=True)
            out, tot = counter.count_graph(g, cfg=cfg)
            np.testing.assert_array_equal(out, ref)
            assert tot == rtot
        off, ind = counter.to_device(g)
        out, _ = counter.count(off, ind)
        lcc = counter.lcc(off, out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(lcc.cpu().numpy().view(np.int64), oracle.lcc(g.off, ref).view(np.int64))


def test_rmat_s24_sampled(counter):
    """R-MAT scale 24, edge factor 16, seed 1 (the benchmark graph: 16.8M
    vertices, 520.8M arcs).  A full CPU count takes hours, so the per-vertex
    counts are checked bit for bit against the oracle's vertex-list row sums
    (clusteringCount.cpp:125-155, OpenMP) on the 16 highest-degree vertices
    and 2^16 seeded uniform vertices, their FP64 LCC bit for bit, and every
    vertex against the size-independent properties: out[v] even, 0 when
    d(v) < 2, at most d(d-1); the int64 total = the sum of out = 6T, equal
    to the global-count mode's."""
    import os

    import torch

    threads = max(1, min(16, os.cpu_count() or 1))
    g = G.rmat(24, 16, seed=1)
    print(f"case R-MAT s24 n={g.n} nnz={g.nnz}", flush=True)
    off, ind = counter.to_device(g)
    out, total = counter.count(off, ind)
    lcc = counter.lcc(off, out)
    tot_global = counter.count_total(off, ind)
    torch.cuda.synchronize()
    h = out.cpu().numpy().astype(np.int64)
    hl = lcc.cpu().numpy()
    tot, totg = int(total.cpu().item()), int(tot_global.cpu().item())
    del off, ind, out, lcc
    d = g.degrees()
    assert (h >= 0).all() and (h % 2 == 0).all() and (h <= d * (d - 1)).all() and (h[d < 2] == 0).all()
    assert tot == int(h.sum()) == totg and tot % 6 == 0
    hubs = np.argsort(-d, kind="stable")[:16]
    rng = np.random.default_rng(24)
    vl = np.concatenate([hubs, rng.integers(0, g.n, 1 << 16)]).astype(np.int32)
    print(f"case R-MAT s24: oracle on {vl.shape[0]} listed vertices, {threads} threads", flush=True)
    ref, rtot = oracle.vertex_list_omp(vl, g.off, g.ind, threads=threads)
    np.testing.assert_array_equal(h[vl], ref)
    assert int(h[vl].sum()) == rtot
    dv = d[vl]
    off_s = np.zeros(vl.shape[0] + 1, np.int64)
    np.cumsum(dv, out=off_s[1:])
    assert off_s[-1] < 2 ** 31
    exp = oracle.lcc(off_s.astype(np.int32), ref.astype(np.int32))
    np.testing.assert_array_equal(hl[vl].view(np.int64), exp.view(np.int64))


def test_oriented_only_for_algo_0(counter):
    """Other algos count the full lists whatever cfg.oriented says."""
    g = G.rmat(11, 16, seed=1)
    for algo in (1, 2, 3):
        _check(counter, g, cfg=dict(oriented=1, algo=algo))


def test_oriented_plan(counter):
    """tc_oriented_plan resolves the automatic choices by size and mean
    degree, keeps explicit ones, and reports the edge scheme as such."""
    plan = counter.oriented_plan
    # R-MAT s24 ef16 (2^24 vertices, 2^29 arcs): the table, in-neighbours, 8 per wave, long rows by workgroups
    assert plan(1 << 24, 1 << 29) == dict(lookup=1, side=1, groups=8, long_rows=1)
    # 2^31 arcs: out-neighbour partners decided on the device, one per wave
    assert plan(1 << 26, 1 << 31) == dict(lookup=1, side=2, groups=1, long_rows=1)
    # dense from 2^24 arcs takes the table; sparse of the same size the walked array
    assert plan(1 << 20, 1 << 25)["lookup"] == 1
    p = plan(1 << 22, 1 << 25)
    assert (p["lookup"], p["side"]) == (0, 2)
    # small graphs and batches: the walked array with in-neighbour partners, groups by mean degree
    assert plan(1 << 14, 1 << 18) == dict(lookup=0, side=1, groups=16, long_rows=0)
    assert plan(1 << 21, 1 << 25, batch=True) == dict(lookup=0, side=1, groups=16, long_rows=0)
    assert plan(1 << 18, 1 << 25, batch=True)["groups"] == 8
    assert plan(1 << 20, 1 << 22)["groups"] == 32
    # explicit settings are kept
    assert plan(1 << 24, 1 << 29, cfg=dict(oriented=1, row_groups=8, lookup=1, long_rows=1)) == \
        dict(lookup=0, side=0, groups=8, long_rows=0)
    assert plan(1 << 24, 1 << 29, cfg=dict(oriented=-1)) == dict(lookup=-1, side=-1, groups=0, long_rows=0)
    with pytest.raises(T.TCError):
        plan(-1, 0)


@pytest.mark.parametrize("v", VARIANTS, ids=lambda v: "x".join(map(str, v)))
@pytest.mark.parametrize("small_max", [16, 300, 2048, 0])
@pytest.mark.parametrize("ranges", [1, 32])
def test_variants(counter, v, small_max, ranges):
    g = G.rmat(13, 16, seed=21)
    cfg = dict(algo=v[0], wg_size=v[1], items=v[2], ring=v[3], chunk=v[4], depth=v[5], small_max=small_max,
               ranges_per_wg=ranges, oriented=-1)  # the edge-list kernels these variants size
    _check(counter, g, cfg)


_SG = []


def _search_graphs():
    if not _SG:
        _SG.extend([G.rmat(13, 32, seed=3), G.rmat(13, 64, seed=4), GR.wheel(3000)])
    return _SG


@pytest.mark.parametrize("algo", [0, 3])
@pytest.mark.parametrize("ratio", [1, 4, 64, -1])
@pytest.mark.parametrize("ranges", [1, 4096])
def test_search_ratio(counter, algo, ratio, ranges):
    """Edges with the longer list at least ratio times the shorter (and at
    least 512 long) are counted by lookups; ratio 1 sends every such edge
    there, -1 none; 4096 ranges per workgroup split long edges across
    ranges.  R-MAT 13/64 has 14 hubs of degree above 2039, so hub-hub edges
    have a shorter list longer than one wave's LDS slice (algo 3 walks it
    in windows)."""
    for g in _search_graphs():
        _check(counter, g, cfg=dict(algo=algo, search_ratio=ratio, ranges_per_wg=ranges, oriented=-1))


@pytest.mark.parametrize("algo", [0, 3])
@pytest.mark.parametrize("probe", [1, 4, -1])
@pytest.mark.parametrize("search", [-1, 64])
def test_probe_ratio(counter, algo, probe, search):
    """Streamed edges with B at least probe times longer than A are merged
    from A's side; probe 1 sends every streamed edge there, -1 none."""
    for g in _search_graphs():
        _check(counter, g, cfg=dict(algo=algo, probe_ratio=probe, search_ratio=search, ranges_per_wg=4096,
                                    oriented=-1))


def _list_expect(g, vl):
    """Oracle vertex-list values; -1 for ids outside the graph."""
    vl = np.asarray(vl, np.int32)
    ok = (vl >= 0) & (vl < g.n)
    tri, _ = oracle.count_vertex_list(vl[ok], g.off, g.ind)
    exp = np.full(vl.shape[0], -1, np.int64)
    exp[ok] = oracle.list_sums(vl[ok], g.off, tri)
    return exp, int(exp[ok].sum())


@pytest.mark.parametrize("cfg", [None, dict(row_groups=32), dict(oriented=-1), dict(oriented=-1, search_ratio=1),
                                 dict(oriented=-1, search_ratio=-1), dict(algo=3), dict(algo=2), dict(algo=1),
                                 dict(oriented=-1, small_max=2048)], ids=str)
def test_vertex_list(counter, cfg):
    """Lists of an eighth of the vertices or more take the oriented count
    and a gather, shorter ones (and oriented = -1) the listed vertices'
    arcs; both give the oracle's vertex-list values."""
    import torch

    g = G.rmat(13, 32, seed=3)
    off, ind = counter.to_device(g)
    rng = np.random.default_rng(5)
    lists = [np.array([0, 1, 2, 0, 4095, 8191, 17], np.int32),
             rng.integers(0, g.n, 3000).astype(np.int32),
             np.arange(g.n, dtype=np.int32)[::-1].copy(),
             np.array([-3, 5, g.n, 5, 2**31 - 1], np.int32),
             np.zeros(0, np.int32)]
    for vl in lists:
        print(f"case vertex list len={vl.shape[0]} cfg={cfg}", flush=True)
        exp, etot = _list_expect(g, vl)
        out, tot = counter.count_vertices(off, ind, torch.from_numpy(vl).to(off.device), cfg=cfg)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().astype(np.int64), exp)
        assert int(tot.item()) == etot


@pytest.mark.parametrize("cfg", [None, dict(oriented=-1, search_ratio=1), dict(algo=3), dict(algo=2), dict(algo=1)],
                         ids=str)
def test_vertex_list_shadow(counter, cfg):
    """Shadow-vertex counts (a common neighbour x adds 1 + [x > lim]) of
    short and long lists at several lim, including -1 (every value doubled)
    and nv - 1 (the plain count), against the oracle's restatement."""
    import torch

    g = G.rmat(13, 32, seed=3)
    off, ind = counter.to_device(g)
    rng = np.random.default_rng(9)
    lists = [np.array([0, 1, 2, 0, 4095, 8191, 17], np.int32), rng.integers(0, g.n, 3000).astype(np.int32),
             np.arange(g.n, dtype=np.int32)[::-1].copy(), np.array([-3, 5, g.n, 5], np.int32),
             np.zeros(0, np.int32)]
    for vl in lists:
        for lim in (-1, 0, 100, g.n // 2, g.n - 1):
            print(f"case shadow list len={vl.shape[0]} lim={lim} cfg={cfg}", flush=True)
            ok = (vl >= 0) & (vl < g.n)
            exp = np.full(vl.shape[0], -1, np.int64)
            exp[ok], etot = oracle.vertex_list_lim(vl[ok], g.off, g.ind, lim)
            out, tot = counter.count_vertices(off, ind, torch.from_numpy(vl).to(off.device), cfg=cfg, lim=lim)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(out.cpu().numpy().astype(np.int64), exp)
            assert int(tot.item()) == etot


def test_partitions_sum_to_the_whole(counter):
    import torch

    g = G.rmat(14, 16, seed=8)
    off, ind = counter.to_device(g)
    full, ftot = counter.count_graph(g)
    part = (np.arange(g.n) * 2654435761 % 5).astype(np.int32)  # 5 partitions
    tot = 0
    for p in range(5):
        vl = G.partition_members(part, p)
        out, t = counter.count_vertices(off, ind, torch.from_numpy(vl).to(off.device))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), full[vl])
        tot += int(t.item())
    assert tot == ftot


def test_simple_kernel_agrees(counter):
    g = G.rmat(12, 16, seed=4)
    _check(counter, g, dict(algo=1))


def test_repeat_deterministic(counter):
    g = G.rmat(14, 16, seed=8)
    a, ta = counter.count_graph(g)
    b, tb = counter.count_graph(g)
    np.testing.assert_array_equal(a, b)
    assert ta == tb


def test_ind_at_every_alignment(counter):
    """ind handed over as a view at every word offset of a 64-byte block and
    ending flush with its allocation: the streamed loads read aligned 64-byte
    chunks around the lists, which must give the same counts at every shift."""
    import torch

    g = G.rmat(13, 24, seed=9)
    ref, rtot = oracle.per_vertex(g.off, g.ind)
    off = torch.from_numpy(g.off).to("cuda")
    for k in range(16):
        print(f"case ind offset {k} words", flush=True)
        buf = torch.full((g.nnz + k,), -1, dtype=torch.int32, device="cuda")
        buf[k:] = torch.from_numpy(g.ind).to("cuda")
        out, tot = counter.count(off, buf[k:])
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
        assert int(tot.item()) == rtot


def test_lcc_bit_exact(counter):
    import torch

    g = G.rmat(12, 16, seed=2)
    off, ind = counter.to_device(g)
    out, _ = counter.count(off, ind)
    lcc = counter.lcc(off, out)
    torch.cuda.synchronize()
    ref_out, _ = oracle.per_vertex(g.off, g.ind)
    ref = oracle.lcc(g.off, ref_out)
    got = lcc.cpu().numpy()
    assert got.view(np.int64).tolist() == ref.view(np.int64).tolist()


def test_kernel_call_shim(counter):
    import torch

    g = G.erdos_renyi(2000, 20000, seed=5)
    off, ind = counter.to_device(g)
    out = torch.full((g.n,), -7, dtype=torch.int32, device=off.device)
    assert T.kernel_call(64, 256, g.n, off, ind, out, 4, 64, 2) == 0
    torch.cuda.synchronize()
    ref, _ = oracle.per_vertex(g.off, g.ind)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    # launch shapes the reference rejects (param.cu:225-247) are errors here
    assert T.kernel_call(64, 100, g.n, off, ind, out, 4, 25, 2) == T.TC_EINVAL
    assert T.kernel_call(64, 256, g.n, off, ind, out, 3, 85, 1) == T.TC_EINVAL
    assert T.kernel_call(64, 256, g.n, off, ind, out, 4, 32, 2) == T.TC_EINVAL


def _count64(counter, g, cfg=None):
    import torch

    off = torch.from_numpy(g.off.astype(np.int64)).cuda()
    ind = torch.from_numpy(g.ind.astype(np.int64)).cuda()
    out, tot = counter.count(off, ind, cfg=cfg)
    torch.cuda.synchronize()
    assert out.dtype == torch.int64
    return out.cpu().numpy(), int(tot.item())


@pytest.mark.parametrize("cfg", [None, dict(row_groups=1), dict(row_groups=32), dict(oriented=2)], ids=str)
def test_int64_csr(counter, cfg):
    """The int64 CSR path (kernelCall<int64_t>, the reference main's type):
    int64 counts identical to the oracle on closed forms, random graphs and
    the long-row and tied-hub cases."""
    k = 70000
    leaves = np.arange(2, k + 2)
    tied = G.from_edges(k + 2, np.concatenate([[0], np.zeros(k, np.int64), np.ones(k, np.int64)]),
                        np.concatenate([[1], leaves, leaves]), name=f"k2_join_{k}")
    for g in GR.small_cases() + GR.random_cases() + [GR.complete(1800), tied]:
        print(f"case int64 {g.name} n={g.n} nnz={g.nnz} cfg={cfg}", flush=True)
        out, tot = _count64(counter, g, cfg)
        ref, rtot = _ref(g)
        np.testing.assert_array_equal(out, ref)
        assert tot == rtot
        if cfg is None:
            import torch

            off = torch.from_numpy(g.off.astype(np.int64)).cuda()
            ind = torch.from_numpy(g.ind.astype(np.int64)).cuda()
            assert int(counter.count_total(off, ind).item()) == rtot


def test_int64_beyond_2_31_arcs(counter):
    """An int64 CSR with more than 2^31 arcs, built on the device: disjoint
    K_64 cliques (every vertex in C(63, 2) triangles), 2^31 + 64512 arcs.
    Closed form out[v] = 63 * 62 = 3906; the total is 3906 nv."""
    import torch

    k = 64
    ncl = 2**31 // (k * (k - 1)) + 16
    nv, nnz = ncl * k, ncl * k * (k - 1)
    assert nnz > 2**31
    dev = torch.device("cuda")
    off = torch.arange(nv + 1, dtype=torch.int64, device=dev) * (k - 1)
    j = torch.arange(k - 1, dtype=torch.int64, device=dev)
    i = torch.arange(k, dtype=torch.int64, device=dev).unsqueeze(1)
    local = j + (j >= i).to(torch.int64)  # row i of a clique: the other 63 members, ascending
    ind = (torch.arange(ncl, dtype=torch.int64, device=dev).view(-1, 1, 1) * k + local.unsqueeze(0)).reshape(-1)
    del local
    assert ind.numel() == nnz
    print(f"case int64 cliques nv={nv} nnz={nnz}", flush=True)
    out, tot = counter.count(off, ind)
    torch.cuda.synchronize()
    assert int(tot.item()) == nv * (k - 1) * (k - 2)
    assert bool((out == (k - 1) * (k - 2)).all().item())
    del out
    assert int(counter.count_total(off, ind).item()) == nv * (k - 1) * (k - 2)  # tc_count_total_i64
    del off, ind
    torch.cuda.empty_cache()


def test_int64_errors(counter):
    import torch

    g = G.rmat(10, 8, seed=2)
    off = torch.from_numpy(g.off.astype(np.int64)).cuda()
    ind = torch.from_numpy(g.ind.astype(np.int64)).cuda()
    for bad in (dict(oriented=-1), dict(algo=3)):
        with pytest.raises(T.TCError):
            counter.count(off, ind, cfg=bad)
    with pytest.raises(TypeError):
        counter.count(off, ind.to(torch.int32))


def test_kernel_call_shim_int64(counter):
    import torch

    g = G.erdos_renyi(2000, 20000, seed=5)
    off = torch.from_numpy(g.off.astype(np.int64)).cuda()
    ind = torch.from_numpy(g.ind.astype(np.int64)).cuda()
    out = torch.full((g.n,), -7, dtype=torch.int64, device=off.device)
    assert T.kernel_call64(64, 256, g.n, off, ind, out, 4, 64, 2) == 0
    torch.cuda.synchronize()
    ref, _ = oracle.per_vertex(g.off, g.ind)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert T.kernel_call64(64, 100, g.n, off, ind, out, 4, 25, 2) == T.TC_EINVAL
    assert T.kernel_call64(64, 256, g.n, off, ind, out, 4, 64, 7) == T.TC_EINVAL


def test_validate(counter):
    import torch

    g = G.rmat(10, 8, seed=3)
    off, ind = counter.to_device(g)
    assert counter.validate(off, ind) == 0
    bad = g.ind.copy()
    r = int(np.argmax(np.diff(g.off) >= 3))
    k = int(g.off[r])
    bad[k], bad[k + 1] = bad[k + 1], bad[k]  # unsorted row
    assert counter.validate(off, torch.from_numpy(bad).to(off.device)) == T.TC_EGRAPH
    sl = g.ind.copy()
    sl[k] = r  # self-loop (and a missing reverse)
    assert counter.validate(off, torch.from_numpy(sl).to(off.device)) == T.TC_EGRAPH
    assert "self-loop" in counter.last_error()


def test_global_count_mode(counter):
    for g in [G.rmat(14, 16, seed=6), GR.complete(40), G.geometric(30000, 6.0, seed=1), GR.empty(3)]:
        off, ind = counter.to_device(g)
        tot = counter.count_total(off, ind)
        _, rtot = oracle.per_vertex(g.off, g.ind)
        assert int(tot.cpu().item()) == rtot, g.name


def _hub_graphs():
    """Graphs with rows of more than 4096 arcs, which the orientation lists
    and hands out first, one row per wave: two hubs of 6001 arcs, and 50
    hubs of 4149 arcs (a K_50 whose vertices all join the same 4100 leaves,
    so every leaf closes C(50, 2) triangles)."""
    k, m = 50, 4100
    a = np.repeat(np.arange(k), m)
    b = np.tile(np.arange(k, k + m), k)
    ii, jj = np.triu_indices(k, 1)
    hubs = G.from_edges(k + m, np.concatenate([a, ii]), np.concatenate([b, jj]), f"hubs{k}x{m}")
    return [GR.two_hubs(6000), hubs]


def test_hub_rows_every_entry_point(counter):
    """Hub rows through every oriented entry point: per-vertex, total only,
    a vertex list of every vertex (the oriented count and a gather) and a
    batch in which they sit between ordinary graphs."""
    import torch

    graphs = _hub_graphs()
    for g in graphs:
        _check(counter, g)
        ref, rtot = _ref(g)
        off, ind = counter.to_device(g)
        assert int(counter.count_total(off, ind).cpu().item()) == rtot, g.name
        vl = np.arange(g.n, dtype=np.int32)[::-1].copy()
        out, tot = counter.count_vertices(off, ind, torch.from_numpy(vl).to(off.device))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy(), ref[vl], err_msg=g.name)
        assert int(tot.item()) == rtot, g.name
    b = G.pack([G.rmat(12, 16, seed=2)] + graphs + [GR.complete(60), GR.empty(4)])
    print(f"case hub batch graphs={b.count}", flush=True)
    out, totals = counter.count_batch(b)
    for gi in range(b.count):
        gr = b.graph(gi)
        ref, rtot = oracle.per_vertex(gr.off, gr.ind)
        vb = b.v_base(gi)
        np.testing.assert_array_equal(out[vb: vb + gr.n], ref, err_msg=f"graph {gi}")
        assert totals[gi] == rtot


@pytest.mark.parametrize("cfg", [None, dict(oriented=1), dict(oriented=2), dict(oriented=-1)], ids=str)
def test_batch_equals_alone(counter, cfg):
    """Every graph of a batch gets the counts it gets alone: the default
    (in-neighbour partners for a batch), each partner side explicitly, and
    the edge scheme."""
    b = G.gen_batch(G=96, n_lo=16, n_hi=3000, seed=3)
    print(f"case batch cfg={cfg}", flush=True)
    out, totals = counter.count_batch(b, cfg=cfg)
    for g in range(b.count):
        gr = b.graph(g)
        ref, rtot = oracle.per_vertex(gr.off, gr.ind)
        vb = b.v_base(g)
        np.testing.assert_array_equal(out[vb: vb + gr.n], ref, err_msg=f"graph {g}")
        assert totals[g] == rtot


def test_errors(counter):
    import torch

    off = torch.zeros(4, dtype=torch.int32, device="cuda")
    ind = torch.zeros(0, dtype=torch.int32, device="cuda")
    with pytest.raises(T.TCError):
        counter.count(off, ind, cfg=dict(wg_size=100))
    with pytest.raises(T.TCError):
        counter.count(off, ind, cfg=dict(ranges_per_wg=100000))


def test_tc_main_partition_run(tmp_path):
    """tc_main <graph> 1 <blocks> <bs> <tpi> <K>: the partitions of
    <graph>.part.K counted as vertex lists sum to the whole-graph total."""
    import os
    import subprocess

    g = G.rmat(12, 16, seed=6)
    base = str(tmp_path / "g")
    g.write_info(base + ".info")
    part = (np.arange(g.n) % 3).astype(np.int32)
    with open(base + ".part.3", "w") as f:
        f.write(f"{g.n}\n" + "\n".join(str(int(x)) for x in part) + "\n")
    exe = os.path.join(os.path.dirname(T.__file__), "_lib", "tc_main")
    print("case tc_main partition run", flush=True)
    r = subprocess.run([exe, base, "1", "8", "256", "4", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nv, total = r.stdout.split()[:2]
    _, rtot = _ref(g)
    assert int(nv) == g.n and int(total) == rtot
    r = subprocess.run([exe, base, "1", "8", "256", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and int(r.stdout.split()[1]) == rtot


def test_tc_main_sweep(tmp_path):
    """tc_main <graph> 1 <blocks> (no block size): the reference's sweep file
    <graph>.o.<blocks> beside the graph (gpurun.cu:146-172), one row
    "bs tpi allocMs execMs sumMs total" per legacy shape in the reference's
    order (param.cu:4-5), every total the oracle's; --variants writes one
    row per oriented partners-per-wave setting and edge-scheme variant."""
    import os
    import subprocess

    g = G.rmat(11, 16, seed=9)
    base = str(tmp_path / "g")
    g.write_info(base + ".info")
    exe = os.path.join(os.path.dirname(T.__file__), "_lib", "tc_main")
    _, rtot = _ref(g)
    print("case tc_main sweep (reference rows)", flush=True)
    r = subprocess.run([exe, base, "1", "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [line.split("\t") for line in open(base + ".o.8").read().splitlines()]  # next to the graph
    shapes = [(b, t) for b in (32, 64, 96, 128, 192, 256, 512, 1024) for t in (1, 2, 4, 8, 16, 32)]
    assert [(int(row[0]), int(row[1])) for row in rows] == shapes
    assert all(len(row) == 6 and int(row[5]) == rtot for row in rows)
    assert all(abs(float(row[4]) - float(row[2]) - float(row[3])) < 1e-3 * max(1.0, float(row[4])) for row in rows)
    print("case tc_main sweep --variants", flush=True)
    r = subprocess.run([exe, "--variants", base, "1", "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [line.split("\t") for line in open(base + ".o.8").read().splitlines()]
    assert {row[0] for row in rows} == {"oriented", "edges"} and len(rows) == 14
    assert all(int(row[-1]) == rtot for row in rows)


def test_tc_main_int64(tmp_path):
    """tc_main --int64: the graph read as the reference binary reads it
    (readGraph<int64_t>, run<int64_t>: main.cu:36) and counted by
    tc_count_i64; .info and the int64 binary cache, single run and sweep,
    totals the oracle's."""
    import os
    import subprocess

    g = G.rmat(12, 16, seed=10)
    base = str(tmp_path / "g")
    g.write_info(base + ".info")
    G.write_bin64(g, base + ".bin")
    exe = os.path.join(os.path.dirname(T.__file__), "_lib", "tc_main")
    _, rtot = _ref(g)
    for spec in (base, base + ".bin"):
        print(f"case tc_main --int64 {os.path.basename(spec)}", flush=True)
        r = subprocess.run([exe, "--int64", spec, "1", "8", "256", "4"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        nv, total = r.stdout.split()[:2]
        assert int(nv) == g.n and int(total) == rtot
    print("case tc_main --int64 sweep", flush=True)
    r = subprocess.run([exe, "--int64", base, "1", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [line.split("\t") for line in open(base + ".o.4").read().splitlines()]
    assert len(rows) == 48 and all(int(row[5]) == rtot for row in rows)
    # int64 inputs are not normalised: a one-way arc is refused before any launch
    bad = tmp_path / "bad.info"
    bad.write_text("4\t3\n0\t1\t2\t3\n1\t2\t0\n")
    r = subprocess.run([exe, "--int64", str(tmp_path / "bad"), "1", "8", "256", "4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1 and "not a simple undirected CSR" in r.stderr
