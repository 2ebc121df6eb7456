"""Serving admission policy on the CPU with a fake HBM budget
(tch_admission_sim: batches formed first in first out, one first-fit arena
block per batch)."""
import numpy as np

from gputrianglecounting_amd import serve as S

KB = 1024


def _blk(share_sum):
    return -(-(S.BATCH_FIXED + share_sum) // 256) * 256


def test_fifo_and_release():
    # each 3 KiB request alone takes a 4864-byte block: two fit in 10 KiB
    assert _blk(3 * KB) == 4864
    at, bo, n, peak = S.admission_sim(10 * KB, [3 * KB, 3 * KB, 3 * KB, -1, -2, -3])
    assert at.tolist() == [0, 1, 3] and n == 3
    assert bo.tolist() == [0, 1, 2]
    assert peak == 2 * 4864


def test_waiting_requests_are_batched_together():
    # r0 fills most of the arena; r1 and r2 wait, then leave as ONE batch
    script = [8 * KB, 3 * KB, 3 * KB, -1]
    at, bo, n, _ = S.admission_sim(12 * KB, script, max_batch=4)
    assert at.tolist() == [0, 3, 3] and bo.tolist() == [0, 1, 1]
    # with one graph per batch they leave as two batches
    at, bo, n, _ = S.admission_sim(12 * KB, script, max_batch=1)
    assert at.tolist() == [0, 3, 3] and bo.tolist() == [0, 1, 2]


def test_head_of_line_blocks_later_small_requests():
    # a large request at the head is not overtaken by a small one behind it
    at, _, n, _ = S.admission_sim(11 * KB, [6 * KB, 6 * KB, 1 * KB, -1])
    assert at.tolist() == [0, 3, 3]


def test_coalescing_makes_room_for_a_large_block():
    script = [3 * KB, 3 * KB, 3 * KB, -1, -2, 7 * KB]
    at, _, n, _ = S.admission_sim(3 * 4864, script)
    assert at.tolist() == [0, 1, 2, 5]


def test_batch_block_released_only_when_all_requests_complete():
    script = [8 * KB, 2 * KB, 2 * KB, -1, -2, 8 * KB, -3, -4]
    at, bo, _, _ = S.admission_sim(12 * KB, script, max_batch=8)
    # r1, r2 share a batch; r3 needs r0's block and that batch's block gone
    assert bo[1] == bo[2] and at[1] == 3
    assert at[3] == 6


def test_never_fits_is_rejected():
    _, _, n, _ = S.admission_sim(4 * KB, [8 * KB])
    assert n == -7  # TCH_ENOMEM


def test_completing_a_request_not_in_flight_is_an_error():
    # request 1 still waits (two 2816-byte blocks exceed 4 KiB): not in flight
    _, _, n, _ = S.admission_sim(4 * KB, [1 * KB, 1 * KB, -2])
    assert n == -1  # TCH_EINVAL
    # completing the same request twice
    _, _, n, _ = S.admission_sim(8 * KB, [1 * KB, -1, -1])
    assert n == -1


class _Model:
    """Independent restatement of the policy: first-fit arena with
    coalescing, FIFO batches of the longest prefix fitting the largest free
    block."""

    def __init__(self, budget, max_batch):
        self.free = [(0, budget - budget % 256)] if budget >= 256 else []
        self.max_batch = max_batch
        self.q = []
        self.blocks = []  # [off, bytes, live]
        self.at, self.bo, self.state = [], [], []
        self.used = self.peak = 0

    def _alloc(self, b):
        for i, (o, ln) in enumerate(self.free):
            if ln >= b:
                self.free[i:i + 1] = [(o + b, ln - b)] if ln > b else []
                self.used += b
                self.peak = max(self.peak, self.used)
                return o
        raise AssertionError("no room")

    def _release(self, o, b):
        self.used -= b
        self.free.append((o, b))
        self.free.sort()
        merged = []
        for fo, fl in self.free:
            if merged and merged[-1][0] + merged[-1][1] == fo:
                merged[-1] = (merged[-1][0], merged[-1][1] + fl)
            else:
                merged.append((fo, fl))
        self.free = merged

    def step(self, i, x):
        if x > 0:
            self.q.append((len(self.state), x))
            self.state.append(0)
            self.at.append(-1)
            self.bo.append(-1)
        else:
            r = -x - 1
            assert self.state[r] == 1
            self.state[r] = 2
            blk = self.blocks[self.bo[r]]
            blk[2] -= 1
            if blk[2] == 0:
                self._release(blk[0], blk[1])
        while self.q:
            room = max((ln for _, ln in self.free), default=0)
            s = k = 0
            for _, sh in self.q:
                if k >= self.max_batch or _blk(s + sh) > room:
                    break
                s += sh
                k += 1
            if k == 0:
                break
            b = _blk(s)
            o = self._alloc(b)
            for r, _ in self.q[:k]:
                self.state[r] = 1
                self.at[r] = i
                self.bo[r] = len(self.blocks)
            self.blocks.append([o, b, k])
            self.q = self.q[k:]


def test_random_scripts_match_independent_model():
    rng = np.random.default_rng(7)
    for trial in range(40):
        budget = int(rng.integers(8, 64)) * KB
        max_batch = int(rng.integers(1, 6))
        m = _Model(budget, max_batch)
        script = []
        for i in range(int(rng.integers(20, 120))):
            live = [r for r, st in enumerate(m.state) if st == 1]
            if live and rng.random() < 0.45:
                x = -(live[int(rng.integers(len(live)))] + 1)
            else:
                x = int(rng.integers(16, budget - S.BATCH_FIXED - 256))
            script.append(x)
            m.step(i, x)
        at, bo, n, peak = S.admission_sim(budget, script, max_batch=max_batch)
        assert at.tolist() == m.at and bo.tolist() == m.bo
        assert n == sum(1 for a in m.at if a >= 0)
        assert peak == m.peak <= budget
        # first in first out: admission steps never decrease with request id
        adm = [a for a in m.at if a >= 0]
        assert adm == sorted(adm)


def test_block_layout_within_share_bound():
    rng = np.random.default_rng(3)
    for _ in range(300):
        G = int(rng.integers(1, 40))
        nv = rng.integers(0, 5000, G)
        nnz = rng.integers(0, 40000, G)
        exact = S.batch_block_bytes(G, int(nv.sum()), int(nnz.sum()))
        bound = -(-(S.BLOCK_FIXED + sum(S.request_share(int(a), int(b)) for a, b in zip(nv, nnz))) // 256) * 256
        assert exact % 256 == 0 and exact <= bound


def test_request_bytes_alignment():
    assert S.request_bytes(1000, 20000) % 256 == 0
    assert S.request_bytes(1000, 20000) >= 4 * (1001 + 20000 + 1000) + 16 + 8
    assert S.request_bytes(1000, 20000) == -(-(S.BLOCK_FIXED + S.request_share(1000, 20000)) // 256) * 256


def _orient_batch_scratch(nv, nnz):
    """Restatement of the scratch a batched count on the oriented scheme lays
    out (tc_capi.hip layout(): widened offsets, graph bases, counters, N+
    bounds, the orientation, piece rows, hub rows, degrees, long rows), each
    region rounded to 256 bytes."""
    r = lambda b: -(-b // 256) * 256  # noqa: E731
    return (r(8 * (nv + 1)) + r(4 * (nv + 1)) + r(8 * 544) + r(96) + r(16 * (nv + 1)) + r(4 * (nnz + 1))
            + r(4 * (nnz // 256 + 2 + 65536)) + r(4 * (nnz // 4096 + 1)) + r(4 * (nv + 1)) + r(4 * (nnz // 512 + 2)))


def test_scratch_bound_covers_the_count():
    """Every batch block's scratch region holds the batched count's scratch:
    the bound of csrc/host/admission.h against the layout restated, at the
    sizes where its rounding and integer divisions are tightest."""
    rng = np.random.default_rng(5)
    sizes = [(0, 0), (1, 0), (0, 255), (1, 511), (255, 4095), (10 ** 6, 3 * 10 ** 7), (2 ** 31 - 2, 2 ** 31 - 1)]
    sizes += [(int(a), int(b)) for a, b in zip(rng.integers(0, 10 ** 7, 500), rng.integers(0, 10 ** 9, 500))]
    for nv, nnz in sizes:
        blk = S.batch_block_bytes(1, nv, nnz)
        data = S.batch_block_bytes(1, 0, 0) - (-(-(S.SCRATCH_FIXED + S.scratch_share(0, 0)) // 256) * 256)
        assert _orient_batch_scratch(nv, nnz) <= S.SCRATCH_FIXED + S.scratch_share(nv, nnz), (nv, nnz)
        assert blk >= data


def _pack_block(graphs):
    import ctypes as C

    from gputrianglecounting_amd.graph import host_lib

    L = host_lib()
    L.tch_pack_block.restype = C.c_int64
    P = C.POINTER(C.c_int32)
    L.tch_pack_block.argtypes = [C.c_int32, P, C.POINTER(C.c_int64), C.POINTER(P), C.POINTER(P), C.c_void_p, C.c_int64]
    G = len(graphs)
    nv = np.array([g.n for g in graphs], np.int32)
    nnz = np.array([g.nnz for g in graphs], np.int64)
    offs = [np.ascontiguousarray(g.off, np.int32) for g in graphs]
    inds = [np.ascontiguousarray(g.ind, np.int32) if g.nnz else np.zeros(1, np.int32) for g in graphs]
    offp = (P * G)(*[o.ctypes.data_as(P) for o in offs])
    indp = (P * G)(*[i.ctypes.data_as(P) for i in inds])
    size = S.batch_block_bytes(G, int(nv.sum()), int(nnz.sum()))
    blk = np.full(size, 0xAB, np.uint8)
    n = L.tch_pack_block(G, nv.ctypes.data_as(P), nnz.ctypes.data_as(C.POINTER(C.c_int64)), offp, indp,
                         blk.ctypes.data, size)
    return blk, int(n), size


def test_server_block_matches_batch_layout():
    """The block the serving loop uploads holds exactly tch_pack's table,
    offsets and ids (the batched counter's input) at 256-aligned regions."""
    from gputrianglecounting_amd import graph as G

    graphs = [G.rmat(8, 8, seed=1), G.from_csr([0, 0], []), G.erdos_renyi(300, 1500, seed=2),
              G.geometric(500, 6.0, seed=3), G.from_csr([0], [])]
    blk, n, size = _pack_block(graphs)
    b = G.pack(graphs)
    Gn = len(graphs)
    tab_bytes = 16 * (Gn + 1)
    off_at = -(-tab_bytes // 256) * 256
    off_bytes = 4 * (b.total_vertices + Gn)
    ind_at = off_at + -(-off_bytes // 256) * 256
    assert n == ind_at + 4 * b.total_arcs and n <= size
    np.testing.assert_array_equal(blk[:tab_bytes].view(np.int64), b.tab)
    np.testing.assert_array_equal(blk[off_at:off_at + off_bytes].view(np.int32), b.off)
    np.testing.assert_array_equal(blk[ind_at:n].view(np.int32), b.ind)
    # every graph's arcs end where the next graph's begin: voff[v+1] closes
    # the last row of each graph
    for g in range(Gn):
        assert b.tab[2 * g + 1] + graphs[g].nnz == b.tab[2 * g + 3]


def test_server_block_parallel_copy():
    """A block of more than 16 MiB is packed by several threads; every byte
    matches tch_pack's batch."""
    from gputrianglecounting_amd import graph as G

    graphs = [G.rmat(17, 16, seed=4), G.from_csr([0, 0], []), G.erdos_renyi(200000, 1500000, seed=5),
              G.rmat(12, 8, seed=6)]
    blk, n, size = _pack_block(graphs)
    b = G.pack(graphs)
    assert 4 * (b.total_arcs + b.total_vertices) > (16 << 20)
    Gn = len(graphs)
    tab_bytes = 16 * (Gn + 1)
    off_at = -(-tab_bytes // 256) * 256
    off_bytes = 4 * (b.total_vertices + Gn)
    ind_at = off_at + -(-off_bytes // 256) * 256
    assert n == ind_at + 4 * b.total_arcs and n <= size
    np.testing.assert_array_equal(blk[:tab_bytes].view(np.int64), b.tab)
    np.testing.assert_array_equal(blk[off_at:off_at + off_bytes].view(np.int32), b.off)
    np.testing.assert_array_equal(blk[ind_at:n].view(np.int32), b.ind)


def test_copy_pool_concurrent_callers():
    """The serving loop's copy pool (large submits are split over its
    threads): several callers at once, many rounds, buffers from a few bytes
    to several MiB (pieces of 1 MiB, the 2 MiB inline threshold on both
    sides); every destination byte equals its source."""
    import ctypes as C

    from gputrianglecounting_amd import graph as G

    L = G.host_lib()
    L.tch_copy_pool_run.restype = C.c_int
    L.tch_copy_pool_run.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]
    rng = np.random.default_rng(5)
    sizes = [1, 255, 4096, (2 << 20) - 4, (2 << 20) + 12, 5 << 20, (3 << 20) + 7]
    srcs = [rng.integers(0, 256, s, dtype=np.uint8) for s in sizes]
    for threads, callers in ((7, 4), (0, 2), (3, 1)):
        dsts = [[np.zeros(s, np.uint8) for s in sizes] for _ in range(callers)]
        dp = (C.c_void_p * (callers * len(sizes)))(*[d.ctypes.data for row in dsts for d in row])
        sp = (C.c_void_p * len(sizes))(*[a.ctypes.data for a in srcs])
        bp = (C.c_int64 * len(sizes))(*sizes)
        assert L.tch_copy_pool_run(threads, callers, 6, len(sizes), dp, sp, bp) == 0
        for row in dsts:
            for d, a in zip(row, srcs):
                np.testing.assert_array_equal(d, a)
