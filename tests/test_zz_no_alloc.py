"""tc_count_noalloc_i32: the count that allocates nothing, with its scratch
in the library's fixed per-GPU array, against tc_count_i32 (which grows a
context-owned scratch) and the CPU oracle, bit for bit.  Both entry points
are called through the C-ABI with ctypes.

Two calls in a row, the second on a smaller graph: the array keeps the
first call's contents, so the second shows that every counter, flag and
offset the kernels read is cleared or rewritten by the call itself."""
import ctypes as C

import numpy as np
import pytest

import oracle
from gputrianglecounting_amd import graph as G
from gputrianglecounting_amd import tc as T

pytestmark = pytest.mark.gpu


def _count(L, h, fn, off, ind, torch):
    nv, nnz = int(off.shape[0]) - 1, int(ind.shape[0])
    out = torch.full((max(nv, 1),), -7, dtype=torch.int32, device=off.device)
    total = torch.full((1,), -7, dtype=torch.int64, device=off.device)
    rc = fn(h, nv, nnz, off.data_ptr(), ind.data_ptr() if nnz else None, out.data_ptr(), total.data_ptr(), None,
            C.c_void_p(int(torch.cuda.current_stream().cuda_stream)))
    assert rc == T.TC_OK, L.tc_last_error(h).decode()
    torch.cuda.synchronize()
    return out[:nv].cpu().numpy(), int(total.cpu().item())


def test_no_alloc_matches_count_i32_twice():
    import torch

    L = T.lib()
    h = C.c_void_p()
    assert L.tc_ctx_create(0, C.byref(h)) == T.TC_OK
    try:
        # first a larger graph, then a smaller one on the same array
        for g in (G.rmat(17, 16, seed=3), G.erdos_renyi(16384, 131072, seed=1)):
            print(f"case no-alloc {g.name} n={g.n} nnz={g.nnz}", flush=True)
            off = torch.from_numpy(g.off).cuda()
            ind = torch.from_numpy(g.ind).cuda()
            a, ta = _count(L, h, L.tc_count_noalloc_i32, off, ind, torch)
            b, tb = _count(L, h, L.tc_count_i32, off, ind, torch)
            ref, rtot = oracle.per_vertex(g.off, g.ind)
            np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(a, ref)
            assert ta == tb == rtot
    finally:
        L.tc_ctx_destroy(h)


def test_no_alloc_rejects_graphs_beyond_the_array():
    """A graph whose scratch exceeds TC_NOALLOC_SCRATCH_BYTES (64 MiB) is
    refused on the host with TC_ERANGE before anything is launched; other
    schemes than the oriented one are refused with TC_EINVAL."""
    import torch

    L = T.lib()
    h = C.c_void_p()
    assert L.tc_ctx_create(0, C.byref(h)) == T.TC_OK
    try:
        nv = 2_500_000  # 28 (nv + 1) bytes of offsets, orientation rows and degrees: 70 MB
        off = torch.zeros(nv + 1, dtype=torch.int32, device="cuda")
        out = torch.zeros(nv, dtype=torch.int32, device="cuda")
        rc = L.tc_count_noalloc_i32(h, nv, 0, off.data_ptr(), None, out.data_ptr(), None, None, None)
        assert rc == T.TC_ERANGE
        assert b"no-allocation array" in L.tc_last_error(h)
        g = G.erdos_renyi(1024, 4096, seed=2)
        d_off, d_ind = torch.from_numpy(g.off).cuda(), torch.from_numpy(g.ind).cuda()
        o = torch.zeros(g.n, dtype=torch.int32, device="cuda")
        rc = L.tc_count_noalloc_i32(h, g.n, g.nnz, d_off.data_ptr(), d_ind.data_ptr(), o.data_ptr(), None,
                                    C.byref(T.TcCfg.make(oriented=-1)), None)
        assert rc == T.TC_EINVAL
    finally:
        L.tc_ctx_destroy(h)
