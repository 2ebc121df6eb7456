"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

``liboracle.so`` is the C restatement in tc_oracle.c; ``_ref/libtcref.so`` is
the reference's own CPU code compiled by oracle/Makefile when the reference
tree is present.  Only tests/, the smoke check and bench.py's cpu_baseline
leg load this module, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_o = None
_r = None


def _p32(a):
    return a.ctypes.data_as(_I32P)


def _p64(a):
    return a.ctypes.data_as(_I64P)


def lib():
    global _o
    if _o is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = C.CDLL(path)
        L.oracle_count_intersections.restype = C.c_int64
        L.oracle_count_intersections.argtypes = [C.c_int32, C.c_int32, _I32P, C.c_int32, C.c_int32, _I32P]
        L.oracle_count_arcs.restype = None
        L.oracle_count_arcs.argtypes = [C.c_int32, _I32P, _I32P, _I64P, _I64P]
        L.oracle_count_vertex_list.restype = None
        L.oracle_count_vertex_list.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I64P, _I64P]
        L.oracle_per_vertex.restype = C.c_int
        L.oracle_per_vertex.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I64P]
        L.oracle_per_vertex_omp.restype = C.c_int
        L.oracle_per_vertex_omp.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I64P, C.c_int32]
        L.oracle_count_arc_sample.restype = C.c_int64
        L.oracle_count_arc_sample.argtypes = [_I32P, _I32P, _I32P, _I64P, C.c_int64]
        L.oracle_vertex_list_lim.restype = None
        L.oracle_vertex_list_lim.argtypes = [C.c_int32, _I32P, _I32P, _I32P, C.c_int32, _I64P, _I64P]
        L.oracle_vertex_list_omp.restype = None
        L.oracle_vertex_list_omp.argtypes = [C.c_int32, _I32P, _I64P, _I32P, _I32P, _I64P, _I64P, C.c_int32]
        L.oracle_lcc.restype = None
        L.oracle_lcc.argtypes = [C.c_int32, _I32P, _I32P, C.POINTER(C.c_double)]
        _o = L
    return _o


def ref_lib():
    """The compiled reference CPU code, or None when it was not built."""
    global _r
    if _r is None:
        path = os.path.join(HERE, "_ref", "libtcref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.tcref_count_intersections.restype = C.c_int64
        L.tcref_count_intersections.argtypes = [C.c_int32, C.c_int32, _I32P, C.c_int32, C.c_int32, _I32P]
        L.tcref_count_all.restype = None
        L.tcref_count_all.argtypes = [C.c_int32, _I32P, _I32P, _I64P, _I64P]
        L.tcref_count_vertex_list.restype = None
        L.tcref_count_vertex_list.argtypes = [C.c_int32, _I32P, _I32P, _I32P, _I64P, _I64P]
        L.tcref_read_info.restype = C.c_int
        L.tcref_read_info.argtypes = [C.c_char_p, _I32P, _I32P, _I32P, C.c_int64, _I32P, C.c_int64]
        L.tcref_read_metis.restype = C.c_int
        L.tcref_read_metis.argtypes = [C.c_char_p, _I32P, _I64P, _I32P, C.c_int64, _I32P, C.c_int64]
        L.tcref_count_arc_sample.restype = C.c_int64
        L.tcref_count_arc_sample.argtypes = [_I32P, _I32P, _I32P, _I64P, C.c_int64]
        L.tcref_write_info.restype = None
        L.tcref_write_info.argtypes = [C.c_char_p, C.c_int32, _I32P, _I32P]
        _r = L
    return _r


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def count_arcs(off, ind):
    """Per-arc counts triNE (int64, CSR order) and total (clusteringCount.cpp:50-70)."""
    off, ind = _c(off, np.int32), _c(ind, np.int32)
    tri = np.zeros(max(ind.shape[0], 1), np.int64)
    tot = np.zeros(1, np.int64)
    lib().oracle_count_arcs(off.shape[0] - 1, _p32(off), _p32(ind), _p64(tri), _p64(tot))
    return tri[: ind.shape[0]], int(tot[0])


def per_vertex(off, ind, threads: int = 0):
    """Per-vertex row sums (2·T(v) on a simple graph) and the int64 total."""
    off, ind = _c(off, np.int32), _c(ind, np.int32)
    n = off.shape[0] - 1
    out = np.zeros(max(n, 1), np.int32)
    tot = np.zeros(1, np.int64)
    if threads:
        rc = lib().oracle_per_vertex_omp(n, _p32(off), _p32(ind), _p32(out), _p64(tot), threads)
    else:
        rc = lib().oracle_per_vertex(n, _p32(off), _p32(ind), _p32(out), _p64(tot))
    if rc != 0:
        raise OverflowError("a per-vertex count exceeds int32")
    return out[:n], int(tot[0])


def lcc(off, out):
    off, out = _c(off, np.int32), _c(out, np.int32)
    n = off.shape[0] - 1
    res = np.zeros(max(n, 1), np.float64)
    lib().oracle_lcc(n, _p32(off), _p32(out), res.ctypes.data_as(C.POINTER(C.c_double)))
    return res[:n]


def intersections(ai, a, bi, b, use_ref=False):
    a, b = _c(a, np.int32), _c(b, np.int32)
    L = ref_lib() if use_ref else lib()
    fn = L.tcref_count_intersections if use_ref else L.oracle_count_intersections
    return int(fn(int(ai), a.shape[0], _p32(a), int(bi), b.shape[0], _p32(b)))


def ref_count_arcs(off, ind):
    """The reference's own count_all_trianglesCPU<int32_t> (needs oracle/_ref)."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref not built")
    off, ind = _c(off, np.int32), _c(ind, np.int32)
    tri = np.zeros(max(ind.shape[0], 1), np.int64)
    tot = np.zeros(1, np.int64)
    L.tcref_count_all(off.shape[0] - 1, _p32(off), _p32(ind), _p64(tri), _p64(tot))
    return tri[: ind.shape[0]], int(tot[0])


def ref_count_vertex_list(vertices, off, ind):
    """Reference vertex-list overload (clusteringCount.cpp:125-155)."""
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref not built")
    off, ind, vl = _c(off, np.int32), _c(ind, np.int32), _c(vertices, np.int32)
    narcs = int(sum(int(off[v + 1]) - int(off[v]) for v in vl))
    tri = np.zeros(max(narcs, 1), np.int64)
    tot = np.zeros(1, np.int64)
    L.tcref_count_vertex_list(vl.shape[0], _p32(vl), _p32(off), _p32(ind), _p64(tri), _p64(tot))
    return tri[:narcs], int(tot[0])


def count_vertex_list(vertices, off, ind):
    """The C restatement of the vertex-list overload (clusteringCount.cpp:125-155):
    per-arc counts of the listed rows in list order, and their total."""
    off, ind, vl = _c(off, np.int32), _c(ind, np.int32), _c(vertices, np.int32)
    narcs = int(sum(int(off[v + 1]) - int(off[v]) for v in vl))
    tri = np.zeros(max(narcs, 1), np.int64)
    tot = np.zeros(1, np.int64)
    lib().oracle_count_vertex_list(vl.shape[0], _p32(vl), _p32(off), _p32(ind), _p64(tri), _p64(tot))
    return tri[:narcs], int(tot[0])


def vertex_list_omp(vertices, off, ind, threads: int = 16):
    """Per-list-position row sums of the vertex-list overload
    (clusteringCount.cpp:125-155) with the listed rows' arcs spread over
    ``threads`` OpenMP threads (int64 per position), and their total."""
    off, ind, vl = _c(off, np.int32), _c(ind, np.int32), _c(vertices, np.int32)
    lens = off[vl.astype(np.int64) + 1].astype(np.int64) - off[vl].astype(np.int64)
    pre = np.zeros(vl.shape[0] + 1, np.int64)
    np.cumsum(lens, out=pre[1:])
    out = np.zeros(max(vl.shape[0], 1), np.int64)
    tot = np.zeros(1, np.int64)
    lib().oracle_vertex_list_omp(vl.shape[0], _p32(vl), _p64(pre), _p32(off), _p32(ind), _p64(out), _p64(tot),
                                 int(threads))
    return out[: vl.shape[0]], int(tot[0])


def vertex_list_lim(vertices, off, ind, lim):
    """Per-list-position sums of the vertex-list GPU kernel
    (clusteringCount.cu:266-312) with the shadow-vertex intersection
    (clusteringCount.cu:150-177): each common neighbour x adds 1 + [x > lim]."""
    off, ind, vl = _c(off, np.int32), _c(ind, np.int32), _c(vertices, np.int32)
    out = np.zeros(max(vl.shape[0], 1), np.int64)
    tot = np.zeros(1, np.int64)
    lib().oracle_vertex_list_lim(vl.shape[0], _p32(vl), _p32(off), _p32(ind), int(lim), _p64(out), _p64(tot))
    return out[: vl.shape[0]], int(tot[0])


def list_sums(vertices, off, tri):
    """Per-list-position sums of vertex-list per-arc counts (the value the
    reference's vertex-list kernel stores at outPutTriangles[i])."""
    off = np.asarray(off, np.int64)
    lens = np.array([off[v + 1] - off[v] for v in vertices], np.int64)
    cs = np.concatenate([[0], np.cumsum(tri, dtype=np.int64)])
    ends = np.cumsum(lens)
    return cs[ends] - cs[ends - lens]


def row_sums(off, tri):
    off = np.asarray(off, np.int64)
    cs = np.concatenate([[0], np.cumsum(tri, dtype=np.int64)])
    return cs[off[1:]] - cs[off[:-1]]


def count_arc_sample(off, ind, arcs, use_ref=True):
    """Sum of per-arc intersections over the arc indices ``arcs`` (visited in
    the given order) with the reference code (``use_ref``) or the port."""
    off, ind = _c(off, np.int32), _c(ind, np.int32)
    arcs = _c(arcs, np.int64)
    src = _c(np.searchsorted(off, arcs, side="right") - 1, np.int32)
    L = ref_lib() if use_ref else lib()
    if L is None:
        raise RuntimeError("oracle/_ref not built")
    fn = L.tcref_count_arc_sample if use_ref else L.oracle_count_arc_sample
    return int(fn(_p32(off), _p32(ind), _p32(src), _p64(arcs), arcs.shape[0]))
