"""ctypes view of the host graph driver (``libtc_host.so``, include/tc_host.h).

Graphs come back as :class:`Graph` objects holding numpy copies of the CSR
(``off`` int32[n+1], ``ind`` int32[nnz]).  Reference counterparts:
``readGraph`` (src/graphRead.cpp:2-26) for :func:`read_info`, the METIS
reader and ``writeGraph`` (src/csrdatacreate.cpp:3-49, 86-114) for
:func:`read_metis` / :meth:`Graph.write_info`.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")

_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)

_host = None


class HostError(RuntimeError):
    """A host-driver call failed; ``code`` is the TCH_E* value."""

    def __init__(self, msg: str, code: int = -1):
        super().__init__(msg)
        self.code = code


def host_lib() -> C.CDLL:
    """Load ``libtc_host.so`` (built by ``__graft_entry__.build()``)."""
    global _host
    if _host is not None:
        return _host
    path = os.path.join(LIB_DIR, "libtc_host.so")
    if not os.path.exists(path):
        raise HostError(f"{path} is missing: build the package first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "tch_last_error": (C.c_char_p, []),
        "tch_graph_n": (C.c_int32, [vp]),
        "tch_graph_nnz": (C.c_int64, [vp]),
        "tch_graph_off": (_I32P, [vp]),
        "tch_graph_ind": (_I32P, [vp]),
        "tch_graph_free": (None, [vp]),
        "tch_from_edges": (vp, [C.c_int32, C.c_int64, _I32P, _I32P]),
        "tch_from_csr": (vp, [C.c_int32, C.c_int64, _I32P, _I32P]),
        "tch_normalize": (vp, [vp]),
        "tch_gen_er": (vp, [C.c_int32, C.c_int64, C.c_uint64]),
        "tch_gen_rmat": (vp, [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int32]),
        "tch_gen_rgg": (vp, [C.c_int32, C.c_double, C.c_uint64]),
        "tch_check_simple": (C.c_int, [C.c_int32, C.c_int64, _I32P, _I32P]),
        "tch_read_info": (vp, [C.c_char_p]),
        "tch_write_info": (C.c_int, [vp, C.c_char_p]),
        "tch_read_metis": (vp, [C.c_char_p]),
        "tch_write_bin": (C.c_int, [vp, C.c_char_p]),
        "tch_read_bin": (vp, [C.c_char_p]),
        "tch_gen_batch": (vp, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint64]),
        "tch_pack": (vp, [C.c_int32, C.POINTER(vp)]),
        "tch_batch_count": (C.c_int32, [vp]),
        "tch_batch_tab": (_I64P, [vp]),
        "tch_batch_off": (_I32P, [vp]),
        "tch_batch_ind": (_I32P, [vp]),
        "tch_batch_kind": (_I32P, [vp]),
        "tch_batch_graph": (vp, [vp, C.c_int32]),
        "tch_batch_free": (None, [vp]),
        "tch_read_partition": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(_I32P)]),
        "tch_partition_members": (C.c_int, [C.c_int32, _I32P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(_I32P)]),
        "tch_free": (None, [vp]),
        "tch_graph64_n": (C.c_int64, [vp]),
        "tch_graph64_nnz": (C.c_int64, [vp]),
        "tch_graph64_off": (_I64P, [vp]),
        "tch_graph64_ind": (_I64P, [vp]),
        "tch_graph64_free": (None, [vp]),
        "tch_read_info64": (vp, [C.c_char_p]),
        "tch_read_bin64": (vp, [C.c_char_p]),
        "tch_write_bin64": (C.c_int, [vp, C.c_char_p]),
        "tch_widen": (vp, [vp]),
        "tch_check_simple64": (C.c_int, [C.c_int64, C.c_int64, _I64P, _I64P]),
        "tch_peek_sizes": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _host = L
    return L


def _err() -> str:
    return host_lib().tch_last_error().decode()


def _ptr32(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


@dataclass
class Graph:
    """A CSR graph in host memory: ``off`` int32[n+1], ``ind`` int32[nnz]."""

    off: np.ndarray
    ind: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.off.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.ind.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.off.astype(np.int64))

    def write_info(self, path: str) -> None:
        """Write the reference ``.info`` text layout (csrdatacreate.cpp:86-114)."""
        h = _to_handle(self)
        try:
            rc = host_lib().tch_write_info(h, path.encode())
            if rc != 0:
                raise HostError(_err(), rc)
        finally:
            host_lib().tch_graph_free(h)

    def write_bin(self, path: str) -> None:
        h = _to_handle(self)
        try:
            rc = host_lib().tch_write_bin(h, path.encode())
            if rc != 0:
                raise HostError(_err(), rc)
        finally:
            host_lib().tch_graph_free(h)

    def check_simple(self) -> int:
        """0 if the CSR is a simple symmetric graph, else a TCH_E* code."""
        return int(host_lib().tch_check_simple(self.n, self.nnz, _ptr32(self.off), _ptr32(self.ind)))

    def normalized(self) -> "Graph":
        h = _to_handle(self)
        try:
            return _from_handle(host_lib().tch_normalize(h), self.name)
        finally:
            host_lib().tch_graph_free(h)


def _from_handle(h, name: str = "") -> Graph:
    L = host_lib()
    if not h:
        raise HostError(_err())
    try:
        n = L.tch_graph_n(h)
        nnz = L.tch_graph_nnz(h)
        off = np.ctypeslib.as_array(L.tch_graph_off(h), shape=(n + 1,)).copy()
        ind = np.ctypeslib.as_array(L.tch_graph_ind(h), shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    finally:
        L.tch_graph_free(h)
    return Graph(off, ind, name)


def _to_handle(g: Graph):
    off = np.ascontiguousarray(g.off, dtype=np.int32)
    ind = np.ascontiguousarray(g.ind, dtype=np.int32)
    h = host_lib().tch_from_csr(g.n, g.nnz, _ptr32(off), _ptr32(ind))
    if not h:
        raise HostError(_err())
    return h


def from_edges(n: int, src, dst, name: str = "") -> Graph:
    """Simple symmetric CSR from an edge list (self-loops/duplicates dropped)."""
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    return _from_handle(host_lib().tch_from_edges(int(n), s.shape[0], _ptr32(s), _ptr32(d)), name)


def from_csr(off, ind, name: str = "") -> Graph:
    """Wrap an existing CSR (no normalisation)."""
    return Graph(np.ascontiguousarray(off, dtype=np.int32), np.ascontiguousarray(ind, dtype=np.int32), name)


def erdos_renyi(n: int, m: int, seed: int = 1) -> Graph:
    """G(n, m draws): m uniform (u, v) draws, simple symmetric CSR."""
    return _from_handle(host_lib().tch_gen_er(n, m, seed & (2**64 - 1)), f"er_n{n}_m{m}_s{seed}")


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1, n_keep: int = 0) -> Graph:
    """R-MAT with ``edge_factor * 2^scale`` draws (vertex 0 is the hub)."""
    m = edge_factor * (n_keep if n_keep > 0 else (1 << scale))
    return _from_handle(host_lib().tch_gen_rmat(scale, m, a, b, c, seed & (2**64 - 1), n_keep),
                        f"rmat_s{scale}_ef{edge_factor}_s{seed}")


def geometric(n: int, mean_degree: float = 6.0, seed: int = 1) -> Graph:
    """2D random geometric graph, r^2 = mean_degree / (pi n)."""
    return _from_handle(host_lib().tch_gen_rgg(n, mean_degree, seed & (2**64 - 1)), f"rgg_n{n}_s{seed}")


def read_info(path_no_ext: str) -> Graph:
    """Reference ``.info`` reader (graphRead.cpp:2-26); appends ``.info``."""
    return _from_handle(host_lib().tch_read_info(path_no_ext.encode()), os.path.basename(path_no_ext))


def read_metis(path: str) -> Graph:
    """METIS/DIMACS10 adjacency reader (csrdatacreate.cpp:3-49); rows sorted, nothing dropped."""
    return _from_handle(host_lib().tch_read_metis(path.encode()), os.path.basename(path))


def _take_i32(ptr, n: int) -> np.ndarray:
    L = host_lib()
    try:
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.int32)
    finally:
        L.tch_free(C.cast(ptr, C.c_void_p))


def read_partition(path_no_ext: str, k: int) -> np.ndarray:
    """The integers of ``<path_no_ext>.part.<k>`` after its count header
    (graphRead.cpp:28-45)."""
    n = C.c_int32(0)
    ptr = _I32P()
    rc = host_lib().tch_read_partition(path_no_ext.encode(), k, C.byref(n), C.byref(ptr))
    if rc != 0:
        raise HostError(_err())
    return _take_i32(ptr, n.value)


def partition_members(part, p: int) -> np.ndarray:
    """Vertices v with part[v] == p, ascending (part: a METIS partition vector)."""
    part = np.ascontiguousarray(part, np.int32)
    cnt = C.c_int32(0)
    ptr = _I32P()
    rc = host_lib().tch_partition_members(part.shape[0], _ptr32(part), p, C.byref(cnt), C.byref(ptr))
    if rc != 0:
        raise HostError(_err())
    return _take_i32(ptr, cnt.value)


def _from_handle64(h):
    L = host_lib()
    if not h:
        raise HostError(_err())
    try:
        n = L.tch_graph64_n(h)
        nnz = L.tch_graph64_nnz(h)
        off = np.ctypeslib.as_array(L.tch_graph64_off(h), shape=(n + 1,)).copy()
        ind = np.ctypeslib.as_array(L.tch_graph64_ind(h), shape=(nnz,)).copy() if nnz else np.zeros(0, np.int64)
    finally:
        L.tch_graph64_free(h)
    return off, ind


def read_info64(path_no_ext: str):
    """``.info`` read as the reference binary reads it (readGraph<int64_t>,
    graphRead.cpp:2-26, run by main.cu:36): int64 ``(off, ind)``, any arc count."""
    return _from_handle64(host_lib().tch_read_info64(path_no_ext.encode()))


def read_bin64(path: str):
    """int64 ``(off, ind)`` from the int64 binary cache, or the int32 one widened."""
    return _from_handle64(host_lib().tch_read_bin64(path.encode()))


def write_bin64(g: Graph, path: str) -> None:
    """The int64 binary cache ("TCCSR064") of ``g``."""
    h = _to_handle(g)
    L = host_lib()
    w = L.tch_widen(h)
    L.tch_graph_free(h)
    if not w:
        raise HostError(_err())
    try:
        rc = L.tch_write_bin64(w, path.encode())
        if rc != 0:
            raise HostError(_err(), rc)
    finally:
        L.tch_graph64_free(w)


def check_simple64(off, ind) -> int:
    """tch_check_simple on an int64 CSR: 0 if simple and symmetric, else a TCH_E* code."""
    o = np.ascontiguousarray(off, dtype=np.int64)
    i = np.ascontiguousarray(ind, dtype=np.int64)
    return int(host_lib().tch_check_simple64(o.shape[0] - 1, i.shape[0], o.ctypes.data_as(_I64P),
                                             i.ctypes.data_as(_I64P)))


def peek_sizes(path: str):
    """(n, nnz) from the header of ``path + '.info'``, or of ``path`` if it ends in ``.bin``."""
    n, nnz = C.c_int64(), C.c_int64()
    rc = host_lib().tch_peek_sizes(path.encode(), C.byref(n), C.byref(nnz))
    if rc != 0:
        raise HostError(_err(), rc)
    return int(n.value), int(nnz.value)


def read_bin(path: str) -> Graph:
    return _from_handle(host_lib().tch_read_bin(path.encode()), os.path.basename(path))


@dataclass
class Batch:
    """Many graphs packed into one concatenated CSR (see include/tc_host.h)."""

    tab: np.ndarray  # int64 [(G+1) * 2]: v_base, arc_base
    off: np.ndarray  # int32 [sum(n_g + 1)]
    ind: np.ndarray  # int32 [sum(nnz_g)]
    kind: np.ndarray  # int32 [G]

    @property
    def count(self) -> int:
        return int(self.kind.shape[0])

    def v_base(self, g: int) -> int:
        return int(self.tab[2 * g])

    def arc_base(self, g: int) -> int:
        return int(self.tab[2 * g + 1])

    def graph(self, g: int) -> Graph:
        vb, ab = self.v_base(g), self.arc_base(g)
        n = int(self.tab[2 * g + 2]) - vb
        nnz = int(self.tab[2 * g + 3]) - ab
        return Graph(self.off[vb + g: vb + g + n + 1].copy(), self.ind[ab: ab + nnz].copy(), f"batch[{g}]")

    @property
    def total_vertices(self) -> int:
        return int(self.tab[-2])

    @property
    def total_arcs(self) -> int:
        return int(self.tab[-1])


def _batch_from_handle(h) -> Batch:
    L = host_lib()
    if not h:
        raise HostError(_err())
    try:
        G = L.tch_batch_count(h)
        tab = np.ctypeslib.as_array(L.tch_batch_tab(h), shape=(2 * (G + 1),)).copy()
        nv, na = int(tab[-2]), int(tab[-1])
        off = np.ctypeslib.as_array(L.tch_batch_off(h), shape=(nv + G,)).copy()
        ind = np.ctypeslib.as_array(L.tch_batch_ind(h), shape=(na,)).copy() if na else np.zeros(0, np.int32)
        kind = np.ctypeslib.as_array(L.tch_batch_kind(h), shape=(G,)).copy()
    finally:
        L.tch_batch_free(h)
    return Batch(tab, off, ind, kind)


def gen_batch(G: int = 4096, n_lo: int = 64, n_hi: int = 65536, ef_lo: int = 8, ef_hi: int = 32,
              kinds: int = 7, seed: int = 7) -> Batch:
    """The mixed ER / R-MAT / geometric batch (include/tc_host.h: tch_gen_batch)."""
    return _batch_from_handle(host_lib().tch_gen_batch(G, n_lo, n_hi, ef_lo, ef_hi, kinds, seed & (2**64 - 1)))


def pack(graphs) -> Batch:
    """Pack a list of :class:`Graph` into one batch."""
    hs = [_to_handle(g) for g in graphs]
    try:
        arr = (C.c_void_p * len(hs))(*hs)
        b = _batch_from_handle(host_lib().tch_pack(len(hs), arr))
    finally:
        for h in hs:
            host_lib().tch_graph_free(h)
    b.kind[:] = -1
    return b
