"""ctypes binding of the gfx950 counter (``libtc.so``, include/tc.h).

Device memory and streams come from PyTorch (plumbing only); every count
runs in the HIP kernels of ``libtc.so``.  There is no CPU fallback: if the
library is missing or no GPU is visible, the calls raise.

Reference interface mirrored here:
  kernelCall<int32_t>(blocks, bs, nv, d_off, d_ind, out, tpi, groups, shifter)
      src/clusteringCount.cu:415-467            -> :func:`kernel_call`
  singleParamTestGPURun<T> (H2D, count, total)  src/gpurun.cu:2-56
                                                -> :meth:`TriangleCounter.count_graph`
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .graph import LIB_DIR, Batch, Graph

TC_OK, TC_EINVAL, TC_ERANGE, TC_EHIP, TC_ENOMEM, TC_EGRAPH, TC_ECANCELED = 0, -1, -2, -3, -4, -5, -6

_lib = None


class TCError(RuntimeError):
    def __init__(self, msg: str, code: int):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class TcCfg(C.Structure):
    """Mirror of ``tc_cfg`` (include/tc.h)."""

    _fields_ = [("wg_size", C.c_int32), ("items", C.c_int32), ("ring", C.c_int32),
                ("small_max", C.c_int32), ("ranges_per_wg", C.c_int32), ("algo", C.c_int32),
                ("chunk", C.c_int32), ("depth", C.c_int32), ("search_ratio", C.c_int32),
                ("probe_ratio", C.c_int32), ("oriented", C.c_int32),
                ("row_groups", C.c_int32), ("lookup", C.c_int32), ("long_rows", C.c_int32),
                ("window", C.c_int32), ("orient_wg", C.c_int32), ("piece", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}

    @classmethod
    def make(cls, **kw) -> "TcCfg":
        c = cls()
        for k, v in kw.items():
            setattr(c, k, int(v))
        return c


class TuneResult(C.Structure):
    """Mirror of ``tc_tune_result``."""

    _fields_ = [("cfg", TcCfg), ("ms", C.c_float), ("identical", C.c_int32)]


class CapturedCount:
    """A recorded count call (``tc_graph``): replay with :meth:`launch`."""

    def __init__(self, handle):
        self._h = handle

    def launch(self, stream=None) -> None:
        rc = lib().tc_graph_launch(self._h, _stream(stream))
        if rc != TC_OK:
            raise TCError("tc_graph_launch failed", rc)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lib() -> C.CDLL:
    """Load ``libtc.so``; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(LIB_DIR, "libtc.so")
    if not os.path.exists(path):
        raise TCError(f"{path} is missing: the HIP extension is not built", TC_EINVAL)
    L = C.CDLL(path)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    sig = {
        "tc_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "tc_ctx_destroy": (C.c_int, [vp]),
        "tc_last_error": (C.c_char_p, [vp]),
        "tc_scratch_bytes": (C.c_size_t, [i32, i64]),
        "tc_batch_scratch_bytes": (C.c_size_t, [i32, i64, i64]),
        "tc_ctx_reserve": (C.c_int, [vp, C.c_size_t]),
        "tc_ctx_set_default_cfg": (C.c_int, [vp, C.POINTER(TcCfg)]),
        "tc_ctx_get_default_cfg": (C.c_int, [vp, C.POINTER(TcCfg)]),
        "tc_ctx_set_timing": (C.c_int, [vp, C.c_int]),
        "tc_last_timing": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "tc_count_i32": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, C.POINTER(TcCfg), vp]),
        "tc_count_noalloc_i32": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, C.POINTER(TcCfg), vp]),
        "tc_kernelCall_i32": (C.c_int, [C.c_uint, C.c_uint, i32, vp, vp, vp, i32, i32, i32]),
        "tc_count_i64": (C.c_int, [vp, i64, i64, vp, vp, vp, vp, C.POINTER(TcCfg), vp]),
        "tc_count_total_i64": (C.c_int, [vp, i64, i64, vp, vp, vp, vp]),
        "tc_kernelCall_i64": (C.c_int, [C.c_uint, C.c_uint, i64, vp, vp, vp, i64, i64, i64]),
        "tc_lcc_f64": (C.c_int, [vp, i32, vp, vp, vp, vp]),
        "tc_validate_i32": (C.c_int, [vp, i32, i64, vp, vp, C.POINTER(C.c_int), vp]),
        "tc_batch_count_i32": (C.c_int, [vp, i32, vp, i64, i64, vp, vp, vp, vp, C.POINTER(TcCfg), vp]),
        "tc_count_total_i32": (C.c_int, [vp, i32, i64, vp, vp, vp, vp]),
        "tc_count_vertices_i32": (C.c_int, [vp, i32, i64, vp, vp, i32, vp, vp, vp, vp, vp]),
        "tc_count_vertices_lim_i32": (C.c_int, [vp, i32, i64, vp, vp, i32, vp, i32, vp, vp, vp, vp]),
        "tc_device_cus": (C.c_int, [vp]),
        "tc_last_edge_work": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "tc_oriented_plan": (C.c_int, [vp, i64, i64, C.c_int32, C.POINTER(TcCfg), C.POINTER(C.c_int32)]),
        "tc_copy_i32": (C.c_int, [vp, vp, vp, i64, vp]),
        "tc_batch_count_pergraph_i32": (C.c_int, [vp, i32, C.POINTER(C.c_int64), vp, vp, vp, vp, C.POINTER(TcCfg), vp]),
        "tc_count_capture_i32": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, C.POINTER(TcCfg), C.POINTER(vp)]),
        "tc_batch_capture_i32": (C.c_int, [vp, i32, vp, i64, i64, vp, vp, vp, vp, C.POINTER(TcCfg), C.POINTER(vp)]),
        "tc_graph_launch": (C.c_int, [vp, vp]),
        "tc_graph_destroy": (C.c_int, [vp]),
        "tc_autotune_i32": (C.c_int, [vp, i32, i64, vp, vp, vp, C.c_int, C.c_int, C.POINTER(TuneResult), C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(TcCfg)]),
        "tc_autotune_batch_i32": (C.c_int, [vp, i32, vp, i64, i64, vp, vp, vp, C.c_int, C.c_int,
                                            C.POINTER(TuneResult), C.c_int, C.POINTER(C.c_int), C.POINTER(TcCfg)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


# Every symbol include/tc.h declares (checked by the CPU test suite).
EXPORTS = ["tc_ctx_create", "tc_ctx_destroy", "tc_last_error", "tc_scratch_bytes", "tc_ctx_reserve",
           "tc_ctx_set_default_cfg", "tc_ctx_get_default_cfg", "tc_ctx_set_timing", "tc_last_timing",
           "tc_count_i32", "tc_count_noalloc_i32", "tc_kernelCall_i32", "tc_count_i64", "tc_kernelCall_i64", "tc_count_total_i64", "tc_lcc_f64",
           "tc_validate_i32",
           "tc_batch_count_i32",
           "tc_batch_scratch_bytes", "tc_count_total_i32", "tc_count_vertices_i32",
           "tc_count_vertices_lim_i32", "tc_device_cus", "tc_last_edge_work", "tc_oriented_plan", "tc_copy_i32",
           "tc_batch_count_pergraph_i32", "tc_count_capture_i32", "tc_batch_capture_i32", "tc_graph_launch",
           "tc_graph_destroy", "tc_autotune_i32", "tc_autotune_batch_i32", "tc_server_create", "tc_server_destroy",
           "tc_server_submit", "tc_server_wait", "tc_server_cancel", "tc_server_get_stats"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise TCError("no GPU visible: the HIP counter has no CPU fallback", TC_EHIP)
    return torch


def _p(t) -> int | None:
    return None if t is None else int(t.data_ptr())


def _stream(stream):
    if stream is None:
        return C.c_void_p(int(_torch().cuda.current_stream().cuda_stream))
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(int(stream.cuda_stream))


def _cfgp(cfg):
    if cfg is None:
        return None
    if isinstance(cfg, dict):
        cfg = TcCfg.make(**cfg)
    return C.byref(cfg)


@dataclass
class Timing:
    prep_ms: float
    count_ms: float
    finish_ms: float
    total_ms: float


class TriangleCounter:
    """A ``tc_ctx`` bound to one GPU."""

    def __init__(self, device: int = 0):
        torch = _torch()
        self.device = device
        self._torch = torch
        with torch.cuda.device(device):
            h = C.c_void_p()
            rc = lib().tc_ctx_create(device, C.byref(h))
            if rc != TC_OK:
                raise TCError("tc_ctx_create failed", rc)
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != TC_OK:
            raise TCError(f"{what}: {lib().tc_last_error(self._h).decode()}", rc)

    @property
    def cus(self) -> int:
        return int(lib().tc_device_cus(self._h))

    # ------------------------------------------------------------ config
    def default_cfg(self) -> dict:
        c = TcCfg()
        self._check(lib().tc_ctx_get_default_cfg(self._h, C.byref(c)), "get cfg")
        return c.as_dict()

    def set_default_cfg(self, **kw) -> None:
        self._check(lib().tc_ctx_set_default_cfg(self._h, C.byref(TcCfg.make(**kw))), "set cfg")

    def set_timing(self, on: bool = True) -> None:
        self._check(lib().tc_ctx_set_timing(self._h, 1 if on else 0), "timing")

    def last_timing(self) -> Timing:
        ms = (C.c_float * 4)()
        self._check(lib().tc_last_timing(self._h, ms), "timing")
        return Timing(*[float(x) for x in ms])

    def last_edge_work(self):
        """(E, W) of the last edge-scheme count: edges intersected and their total merge length ((0, 0) after an oriented count)."""
        buf = (C.c_int64 * 2)()
        self._check(lib().tc_last_edge_work(self._h, buf), "edge work")
        return int(buf[0]), int(buf[1])

    def oriented_plan(self, nv: int, nnz: int, batch: bool = False, cfg=None) -> dict:
        """The oriented count's resolved launch plan (tc_oriented_plan):
        lookup, partner side, partners per wave, long rows by workgroups."""
        buf = (C.c_int32 * 4)()
        self._check(lib().tc_oriented_plan(self._h, nv, nnz, 1 if batch else 0, _cfgp(cfg), buf), "oriented plan")
        return dict(zip(("lookup", "side", "groups", "long_rows"), (int(x) for x in buf)))

    def reserve(self, nbytes: int) -> None:
        self._check(lib().tc_ctx_reserve(self._h, nbytes), "reserve")

    # ------------------------------------------------------------ device API
    def count(self, off, ind, out=None, total=None, cfg=None, stream=None):
        """Per-vertex counts of a device CSR (int32 torch tensors; an int64
        pair goes to :meth:`count64`).

        Returns ``(out, total)``: ``out`` int32[nv] = 2·T(v), ``total`` int64[1]."""
        torch = self._torch
        if off.dtype == torch.int64 and ind.dtype == torch.int64:
            return self.count64(off, ind, out, total, cfg, stream)
        nv = int(off.shape[0]) - 1
        nnz = int(ind.shape[0])
        if off.dtype != torch.int32 or ind.dtype != torch.int32:
            raise TypeError("off and ind must both be int32 or both int64")
        if out is None:
            out = torch.empty(max(nv, 0), dtype=torch.int32, device=off.device)
        if total is None:
            total = torch.empty(1, dtype=torch.int64, device=off.device)
        rc = lib().tc_count_i32(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, _p(out) if nv else None,
                                _p(total), _cfgp(cfg), _stream(stream))
        self._check(rc, "tc_count_i32")
        return out, total

    def count_noalloc(self, off, ind, out, total=None, cfg=None, stream=None):
        """:meth:`count` through ``tc_count_noalloc_i32``: the library's fixed
        per-GPU scratch array, no device allocation.  ``out`` (and ``total``)
        must be given; graphs beyond the array raise TCError (TC_ERANGE)."""
        nv = int(off.shape[0]) - 1
        nnz = int(ind.shape[0])
        rc = lib().tc_count_noalloc_i32(self._h, nv, nnz, _p(off), _p(ind) if nnz else None,
                                        _p(out) if nv else None, _p(total), _cfgp(cfg), _stream(stream))
        self._check(rc, "tc_count_noalloc_i32")
        return out, total

    def count64(self, off, ind, out=None, total=None, cfg=None, stream=None):
        """Per-vertex counts of an int64 device CSR (tc_count_i64, the
        reference's kernelCall<int64_t>).  Returns ``(out, total)`` with
        ``out`` int64[nv] = 2·T(v)."""
        torch = self._torch
        if off.dtype != torch.int64 or ind.dtype != torch.int64:
            raise TypeError("off and ind must be int64")
        nv = int(off.shape[0]) - 1
        nnz = int(ind.shape[0])
        if out is None:
            out = torch.empty(max(nv, 0), dtype=torch.int64, device=off.device)
        if total is None:
            total = torch.empty(1, dtype=torch.int64, device=off.device)
        rc = lib().tc_count_i64(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, _p(out) if nv else None,
                                _p(total), _cfgp(cfg), _stream(stream))
        self._check(rc, "tc_count_i64")
        return out, total

    def count_total(self, off, ind, total=None, stream=None):
        """Global-count mode: int64[1] = 6·T (degree-ordered orientation)."""
        torch = self._torch
        nv = int(off.shape[0]) - 1
        nnz = int(ind.shape[0])
        if total is None:
            total = torch.empty(1, dtype=torch.int64, device=off.device)
        if off.dtype == torch.int64 and ind.dtype == torch.int64:
            rc = lib().tc_count_total_i64(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, _p(total),
                                          _stream(stream))
            self._check(rc, "tc_count_total_i64")
            return total
        rc = lib().tc_count_total_i32(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, _p(total),
                                      _stream(stream))
        self._check(rc, "tc_count_total_i32")
        return total

    def count_vertices(self, off, ind, vertices, out=None, total=None, cfg=None, stream=None, lim=None):
        """Vertex-list count (the reference's vertex-list kernelCall overload):
        ``out[i]`` = the per-vertex count of ``vertices[i]`` (-1 for an id
        outside the graph), ``total`` int64[1] = their sum.  With ``lim``
        (the reference's shadow-vertex intersection, clusteringCount.cu:150-177)
        a common neighbour x counts 1 + [x > lim]."""
        torch = self._torch
        nv = int(off.shape[0]) - 1
        nnz = int(ind.shape[0])
        nl = int(vertices.shape[0])
        if vertices.dtype != torch.int32:
            raise TypeError("vertices must be int32")
        if out is None:
            out = torch.empty(max(nl, 0), dtype=torch.int32, device=off.device)
        if total is None:
            total = torch.empty(1, dtype=torch.int64, device=off.device)
        if lim is not None:
            rc = lib().tc_count_vertices_lim_i32(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, nl,
                                                 _p(vertices) if nl else None, int(lim), _p(out) if nl else None,
                                                 _p(total), _cfgp(cfg), _stream(stream))
            self._check(rc, "tc_count_vertices_lim_i32")
            return out, total
        rc = lib().tc_count_vertices_i32(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, nl,
                                         _p(vertices) if nl else None, _p(out) if nl else None, _p(total),
                                         _cfgp(cfg), _stream(stream))
        self._check(rc, "tc_count_vertices_i32")
        return out, total

    def lcc(self, off, out, lcc=None, stream=None):
        torch = self._torch
        nv = int(off.shape[0]) - 1
        if lcc is None:
            lcc = torch.empty(max(nv, 0), dtype=torch.float64, device=off.device)
        rc = lib().tc_lcc_f64(self._h, nv, _p(off), _p(out) if nv else None, _p(lcc) if nv else None,
                              _stream(stream))
        self._check(rc, "tc_lcc_f64")
        return lcc

    def validate(self, off, ind, stream=None) -> int:
        """0 if the device CSR is a simple undirected graph, else TC_EGRAPH."""
        st = C.c_int(-99)
        nv = int(off.shape[0]) - 1
        rc = lib().tc_validate_i32(self._h, nv, int(ind.shape[0]), _p(off), _p(ind) if ind.numel() else None,
                                   C.byref(st), _stream(stream))
        self._check(rc, "tc_validate_i32")
        return int(st.value)

    def last_error(self) -> str:
        return lib().tc_last_error(self._h).decode()

    def batch_count(self, tab, off_cat, ind_cat, nv_total: int, nnz_total: int, out=None, totals=None,
                    cfg=None, stream=None):
        torch = self._torch
        G = int(tab.shape[0]) // 2 - 1
        if out is None:
            out = torch.empty(max(nv_total, 0), dtype=torch.int32, device=tab.device)
        if totals is None:
            totals = torch.empty(G, dtype=torch.int64, device=tab.device)
        rc = lib().tc_batch_count_i32(self._h, G, _p(tab), nv_total, nnz_total, _p(off_cat),
                                      _p(ind_cat) if nnz_total else None, _p(out) if nv_total else None,
                                      _p(totals), _cfgp(cfg), _stream(stream))
        self._check(rc, "tc_batch_count_i32")
        return out, totals

    def batch_count_pergraph(self, tab_host, off_cat, ind_cat, out, totals=None, cfg=None, stream=None):
        """One tc_count_i32 per graph (the per-graph-launch baseline)."""
        tab_host = np.ascontiguousarray(tab_host, dtype=np.int64)
        G = tab_host.shape[0] // 2 - 1
        rc = lib().tc_batch_count_pergraph_i32(self._h, G, tab_host.ctypes.data_as(C.POINTER(C.c_int64)),
                                               _p(off_cat), _p(ind_cat), _p(out), _p(totals), _cfgp(cfg),
                                               _stream(stream))
        self._check(rc, "tc_batch_count_pergraph_i32")
        return out, totals

    def capture_count(self, off, ind, out, total, cfg=None) -> CapturedCount:
        h = C.c_void_p()
        nv, nnz = int(off.shape[0]) - 1, int(ind.shape[0])
        rc = lib().tc_count_capture_i32(self._h, nv, nnz, _p(off), _p(ind) if nnz else None, _p(out), _p(total),
                                        _cfgp(cfg), C.byref(h))
        self._check(rc, "tc_count_capture_i32")
        return CapturedCount(h)

    def capture_batch(self, tab, off_cat, ind_cat, nv_total, nnz_total, out, totals, cfg=None) -> CapturedCount:
        h = C.c_void_p()
        G = int(tab.shape[0]) // 2 - 1
        rc = lib().tc_batch_capture_i32(self._h, G, _p(tab), nv_total, nnz_total, _p(off_cat), _p(ind_cat),
                                        _p(out), _p(totals), _cfgp(cfg), C.byref(h))
        self._check(rc, "tc_batch_capture_i32")
        return CapturedCount(h)

    def _tune(self, fn, args, reps, set_default):
        res = (TuneResult * 256)()
        nw = C.c_int(0)
        best = TcCfg()
        rc = fn(self._h, *args, reps, 1 if set_default else 0, res, 256, C.byref(nw), C.byref(best))
        rows = [dict(r.cfg.as_dict(), ms=float(r.ms), identical=bool(r.identical)) for r in res[: nw.value]]
        self._check(rc, "autotune")
        return best.as_dict(), rows

    def autotune(self, off, ind, out, reps=3, set_default=True):
        """Time every candidate configuration; returns (best cfg, table)."""
        nv, nnz = int(off.shape[0]) - 1, int(ind.shape[0])
        return self._tune(lib().tc_autotune_i32, (nv, nnz, _p(off), _p(ind), _p(out)), reps, set_default)

    def autotune_batch(self, tab, off_cat, ind_cat, nv_total, nnz_total, out, reps=3, set_default=True):
        G = int(tab.shape[0]) // 2 - 1
        return self._tune(lib().tc_autotune_batch_i32,
                          (G, _p(tab), nv_total, nnz_total, _p(off_cat), _p(ind_cat), _p(out)), reps, set_default)

    # ------------------------------------------------------------ host convenience
    def to_device(self, g: Graph):
        torch = self._torch
        dev = torch.device("cuda", self.device)
        return (torch.from_numpy(np.ascontiguousarray(g.off)).to(dev),
                torch.from_numpy(np.ascontiguousarray(g.ind)).to(dev))

    def count_graph(self, g: Graph, cfg=None):
        """H2D, count, D2H: ``(out int32[n], total int)`` like singleParamTestGPURun."""
        off, ind = self.to_device(g)
        out, total = self.count(off, ind, cfg=cfg)
        self._torch.cuda.synchronize(self.device)
        return out.cpu().numpy(), int(total.cpu().item())

    def batch_to_device(self, b: Batch):
        torch = self._torch
        dev = torch.device("cuda", self.device)
        return (torch.from_numpy(b.tab).to(dev), torch.from_numpy(b.off).to(dev), torch.from_numpy(b.ind).to(dev))

    def count_batch(self, b: Batch, cfg=None):
        tab, off, ind = self.batch_to_device(b)
        out, totals = self.batch_count(tab, off, ind, b.total_vertices, b.total_arcs, cfg=cfg)
        self._torch.cuda.synchronize(self.device)
        return out.cpu().numpy(), totals.cpu().numpy()


def kernel_call64(numberBlocks: int, numberThreads: int, nv: int, d_off, d_ind, out, threads_per_block: int,
                  number_blocks: int, shifter: int) -> int:
    """``kernelCall<int64_t>`` (clusteringCount.cu:518-524) with the reference argument list."""
    return int(lib().tc_kernelCall_i64(numberBlocks, numberThreads, nv, _p(d_off), _p(d_ind), _p(out),
                                       threads_per_block, number_blocks, shifter))


def kernel_call(numberBlocks: int, numberThreads: int, nv: int, d_off, d_ind, out, threads_per_block: int,
                number_blocks: int, shifter: int) -> int:
    """``kernelCall<int32_t>`` with the reference argument list (clusteringCount.cu:415-467).

    Returns the C return code (0 on success, TC_EINVAL for a launch shape the
    reference would not accept)."""
    return int(lib().tc_kernelCall_i32(numberBlocks, numberThreads, nv, _p(d_off), _p(d_ind), _p(out),
                                       threads_per_block, number_blocks, shifter))
