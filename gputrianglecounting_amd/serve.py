"""ctypes view of the serving loop (``tc_server_*`` in include/tc.h).

A :class:`Server` batches submitted graphs first-in first-out as they
arrive, holds each batch in a device arena of a fixed HBM budget while it
runs, and overlaps batches on several launch slots.  ``submit`` returns at once;
``wait`` blocks for one request's per-vertex counts."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import tc as T
from .graph import Graph, host_lib

_I32P = C.POINTER(C.c_int32)


class ServerCfg(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("max_batch_graphs", C.c_int32),
                ("reserved", C.c_int32), ("hbm_budget", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "rejected", "batches", "bytes_in_use",
                                         "peak_bytes", "waiting", "in_flight", "canceled")]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_bound = False


def _bind():
    global _bound
    if _bound:
        return T.lib()
    L = T.lib()
    vp = C.c_void_p
    L.tc_server_create.restype = C.c_int
    L.tc_server_create.argtypes = [C.POINTER(ServerCfg), C.POINTER(vp)]
    L.tc_server_destroy.restype = C.c_int
    L.tc_server_destroy.argtypes = [vp]
    L.tc_server_submit.restype = C.c_int64
    L.tc_server_submit.argtypes = [vp, C.c_int32, C.c_int64, vp, vp]
    L.tc_server_wait.restype = C.c_int
    L.tc_server_wait.argtypes = [vp, C.c_int64, _I32P, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_double)]
    L.tc_server_cancel.restype = C.c_int
    L.tc_server_cancel.argtypes = [vp, C.c_int64]
    L.tc_server_get_stats.restype = C.c_int
    L.tc_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    _bound = True
    return L


@dataclass
class Result:
    out: np.ndarray
    total: int
    latency_ms: float


class Server:
    def __init__(self, hbm_budget: int = 1 << 30, slots: int = 2, max_batch_graphs: int = 256, device: int = 0):
        import torch

        if not torch.cuda.is_available():
            raise T.TCError("no GPU visible: the serving loop has no CPU fallback", T.TC_EHIP)
        L = _bind()
        cfg = ServerCfg(device, slots, max_batch_graphs, 0, hbm_budget)
        h = C.c_void_p()
        rc = L.tc_server_create(C.byref(cfg), C.byref(h))
        if rc != T.TC_OK:
            raise T.TCError("tc_server_create failed", rc)
        self._h = h
        self._nv = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _bind().tc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, g: Graph) -> int:
        off = np.ascontiguousarray(g.off, np.int32)
        ind = np.ascontiguousarray(g.ind, np.int32)
        rid = _bind().tc_server_submit(self._h, g.n, g.nnz, off.ctypes.data, ind.ctypes.data if g.nnz else None)
        if rid < 0:
            raise T.TCError("tc_server_submit rejected the request", int(rid))
        self._nv[rid] = g.n
        return int(rid)

    def cancel(self, rid: int) -> bool:
        """Cancels a request still waiting for admission (True); False when it
        was admitted already and completes as usual."""
        rc = _bind().tc_server_cancel(self._h, rid)
        if rc < 0:
            raise T.TCError(f"tc_server_cancel({rid}) failed", rc)
        if rc == T.TC_OK:
            del self._nv[rid]
            _bind().tc_server_wait(self._h, rid, None, None, 0, None)  # collect the canceled request
        return rc == T.TC_OK

    def wait(self, rid: int, timeout_ms: int = -1) -> Result | None:
        n = self._nv[rid]
        out = np.empty(max(n, 1), np.int32)
        tot = C.c_int64(0)
        lat = C.c_double(0)
        rc = _bind().tc_server_wait(self._h, rid, out.ctypes.data_as(_I32P), C.byref(tot), timeout_ms, C.byref(lat))
        if rc == 1:
            return None
        del self._nv[rid]
        if rc != T.TC_OK:
            raise T.TCError(f"request {rid} failed", rc)
        return Result(out[:n], int(tot.value), float(lat.value))

    def stats(self) -> dict:
        st = ServerStats()
        _bind().tc_server_get_stats(self._h, C.byref(st))
        return st.as_dict()


def admission_sim(budget: int, script, max_batch: int = 1) -> tuple[np.ndarray, np.ndarray, int, int]:
    """Run the admission policy without a device (tch_admission_sim).

    script[i] > 0 submits a request with that many bytes of batch-block
    share; script[i] < 0 completes request -script[i]-1.  Returns the step at
    which each request was admitted (-1: never), its batch number, the number
    admitted (or a negative TCH_E* code) and the arena's peak bytes."""
    L = host_lib()
    L.tch_admission_sim.restype = C.c_int
    L.tch_admission_sim.argtypes = [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_int64), _I32P, _I32P,
                                    C.POINTER(C.c_int64)]
    sc = np.ascontiguousarray(script, np.int64)
    nreq = int((sc > 0).sum())
    at = np.full(max(nreq, 1), -2, np.int32)
    bo = np.full(max(nreq, 1), -2, np.int32)
    peak = C.c_int64(0)
    n = L.tch_admission_sim(budget, max_batch, sc.shape[0], sc.ctypes.data_as(C.POINTER(C.c_int64)),
                            at.ctypes.data_as(_I32P), bo.ctypes.data_as(_I32P), C.byref(peak))
    return at[:nreq], bo[:nreq], int(n), int(peak.value)


def request_bytes(nv: int, nnz: int) -> int:
    """Arena bytes of one request batched alone."""
    L = host_lib()
    L.tch_request_bytes.restype = C.c_int64
    L.tch_request_bytes.argtypes = [C.c_int64, C.c_int64]
    return int(L.tch_request_bytes(nv, nnz))


def scratch_share(nv: int, nnz: int) -> int:
    """A request's share of its batch's count scratch (csrc/host/admission.h)."""
    return 32 * nv + 4 * nnz + (nnz + 31) // 32


def request_share(nv: int, nnz: int) -> int:
    """A request's share of a batch block: one table row, nv+1 offsets, nnz
    ids, nv counts, one total, and its share of the count's scratch."""
    return 28 + 8 * nv + 4 * nnz + scratch_share(nv, nnz)


BATCH_FIXED = 7 * 256  # a block's table row and region rounding (the admission policy's fixed part)
SCRATCH_FIXED = 1060 * 256  # the scratch's fixed part (piece-list slack, counters, rounding)
BLOCK_FIXED = BATCH_FIXED + SCRATCH_FIXED  # what a serving block adds to its requests' shares


def batch_block_bytes(G: int, nv_total: int, nnz_total: int) -> int:
    L = host_lib()
    L.tch_batch_block_bytes.restype = C.c_int64
    L.tch_batch_block_bytes.argtypes = [C.c_int32, C.c_int64, C.c_int64]
    return int(L.tch_batch_block_bytes(G, nv_total, nnz_total))
