"""The C-ABI libraries load on the CPU and export every symbol their
headers declare (no compute calls: there is no GPU here)."""
import ctypes as C
import os
import re

import pytest

from gputrianglecounting_amd import graph as G
from gputrianglecounting_amd import tc as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tch?_[a-zA-Z0-9_]+)\s*\(", text)))


def test_libtc_exports_everything_declared():
    L = T.lib()
    names = _declared("tc.h")
    assert "tc_count_i32" in names and "tc_kernelCall_i32" in names
    for n in names:
        assert hasattr(L, n), f"libtc.so lacks {n}"
    assert sorted(T.EXPORTS) == names


def test_libtc_host_exports_everything_declared():
    L = G.host_lib()
    names = _declared("tc_host.h")
    assert "tch_read_info" in names
    for n in names:
        assert hasattr(L, n), f"libtc_host.so lacks {n}"


def test_host_only_entry_points():
    L = T.lib()
    # sizing needs no device
    assert L.tc_scratch_bytes(1000, 20000) > 20000 * 4
    assert L.tc_batch_scratch_bytes(4, 1000, 20000) >= L.tc_scratch_bytes(1000, 20000)
    # argument checks happen before any HIP call
    assert L.tc_count_i32(None, 0, 0, None, None, None, None, None, None) == T.TC_EINVAL
    assert L.tc_kernelCall_i32(1, 100, 0, None, None, None, 1, 100, 0) == T.TC_EINVAL
    assert L.tc_last_error(None) == b"null context"


def test_no_cpu_fallback_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(T.TCError):
        T.TriangleCounter(0)
