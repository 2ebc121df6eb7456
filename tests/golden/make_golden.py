"""Write the golden vectors under tests/golden/ from the reference's own CPU
code (oracle/_ref/libtcref.so, compiled from the reference sources by
oracle/Makefile).  Run in a container where the reference tree exists:

    make -C oracle && python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  arcs_<case>.npz   small CSRs, some deliberately not simple (self-loops,
                    duplicate arcs, missing reverse arcs), with the per-arc
                    counts triNE and total of count_all_trianglesCPU<int32_t>
                    (clusteringCount.cpp:50-70)
  vertex_<case>.npz seeded generator graphs: a SHA-256 of their CSR, the
                    per-vertex row sums of triNE (int32) and the total
  metis_sample.graph / metis_sample.npz   a METIS file and the CSR the
                    reference readGraph (csrdatacreate.cpp:3-49) makes of it
  info_sample.info  the bytes writeGraph (csrdatacreate.cpp:86-114) writes
                    for the K4 + pendant graph stored in info_sample.npz
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gputrianglecounting_amd import graph as G  # noqa: E402


def csr_hash(off, ind) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, np.int32).tobytes())
    h.update(np.ascontiguousarray(ind, np.int32).tobytes())
    return h.hexdigest()


def raw_csr(rows):
    off = np.zeros(len(rows) + 1, np.int32)
    for i, r in enumerate(rows):
        off[i + 1] = off[i] + len(r)
    ind = np.array([x for r in rows for x in r], np.int32)
    return off, ind


def arc_cases():
    rng = np.random.default_rng(2024)
    cases = {}
    cases["k4"] = raw_csr([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    cases["selfloops"] = raw_csr([[0, 1, 2], [0, 1, 2], [0, 1, 2, 3], [2, 3]])
    cases["duplicates"] = raw_csr([[1, 1, 2], [0, 0, 2, 2], [0, 1, 1, 3], [2]])
    cases["directed"] = raw_csr([[1, 2], [2], [0, 3], [1], []])
    cases["isolated"] = raw_csr([[], [2], [1], []])
    # random rows, sorted, with self-loops and duplicates allowed
    for t in range(4):
        n = int(rng.integers(20, 60))
        rows = []
        for v in range(n):
            k = int(rng.integers(0, 9))
            rows.append(sorted(rng.integers(0, n, size=k).tolist()))
        cases[f"rand{t}"] = raw_csr(rows)
    return cases


def main() -> None:
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libtcref.so is not built (needs the reference tree)")
    for name, (off, ind) in arc_cases().items():
        tri, tot = oracle.ref_count_arcs(off, ind)
        np.savez(os.path.join(HERE, f"arcs_{name}.npz"), off=off, ind=ind, tri=tri, total=np.int64(tot))
    gens = {
        "er_n16384_m131072_s1": lambda: G.erdos_renyi(16384, 131072, seed=1),
        "rmat_s12_ef16_s1": lambda: G.rmat(12, 16, seed=1),
        "rgg_n20000_d6_s1": lambda: G.geometric(20000, 6.0, seed=1),
    }
    for name, fn in gens.items():
        g = fn()
        tri, tot = oracle.ref_count_arcs(g.off, g.ind)
        per = oracle.row_sums(g.off, tri).astype(np.int32)
        np.savez(os.path.join(HERE, f"vertex_{name}.npz"), sha256=np.array(csr_hash(g.off, g.ind)), n=np.int64(g.n),
                 nnz=np.int64(g.nnz), out=per, total=np.int64(tot))
    # METIS sample: comment lines, an empty row, unsorted rows
    metis = "% sample\n% two comments\n6 7\n2 3\n1 3 4\n2 1\n2 5 6\n4\n4\n"
    mpath = os.path.join(HERE, "metis_sample.graph")
    with open(mpath, "w") as f:
        f.write(metis)
    L = oracle.ref_lib()
    nv, nnz = C.c_int32(0), C.c_int64(0)
    L.tcref_read_metis(mpath.encode(), C.byref(nv), C.byref(nnz), None, 0, None, 0)
    o = np.zeros(nv.value + 1, np.int32)
    i = np.zeros(max(nnz.value, 1), np.int32)
    rc = L.tcref_read_metis(mpath.encode(), C.byref(nv), C.byref(nnz), oracle._p32(o), o.shape[0], oracle._p32(i),
                            i.shape[0])
    assert rc == 0
    np.savez(os.path.join(HERE, "metis_sample.npz"), off=o, ind=i[: nnz.value])
    # .info writer bytes
    off, ind = raw_csr([[1, 2, 3], [0, 2, 3], [0, 1, 3, 4], [0, 1, 2], [2]])
    ipath = os.path.join(HERE, "info_sample.info")
    L.tcref_write_info(ipath.encode(), off.shape[0] - 1, oracle._p32(off), oracle._p32(ind))
    np.savez(os.path.join(HERE, "info_sample.npz"), off=off, ind=ind)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
