"""Graph cases shared by the CPU and GPU tests (small closed forms and seeded
random graphs from the host generators)."""
from __future__ import annotations

import numpy as np

from gputrianglecounting_amd import graph as G


def complete(n):
    s, d = np.triu_indices(n, 1)
    return G.from_edges(n, s, d, f"K{n}")


def cycle(n):
    s = np.arange(n)
    return G.from_edges(n, s, (s + 1) % n, f"C{n}")


def wheel(k):
    """Hub 0 joined to a k-cycle 1..k."""
    rim = np.arange(1, k + 1)
    s = np.concatenate([np.zeros(k, int), rim])
    d = np.concatenate([rim, 1 + (rim % k)])
    return G.from_edges(k + 1, s, d, f"W{k}")


def friendship(k):
    """k triangles sharing vertex 0."""
    a = 1 + 2 * np.arange(k)
    s = np.concatenate([np.zeros(k, int), np.zeros(k, int), a])
    d = np.concatenate([a, a + 1, a + 1])
    return G.from_edges(2 * k + 1, s, d, f"F{k}")


def complete_bipartite(a, b):
    s, d = np.meshgrid(np.arange(a), a + np.arange(b), indexing="ij")
    return G.from_edges(a + b, s.ravel(), d.ravel(), f"K{a},{b}")


def star(k):
    return G.from_edges(k + 1, np.zeros(k, int), np.arange(1, k + 1), f"S{k}")


def petersen():
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    e = np.array(outer + spokes + inner)
    return G.from_edges(10, e[:, 0], e[:, 1], "petersen")


def grid(r, c):
    idx = np.arange(r * c).reshape(r, c)
    s = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    d = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    return G.from_edges(r * c, s, d, f"grid{r}x{c}")


def two_hubs(m):
    """Vertices 0 and 1 adjacent to each other and to m shared leaves 2..m+1:
    one edge (0,1) with |N(0)|+|N(1)| = 2m+2 and m triangles."""
    leaves = np.arange(2, m + 2)
    s = np.concatenate([[0], np.zeros(m, int), np.ones(m, int)])
    d = np.concatenate([[1], leaves, leaves])
    return G.from_edges(m + 2, s, d, f"twohubs{m}")


def empty(n):
    return G.from_csr(np.zeros(n + 1, np.int32), np.zeros(0, np.int32), f"empty{n}")


def small_cases():
    return [complete(3), complete(4), complete(12), complete(64), cycle(3), cycle(4), cycle(17), wheel(4),
            wheel(9), friendship(1), friendship(7), complete_bipartite(3, 5), star(10), petersen(), grid(6, 7),
            two_hubs(50), empty(1), empty(5)]


def random_cases():
    return [G.erdos_renyi(500, 4000, seed=3), G.erdos_renyi(16384, 131072, seed=1), G.geometric(20000, 6.0, seed=2),
            G.rmat(10, 16, seed=5), G.rmat(12, 16, seed=9), G.rmat(14, 8, seed=11), two_hubs(9000),
            complete(300)]
