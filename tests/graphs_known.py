"""Graphs with closed-form diameter and distance sum S (ordered pairs)."""
from __future__ import annotations

import itertools

import numpy as np


def complete(n):
    adj = np.array([[u for u in range(n) if u != v] for v in range(n)], dtype=np.int32)
    return adj, 1, n * (n - 1)


def cycle(n):
    adj = np.array([[(v - 1) % n, (v + 1) % n] for v in range(n)], dtype=np.int32)
    return adj, n // 2, n * ((n * n) // 4)


def hypercube(k):
    n = 1 << k
    adj = np.array([[v ^ (1 << b) for b in range(k)] for v in range(n)], dtype=np.int32)
    return adj, k, n * k * (1 << (k - 1))


def petersen():
    outer = [(i, (i + 1) % 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    adj = [[] for _ in range(10)]
    for a, b in outer + inner + spokes:
        adj[a].append(b)
        adj[b].append(a)
    # diameter 2, ASPL 5/3 -> S = 10*9*5/3 = 150
    return np.array(adj, dtype=np.int32), 2, 150


def hoffman_singleton():
    """Robertson's pentagon/pentagram construction: n=50, d=7, diameter 2,
    ASPL 91/49 (a Moore graph)."""
    adj = [[] for _ in range(50)]

    def P(h, i):
        return 10 * h + i

    def Q(j, i):
        return 10 * j + 5 + i

    def add(a, b):
        adj[a].append(b)
        adj[b].append(a)

    for h in range(5):
        for i in range(5):
            add(P(h, i), P(h, (i + 1) % 5))
            add(Q(h, i), Q(h, (i + 2) % 5))
    for h, j, i in itertools.product(range(5), range(5), range(5)):
        add(P(h, i), Q(j, (h * j + i) % 5))
    return np.array(adj, dtype=np.int32), 2, 50 * 49 * 91 // 49


def torus(k):
    """k x k torus (4-regular for k>=3): diameter 2*(k//2)."""
    n = k * k
    adj = []
    for x in range(k):
        for y in range(k):
            adj.append([((x + 1) % k) * k + y, ((x - 1) % k) * k + y,
                        x * k + (y + 1) % k, x * k + (y - 1) % k])
    # per-axis distance sum on a cycle C_k is k*floor(k^2/4)/k per vertex
    per_axis = (k * k) // 4  # sum of cycle distances from one vertex
    S = n * (2 * k * per_axis)
    return np.array(adj, dtype=np.int32), 2 * (k // 2), S


def mesh(p, q):
    """p x q mesh padded to degree 4 with self-loops: diameter p+q-2."""
    n = p * q
    adj = []
    for x in range(p):
        for y in range(q):
            nb = []
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                X, Y = x + dx, y + dy
                if 0 <= X < p and 0 <= Y < q:
                    nb.append(X * q + Y)
            nb += [x * q + y] * (4 - len(nb))
            adj.append(nb)
    # path distance sums: sum over pairs |x-x'| + |y-y'|
    sx = sum(abs(a - b) for a in range(p) for b in range(p))
    sy = sum(abs(a - b) for a in range(q) for b in range(q))
    S = sx * q * q + sy * p * p
    return np.array(adj, dtype=np.int32), p + q - 2, S


def all_known():
    out = [("K8", *complete(8)), ("K33", *complete(33)), ("C9", *cycle(9)), ("C64", *cycle(64)),
           ("C101", *cycle(101)), ("Q3", *hypercube(3)), ("Q7", *hypercube(7)),
           ("Q10", *hypercube(10)), ("petersen", *petersen()), ("HS", *hoffman_singleton()),
           ("T7", *torus(7)), ("T16", *torus(16)), ("M3x3", *mesh(3, 3)), ("M5x9", *mesh(5, 9)),
           ("M20x20", *mesh(20, 20))]
    return out
