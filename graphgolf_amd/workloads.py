"""Seeded workload generators shared by the benchmark and the tests.

Random 2-opt moves of general graphs (the edge swap of general/sa.c:217-369,
uniform lines) and of grid graphs (the length-limited swap of
grid/sa.c:137-221), and the shapes of the uneven batch of 512 graphs.
"""
from __future__ import annotations

import numpy as np

from . import graphs as G


def two_opt(edges: np.ndarray, nodes: int, rng: np.random.Generator, moves: int = 1) -> np.ndarray:
    """Random 2-opt swaps (a,b),(c,d) -> (a,d),(c,b) keeping the graph simple."""
    e = edges.copy()
    present = set(map(tuple, np.sort(e, axis=1).tolist()))
    done = 0
    while done < moves:
        i, j = rng.integers(0, len(e), 2)
        if i == j:
            continue
        a, b = e[i]
        c, d = e[j]
        if len({a, b, c, d}) < 4:
            continue
        n1, n2 = tuple(sorted((a, d))), tuple(sorted((c, b)))
        if n1 in present or n2 in present:
            continue
        present.discard(tuple(sorted((a, b))))
        present.discard(tuple(sorted((c, d))))
        present.add(n1)
        present.add(n2)
        e[i] = (a, d)
        e[j] = (c, b)
        done += 1
    return e


def grid_two_opt(edges: np.ndarray, height: int, length: int, rng: np.random.Generator,
                 moves: int = 1) -> np.ndarray:
    """Random 2-opt swaps of a grid graph keeping it simple and every edge
    within `length` Manhattan steps (vertex v at (v // height, v % height)).
    The second line is drawn near the first so that most draws are valid."""
    e = edges.copy()
    present = set(map(tuple, np.sort(e, axis=1).tolist()))
    by_vertex = {}
    for k, (u, v) in enumerate(e.tolist()):
        by_vertex.setdefault(u, []).append(k)
        by_vertex.setdefault(v, []).append(k)

    def dist(u, v):
        return abs(u // height - v // height) + abs(u % height - v % height)

    width = len(by_vertex) // height
    ball = [(dx, dy) for dx in range(-length, length + 1) for dy in range(-length, length + 1)
            if 0 < abs(dx) + abs(dy) <= length]
    done = 0
    while done < moves:
        i = int(rng.integers(0, len(e)))
        a, b = (int(x) for x in e[i])
        dx, dy = ball[int(rng.integers(0, len(ball)))]
        cx, cy = a // height + dx, a % height + dy
        if not (0 <= cx < width and 0 <= cy < height):
            continue
        c0 = cx * height + cy
        j = by_vertex[c0][int(rng.integers(0, len(by_vertex[c0])))]
        if i == j:
            continue
        c, d = (int(x) for x in e[j])
        if len({a, b, c, d}) < 4:
            continue
        n1, n2 = ((a, d), (c, b)) if rng.integers(0, 2) == 0 else ((a, c), (b, d))
        if dist(*n1) > length or dist(*n2) > length:
            continue
        k1, k2 = tuple(sorted(n1)), tuple(sorted(n2))
        if k1 in present or k2 in present:
            continue
        present -= {tuple(sorted((a, b))), tuple(sorted((c, d)))}
        present |= {k1, k2}
        for k, old, new in ((i, (a, b), n1), (j, (c, d), n2)):
            for x in old:
                by_vertex[x].remove(k)
            for x in new:
                by_vertex[x].append(k)
            e[k] = new
        done += 1
    return e


def grid_two_opt_variants(edges: np.ndarray, height: int, length: int,
                          rng: np.random.Generator, count: int) -> list:
    """`count` copies of a grid graph, each one valid random 2-opt away from it
    (same move rules as grid_two_opt)."""
    present = set(map(tuple, np.sort(edges, axis=1).tolist()))
    by_vertex = {}
    for k, (u, v) in enumerate(edges.tolist()):
        by_vertex.setdefault(u, []).append(k)
        by_vertex.setdefault(v, []).append(k)
    width = len(by_vertex) // height
    ball = [(dx, dy) for dx in range(-length, length + 1) for dy in range(-length, length + 1)
            if 0 < abs(dx) + abs(dy) <= length]

    def dist(u, v):
        return abs(u // height - v // height) + abs(u % height - v % height)

    out = []
    while len(out) < count:
        i = int(rng.integers(0, len(edges)))
        a, b = (int(x) for x in edges[i])
        dx, dy = ball[int(rng.integers(0, len(ball)))]
        cx, cy = a // height + dx, a % height + dy
        if not (0 <= cx < width and 0 <= cy < height):
            continue
        c0 = cx * height + cy
        j = by_vertex[c0][int(rng.integers(0, len(by_vertex[c0])))]
        c, d = (int(x) for x in edges[j])
        if i == j or len({a, b, c, d}) < 4:
            continue
        n1, n2 = ((a, d), (c, b)) if rng.integers(0, 2) == 0 else ((a, c), (b, d))
        if dist(*n1) > length or dist(*n2) > length:
            continue
        if tuple(sorted(n1)) in present or tuple(sorted(n2)) in present:
            continue
        e = edges.copy()
        e[i] = n1
        e[j] = n2
        out.append(e)
    return out


def c3_shapes(count: int, seed: int):
    """(nodes, degree, seed) triples of the uneven batch: n ~ U[1024,16384],
    d in {8,12,16,24,32}, n*d even."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        d = int(rng.choice([8, 12, 16, 24, 32]))
        n = int(rng.integers(1024, 16385))
        if (n * d) % 2:
            n += 1
        out.append((n, d, 1000 + k))
    return out


def c3_batch(count: int = 512, seed: int = 7) -> list:
    """The uneven batch: `count` graphs n ~ U[1024,16384], d in
    {8,12,16,24,32}, each a seeded random regular graph after one 2-opt.
    Returns adjacency tables."""
    out = []
    for (n, d, s) in c3_shapes(count, seed):
        e = two_opt(G.random_regular(n, d, s), n, np.random.default_rng(s), 1)
        out.append(G.edges_to_adj(n, d, e))
    return out


def grid_batch(width: int = 100, height: int = 100, degree: int = 4, length: int = 2,
               seed: int = 1, count: int = 256, move_seed: int = 4):
    """A seeded random grid graph and `count` length-respecting 2-opt
    proposals of it.  Returns (base edges, list of proposal edge lists)."""
    e0 = G.random_grid(width, height, degree, length, seed)
    return e0, grid_two_opt_variants(e0, height, length, np.random.default_rng(move_seed), count)
