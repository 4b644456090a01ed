"""Seeded graph construction and the reference's graph file formats.

Edge lists are int32 arrays [lines, 2]; adjacency tables int32 [nodes, degree]
in edge order (general/sa.c:61-74).  Generators run in libggeval.so (host
C++), so the same seed gives the same graph everywhere.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def random_regular(nodes: int, degree: int, seed: int) -> np.ndarray:
    """Connected random d-regular simple graph (configuration-model pairing,
    self-loops and multi-edges rejected)."""
    e = np.zeros((nodes * degree // 2, 2), dtype=np.int32)
    L.check(L.lib().gg_random_regular(nodes, degree, seed, L.ptr(e)))
    return e


def random_symmetric(nodes: int, degree: int, groups: int, seed: int) -> np.ndarray:
    """Connected random d-regular graph invariant under v -> v + nodes/groups (mod nodes)."""
    e = np.zeros((nodes * degree // 2, 2), dtype=np.int32)
    L.check(L.lib().gg_random_symmetric(nodes, degree, groups, seed, L.ptr(e)))
    return e


def random_grid(width: int, height: int, degree: int, length: int, seed: int) -> np.ndarray:
    """Connected random grid graph: vertex v = x*height + y, Manhattan edge length <= length."""
    n = width * height
    e = np.zeros((n * degree // 2, 2), dtype=np.int32)
    L.check(L.lib().gg_random_grid(width, height, degree, length, seed, L.ptr(e)))
    return e


def random_grid_symmetric(width: int, height: int, degree: int, length: int, groups: int,
                          seed: int) -> np.ndarray:
    """groups-fold rotationally symmetric grid graph (create_symmetric_edge,
    grid/main.c:393-445) grown from a seeded base lattice."""
    n = width * height
    e = np.zeros((n * degree // 2, 2), dtype=np.int32)
    L.check(L.lib().gg_random_grid_symmetric(width, height, degree, length, groups, seed,
                                             L.ptr(e)))
    return e


def grid_symmetric_from_base(width: int, height: int, degree: int, length: int, groups: int,
                             seed: int, base_edges) -> np.ndarray:
    """create_symmetric_edge (grid/main.c:393-445) from a base-grid edge list
    (vertex v at (v // base_height, v % base_height))."""
    n = width * height
    e = np.zeros((n * degree // 2, 2), dtype=np.int32)
    b = L.c32(base_edges).reshape(-1, 2)
    if b.shape[0] * groups != e.shape[0]:
        raise ValueError(f"{b.shape[0]} base lines x {groups} groups != {e.shape[0]} lines")
    L.check(L.lib().gg_grid_symmetric_from_base(width, height, degree, length, groups, seed,
                                                L.ptr(b), L.ptr(e)))
    return e


def grid_lattice(width: int, height: int, degree: int, length: int, seed: int) -> np.ndarray:
    """The reference's random lattice (create_lattice, grid/main.c:232-330)
    after srandom(seed)."""
    e = np.zeros((width * height * degree // 2, 2), dtype=np.int32)
    L.check(L.lib().gg_grid_lattice(width, height, degree, length, seed, L.ptr(e)))
    return e


def grid_lower_bound(width: int, height: int, degree: int, length: int):
    """(diameter, ASPL) lower bounds of a grid graph (grid/main.c:480-529)."""
    d = ctypes.c_int(0)
    a = ctypes.c_double(0)
    L.check(L.lib().gg_grid_lower_bound(width, height, degree, length, ctypes.byref(d),
                                        ctypes.byref(a)))
    return d.value, a.value


def edges_to_adj(nodes: int, degree: int, edges) -> np.ndarray:
    e = L.c32(edges).reshape(-1, 2)
    adj = np.zeros((nodes, degree), dtype=np.int32)
    L.check(L.lib().gg_edges_to_adj(nodes, e.shape[0], degree, L.ptr(e), L.ptr(adj)))
    return adj


def max_degree(nodes: int, edges) -> int:
    return int(np.bincount(np.asarray(edges).ravel(), minlength=nodes).max())


# ---- Reference file formats -----------------------------------------------

def read_edges(path: str) -> np.ndarray:
    """General edge file: one "a b" pair per line (general/main.c:112-126)."""
    with open(path) as f:
        vals = [int(t) for t in f.read().split()]
    return np.asarray(vals, dtype=np.int32).reshape(-1, 2)


def write_edges(path: str, edges) -> None:
    """general/main.c:311-315."""
    with open(path, "w") as f:
        for a, b in np.asarray(edges).reshape(-1, 2):
            f.write(f"{int(a)} {int(b)}\n")


def read_grid_edges(path: str):
    """Grid edge file: "x,y x,y" per line (grid/main.c:353-382).  Returns
    (edges with v = x*height + y, width, height)."""
    pts = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            a, b = line.split()
            x0, y0 = (int(t) for t in a.split(","))
            x1, y1 = (int(t) for t in b.split(","))
            pts.append((x0, y0, x1, y1))
    arr = np.asarray(pts, dtype=np.int64)
    width = int(max(arr[:, 0].max(), arr[:, 2].max())) + 1
    height = int(max(arr[:, 1].max(), arr[:, 3].max())) + 1
    e = np.stack([arr[:, 0] * height + arr[:, 1], arr[:, 2] * height + arr[:, 3]], axis=1)
    return e.astype(np.int32), width, height


def write_grid_edges(path: str, edges, height: int) -> None:
    """grid/common.c:163-168 / grid/main.c:579-584."""
    with open(path, "w") as f:
        for a, b in np.asarray(edges).reshape(-1, 2):
            f.write(f"{int(a) // height},{int(a) % height} {int(b) // height},{int(b) % height}\n")


# ---- Grid symmetry helpers (grid/common.c:63-86, :312-341) -----------------

def rotate(v: int, height: int, width: int, groups: int, degree: int) -> int:
    w, h = v // height, v % height
    if groups == 2:
        return (width - w - 1) * height + (height - h - 1)
    if degree == 90:
        return h * height + (height - w - 1)
    if degree == 180:
        return (height - w - 1) * height + (height - h - 1)
    return (height - h - 1) * height + w


def rotate_hash(nodes: int, height: int, width: int, groups: int) -> np.ndarray:
    """Permutation mapping the base quadrant to 0..B-1 and its rotated
    images to the following blocks."""
    B = nodes // groups
    h = np.zeros(nodes, dtype=np.int32)
    if groups == 1:
        h[:] = np.arange(nodes)
        return h
    bh = height // 2
    for i in range(B):
        j = (i // bh) * height + (i % bh)
        h[j] = i
        if groups == 2:
            h[rotate(j, height, width, 2, 180)] = i + B
        else:
            for k, deg in enumerate((90, 180, 270), start=1):
                h[rotate(j, height, width, 4, deg)] = i + k * B
    return h


def moore_bound(nodes: int, degree: int):
    """Lower bounds of diameter and ASPL (general/main.c:244-265)."""
    diam, n, r = -1, 1, 1
    aspl = 0.0
    while True:
        tmp = n + degree * (degree - 1) ** (r - 1)
        if tmp >= nodes:
            break
        n = tmp
        aspl += r * degree * (degree - 1) ** (r - 1)
        diam = r
        r += 1
    diam += 1
    aspl += diam * (nodes - n)
    aspl /= (nodes - 1)
    return diam, aspl
