"""Incremental re-scoring of 2-opt moves (include/gg_incr.h).

The scorer keeps one simple d-regular graph (d <= 32) on the device with
every vertex's distance sum and eccentricity, and scores a move -- two edges
removed, two added between their endpoints -- by re-running the BFS only from
the sources whose distances the move can change.  Every result is the exact
(connected, diameter, S) of the moved graph.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .evaluation import Result, _result


class IncrementalScorer:
    """`IncrementalScorer(adj)`; `score(rem, add)` -> (Result, affected sources);
    then `accept()` or `reject()`."""

    def __init__(self, adj):
        a = L.c32(adj)
        if a.ndim != 2:
            raise ValueError("adjacency must be nodes x degree")
        self.nodes, self.degree = a.shape
        h = L.lib().gg_incr_create(self.nodes, self.degree, L.ptr(a))
        if not h:
            raise L.GGError(L.GG_ERR_INVALID, (L.lib().gg_last_error() or b"").decode())
        self._h = ctypes.c_void_p(h)

    def state(self) -> Result:
        r = L.GGResult()
        L.check(L.lib().gg_incr_state(self._h, ctypes.byref(r)))
        return _result(r, self.nodes)

    def score(self, rem, add):
        rm = np.ascontiguousarray(np.asarray(rem, dtype=np.int32).ravel())
        ad = np.ascontiguousarray(np.asarray(add, dtype=np.int32).ravel())
        if rm.size != 4 or ad.size != 4:
            raise ValueError("rem and add hold two edges each")
        r = L.GGResult()
        aff = ctypes.c_int32(0)
        L.check(L.lib().gg_incr_eval(self._h, L.ptr(rm), L.ptr(ad), ctypes.byref(r),
                                     ctypes.byref(aff)))
        return _result(r, self.nodes), int(aff.value)

    def accept(self) -> None:
        L.check(L.lib().gg_incr_accept(self._h))

    def reject(self) -> None:
        L.check(L.lib().gg_incr_reject(self._h))

    def adjacency(self) -> np.ndarray:
        a = np.zeros((self.nodes, self.degree), dtype=np.int32)
        L.check(L.lib().gg_incr_adjacency(self._h, L.ptr(a)))
        return a

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().gg_incr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
