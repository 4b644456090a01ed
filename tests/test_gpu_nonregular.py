"""The non-regular variants' drop-in shims (libgg_nonregular_{general,grid}.so,
include/gg_nonregular.h) on the GPU against the variants' own golden
vectors, called through the C ABI exactly as a variant's sa.c would."""
import csv
import ctypes

import numpy as np
import pytest

from cases import GOLDEN_NONREGULAR, build_nonregular, nonregular_cases
from graphgolf_amd import _lib as L
from graphgolf_amd import evaluation as E

pytestmark = pytest.mark.gpu

with open(GOLDEN_NONREGULAR) as _f:
    GOLD = {r["id"]: r for r in csv.DictReader(_f)}

_libs = {}


def shim(kind):
    if kind not in _libs:
        ip = ctypes.POINTER(ctypes.c_int)
        if kind == "general":
            so = ctypes.CDLL(L.NONREGULAR_GENERAL_PATH)
            so.evaluation.argtypes = [ctypes.c_int] * 5 + [ip, ip, ctypes.POINTER(ctypes.c_double),
                                                           ctypes.c_int, ip, ctypes.c_int]
        else:
            so = ctypes.CDLL(L.NONREGULAR_GRID_PATH)
            so.evaluation.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ip,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_bool, ip]
        so.evaluation.restype = ctypes.c_bool
        _libs[kind] = so
    return _libs[kind]


@pytest.mark.parametrize("case", nonregular_cases(), ids=lambda c: c["id"])
def test_gpu_nonregular_shim_matches_reference(case):
    row = GOLD[case["id"]]
    adj, lens, kw = build_nonregular(case)
    n, w = adj.shape
    ip = ctypes.POINTER(ctypes.c_int)
    a = np.ascontiguousarray(adj, dtype=np.int32)
    ln = np.ascontiguousarray(lens, dtype=np.int32)
    diam = ctypes.c_int(-7)
    aspl = ctypes.c_double(-7.0)
    if case["kind"] == "nr_general":
        ok = shim("general").evaluation(n, kw["based_nodes"], kw["groups"], int(ln.sum()) // 2, w,
                                        a.ctypes.data_as(ip), ctypes.byref(diam),
                                        ctypes.byref(aspl), kw["added_centers"],
                                        ln.ctypes.data_as(ip), case["algo"])
        mirror = E.evaluation_nonregular(n, kw["based_nodes"], kw["groups"], int(ln.sum()) // 2, w,
                                         a, ln, kw["added_centers"], case["algo"])
    else:
        rh = np.ascontiguousarray(kw["rotate_hash"], dtype=np.int32)
        ok = shim("grid").evaluation(n, w, ln.ctypes.data_as(ip), kw["groups"], a.ctypes.data_as(ip),
                                     kw["based_nodes"], kw["height"], kw["based_height"],
                                     ctypes.byref(diam), ctypes.byref(aspl), case["bfs"],
                                     rh.ctypes.data_as(ip))
        mirror = E.grid_evaluation_nonregular(n, w, ln, kw["groups"], a, kw["based_nodes"],
                                              kw["height"], kw["based_height"], case["bfs"], rh)
    assert int(ok) == int(row["ok"]) == int(mirror[0])
    if ok:
        assert diam.value == int(row["diameter"]) == mirror[1]
        assert aspl.value == float.fromhex(row["aspl_hex"]) == mirror[2]
    elif case["kind"] == "nr_grid":
        assert diam.value == int(row["diameter"]) == mirror[1]
        assert aspl.value == -7.0  # not written, as in the reference
    else:
        assert diam.value == 2147483647 and mirror[1] == 2147483647
