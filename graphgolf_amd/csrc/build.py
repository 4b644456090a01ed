"""Builds the evaluator libraries in-tree (gfx950).

    python -m graphgolf_amd.csrc.build        # or __graft_entry__.build()

Outputs, next to the package so they travel with the tree:
    graphgolf_amd/libggeval.so     kernels + engine + host SA/graph code (C ABI: include/gg_eval.h)
    graphgolf_amd/libgg_general.so reference ABI shim: general evaluation()  (include/gg_general.h)
    graphgolf_amd/libgg_grid.so    reference ABI shim: grid evaluation()     (include/gg_grid.h)
    graphgolf_amd/libgg_nonregular_{general,grid}.so  non-regular variants' evaluation()
                                   (include/gg_nonregular.h)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GG_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instruction forms (v_pk_*_f32) are disabled for all device code:
# the compiler's float paths (the MLP) then use scalar FP32 instructions.
# hipcc's host pass warns that it does not know the feature.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INC, "-I", HERE,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-function"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-I", INC, "-I", HERE, "-Wall",
              "-march=x86-64-v3"]

HIP_SOURCES = ["gg_eval.hip", "gg_wide.hip", "gg_mlp.hip", "gg_bsa.hip", "gg_incr.hip"]
HOST_SOURCES = ["gg_host.cpp", "gg_sa_sym.cpp", "gg_layout.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r.stdout


def build(verbose: bool = False) -> None:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(HERE, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
    for src in HOST_SOURCES:
        path = os.path.join(HERE, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        jobs.append(["g++", *HOST_FLAGS, "-c", path, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            sys.stdout.write(o)
    core = os.path.join(PKG, "libggeval.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", core + ".tmp", *objs])
    os.replace(core + ".tmp", core)
    for shim, lib, defs in (("gg_general_shim.cpp", "libgg_general.so", []),
                            ("gg_grid_shim.cpp", "libgg_grid.so", []),
                            ("gg_nonregular_shim.cpp", "libgg_nonregular_general.so", []),
                            ("gg_nonregular_shim.cpp", "libgg_nonregular_grid.so",
                             ["-DGG_NONREGULAR_GRID"])):
        out = os.path.join(PKG, lib)
        _run(["g++", *HOST_FLAGS, *defs, "-shared", "-o", out + ".tmp", os.path.join(HERE, shim),
              "-L", PKG, "-lggeval", "-Wl,-rpath,$ORIGIN"])
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
