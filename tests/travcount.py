"""ctypes wrapper of tests/cpp/trav_count.cpp: the instrumented host walk of
the product's flattened BVH (test infrastructure)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "trav_count.cpp")
LIB = os.path.join(ROOT, "tests", "cpp", "_build", "libtrav_count.so")
_lib = None


def build():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                    "-pthread", "-o", tmp, SRC, os.path.join(ROOT, "xitils_amd", "csrc", "bvh.cpp")], check=True)
    os.replace(tmp, LIB)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing (run __graft_entry__.build())")
        L = C.CDLL(LIB)
        L.tc_build.restype = C.c_void_p
        L.tc_build.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.tc_free.argtypes = [C.c_void_p]
        L.tc_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p]
        _lib = L
    return _lib


def world_triangles(desc) -> np.ndarray:
    """(n, 9) float32 world-space triangles in global prim order (object
    order, then triangle order), for scenes without instance transforms."""
    out = []
    for o in desc.objects:
        if o.xf is not None:
            raise ValueError("the host walk takes scenes without instance transforms")
        out.append(np.asarray(o.positions, np.float32)[np.asarray(o.indices)].reshape(-1, 9))
    return np.ascontiguousarray(np.concatenate(out), np.float32)


class HostBVH:
    def __init__(self, desc):
        self.tris = world_triangles(desc)
        info = np.zeros(3, np.uint32)
        self.h = C.c_void_p(lib().tc_build(self.tris.ctypes.data, self.tris.shape[0], info.ctypes.data))
        self.nodes, self.depth, self.top_nodes = (int(v) for v in info)

    def __del__(self):
        try:
            lib().tc_free(self.h)
        except Exception:
            pass

    def trace(self, rays, any_hit=False, per_ray=False):
        """-> (hits (n, 4) or occluded (n,), {"nodes": .., "tris": ..}, per-ray counts or None)"""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        occ = np.zeros(n, np.uint8)
        counts = np.zeros(2, np.uint64)
        pr = np.zeros((n, 2), np.uint32) if per_ray else None
        lib().tc_trace(self.h, rays.ctypes.data, n, int(any_hit), hits.ctypes.data, occ.ctypes.data,
                       counts.ctypes.data, None if pr is None else pr.ctypes.data)
        return (occ if any_hit else hits), {"nodes": int(counts[0]), "tris": int(counts[1])}, pr
