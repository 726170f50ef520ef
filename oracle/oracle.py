"""ctypes wrapper of liboracle.so, the CPU restatement (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; the product path never does.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int
_i64 = ctypes.c_int64

SIGS = {
    "orc_philox4x32_10": (None, [_vp, _vp, _vp]),
    "orc_philox4x64_10": (None, [_vp, _vp, _vp]),
    "orc_uniform1": (_u64, [_u64, _u64]),
    "orc_cc": (_u32, [_u32, _u64, _vp, _i32, _vp]),
    "orc_rmat": (_u64, [_u32, _u64, _u32, _u32, _u32, _u32, _vp]),
    "orc_stoer_wagner": (_i64, [_u32, _vp]),
    "orc_brute_mincut": (_i64, [_u32, _vp]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        for name, (res, args) in SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox4x32(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def philox4x64(ctr, key):
    c = np.asarray(ctr, dtype=np.uint64)
    k = np.asarray(key, dtype=np.uint64)
    o = np.zeros(4, dtype=np.uint64)
    lib().orc_philox4x64_10(_p(c), _p(k), _p(o))
    return o


def uniform1(x: int, W: int) -> int:
    return int(lib().orc_uniform1(x, W))


def cc(edges: np.ndarray, n: int, mode: int = 0):
    e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1)
    labels = np.empty(n, dtype=np.uint32)
    nc = lib().orc_cc(n, e.size // 2, _p(e), mode, _p(labels))
    return labels, int(nc)


def rmat(scale: int, m: int, seed: int, thresholds):
    out = np.empty(2 * m, dtype=np.uint32)
    ta, tab, tabc = thresholds
    d = lib().orc_rmat(scale, m, seed, ta, tab, tabc, _p(out))
    return out.reshape(m, 2), int(d)


def stoer_wagner(adj: np.ndarray) -> int:
    a = np.ascontiguousarray(adj, dtype=np.int64)
    return int(lib().orc_stoer_wagner(a.shape[0], _p(a)))


def brute_mincut(adj: np.ndarray) -> int:
    a = np.ascontiguousarray(adj, dtype=np.int64)
    return int(lib().orc_brute_mincut(a.shape[0], _p(a)))


def edges_to_dense(n: int, edges: np.ndarray) -> np.ndarray:
    """(m, 3) [u, v, w] -> symmetric int64 matrix; parallel edges summed, loops dropped."""
    A = np.zeros((n, n), dtype=np.int64)
    e = np.asarray(edges, dtype=np.int64)
    if e.size:
        u, v, w = e[:, 0], e[:, 1], e[:, 2]
        keep = u != v
        np.add.at(A, (u[keep], v[keep]), w[keep])
        np.add.at(A, (v[keep], u[keep]), w[keep])
    return A
