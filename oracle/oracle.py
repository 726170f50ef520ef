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
    "orc_cc_mt": (_u32, [_u32, _u64, _vp, _i32, _vp, _i32]),
    "orc_rmat": (_u64, [_u32, _u64, _u32, _u32, _u32, _u32, _vp]),
    "orc_stoer_wagner": (_i64, [_u32, _vp]),
    "orc_brute_mincut": (_i64, [_u32, _vp]),
    "orc_ks_target": (_u32, [_u32]),
    "orc_ks_min_success": (ctypes.c_double, [_u32]),
    "orc_ks_trials": (_u32, [_u32, ctypes.c_double]),
    "orc_intermediate_size": (_u32, [_u32, _u64]),
    "orc_sqrt_trials": (_u32, [_u32, _u64, ctypes.c_double]),
    "orc_approx_trials": (_u32, [ctypes.c_double]),
    "orc_ks_trial": (_i64, [_u32, _vp, _u32, _u32]),
    "orc_min_degree": (_i64, [_u32, _vp]),
    "orc_sample_contract": (_u32, [_u32, _u64, _vp, _u32, _u32, _u32, _vp]),
    "orc_sample_contract_rule": (_u32, [_u32, _u64, _vp, _u32, _u32, _u32, _u32, _vp]),
    "orc_hc_draws": (_u32, [_u32]),
    "orc_densify": (_u32, [_u32, _u64, _vp, _vp, _vp]),
    "orc_sqrt_trial": (_i64, [_u32, _u64, _vp, _u32, _u32, _u32]),
    "orc_sqrt_trial_rule": (_i64, [_u32, _u64, _vp, _u32, _u32, _u32, _u32]),
    "orc_ks_first_pick": (None, [_u32, _vp, _u32, _u32, _u32, _vp, _vp]),
    "orc_keep_threshold": (_u64, [_u32, _u64]),
    "orc_approx_cut": (None, [_u32, _u64, _vp, ctypes.c_double, _u32, _vp, _vp]),
    "orc_race_key": (_u64, [_u64, _u64]),
    "orc_sparse_contract": (_vp, [_u32, _u64, _vp, _u32, _u32, _u32, _u32, _vp]),
    "orc_ks_sparse_trial": (_i64, [_u32, _u64, _vp, _u32, _u32]),
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


def cc_mt(edges: np.ndarray, n: int, mode: int = 0, threads: int = 0):
    """cc() on `threads` threads (0: every online core); the same output."""
    e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1)
    labels = np.empty(n, dtype=np.uint32)
    nc = lib().orc_cc_mt(n, e.size // 2, _p(e), mode, _p(labels), threads)
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


def ks_target(s: int) -> int:
    return int(lib().orc_ks_target(s))


def ks_min_success(s: int) -> float:
    return float(lib().orc_ks_min_success(s))


def ks_trials(n: int, p: float) -> int:
    return int(lib().orc_ks_trials(n, p))


def intermediate_size(n: int, m: int) -> int:
    return int(lib().orc_intermediate_size(n, m))


def sqrt_trials(n: int, m: int, p: float) -> int:
    return int(lib().orc_sqrt_trials(n, m, p))


def approx_trials(p: float) -> int:
    return int(lib().orc_approx_trials(p))


def min_degree(adj) -> int:
    a = np.ascontiguousarray(adj, dtype=np.int64)
    return int(lib().orc_min_degree(a.shape[0], _p(a)))


def ks_trial(adj, seed: int, trial: int) -> int:
    """Raw cut of one dense Karger-Stein trial (min(bound, leaves))."""
    a = np.ascontiguousarray(adj, dtype=np.int64)
    return int(lib().orc_ks_trial(a.shape[0], _p(a), seed, trial))


def ks_first_pick(adj, seed: int, trial: int, nid: int = 1):
    a = np.ascontiguousarray(adj, dtype=np.int64)
    x = np.zeros(1, np.uint32)
    y = np.zeros(1, np.uint32)
    lib().orc_ks_first_pick(a.shape[0], _p(a), seed, trial, nid, _p(x), _p(y))
    return int(x[0]), int(y[0])


def _edges3(edges):
    return np.ascontiguousarray(np.asarray(edges, dtype=np.uint64).reshape(-1, 3))


def sample_contract(n: int, edges, t: int, seed: int, trial: int, draws_per_round: int = 0):
    """draws_per_round: 0 = |V| per round (low-concurrency rule), else that many."""
    e = _edges3(edges)
    lab = np.empty(n, dtype=np.uint32)
    nv = lib().orc_sample_contract_rule(n, e.shape[0], _p(e), t, seed, trial, draws_per_round, _p(lab))
    return lab, int(nv)


def hc_draws(n0: int) -> int:
    """Draws per round of the high-concurrency rule: (unsigned) powf(n0, 1.05f)."""
    return int(lib().orc_hc_draws(n0))


def densify(n: int, edges, labels):
    e = _edges3(edges)
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    k = int(np.count_nonzero(lab == np.arange(n, dtype=np.uint32)))
    out = np.zeros((k, k), dtype=np.int64)
    lib().orc_densify(n, e.shape[0], _p(e), _p(lab), _p(out))
    return out


def sqrt_trial(n: int, edges, t: int, seed: int, trial: int, draws_per_round: int = 0) -> int:
    e = _edges3(edges)
    return int(lib().orc_sqrt_trial_rule(n, e.shape[0], _p(e), t, seed, trial, draws_per_round))


def race_key(x: int, w: int) -> int:
    return int(lib().orc_race_key(x, w))


def sparse_contract(n: int, edges, tgt: int, seed: int, trial: int, nid: int):
    """Child edge list ((m', 3) uint64) of one sparse contraction, or None if disconnected."""
    e = _edges3(edges)
    mc = ctypes.c_uint64()
    ptr = lib().orc_sparse_contract(n, e.shape[0], _p(e), tgt, seed, trial, nid, ctypes.byref(mc))
    if not ptr:
        return None
    buf = (ctypes.c_uint64 * (3 * max(mc.value, 1))).from_address(ptr)
    out = np.frombuffer(buf, dtype=np.uint64, count=3 * mc.value).reshape(-1, 3).copy()
    ctypes.CDLL(None).free(ctypes.c_void_p(ptr))
    return out


def ks_sparse_trial(n: int, edges, seed: int, trial: int) -> int:
    e = _edges3(edges)
    return int(lib().orc_ks_sparse_trial(n, e.shape[0], _p(e), seed, trial))


def keep_threshold(i: int, w: int) -> int:
    return int(lib().orc_keep_threshold(i, w))


def approx_cut(n: int, edges, p: float, seed: int):
    e = _edges3(edges)
    a = np.zeros(1, np.uint64)
    t = np.zeros(1, np.uint32)
    lib().orc_approx_cut(n, e.shape[0], _p(e), p, seed, _p(a), _p(t))
    return int(a[0]), int(t[0])


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


# ---- the reference's own Karger-Stein library (oracle/_ref, built from the
# reference sources by build_ref.sh when the reference tree is available) ----
REF_LIB = HERE / "_ref" / "libref.so"
_ref = None


def ref_available() -> bool:
    return REF_LIB.exists()


def ref():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(str(REF_LIB))
        L.ref_ks_target.restype = _i32
        L.ref_ks_target.argtypes = [_i32]
        L.ref_ks_min_success.restype = ctypes.c_double
        L.ref_ks_min_success.argtypes = [_i32]
        L.ref_ks_trials.restype = _i32
        L.ref_ks_trials.argtypes = [_i32, ctypes.c_double]
        L.ref_deterministic_mincut.restype = _u64
        L.ref_deterministic_mincut.argtypes = [_i32, ctypes.c_long, _vp]
        L.ref_ks_mincut.restype = _u64
        L.ref_ks_mincut.argtypes = [_i32, ctypes.c_long, _vp, ctypes.c_double, _i32]
        _ref = L
    return _ref


def ref_deterministic_mincut(n: int, edges) -> int:
    e = np.ascontiguousarray(edges, dtype=np.uint64).reshape(-1)
    return int(ref().ref_deterministic_mincut(n, e.size // 3, _p(e)))


def ref_ks_mincut(n: int, edges, p: float, seed: int) -> int:
    e = np.ascontiguousarray(edges, dtype=np.uint64).reshape(-1)
    return int(ref().ref_ks_mincut(n, e.size // 3, _p(e), p, seed))
