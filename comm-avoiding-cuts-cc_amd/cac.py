"""ctypes binding of libcac.so (the C ABI declared in include/cac.h).

This is the host-side mirror used by the tests and bench.py; the C++ command
line tools link the same library directly.  Loading fails loudly when the
library has not been built: there is no CPU fallback anywhere on this path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libcac.so"
# Launch shapes written by autotune.py; every Context applies them when present.
TUNING_PATH = PKG_DIR / "tuning" / "gfx950.json"
TUNE_KS_CONTRACT_BLOCK = 1
TUNE_KS_CONTRACT_KEYS = (512, 1024, 2048, 4096)
TUNE_SAMPLE_OCCUPANCY = 2
TUNE_SAMPLE_OCCUPANCY_KEYS = (16, 20, 24, 32)
TUNE_WIDE_SAMPLING = 3  # key 0: log2 of the smallest edge count run in the wide form (0 built-in, 63 never)
TUNE_KS_COMPACT_BLOCK = 4
TUNE_KS_COMPACT_KEYS = (512, 1024, 2048, 4096, 8192)

CAC_OK = 0
CAC_EINVAL = -1
CAC_ENOMEM = -2
CAC_EDEVICE = -3
CAC_ENOTCONN = -4
CAC_ERANGE = -5
CAC_EIO = -6
LABEL_MIN_ID = 0
LABEL_RANK = 1
CC_STATIC_MAX_N = 1 << 25  # CAC_CC_STATIC_MAX_N
SAMPLE_LC = 0  # |V| draws per round (the reference's low-concurrency rule)
SAMPLE_HC = 1  # floor(n^1.05) draws per round (its high-concurrency rule)

# Edge records, identical to cac_uedge / cac_edge.
UEDGE = np.dtype([("u", "<u4"), ("v", "<u4")])
EDGE = np.dtype([("u", "<u4"), ("v", "<u4"), ("w", "<u8")])
# cac_graph_desc (24 bytes): one graph of a batch.
GRAPH_DESC = np.dtype([("edge_offset", "<u8"), ("m", "<u8"), ("n", "<u4"), ("trials", "<u4")])
# cac_query (48 bytes) and cac_query_result (64 bytes) of the serving loop.
QUERY_CC = 0
QUERY_MINCUT = 1
QUERY = np.dtype([("kind", "<i4"), ("n", "<u4"), ("m", "<u8"), ("edges", "<u8"), ("p", "<f8"), ("seed", "<u4"),
                  ("trials", "<u4"), ("labels_out", "<u8"), ("graph", "<u8")])
QUERY_RESULT = np.dtype([("id", "<u8"), ("status", "<i4"), ("ncomp", "<u4"), ("cut", "<i8"), ("t", "<u4"),
                         ("trials", "<u4"), ("submit_s", "<f8"), ("admit_s", "<f8"), ("done_s", "<f8"),
                         ("batch_size", "<u4"), ("reserved", "<u4")])

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int
_pu32 = ctypes.POINTER(ctypes.c_uint32)
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "cac_ctx_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "cac_ctx_destroy": (_i32, [_vp]),
    "cac_strerror": (ctypes.c_char_p, [_i32]),
    "cac_last_error": (ctypes.c_char_p, [_vp]),
    "cac_ctx_stream": (_vp, [_vp]),
    "cac_ctx_synchronize": (_i32, [_vp]),
    "cac_dev_alloc": (_i32, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "cac_dev_free": (_i32, [_vp, _vp]),
    "cac_memcpy_h2d": (_i32, [_vp, _vp, _vp, ctypes.c_size_t]),
    "cac_memcpy_d2h": (_i32, [_vp, _vp, _vp, ctypes.c_size_t]),
    "cac_memcpy_d2d": (_i32, [_vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "cac_tune_set": (_i32, [_vp, _i32, _u32, ctypes.c_int32]),
    "cac_tune_get": (_i32, [_vp, _i32, _u32, ctypes.POINTER(ctypes.c_int32)]),
    "cac_prof_enable": (_i32, [_vp, _i32]),
    "cac_prof_reset": (_i32, [_vp]),
    "cac_prof_count": (_i32, [_vp]),
    "cac_prof_get": (_i32, [_vp, _i32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_float)]),
    "cac_cc": (_i32, [_vp, _vp, _u64, _u32, _i32, _pu32, _pu32]),
    "cac_cc_dev": (_i32, [_vp, _vp, _u64, _u32, _i32, _vp, _pu32, _vp]),
    "cac_cc_static_dev": (_i32, [_vp, _vp, _u64, _u32, _i32, _vp, _pu32, _vp]),
    "cac_cc_csr_dev": (_i32, [_vp, _vp, _vp, _u32, _i32, _vp, _pu32, _vp]),
    "cac_cc_csr_graph_dev": (_i32, [_vp, _vp, _vp, _u32, _i32, _vp, _pu32, _vp]),
    "cac_csr_build_dev": (_i32, [_vp, _vp, _u64, _u32, _vp, _vp, _pu64, _vp]),
    "cac_gen_rmat_dev": (_i32, [_vp, _u32, _u64, _u32, _u32, _u32, _u32, _vp, _pu64]),
    "cac_ks_dense": (_i32, [_vp, _vp, _u32, _u32, _u32, _u32, _pi64, _pi64]),
    "cac_ks_dense_dev": (_i32, [_vp, _vp, _u32, _u32, _u32, _u32, _pi64, _pi64, _vp]),
    "cac_ks_sparse": (_i32, [_vp, _vp, _u64, _u32, _u32, _u32, _u32, _pi64, _pi64]),
    "cac_karger_stein": (_i32, [_vp, _vp, _u64, _u32, ctypes.c_double, _u32, _pi64, _pu32]),
    "cac_race_key": (_u64, [_u64, _u64]),
    "cac_sparse_contract": (_i32, [_vp, _vp, _u64, _u32, _u32, _u32, _u32, _u32, _vp, _u64, _pu64]),
    "cac_ks_trials": (_u32, [_u32, ctypes.c_double]),
    "cac_intermediate_size": (_u32, [_u32, _u64]),
    "cac_sqrt_trials": (_u32, [_u32, _u64, ctypes.c_double]),
    "cac_set_sampling_rule": (_i32, [_vp, _i32]),
    "cac_hc_draws": (_u32, [_u32]),
    "cac_sample_contract": (_i32, [_vp, _vp, _u64, _u32, _u32, _u32, _u32, _pu32, _pu32, _vp, _u64, _pu64]),
    "cac_sqrt_cut_trials": (_i32, [_vp, _vp, _u64, _u32, _u32, _u32, _u32, _u32, _pi64]),
    "cac_mincut": (_i32, [_vp, _vp, _u64, _u32, ctypes.c_double, _u32, _pi64, _pu32]),
    "cac_mincut_dev": (_i32, [_vp, _vp, _u64, _u32, ctypes.c_double, _u32, _pi64, _pu32, _vp]),
    "cac_approx_trials": (_u32, [ctypes.c_double]),
    "cac_approx_cut": (_i32, [_vp, _vp, _u64, _u32, ctypes.c_double, _u32, _pu64, _pu32]),
    "cac_gen_complete_dev": (_i32, [_vp, _u32, _u32, _u32, _vp]),
    "cac_gen_er": (_i32, [_u32, _u32, _u32, _u32, _vp, _u64, _pu64]),
    "cac_mincut_batch": (_i32, [_vp, _vp, _vp, _u32, ctypes.c_double, _u32, _pi64, _pu32]),
    "cac_mincut_batch_dev": (_i32, [_vp, _vp, _vp, _u32, ctypes.c_double, _u32, _pi64, _pu32, _vp]),
    "cac_serve_mincut_plan": (_i32, [_u32, _u64, ctypes.c_double, _u32, _pu32, _pu32]),
    "cac_server_create": (_i32, [_i32, _u32, _u64, _u32, ctypes.POINTER(_vp)]),
    "cac_server_submit": (_i32, [_vp, _vp, _pu64]),
    "cac_server_poll": (_i32, [_vp, _vp, _u32, _pu32, _i32]),
    "cac_server_clock": (ctypes.c_double, [_vp]),
    "cac_server_replay": (_i32, [_vp, _vp, _vp, _u32, _vp]),
    "cac_server_stats": (_i32, [_vp, _pu64, _pu64]),
    "cac_server_destroy": (_i32, [_vp]),
    "cac_server_register_graph": (_i32, [_vp, _vp, _u64, _u32, _pu64]),
    "cac_server_unregister_graph": (_i32, [_vp, _u64]),
    "cac_gen_gnp": (_i32, [_u32, ctypes.c_double, _u32, _u32, _u32, _vp, _u64, _pu64]),
    "cac_gen_ws": (_i32, [_u32, _u32, ctypes.c_double, _u32, _u64, _vp, _u64, _pu64]),
    "cac_gen_planted_sizes": (_i32, [_u32, _u32, _u32, _u32, _u32, ctypes.c_double, _u32, _pu32, _pu64]),
    "cac_gen_planted": (_i32, [_u32, _u32, _u32, _u32, _u32, ctypes.c_double, _u32, _u32, _vp, _vp, _u64, _pu64,
                               _pu32]),
    "cac_gen_planted_hard": (_i32, [_u32, _u32, _u32, _u32, _u32, _u32, _u32, _u32, ctypes.c_double, _vp, _vp, _u64,
                                    _pu64, _pu32]),
}

_lib = None


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load libcac.so and declare every signature; raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"libcac.so not built at {p}: run __graft_entry__.build() (make -C comm-avoiding-cuts-cc_amd)")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class CacError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """Device allocation owned by a Context (for the _dev entry points)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        ctx._check(ctx.lib.cac_dev_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        assert host.nbytes <= self.nbytes
        self.ctx._check(self.ctx.lib.cac_memcpy_h2d(self.ctx.handle, self.ptr, _ptr(host), host.nbytes))

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._check(self.ctx.lib.cac_memcpy_d2h(self.ctx.handle, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx._check(self.ctx.lib.cac_dev_free(self.ctx.handle, self.ptr))
            self.ptr = None


class Context:
    """One cac_ctx on one device (not thread-safe, like the C object)."""

    def __init__(self, device: int = 0, tuning: str | os.PathLike | None | bool = None):
        """tuning: a JSON file written by autotune.py (default: TUNING_PATH when it
        exists), or False for the built-in launch shapes."""
        self.lib = load_library()
        h = _vp()
        rc = self.lib.cac_ctx_create(device, ctypes.byref(h))
        if rc != CAC_OK:
            raise CacError(rc, f"cac_ctx_create: {self.lib.cac_strerror(rc).decode()}")
        self.handle = h.value
        self.tuning_file = None
        if tuning is not False:
            path = Path(tuning) if tuning else TUNING_PATH
            if tuning or path.exists():
                self.load_tuning(path)

    def set_sampling_rule(self, rule: int):
        """SAMPLE_LC (default) or SAMPLE_HC for every sampling call on this context."""
        self._check(self.lib.cac_set_sampling_rule(self.handle, rule))

    # ---- tuning -----------------------------------------------------------
    def tune_set(self, knob: int, key: int, value: int):
        self._check(self.lib.cac_tune_set(self.handle, knob, key, value))

    def tune_get(self, knob: int, key: int) -> int:
        v = ctypes.c_int32()
        self._check(self.lib.cac_tune_get(self.handle, knob, key, ctypes.byref(v)))
        return int(v.value)

    def load_tuning(self, path: str | os.PathLike):
        """Apply a table written by autotune.py: {"ks_contract_block": {"512": B, ...},
        "sample_occupancy": {"16": k, ...}, "wide_sampling_log2": v}."""
        import json

        table = json.loads(Path(path).read_text())
        for key, value in table.get("ks_contract_block", {}).items():
            self.tune_set(TUNE_KS_CONTRACT_BLOCK, int(key), int(value))
        for key, value in table.get("sample_occupancy", {}).items():
            self.tune_set(TUNE_SAMPLE_OCCUPANCY, int(key), int(value))
        for key, value in table.get("ks_compact_block", {}).items():
            self.tune_set(TUNE_KS_COMPACT_BLOCK, int(key), int(value))
        if "wide_sampling_log2" in table:
            self.tune_set(TUNE_WIDE_SAMPLING, 0, int(table["wide_sampling_log2"]))
        self.tuning_file = str(path)

    def close(self):
        if self.handle:
            self.lib.cac_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != CAC_OK:
            msg = self.lib.cac_last_error(self.handle).decode() if self.handle else ""
            raise CacError(rc, f"{self.lib.cac_strerror(rc).decode()}: {msg}")

    # ---- device memory / timing ------------------------------------------
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    @property
    def stream(self):
        return self.lib.cac_ctx_stream(self.handle)

    def synchronize(self):
        self._check(self.lib.cac_ctx_synchronize(self.handle))

    def copy_dev(self, dst: int, src: int, nbytes: int):
        """Enqueue a device-to-device copy kernel on the context's stream."""
        self._check(self.lib.cac_memcpy_d2d(self.handle, dst, src, nbytes, None))

    def prof_enable(self, on: bool = True):
        self._check(self.lib.cac_prof_enable(self.handle, 1 if on else 0))

    def prof_reset(self):
        self._check(self.lib.cac_prof_reset(self.handle))

    def prof_records(self):
        out = []
        for i in range(self.lib.cac_prof_count(self.handle)):
            name = ctypes.c_char_p()
            ms = ctypes.c_float()
            self._check(self.lib.cac_prof_get(self.handle, i, ctypes.byref(name), ctypes.byref(ms)))
            out.append((name.value.decode(), float(ms.value)))
        return out

    # ---- connected components --------------------------------------------
    def cc(self, edges: np.ndarray, n: int, mode: int = LABEL_MIN_ID):
        """edges: (m, 2) uint32 or UEDGE records.  Returns (labels, ncomp)."""
        e = np.ascontiguousarray(edges).view(np.uint32).reshape(-1)
        m = e.size // 2
        labels = np.empty(n, dtype=np.uint32)
        nc = ctypes.c_uint32()
        self._check(self.lib.cac_cc(self.handle, _ptr(e), m, n, mode,
                                    labels.ctypes.data_as(_pu32), ctypes.byref(nc)))
        return labels, int(nc.value)

    def cc_dev(self, d_edges: int, m: int, n: int, d_labels: int, mode: int = LABEL_MIN_ID, stream=None) -> int:
        nc = ctypes.c_uint32()
        self._check(self.lib.cac_cc_dev(self.handle, d_edges, m, n, mode, d_labels, ctypes.byref(nc), stream))
        return int(nc.value)

    def cc_static_dev(self, d_edges: int, m: int, n: int, d_labels: int, mode: int = LABEL_MIN_ID,
                      stream=None) -> int:
        """cc_dev without device allocation (static scratch; n <= CC_STATIC_MAX_N)."""
        nc = ctypes.c_uint32()
        self._check(self.lib.cac_cc_static_dev(self.handle, d_edges, m, n, mode, d_labels, ctypes.byref(nc),
                                               stream))
        return int(nc.value)

    def csr_build_dev(self, d_edges: int, m: int, n: int, d_offsets: int, d_cols: int, stream=None) -> int:
        """Symmetric CSR of a device edge list into d_offsets[n+1] / d_cols[2m]; returns nnz."""
        nnz = ctypes.c_uint64()
        self._check(self.lib.cac_csr_build_dev(self.handle, d_edges, m, n, d_offsets, d_cols, ctypes.byref(nnz),
                                               stream))
        return int(nnz.value)

    def cc_csr_dev(self, d_offsets: int, d_cols: int, n: int, d_labels: int, mode: int = LABEL_MIN_ID,
                   stream=None) -> int:
        nc = ctypes.c_uint32()
        self._check(self.lib.cac_cc_csr_dev(self.handle, d_offsets, d_cols, n, mode, d_labels, ctypes.byref(nc),
                                            stream))
        return int(nc.value)

    def cc_csr_graph_dev(self, d_offsets: int, d_cols: int, n: int, d_labels: int, mode: int = LABEL_MIN_ID,
                         stream=None) -> int:
        """cc_csr_dev replayed from a captured HIP graph (captured on first use)."""
        nc = ctypes.c_uint32()
        self._check(self.lib.cac_cc_csr_graph_dev(self.handle, d_offsets, d_cols, n, mode, d_labels,
                                                  ctypes.byref(nc), stream))
        return int(nc.value)

    def cc_csr(self, edges, n: int, mode: int = LABEL_MIN_ID):
        """Host convenience: builds the CSR on the device and runs cc_csr_dev.
        Returns (labels, ncomp, offsets, cols)."""
        e = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
        m = len(e)
        de = self.alloc(max(m, 1) * 8)
        do = self.alloc((n + 1) * 4)
        dc = self.alloc(max(2 * m, 1) * 4)
        dl = self.alloc(max(n, 1) * 4)
        try:
            if m:
                de.upload(e)
            nnz = self.csr_build_dev(de.ptr, m, n, do.ptr, dc.ptr)
            nc = self.cc_csr_dev(do.ptr, dc.ptr, n, dl.ptr, mode)
            lab = dl.download(np.uint32, n)
            off = do.download(np.uint32, n + 1)
            cols = dc.download(np.uint32, nnz) if nnz else np.empty(0, np.uint32)
        finally:
            for b in (de, do, dc, dl):
                b.free()
        return lab, nc, off, cols

    # ---- dense Karger-Stein ------------------------------------------------
    def ks_dense(self, adj: np.ndarray, seed: int, first_trial: int, ntrials: int):
        """Returns (raw cut per trial, min cut)."""
        a = np.ascontiguousarray(adj, dtype=np.int64)
        n = a.shape[0]
        raw = np.empty(ntrials, dtype=np.int64)
        mn = ctypes.c_int64()
        self._check(self.lib.cac_ks_dense(self.handle, _ptr(a), n, seed, first_trial, ntrials,
                                          raw.ctypes.data_as(_pi64), ctypes.byref(mn)))
        return raw, int(mn.value)

    def ks_dense_dev(self, d_adj: int, n: int, seed: int, first_trial: int, ntrials: int, stream=None):
        raw = np.empty(ntrials, dtype=np.int64)
        mn = ctypes.c_int64()
        self._check(self.lib.cac_ks_dense_dev(self.handle, d_adj, n, seed, first_trial, ntrials,
                                              raw.ctypes.data_as(_pi64), ctypes.byref(mn), stream))
        return raw, int(mn.value)

    # ---- sparse Karger-Stein ----------------------------------------------
    def ks_sparse(self, edges, n: int, seed: int, first_trial: int, ntrials: int):
        """Sparse recursion on an edge list.  Returns (raw cut per trial, min cut)."""
        e = to_edges(edges)
        raw = np.empty(ntrials, dtype=np.int64)
        mn = ctypes.c_int64()
        self._check(self.lib.cac_ks_sparse(self.handle, _ptr(e), len(e), n, seed, first_trial, ntrials,
                                           raw.ctypes.data_as(_pi64), ctypes.byref(mn)))
        return raw, int(mn.value)

    def sparse_contract(self, edges, n: int, tgt: int, seed: int, trial: int, nid: int) -> np.ndarray:
        """One sparse contraction; returns the child edges as (m', 3) uint64 [u, v, w]."""
        e = to_edges(edges)
        out = np.empty(max(len(e), 1), dtype=EDGE)
        mo = ctypes.c_uint64()
        self._check(self.lib.cac_sparse_contract(self.handle, _ptr(e), len(e), n, tgt, seed, trial, nid, _ptr(out),
                                                 len(out), ctypes.byref(mo)))
        c = out[: mo.value]
        return np.stack([c["u"].astype(np.uint64), c["v"].astype(np.uint64), c["w"]], axis=1)

    def karger_stein(self, edges, n: int, p: float, seed: int):
        """SequentialKargerSteinCut semantics: returns (cut, trials)."""
        e = to_edges(edges)
        cut = ctypes.c_int64()
        T = ctypes.c_uint32()
        self._check(self.lib.cac_karger_stein(self.handle, _ptr(e), len(e), n, p, seed, ctypes.byref(cut),
                                              ctypes.byref(T)))
        return int(cut.value), int(T.value)

    # ---- sampling contraction and sqrt-cut min cut ------------------------
    def sample_contract(self, edges, n: int, t: int, seed: int, trial: int):
        """Returns (vmap, n_out, merged (k, 3) uint64 [u, v, w])."""
        e = to_edges(edges)
        vmap = np.empty(n, dtype=np.uint32)
        nout = ctypes.c_uint32()
        mout = ctypes.c_uint64()
        cap = len(e)
        merged = np.empty(max(cap, 1), dtype=EDGE)
        self._check(self.lib.cac_sample_contract(self.handle, _ptr(e), len(e), n, t, seed, trial,
                                                 vmap.ctypes.data_as(_pu32), ctypes.byref(nout), _ptr(merged), cap,
                                                 ctypes.byref(mout)))
        mg = merged[: mout.value]
        out = np.stack([mg["u"].astype(np.uint64), mg["v"].astype(np.uint64), mg["w"]], axis=1)
        return vmap, int(nout.value), out

    def sqrt_cut_trials(self, edges, n: int, t: int, seed: int, first_trial: int, ntrials: int) -> np.ndarray:
        e = to_edges(edges)
        raw = np.empty(ntrials, dtype=np.int64)
        self._check(self.lib.cac_sqrt_cut_trials(self.handle, _ptr(e), len(e), n, t, seed, first_trial, ntrials,
                                                 raw.ctypes.data_as(_pi64)))
        return raw

    def mincut(self, edges, n: int, p: float, seed: int):
        """Returns (cut, trials) with SquareRootCut low-concurrency semantics."""
        e = to_edges(edges)
        cut = ctypes.c_int64()
        T = ctypes.c_uint32()
        self._check(self.lib.cac_mincut(self.handle, _ptr(e), len(e), n, p, seed, ctypes.byref(cut), ctypes.byref(T)))
        return int(cut.value), int(T.value)

    def mincut_batch(self, edges, descs: np.ndarray, p: float, seed: int):
        """Batched square-root cut over graphs descs (GRAPH_DESC records) whose
        edges lie back to back in edges.  Returns (cuts, trials) arrays."""
        e = to_edges(edges)
        d = np.ascontiguousarray(descs, dtype=GRAPH_DESC)
        cuts = np.empty(len(d), dtype=np.int64)
        trials = np.empty(len(d), dtype=np.uint32)
        self._check(self.lib.cac_mincut_batch(self.handle, _ptr(e), _ptr(d), len(d), p, seed,
                                              cuts.ctypes.data_as(_pi64), trials.ctypes.data_as(_pu32)))
        return cuts, trials

    def mincut_batch_dev(self, d_edges: int, descs: np.ndarray, p: float, seed: int, stream=None):
        d = np.ascontiguousarray(descs, dtype=GRAPH_DESC)
        cuts = np.empty(len(d), dtype=np.int64)
        trials = np.empty(len(d), dtype=np.uint32)
        self._check(self.lib.cac_mincut_batch_dev(self.handle, d_edges, _ptr(d), len(d), p, seed,
                                                  cuts.ctypes.data_as(_pi64), trials.ctypes.data_as(_pu32), stream))
        return cuts, trials

    def approx_cut(self, edges, n: int, p: float, seed: int):
        """Returns (2^k approximation, trials) with ApproximateCut semantics."""
        e = to_edges(edges)
        a = ctypes.c_uint64()
        T = ctypes.c_uint32()
        self._check(self.lib.cac_approx_cut(self.handle, _ptr(e), len(e), n, p, seed, ctypes.byref(a),
                                            ctypes.byref(T)))
        return int(a.value), int(T.value)

    # ---- generators -------------------------------------------------------
    def gen_complete_dev(self, n: int, seed: int, wmax: int, d_adj: int):
        self._check(self.lib.cac_gen_complete_dev(self.handle, n, seed, wmax, d_adj))

    def gen_rmat_dev(self, scale: int, m: int, seed: int, thresholds, d_edges: int) -> int:
        ta, tab, tabc = thresholds
        used = ctypes.c_uint64()
        self._check(self.lib.cac_gen_rmat_dev(self.handle, scale, m, seed, ta, tab, tabc, d_edges,
                                              ctypes.byref(used)))
        return int(used.value)


class Server:
    """cac_server: worker threads with their own contexts serving CC and min-cut
    queries with footprint-based admission and per-kind launch groups."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget: int = 64 << 30, max_batch: int = 64):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.cac_server_create(device, workers, hbm_budget, max_batch, ctypes.byref(h))
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_create")
        self.handle = h
        self._keep = []

    @staticmethod
    def query(kind: int, edges, n: int, p: float = 0.95, seed: int = 0, labels: np.ndarray | None = None,
              trials: int = 0, graph: int = 0):
        """One QUERY record; returns (record, arrays to keep alive).  trials: 0 = the
        reference's rule for p, k > 0 = a budget of k trials (serve_mincut_plan).
        graph: a handle from register_graph (then edges may be None)."""
        e = to_edges(edges) if edges is not None else np.zeros(0, dtype=EDGE)
        q = np.zeros(1, dtype=QUERY)
        q["kind"], q["n"], q["m"] = kind, n, len(e)
        q["edges"] = e.ctypes.data if len(e) else 0
        q["p"], q["seed"], q["trials"], q["graph"] = p, seed, trials, graph
        q["labels_out"] = labels.ctypes.data if labels is not None else 0
        return q[0], (e, labels)

    def replay(self, queries, arrivals) -> np.ndarray:
        """queries: list of (record, keep) from query(); arrivals: seconds."""
        recs = np.array([q for q, _ in queries], dtype=QUERY)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        out = np.zeros(len(recs), dtype=QUERY_RESULT)
        rc = self.lib.cac_server_replay(self.handle, _ptr(recs), _ptr(arr), len(recs), _ptr(out))
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_replay")
        return out

    def submit(self, record, keep) -> int:
        rec = np.array([record], dtype=QUERY)
        i = ctypes.c_uint64()
        rc = self.lib.cac_server_submit(self.handle, _ptr(rec), ctypes.byref(i))
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_submit")
        self._keep.append(keep)
        return int(i.value)

    def poll(self, cap: int = 256, wait: bool = True) -> np.ndarray:
        out = np.zeros(cap, dtype=QUERY_RESULT)
        k = ctypes.c_uint32()
        rc = self.lib.cac_server_poll(self.handle, _ptr(out), cap, ctypes.byref(k), 1 if wait else 0)
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_poll")
        return out[: k.value]

    def register_graph(self, edges, n: int) -> int:
        """Copies a graph to device memory held by the server; returns its handle."""
        e = to_edges(edges)
        h = ctypes.c_uint64()
        rc = self.lib.cac_server_register_graph(self.handle, _ptr(e), len(e), n, ctypes.byref(h))
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_register_graph")
        return int(h.value)

    def unregister_graph(self, graph: int):
        rc = self.lib.cac_server_unregister_graph(self.handle, graph)
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_unregister_graph")

    def stats(self):
        """(peak admitted footprint in bytes, launch groups run)."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        rc = self.lib.cac_server_stats(self.handle, ctypes.byref(a), ctypes.byref(b))
        if rc != CAC_OK:
            raise CacError(rc, "cac_server_stats")
        return int(a.value), int(b.value)

    def close(self):
        if self.handle:
            self.lib.cac_server_destroy(self.handle)
            self.handle = None
            self._keep = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serve_mincut_plan(n: int, m: int, p: float, trials: int = 0):
    """(t, T) of a served min-cut query (trials = 0: the reference's rule)."""
    t, T = ctypes.c_uint32(), ctypes.c_uint32()
    rc = load_library().cac_serve_mincut_plan(n, m, p, trials, ctypes.byref(t), ctypes.byref(T))
    if rc != CAC_OK:
        raise CacError(rc, "cac_serve_mincut_plan")
    return int(t.value), int(T.value)


def race_key(x: int, w: int) -> int:
    """The sparse contraction's race key of a 64-bit word x and weight w."""
    return int(load_library().cac_race_key(x, w))


def to_edges(edges) -> np.ndarray:
    """(m, 3) [u, v, w] integers (or EDGE records) -> contiguous EDGE records."""
    a = np.asarray(edges)
    if a.dtype == EDGE:
        return np.ascontiguousarray(a)
    a = a.reshape(-1, 3)
    out = np.empty(len(a), dtype=EDGE)
    out["u"] = a[:, 0]
    out["v"] = a[:, 1]
    out["w"] = a[:, 2]
    return out


def hc_draws(n0: int) -> int:
    """Draws per round of the high-concurrency sampling rule for n0 vertices."""
    return int(load_library().cac_hc_draws(n0))


def ks_trials(n: int, p: float) -> int:
    return int(load_library().cac_ks_trials(n, p))


def intermediate_size(n: int, m: int) -> int:
    return int(load_library().cac_intermediate_size(n, m))


def sqrt_trials(n: int, m: int, p: float) -> int:
    return int(load_library().cac_sqrt_trials(n, m, p))


def gen_er(n: int, p: float, seed: int, wmax: int = 1):
    """G(n, p) edge list (EDGE records), redrawn with seed+1, seed+2, ... until
    connected, as the square_root configuration prescribes.  Returns (edges, seed used)."""
    from fractions import Fraction

    lib = load_library()
    thr = int(Fraction(str(p)) * (1 << 32))
    while True:
        m = ctypes.c_uint64()
        lib.cac_gen_er(n, thr, seed, wmax, None, 0, ctypes.byref(m))
        out = np.empty(max(m.value, 1), dtype=EDGE)
        rc = lib.cac_gen_er(n, thr, seed, wmax, _ptr(out), m.value, ctypes.byref(m))
        if rc != CAC_OK:
            raise CacError(rc, "cac_gen_er")
        out = out[: m.value]
        if _connected(n, out):
            return out, seed
        seed += 1


def gen_planted(first: int, count: int, seed: int, nmin: int = 256, nmax: int = 65536, avg_degree: float = 16.0,
                wmax: int = 100, cmax: int = 8):
    """Planted-cut graphs g = first .. first+count-1 (see cac_gen_planted).
    Returns (descs GRAPH_DESC[count], edges EDGE[m_total], planted uint32[count])."""
    lib = load_library()
    ns = np.empty(max(count, 1), dtype=np.uint32)
    bound = ctypes.c_uint64()
    rc = lib.cac_gen_planted_sizes(first, count, seed, nmin, nmax, avg_degree, cmax, ns.ctypes.data_as(_pu32),
                                   ctypes.byref(bound))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_planted_sizes")
    descs = np.empty(max(count, 1), dtype=GRAPH_DESC)
    edges = np.empty(max(bound.value, 1), dtype=EDGE)
    planted = np.empty(max(count, 1), dtype=np.uint32)
    m = ctypes.c_uint64()
    rc = lib.cac_gen_planted(first, count, seed, nmin, nmax, avg_degree, wmax, cmax, _ptr(descs), _ptr(edges),
                             len(edges), ctypes.byref(m), planted.ctypes.data_as(_pu32))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_planted")
    return descs[:count], edges[: m.value], planted[:count]


def gen_planted_hard(first: int, count: int, seed: int, nmin: int = 256, nmax: int = 65536, cycles: int = 8,
                     wmax: int = 100, cmax: int = 8, cut_frac: float = 0.9):
    """Hard planted-cut graphs (see cac_gen_planted_hard): halves of `cycles`
    weighted Hamiltonian cycles, planted cut C = floor(cut_frac * minimum
    weighted degree).  Returns (descs, edges, planted C per graph)."""
    lib = load_library()
    ns = np.empty(max(count, 1), dtype=np.uint32)
    bound = ctypes.c_uint64()
    rc = lib.cac_gen_planted_sizes(first, count, seed, nmin, nmax, 2.0 * cycles, cmax, ns.ctypes.data_as(_pu32),
                                   ctypes.byref(bound))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_planted_sizes")
    descs = np.empty(max(count, 1), dtype=GRAPH_DESC)
    edges = np.empty(max(bound.value, 1), dtype=EDGE)
    planted = np.empty(max(count, 1), dtype=np.uint32)
    m = ctypes.c_uint64()
    rc = lib.cac_gen_planted_hard(first, count, seed, nmin, nmax, cycles, wmax, cmax, cut_frac, _ptr(descs),
                                  _ptr(edges), len(edges), ctypes.byref(m), planted.ctypes.data_as(_pu32))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_planted_hard")
    return descs[:count], edges[: m.value], planted[:count]


def gen_ws(n: int, d: int, seed: int, beta: float = 0.3, weight: int = 100) -> np.ndarray:
    """Watts-Strogatz graph (see cac_gen_ws) as EDGE records, n * d edges."""
    lib = load_library()
    m = ctypes.c_uint64()
    rc = lib.cac_gen_ws(n, d, beta, seed, weight, None, 0, ctypes.byref(m))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_ws")
    out = np.empty(max(m.value, 1), dtype=EDGE)
    rc = lib.cac_gen_ws(n, d, beta, seed, weight, _ptr(out), len(out), ctypes.byref(m))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_ws")
    return out[: m.value]


def gen_gnp(n: int, avg_degree: float, seed: int, stream: int, wmax: int = 100) -> np.ndarray:
    """G(n, avg_degree/(n-1)) as EDGE records (see cac_gen_gnp); may be disconnected."""
    lib = load_library()
    mean = n * avg_degree / 2.0
    cap = int(mean + 8.0 * mean ** 0.5 + 64)
    out = np.empty(max(cap, 1), dtype=EDGE)
    m = ctypes.c_uint64()
    rc = lib.cac_gen_gnp(n, avg_degree, seed, stream, wmax, _ptr(out), cap, ctypes.byref(m))
    if rc == CAC_ERANGE:
        out = np.empty(m.value, dtype=EDGE)
        rc = lib.cac_gen_gnp(n, avg_degree, seed, stream, wmax, _ptr(out), m.value, ctypes.byref(m))
    if rc != CAC_OK:
        raise CacError(rc, "cac_gen_gnp")
    return out[: m.value]


def _connected(n: int, e: np.ndarray) -> bool:
    par = np.arange(n)

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x

    comps = n
    for u, v in zip(e["u"].tolist(), e["v"].tolist()):
        a, b = find(u), find(v)
        if a != b:
            par[a] = b
            comps -= 1
    return comps == 1


def rmat_thresholds(a: float, b: float, c: float):
    """Cumulative u32 thresholds floor(2^32 * x) for the R-MAT quadrant choice,
    computed from decimal fractions exactly (a=0.45 -> 45/100)."""
    from fractions import Fraction

    fa, fb, fc = (Fraction(str(x)) for x in (a, b, c))
    t = lambda f: int(f * (1 << 32))
    return t(fa), t(fa + fb), t(fa + fb + fc)


def declared_symbols(header: str | os.PathLike | None = None):
    """Function names declared in include/cac.h (for the export check)."""
    import re

    h = Path(header) if header else PKG_DIR.parent / "include" / "cac.h"
    text = h.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cac_[a-z0-9_]+)\s*\(", text)))
