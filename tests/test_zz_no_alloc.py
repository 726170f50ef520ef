"""cac_cc_static_dev (connected components without device allocation) gives the
labels and count of cac_cc_dev, called through the C ABI: two calls in a row
on one context, the second on a smaller graph (the static scratch keeps the
first call's contents), in both label forms, on the sampled path (m >= 2^18)
and the plain hook path; sizes above the static scratch are rejected."""
import ctypes
import sys

import numpy as np
import pytest

import cac

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _graph(seed, n, m):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    e[: m // 10, 1] = e[: m // 10, 0]  # loops
    return e


def _both(ctx, e, n, mode):
    m = len(e)
    de = ctx.alloc(max(m, 1) * 8)
    la = ctx.alloc(max(n, 1) * 4)
    lb = ctx.alloc(max(n, 1) * 4)
    try:
        if m:
            de.upload(e)
        nc_a = ctx.cc_dev(de.ptr, m, n, la.ptr, mode)
        nc_b = ctx.cc_static_dev(de.ptr, m, n, lb.ptr, mode)
        return nc_a, la.download(np.uint32, n), nc_b, lb.download(np.uint32, n)
    finally:
        de.free()
        la.free()
        lb.free()


@pytest.mark.parametrize("mode", [cac.LABEL_MIN_ID, cac.LABEL_RANK])
@pytest.mark.parametrize("sizes", [((1 << 21, 3 << 20), (300000, 700000)),   # sampled path, then smaller
                                   ((1 << 20, 1 << 19), (5000, 4000)),        # sampled path, then plain hook
                                   ((70000, 150000), (17, 20))])              # plain hook twice
def test_static_equals_cc_dev(ctx, orc, mode, sizes):
    for k, (n, m) in enumerate(sizes):
        _say(f"cc static call {k}: n {n} m {m} mode {mode}")
        e = _graph(100 * k + n % 97, n, m)
        nc_a, lab_a, nc_b, lab_b = _both(ctx, e, n, mode)
        assert nc_b == nc_a
        assert np.array_equal(lab_b, lab_a)
        if n <= 300000:
            want, wn = orc.cc(e, n, mode)
            assert nc_b == wn and np.array_equal(lab_b, want)


def test_static_empty_and_isolated(ctx):
    for n, m in ((0, 0), (1, 0), (64, 0), (65, 1)):
        e = np.zeros((m, 2), dtype=np.uint32)
        for mode in (cac.LABEL_MIN_ID, cac.LABEL_RANK):
            nc_a, lab_a, nc_b, lab_b = _both(ctx, e, n, mode)
            assert nc_b == nc_a == n and np.array_equal(lab_b, lab_a)


def test_static_rejects_oversize(ctx):
    nc = ctypes.c_uint32()
    lab = ctx.alloc(64)
    e = ctx.alloc(64)
    try:
        rc = ctx.lib.cac_cc_static_dev(ctx.handle, e.ptr, 0, cac.CC_STATIC_MAX_N + 1, cac.LABEL_MIN_ID, lab.ptr,
                                       ctypes.byref(nc), None)
        assert rc == cac.CAC_EINVAL
        assert b"CAC_CC_STATIC_MAX_N" in ctx.lib.cac_last_error(ctx.handle)
    finally:
        lab.free()
        e.free()
