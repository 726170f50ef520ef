"""Dense Karger-Stein at the C3 size, n = 8192 (a 512 MiB int64 matrix):

* a planted graph -- four complete blocks of 2048 vertices with weights
  U[1, 42], joined by a few heavy cross edges, so several cuts lie far below
  the minimum weighted degree and raw per-trial cuts differ -- whose raw trial
  cuts equal the oracle's trials on the same Philox stream, bit for bit;
* the complete graph with weights U[1, 42] (the device generator, checked
  against its Philox definition on sampled entries) over all log2(n)^2 = 169
  trials: every raw cut equals the minimum weighted degree, which the
  closed-form certificate proves is the minimum cut."""
import ctypes
import sys

import numpy as np
import pytest

import cac

pytestmark = pytest.mark.gpu
N = 8192


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _planted_blocks(seed):
    rng = np.random.default_rng(seed)
    A = rng.integers(1, 43, size=(N, N), dtype=np.int64)
    A = np.triu(A, 1)
    b = N // 4
    block = np.arange(N) // b
    A[block[:, None] != block[None, :]] = 0
    # cross edges of weight 1500: 5 between blocks 0-1, 7 between 2-3, 9
    # between 1-2, 11 between 0-3
    for (x, y, count, w) in ((0, 1, 5, 1500), (2, 3, 7, 1500), (1, 2, 9, 1500), (0, 3, 11, 1500)):
        us = x * b + rng.choice(b, count, replace=False)
        vs = y * b + rng.choice(b, count, replace=False)
        A[us, vs] = w
    return A + A.T


def test_c3_planted_trials_bit_exact(ctx, orc):
    A = _planted_blocks(3)
    deg = A.sum(axis=1)
    b = N // 4
    block = np.arange(N) // b
    # the cut of every union of blocks (a cut splitting a block costs about its
    # complete graph's minimum degree)
    cuts = {}
    for mask in range(1, 8):
        side = np.isin(block, [k for k in range(4) if mask >> k & 1])
        cuts[mask] = int(A[side][:, ~side].sum())
    _say(f"c3 planted: min degree {deg.min()}, block-union cuts {sorted(cuts.values())}")
    below = sorted(v for v in cuts.values() if v < deg.min())
    assert len(set(below)) >= 3  # several distinct cuts below the bound every trial starts from
    T = 3
    raw, best = ctx.ks_dense(A, 11, 0, T)
    for t in range(T):
        want = orc.ks_trial(A, 11, t)
        _say(f"  trial {t}: device {raw[t]} oracle {want}")
        assert raw[t] == want
    assert best == raw.min() and best >= min(cuts.values())


def _certificate(delta, wmax, n):
    """The minimum cut of a complete graph with weights in [1, wmax] equals its
    minimum weighted degree delta when, for every 2 <= k <= n/2, a k-set's cut
    k delta - k (k-1) wmax or its k (n - k) cross pairs of weight >= 1 reach delta."""
    k = np.arange(2, n // 2 + 1, dtype=np.int64)
    return bool(np.all((k * delta - k * (k - 1) * wmax >= delta) | (k * (n - k) >= delta)))


def test_c3_complete_all_trials_equal_certificate(ctx, orc):
    trials = 169  # ceil(log2(8192))^2
    d_adj = ctx.alloc(N * N * 8)
    try:
        ctx.gen_complete_dev(N, 0, 42, d_adj.ptr)
        # row sums and sampled entries against the generator's Philox definition
        deg_min = None
        rng = np.random.default_rng(0)
        for r in range(0, N, 1024):
            blk = np.empty((1024, N), dtype=np.int64)
            ctx._check(ctx.lib.cac_memcpy_d2h(ctx.handle, blk.ctypes.data_as(ctypes.c_void_p), d_adj.ptr + r * N * 8,
                                              blk.nbytes))
            s = int(blk.sum(axis=1).min())
            deg_min = s if deg_min is None else min(deg_min, s)
            assert np.all(blk[np.arange(1024), r + np.arange(1024)] == 0)
            for i, j in zip(rng.integers(0, 1024, 64), rng.integers(0, N, 64)):
                u, v = r + int(i), int(j)
                if u == v:
                    continue
                a, c = min(u, v), max(u, v)
                p = a * N + c
                x = int(orc.philox4x32([p & 0xFFFFFFFF, p >> 32, 0, 0], [0, 0x47454E52])[0])
                assert blk[i, j] == 1 + ((42 * x) >> 32)
        assert _certificate(deg_min, 42, N)
        _say(f"c3 complete: min degree {deg_min}, {trials} trials")
        raw, cut = ctx.ks_dense_dev(d_adj.ptr, N, 0, 0, trials)
        assert len(raw) == trials and np.all(raw == deg_min) and cut == deg_min
    finally:
        d_adj.free()
