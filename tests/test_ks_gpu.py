"""Dense Karger-Stein parity on the device: every trial's raw cut from
libcac.so equals the oracle's trial fed the same PRNG stream, bit for bit; final
cuts equal exact Stoer-Wagner on planted-cut graphs."""
import sys

import numpy as np
import pytest
from graphgen import dense, erdos_renyi, necklace, planted

import cac

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,seed", [(2, 0), (5, 1), (64, 2), (128, 3), (129, 4), (150, 5), (260, 6), (520, 7)])
def test_ks_trials_bit_exact(ctx, orc, n, seed):
    rng = np.random.default_rng(seed)
    if n < 10:
        A = dense(n, erdos_renyi(rng, n, 1.0, wmax=9))
    else:
        A = dense(n, planted(rng, n, deg=10, wmax=30, c=int(rng.integers(1, 6))))
    T = 6
    _say(f"ks parity n {n} seed {seed} trials {T}")
    raw, mn = ctx.ks_dense(A, 17, 3, T)
    want = [orc.ks_trial(A, 17, 3 + t) for t in range(T)]
    assert raw.tolist() == want
    assert mn == min(want)
    assert mn >= orc.stoer_wagner(A)


@pytest.mark.parametrize("n", [200, 700])
def test_ks_pr_pass(ctx, orc, n):
    # a vertex of degree 2 makes the bound 2, so almost every contracted weight
    # reaches it and the PR pass of each contraction runs
    rng = np.random.default_rng(21 + n)
    e = erdos_renyi(rng, n - 1, 8.0 / n, wmax=5)
    e = np.concatenate([e, [[n - 1, 0, 1], [n - 1, 1, 1]]])
    A = dense(n, e)
    _say(f"ks pr pass n {n}")
    raw, mn = ctx.ks_dense(A, 5, 0, 8)
    want = [orc.ks_trial(A, 5, t) for t in range(8)]
    assert raw.tolist() == want
    assert mn == orc.stoer_wagner(A)


@pytest.mark.parametrize("n,k,deg_in", [(900, 225, 2.0), (1500, 375, 2.0), (3000, 600, 3.0)])
def test_ks_pr_pass_pins_trials(ctx, orc, n, k, deg_in):
    """Rings of small heavy clusters: the PR pass merges inside clusters and the
    trials' raw cuts differ, so a PR pass that stops early or is skipped changes
    them (checked on the CPU with the oracle's PR pass cut short after its
    first merge, and with it removed).  Children of 900, 1500 and 3000
    vertices take the 128-, 256- and 512-thread contractions."""
    rng = np.random.default_rng(n + k)
    A = dense(n, necklace(rng, n, k, 3 * 10**6, deg_in))
    T = 8
    _say(f"ks pr pass necklace n {n} clusters {k}")
    want = [orc.ks_trial(A, 5, t) for t in range(T)]
    assert len(set(want)) > 1
    raw, mn = ctx.ks_dense(A, 5, 0, T)
    assert raw.tolist() == want


def test_ks_disconnected_is_zero(ctx, orc):
    rng = np.random.default_rng(3)
    h = 150
    e1 = erdos_renyi(rng, h, 0.1, wmax=5)
    e2 = erdos_renyi(rng, h, 0.1, wmax=5)
    e2[:, :2] += h
    A = dense(2 * h, np.concatenate([e1, e2]))
    _say("ks disconnected n 300")
    raw, mn = ctx.ks_dense(A, 1, 0, 3)
    assert mn == 0 and raw.tolist() == [orc.ks_trial(A, 1, t) for t in range(3)]


def test_ks_final_cut_planted(ctx, orc):
    rng = np.random.default_rng(99)
    n = 1000
    c = 4
    A = dense(n, planted(rng, n, deg=16, wmax=100, c=c))
    T = orc.ks_trials(n, 0.95)
    _say(f"ks planted n {n} trials {T}")
    raw, mn = ctx.ks_dense(A, 2, 0, T)
    assert mn == c == orc.stoer_wagner(A)
    assert raw.tolist() == [orc.ks_trial(A, 2, t) for t in range(T)]


def test_ks_rejects_bad_matrix(ctx):
    A = np.zeros((4, 4), dtype=np.int64)
    A[0, 1] = 3
    with pytest.raises(cac.CacError) as ei:
        ctx.ks_dense(A, 0, 0, 1)
    assert ei.value.code == cac.CAC_EINVAL


@pytest.mark.parametrize("n", [300, 600, 1500])
def test_ks_varying_trial_cuts(ctx, orc, n):
    # weighted cycles: many near-minimum cuts, so raw trial cuts differ between
    # trials and bit-exactness pins the whole pick sequence, not just the answer
    rng = np.random.default_rng(n)
    e = np.array([[i, (i + 1) % n, int(rng.integers(10**6, 2 * 10**6))] for i in range(n)], dtype=np.int64)
    A = dense(n, e)
    T = 10
    _say(f"ks varying cycle n {n}")
    raw, mn = ctx.ks_dense(A, 3, 0, T)
    want = [orc.ks_trial(A, 3, t) for t in range(T)]
    assert len(set(want)) > 1 or n < 400
    assert raw.tolist() == want
    raw2, _ = ctx.ks_dense(A, 3, 0, T)  # same stream, same cuts: no run-to-run drift
    assert raw2.tolist() == want


@pytest.mark.parametrize("table", [{512: 1024, 1024: 1024, 2048: 1024, 4096: 1024},
                                   {512: 128, 1024: 512, 2048: 256, 4096: 1024}])
def test_ks_tuned_block_sizes_bit_exact(orc, table):
    # the contraction block size of each row-size class (cac_tune_set) changes
    # the launch shape only: weighted cycles, whose raw cuts differ per trial,
    # give the oracle's cuts under the largest blocks and a mixed table
    n = 1500
    rng = np.random.default_rng(77)
    e = np.array([[i, (i + 1) % n, int(rng.integers(10**6, 2 * 10**6))] for i in range(n)], dtype=np.int64)
    A = dense(n, e)
    T = 6
    _say(f"ks tuned blocks {table} n {n}")
    with cac.Context(0, tuning=False) as c:
        for key, b in table.items():
            c.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, key, b)
        assert {k: c.tune_get(cac.TUNE_KS_CONTRACT_BLOCK, k) for k in table} == table
        raw, _ = c.ks_dense(A, 11, 0, T)
    want = [orc.ks_trial(A, 11, t) for t in range(T)]
    assert raw.tolist() == want


def test_ks_tune_rejects_bad_values():
    with cac.Context(0, tuning=False) as c:
        for key, b in [(1000, 128), (512, 32), (1024, 64), (2048, 300), (4096, 2048), (8192, 1024)]:
            with pytest.raises(cac.CacError):
                c.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, key, b)
        with pytest.raises(cac.CacError):
            c.tune_set(99, 512, 64)
        c.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, 512, 256)
        c.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, 512, 0)
        assert c.tune_get(cac.TUNE_KS_CONTRACT_BLOCK, 512) == 0


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_ks_compaction_block_sizes_bit_exact(orc, block):
    # the compaction block size of every child-size class (cac_tune_set)
    # changes the launch shape only: the oracle's per-trial cuts at n = 1500
    n = 1500
    rng = np.random.default_rng(78)
    e = np.array([[i, (i + 1) % n, int(rng.integers(10**6, 2 * 10**6))] for i in range(n)], dtype=np.int64)
    A = dense(n, e)
    T = 6
    _say(f"ks compaction block {block} n {n}")
    with cac.Context(0, tuning=False) as c:
        for key in cac.TUNE_KS_COMPACT_KEYS:
            c.tune_set(cac.TUNE_KS_COMPACT_BLOCK, key, block)
        assert all(c.tune_get(cac.TUNE_KS_COMPACT_BLOCK, k) == block for k in cac.TUNE_KS_COMPACT_KEYS)
        raw, _ = c.ks_dense(A, 12, 0, T)
    assert raw.tolist() == [orc.ks_trial(A, 12, t) for t in range(T)]


def test_ks_compaction_tune_rejects_bad_values():
    with cac.Context(0, tuning=False) as c:
        for key, b in [(512, 128), (1024, 2048), (3000, 512), (16384, 512), (8192, 300)]:
            with pytest.raises(cac.CacError):
                c.tune_set(cac.TUNE_KS_COMPACT_BLOCK, key, b)
        c.tune_set(cac.TUNE_KS_COMPACT_BLOCK, 8192, 1024)
        assert c.tune_get(cac.TUNE_KS_COMPACT_BLOCK, 8192) == 1024
        c.tune_set(cac.TUNE_KS_COMPACT_BLOCK, 8192, 0)
        assert c.tune_get(cac.TUNE_KS_COMPACT_BLOCK, 8192) == 0
