"""The wide form of a square-root trial: its sampling rounds and densification
run as grid-wide kernels over the edge list instead of one workgroup per trial.
Forced down to small graphs it must give the oracle's raw cuts bit for bit; at
sizes where both forms run, its cuts equal the workgroup-per-trial form's."""
import sys

import numpy as np
import pytest
from graphgen import dense, erdos_renyi, planted

import cac

pytestmark = pytest.mark.gpu
F32_95 = float(np.float32(0.95))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def wide():
    c = cac.Context(0)
    c.tune_set(cac.TUNE_WIDE_SAMPLING, 0, 1)  # every trial of 2 or more edges runs wide
    yield c
    c.close()


@pytest.fixture(scope="module")
def narrow():
    c = cac.Context(0)
    c.tune_set(cac.TUNE_WIDE_SAMPLING, 0, 63)  # never wide
    yield c
    c.close()


def test_wide_knob_range(wide):
    assert wide.tune_get(cac.TUNE_WIDE_SAMPLING, 0) == 1
    for bad in (-1, 64):
        with pytest.raises(cac.CacError) as ei:
            wide.tune_set(cac.TUNE_WIDE_SAMPLING, 0, bad)
        assert ei.value.code == cac.CAC_EINVAL
    with pytest.raises(cac.CacError):
        wide.tune_set(cac.TUNE_WIDE_SAMPLING, 1, 20)
    assert wide.tune_get(cac.TUNE_WIDE_SAMPLING, 0) == 1


@pytest.mark.parametrize("n,seed", [(200, 0), (1000, 1), (3000, 2)])
def test_wide_trials_bit_exact(wide, orc, n, seed):
    rng = np.random.default_rng(seed)
    e = planted(rng, n, deg=12, wmax=40, c=3)
    t = orc.intermediate_size(n, len(e))
    _say(f"wide trials n {n} m {len(e)} t {t}")
    raw = wide.sqrt_cut_trials(e, n, t, 7, 0, 12)
    assert raw.tolist() == [orc.sqrt_trial(n, e, t, 7, k) for k in range(12)]


def test_wide_varying_cuts(wide, orc):
    # weighted cycle with chords: trial cuts differ, so the comparison pins the
    # draw sequence, the draw-prefix search and the densified matrix
    rng = np.random.default_rng(5)
    n = 700
    e = [[i, (i + 1) % n, int(rng.integers(10**6, 2 * 10**6))] for i in range(n)]
    e += [[int(a), int(b), int(rng.integers(1, 10**5))] for a, b in rng.integers(0, n, size=(n // 3, 2)) if a != b]
    e = np.array(e, dtype=np.int64)
    t = orc.intermediate_size(n, len(e))
    _say(f"wide varying n {n} t {t}")
    raw = wide.sqrt_cut_trials(e, n, t, 1, 0, 16)
    want = [orc.sqrt_trial(n, e, t, 1, k) for k in range(16)]
    assert len(set(want)) > 1
    assert raw.tolist() == want


@pytest.mark.parametrize("kind", ["one-weight", "one-weight-loop", "skewed"])
def test_wide_weight_forms(wide, orc, kind):
    """Unit weights (which the workgroup form draws by division in its first
    round), a loop, and nearly all weight on a few edges."""
    rng = np.random.default_rng(len(kind))
    n = 400
    e = erdos_renyi(rng, n, 10.0 / n)
    e[:, 2] = 7
    if kind == "one-weight-loop":
        e = np.concatenate([e, np.array([[5, 5, 7]], dtype=np.int64)])
    if kind == "skewed":
        e[:, 2] = 1
        e[rng.choice(len(e), 9, replace=False), 2] = 1_000_000_000
    t = 60
    _say(f"wide weight forms {kind}: n {n} m {len(e)}")
    raw = wide.sqrt_cut_trials(e, n, t, 11, 0, 6)
    assert raw.tolist() == [orc.sqrt_trial(n, e, t, 11, k) for k in range(6)]


def test_wide_hc_rule(orc):
    n, t = 3000, 400
    rng = np.random.default_rng(n)
    u = rng.integers(0, n, 12 * n)
    v = rng.integers(0, n, 12 * n)
    keep = u != v
    E = np.stack([u[keep], v[keep], rng.integers(1, 30, keep.sum())], 1).astype(np.uint64)
    dpr = orc.hc_draws(n)
    with cac.Context(0) as c:
        c.tune_set(cac.TUNE_WIDE_SAMPLING, 0, 1)
        c.set_sampling_rule(cac.SAMPLE_HC)
        raw = c.sqrt_cut_trials(E, n, t, 9, 0, 4)
        want = [orc.sqrt_trial(n, E, t, 9, k, dpr) for k in range(4)]
        _say(f"wide hc rule n {n}: device {raw.tolist()} oracle {want}")
        assert raw.tolist() == want


def test_wide_disconnected(wide, narrow):
    # two components: sampling stops with more than t super-vertices, cut 0
    rng = np.random.default_rng(3)
    a = erdos_renyi(rng, 300, 0.05, wmax=9)
    b = erdos_renyi(rng, 300, 0.05, wmax=9)
    b[:, :2] += 300
    e = np.concatenate([a, b])
    _say("wide disconnected n 600")
    raw_w = wide.sqrt_cut_trials(e, 600, 40, 2, 0, 5)
    raw_n = narrow.sqrt_cut_trials(e, 600, 40, 2, 0, 5)
    assert raw_w.tolist() == raw_n.tolist()
    cut, _ = wide.mincut(e, 600, F32_95, 0)
    assert cut == 0


def test_wide_mincut_er1000(wide, orc):
    rng = np.random.default_rng(0)
    n = 1000
    e = erdos_renyi(rng, n, 0.05, wmax=1)
    _say(f"wide mincut ER1000 m {len(e)}")
    cut, T = wide.mincut(e, n, F32_95, 0)
    assert T == orc.sqrt_trials(n, len(e), F32_95)
    assert cut == orc.stoer_wagner(dense(n, e))


def test_wide_mid_size_bit_exact(wide, orc):
    # n = 20000, m = 320000: several rounds with large draw counts
    rng = np.random.default_rng(11)
    n, m = 20000, 320000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    E = np.stack([np.minimum(u, v), np.maximum(u, v), rng.integers(1, 100, m)], 1).astype(np.uint64)
    t = orc.intermediate_size(n, m)
    raw = wide.sqrt_cut_trials(E, n, t, 4, 0, 2)
    want = [orc.sqrt_trial(n, E, t, 4, k) for k in range(2)]
    _say(f"wide mid size n {n} m {m} t {t}: device {raw.tolist()} oracle {want}")
    assert raw.tolist() == want


def test_wide_equals_workgroup_form_large(narrow):
    # n = 2^17 (labels beyond LDS), m = 2^21: the built-in threshold (2^20 edges) runs it wide
    rng = np.random.default_rng(17)
    n, m = 1 << 17, 1 << 21
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    E = np.stack([np.minimum(u, v), np.maximum(u, v), rng.integers(1, 100, m)], 1).astype(np.uint64)
    t = min(cac.intermediate_size(n, m), 4096)
    with cac.Context(0) as c:
        c.tune_set(cac.TUNE_WIDE_SAMPLING, 0, 0)
        raw_w = c.sqrt_cut_trials(E, n, t, 6, 0, 4)
    raw_n = narrow.sqrt_cut_trials(E, n, t, 6, 0, 4)
    _say(f"wide vs workgroup form n {n} m {m} t {t}: {raw_w.tolist()} {raw_n.tolist()}")
    assert raw_w.tolist() == raw_n.tolist()


def test_wide_beyond_dense_engine(narrow):
    # t > 8192 (m above 16.7M): the contracted graphs become merged edge lists
    # for the sparse Karger-Stein; the sampling before it runs wide by default
    rng = np.random.default_rng(23)
    n, m = 50_000, 17_000_000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    E = np.stack([np.minimum(u, v), np.maximum(u, v), rng.integers(1, 10, m)], 1).astype(np.uint64)
    t = cac.intermediate_size(n, m)
    assert t > 8192
    with cac.Context(0) as c:
        c.tune_set(cac.TUNE_WIDE_SAMPLING, 0, 0)
        raw_w = c.sqrt_cut_trials(E, n, t, 8, 0, 3)
    raw_n = narrow.sqrt_cut_trials(E, n, t, 8, 0, 3)
    _say(f"beyond dense engine n {n} m {m} t {t}: wide {raw_w.tolist()} workgroup {raw_n.tolist()}")
    assert raw_w.tolist() == raw_n.tolist()


def test_wide_profiled_solve():
    # with profiling on, a wide solve records its timed ranges; every record
    # closes and reads back, and the cuts do not change
    rng = np.random.default_rng(29)
    n, m = 1 << 17, 1 << 21
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    E = np.stack([np.minimum(u, v), np.maximum(u, v), rng.integers(1, 100, m)], 1).astype(np.uint64)
    t = min(cac.intermediate_size(n, m), 4096)
    with cac.Context(0) as c:
        c.tune_set(cac.TUNE_WIDE_SAMPLING, 0, 20)
        want = c.sqrt_cut_trials(E, n, t, 2, 0, 3)
        c.prof_reset()
        c.prof_enable(True)
        got = c.sqrt_cut_trials(E, n, t, 2, 0, 3)
        recs = c.prof_records()
        c.prof_enable(False)
    assert got.tolist() == want.tolist()
    names = [k for k, _ in recs]
    _say(f"profiled wide solve: {len(recs)} records, names {sorted(set(names))}")
    assert "wide_sample" in names and "wide_densify" in names
    assert all(ms >= 0 for _, ms in recs)
