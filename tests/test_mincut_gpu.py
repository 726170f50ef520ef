"""Sampling contraction and square-root min cut on the device, bit-exact against
the oracle: the vertex map and merged edge list of one trial, every trial's raw
cut, and the final cut against exact Stoer-Wagner."""
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


@pytest.mark.parametrize("n,t,seed", [(60, 20, 0), (300, 40, 1), (1000, 318, 2), (2000, 130, 3)])
def test_sample_contract_matches_oracle(ctx, orc, n, t, seed):
    rng = np.random.default_rng(seed)
    e = erdos_renyi(rng, n, min(1.0, 12.0 / n), wmax=50)
    _say(f"sample_contract n {n} t {t} seed {seed}")
    vmap, k, merged = ctx.sample_contract(e, n, t, 9, 4)
    lab, nv = orc.sample_contract(n, e, t, 9, 4)
    assert k == nv == t
    roots = np.flatnonzero(lab == np.arange(n))
    rank = np.zeros(n, dtype=np.int64)
    rank[roots] = np.arange(len(roots))
    assert np.array_equal(vmap, rank[lab])
    M = orc.densify(n, e, lab)
    iu = np.triu_indices(k, 1)
    nz = M[iu] != 0
    want = np.stack([iu[0][nz], iu[1][nz], M[iu][nz]], axis=1).astype(np.uint64)
    assert np.array_equal(merged, want)


@pytest.mark.parametrize("kind", ["one-weight", "one-weight-loop", "one-heavier", "one-weight-large", "skewed"])
def test_sampling_first_round_forms(ctx, orc, kind):
    """A graph whose edges all weigh the same and hold no loop takes the first
    round's draws by division instead of the prefix search; a loop or a second
    weight keeps the search.  Both must give the oracle's labels and cuts.
    "skewed" puts nearly all the weight on a few edges amid weight-1 edges (and
    an edge count that is no multiple of 8), so most draws land on a few
    edges of the prefix search."""
    rng = np.random.default_rng(len(kind))
    n = 400
    e = erdos_renyi(rng, n, 10.0 / n)
    e[:, 2] = 7 if kind != "one-weight-large" else 4_000_000_000
    if kind == "one-weight-loop":
        e = np.concatenate([e, np.array([[5, 5, 7]], dtype=np.int64)])
    if kind == "one-heavier":
        e[len(e) // 2, 2] = 8
    if kind == "skewed":
        e = e[: len(e) - (len(e) % 8) - 3]
        e[:, 2] = 1
        e[rng.choice(len(e), 9, replace=False), 2] = 1_000_000_000
    t = 60
    _say(f"first-round forms {kind}: n {n} m {len(e)} t {t}")
    vmap, k, _ = ctx.sample_contract(e, n, t, 3, 2)
    lab, nv = orc.sample_contract(n, e, t, 3, 2)
    assert k == nv
    roots = np.flatnonzero(lab == np.arange(n))
    rank = np.zeros(n, dtype=np.int64)
    rank[roots] = np.arange(len(roots))
    assert np.array_equal(vmap, rank[lab])
    raw = ctx.sqrt_cut_trials(e, n, t, 11, 0, 6)
    assert raw.tolist() == [orc.sqrt_trial(n, e, t, 11, k) for k in range(6)]


@pytest.mark.parametrize("n,seed", [(200, 0), (1000, 1)])
def test_sqrt_trials_bit_exact(ctx, orc, n, seed):
    rng = np.random.default_rng(seed)
    e = planted(rng, n, deg=12, wmax=40, c=3)
    m = len(e)
    t = orc.intermediate_size(n, m)
    T = 12
    _say(f"sqrt trials n {n} t {t}")
    raw = ctx.sqrt_cut_trials(e, n, t, 7, 0, T)
    want = [orc.sqrt_trial(n, e, t, 7, k) for k in range(T)]
    assert raw.tolist() == want


def test_sqrt_trials_varying_cuts(ctx, orc):
    # weighted cycle with chords: trial cuts differ, so the comparison pins the
    # draw sequence, the contraction prefix and the KS picks
    rng = np.random.default_rng(5)
    n = 700
    e = [[i, (i + 1) % n, int(rng.integers(10**6, 2 * 10**6))] for i in range(n)]
    e += [[int(a), int(b), int(rng.integers(1, 10**5))] for a, b in rng.integers(0, n, size=(n // 3, 2)) if a != b]
    e = np.array(e, dtype=np.int64)
    t = orc.intermediate_size(n, len(e))
    _say(f"sqrt varying n {n} t {t}")
    raw = ctx.sqrt_cut_trials(e, n, t, 1, 0, 16)
    want = [orc.sqrt_trial(n, e, t, 1, k) for k in range(16)]
    assert len(set(want)) > 1
    assert raw.tolist() == want


def test_sqrt_trials_pr_chain_bit_exact(ctx, orc):
    """G(10^5, 48/n) sampled to t = 3100: one super-vertex takes most of the
    weight, so each level-0 child runs about 900 picks, then a PR pass of about
    2200 merges chained through the heavy row (512-thread contraction blocks).
    Raw trials bit-exact against the oracle."""
    rng = np.random.default_rng(11)
    n, m = 100_000, 2_400_000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    E = np.stack([u[keep], v[keep], rng.integers(1, 101, keep.sum())], 1).astype(np.uint64)
    t = orc.intermediate_size(n, len(E))
    assert 2048 < t <= 4096
    raw = ctx.sqrt_cut_trials(E, n, t, 3, 0, 4)
    want = [orc.sqrt_trial(n, E, t, 3, k) for k in range(4)]
    _say(f"pr chain: n {n} m {len(E)} t {t}: device {raw.tolist()} oracle {want}")
    assert raw.tolist() == want


def test_mincut_er1000(ctx, orc):
    # the square_root configuration: ER(1000, 0.05), unit weights, p = 0.95, seed 0
    rng = np.random.default_rng(0)
    n = 1000
    e = erdos_renyi(rng, n, 0.05, wmax=1)
    _say(f"mincut ER1000 m {len(e)}")
    cut, T = ctx.mincut(e, n, F32_95, 0)
    assert T == orc.sqrt_trials(n, len(e), F32_95)
    assert cut == orc.stoer_wagner(dense(n, e))


def test_mincut_planted(ctx, orc):
    rng = np.random.default_rng(12)
    n = 1500
    e = planted(rng, n, deg=16, wmax=100, c=5)
    _say("mincut planted n 1500")
    cut, T = ctx.mincut(e, n, F32_95, 3)
    assert cut == 5


def test_mincut_reference_fixtures(ctx, orc):
    from pathlib import Path

    from graphfile import read_graph

    gold = Path(__file__).resolve().parent / "golden" / "reference_inputs"
    for name, want in (("cycle_4.graph", 200), ("cycle_6.graph", 10000), ("cycle_10.graph", 10),
                       ("cycle_20.graph", 100), ("complete_4.graph", 85)):
        n, e = read_graph(gold / name)
        _say(f"mincut fixture {name}")
        cut, _ = ctx.mincut(e.astype(np.int64), n, F32_95, 0)
        assert cut == want, name


def test_mincut_rejects_bad_ids(ctx):
    with pytest.raises(cac.CacError) as ei:
        ctx.mincut(np.array([[0, 7, 1]]), 5, 0.9, 0)
    assert ei.value.code == cac.CAC_EINVAL


def test_weight_sum_that_wraps_64_bits_is_rejected(ctx):
    # 2^24 + 2^20 edges of weight 2^40 - 1 weigh about 2^64 + 2^60: a plain u64
    # sum wraps to about 2^60 (< 2^61) and would pass; the high-word counter
    # sees the true total
    m = (1 << 24) + (1 << 20)
    e = np.zeros(m, dtype=cac.EDGE)
    e["u"] = np.arange(m, dtype=np.uint64) % 1000
    e["v"] = (e["u"] + 1) % 1000
    e["w"] = (1 << 40) - 1
    wrapped = (m * ((1 << 40) - 1)) % (1 << 64)
    assert wrapped < (1 << 61)
    _say(f"weight-sum wrap: m {m}")
    with pytest.raises(cac.CacError) as ei:
        ctx.mincut(e, 1000, 0.9, 0)
    assert ei.value.code == cac.CAC_ERANGE
    with pytest.raises(cac.CacError) as ei:
        ctx.sample_contract(e, 1000, 100, 0, 0)
    assert ei.value.code == cac.CAC_ERANGE


def test_sqrt_trial_beyond_dense_engine(ctx, orc):
    # m > 16.7M edges gives t = 2 sqrt(m) + 1 > 8192: the contracted graph goes
    # through the sparse Karger-Stein as a merged edge list (bit-exact raw cut)
    rng = np.random.default_rng(1)
    n, m = 9000, 17_000_000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    E = np.stack([np.minimum(u, v), np.maximum(u, v), rng.integers(1, 10, m)], 1).astype(np.uint64)
    t = orc.intermediate_size(n, m)
    assert t > 8192
    print(f"sqrt trial beyond dense: n={n} m={m} t={t}", flush=True)
    raw = ctx.sqrt_cut_trials(E, n, t, 3, 0, 1)
    assert raw[0] == orc.sqrt_trial(n, E, t, 3, 0)


def test_sqrt_trials_beyond_dense_engine_merge(ctx, orc):
    # n = 10^5 contracted to t = 8248: most edges become loops or parallel
    # edges, so the merge (radix sort of every trial's keys together, sum of
    # parallel weights) does real work; four trials of one call, bit-exact
    rng = np.random.default_rng(4)
    n, m = 100_000, 17_000_000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    E = np.stack([np.minimum(u, v), np.maximum(u, v), rng.integers(1, 10, m)], 1).astype(np.uint64)
    t = orc.intermediate_size(n, m)
    assert t > 8192
    raw = ctx.sqrt_cut_trials(E, n, t, 5, 0, 4)
    for k in range(4):
        want = orc.sqrt_trial(n, E, t, 5, k)
        print(f"sqrt trial {k} beyond dense (merge): n={n} m={m} t={t} device {raw[k]} oracle {want}", flush=True)
        assert raw[k] == want


@pytest.mark.parametrize("n,t", [(1000, 318), (3000, 400)])
def test_hc_rule_bit_exact(orc, n, t):
    """The high-concurrency rule (floor(n^1.05) draws per round) on the device:
    labels, merged graph and raw sqrt-cut trials equal the oracle's."""
    rng = np.random.default_rng(n)
    u = rng.integers(0, n, 12 * n)
    v = rng.integers(0, n, 12 * n)
    keep = u != v
    E = np.stack([u[keep], v[keep], rng.integers(1, 30, keep.sum())], 1).astype(np.uint64)
    dpr = orc.hc_draws(n)
    assert cac.hc_draws(n) == dpr
    with cac.Context(0) as c:
        c.set_sampling_rule(cac.SAMPLE_HC)
        for trial in range(3):
            vmap, k, merged = c.sample_contract(E, n, t, 9, trial)
            lab, nv = orc.sample_contract(n, E, t, 9, trial, dpr)
            assert k == nv
            _, rank = np.unique(lab, return_inverse=True)
            assert np.array_equal(vmap, rank.astype(np.uint32))
        raw = c.sqrt_cut_trials(E, n, t, 9, 0, 4)
        want = [orc.sqrt_trial(n, E, t, 9, k, dpr) for k in range(4)]
        _say(f"hc rule n {n} t {t}: device {raw.tolist()} oracle {want}")
        assert raw.tolist() == want
        c.set_sampling_rule(cac.SAMPLE_LC)
        lc = c.sqrt_cut_trials(E, n, t, 9, 0, 4)
        assert lc.tolist() == [orc.sqrt_trial(n, E, t, 9, k) for k in range(4)]
