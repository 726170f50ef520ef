"""Pins the oracle's min-cut restatement: trial-count and size rules against the
reference library (oracle/_ref) and its published closed forms, the pick and draw
distributions against the textbook (chi-square), raw trial cuts against exact
Stoer-Wagner, final cuts against the reference's own KS and deterministic cut."""
import numpy as np
import pytest
from graphgen import dense, erdos_renyi, necklace, planted
from scipy.stats import chisquare

F32_95 = float(np.float32(0.95))  # the reference parses p with std::stof

# closed forms (co_mincut.cpp:18-74, SquareRootCut.cpp:63-83) evaluated for the configs
KS_SIZES_8192 = [8192, 5794, 4098, 2899, 2051, 1452, 1028, 728, 516, 366, 260, 185, 132, 128]
KS_TRIALS_95 = {256: 5, 1024: 8, 4096: 11, 8192: 13, 65536: 18, 10**6: 23}


def test_ks_sizes_and_trials(orc):
    s, seq = 8192, [8192]
    while s > 128:
        s = orc.ks_target(s)
        seq.append(s)
    assert seq == KS_SIZES_8192
    for n, t in KS_TRIALS_95.items():
        assert orc.ks_trials(n, 0.95) == t
    assert orc.ks_trials(100, 0.95) == 1 and orc.ks_trials(5000, 0.0) == 1
    assert orc.ks_min_success(128) == 1.0


def test_sqrt_cut_rules(orc):
    # C1: ER(1000, 0.05) with m = 24975 -> t = 318, T = 59 at p = 0.95 (float)
    assert orc.intermediate_size(1000, 24975) == 318
    assert orc.sqrt_trials(1000, 24975, F32_95) == 59
    assert abs(orc.ks_min_success(318) - 0.5165405273) < 1e-9
    for n, T in ((1024, 158), (4096, 853), (65536, 20427)):
        assert orc.sqrt_trials(n, 8 * n, F32_95) == T
    assert orc.intermediate_size(50, 10**6) == 50
    assert orc.approx_trials(F32_95) == 5


def test_rules_match_reference_library(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    R = orc.ref()
    for s in list(range(120, 400)) + [1000, 4097, 8192, 65536, 10**6]:
        assert orc.ks_target(s) == R.ref_ks_target(s)
        assert orc.ks_min_success(s) == R.ref_ks_min_success(s)
        for p in (0.5, 0.9, F32_95, 0.99):
            assert orc.ks_trials(s, p) == R.ref_ks_trials(s, p)


@pytest.mark.parametrize("seed", range(3))
def test_first_pick_distribution(orc, seed):
    # Karger's rule: an edge is contracted with probability proportional to its weight
    rng = np.random.default_rng(seed)
    n = 7
    e = erdos_renyi(rng, n, 0.6, wmax=9)
    A = dense(n, e)
    counts = {}
    N = 12000
    for t in range(N):
        ab = orc.ks_first_pick(A, 100 + seed, t, 5)
        assert A[ab] > 0
        counts[ab] = counts.get(ab, 0) + 1
    pairs = [(a, b) for a in range(n) for b in range(a) if A[a, b] > 0]
    obs = np.array([counts.get(p, 0) for p in pairs], dtype=float)
    exp = np.array([A[p] for p in pairs], dtype=float)
    exp = exp / exp.sum() * N
    assert chisquare(obs, exp).pvalue > 1e-4


def test_first_draw_distribution(orc):
    # t = n - 1: exactly one union, made by the round's first draw
    rng = np.random.default_rng(11)
    n = 6
    e = erdos_renyi(rng, n, 0.7, wmax=7)
    counts = {}
    N = 12000
    for t in range(N):
        lab, nv = orc.sample_contract(n, e, n - 1, 3, t)
        assert nv == n - 1
        merged = tuple(int(v) for v in np.flatnonzero(lab != np.arange(n)))
        assert len(merged) == 1
        v = merged[0]
        key = (v, int(lab[v]))
        counts[key] = counts.get(key, 0) + 1
    A = dense(n, e)
    pairs = [(a, b) for a in range(n) for b in range(a) if A[a, b] > 0]
    obs = np.array([counts.get(p, 0) for p in pairs], dtype=float)
    exp = np.array([A[p] for p in pairs], dtype=float)
    exp = exp / exp.sum() * N
    assert chisquare(obs, exp).pvalue > 1e-4


@pytest.mark.parametrize("seed", range(4))
def test_ks_trial_raw_cut_bounds(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([150, 260, 400]))
    e = planted(rng, n, deg=12, wmax=20, c=int(rng.integers(1, 6)))
    A = dense(n, e)
    exact = orc.stoer_wagner(A)
    bound = orc.min_degree(A)
    hits = 0
    for t in range(6):
        raw = orc.ks_trial(A, seed, t)
        assert exact <= raw <= bound
        hits += raw == exact
    assert hits >= 2  # per-trial success probability is far above 1/3 here


def test_necklace_trials_vary_and_stay_above_the_cut(orc):
    """The graph family the device test of the PR pass uses: the oracle's raw
    cuts differ between trials (so they pin the order of merges, PR merges
    included), every one lies between the exact minimum cut and the minimum
    weighted degree, and the heavy edges reach that degree (so the PR pass
    merges)."""
    rng = np.random.default_rng(900 + 225)
    A = dense(900, necklace(rng, 900, 225, 3 * 10**6, 2.0))
    exact = orc.stoer_wagner(A)
    bound = orc.min_degree(A)
    assert bound <= 3 * 10**6
    raw = [orc.ks_trial(A, 5, t) for t in range(8)]
    assert len(set(raw)) > 1
    assert all(exact <= r <= bound for r in raw)


@pytest.mark.parametrize("seed", range(3))
def test_final_cut_matches_reference(orc, seed):
    rng = np.random.default_rng(50 + seed)
    n = 300
    c = int(rng.integers(1, 8))
    e = planted(rng, n, deg=10, wmax=100, c=c)
    A = dense(n, e)
    T = orc.ks_trials(n, 0.95)
    best = min(orc.ks_trial(A, 9, t) for t in range(T))
    assert best == orc.stoer_wagner(A) == c
    if orc.ref_available():
        assert orc.ref_deterministic_mincut(n, e) == c
        assert orc.ref_ks_mincut(n, e, 0.95, 9) == c


def test_sample_contract_invariants(orc):
    rng = np.random.default_rng(5)
    n = 400
    e = erdos_renyi(rng, n, 0.05, wmax=30)
    for t_target, trial in ((n, 0), (200, 1), (37, 2), (1, 3)):
        lab, nv = orc.sample_contract(n, e, t_target, 1, trial)
        assert nv == t_target
        assert np.all(lab <= np.arange(n)) and np.array_equal(lab[lab], lab)
        assert len(set(lab.tolist())) == t_target
        # every super-vertex is connected through its own edges
        inner = e[lab[e[:, 0]] == lab[e[:, 1]]]
        from scipy.sparse.csgraph import connected_components
        k, _ = connected_components(dense(n, inner) > 0, directed=False)
        assert k == t_target
        M = orc.densify(n, e, lab)
        assert M.shape == (t_target, t_target)
        assert np.array_equal(M, M.T) and np.all(np.diag(M) == 0)
        assert M.sum() == 2 * e[lab[e[:, 0]] != lab[e[:, 1]], 2].sum()


def test_sqrt_trials_find_planted_cut(orc):
    rng = np.random.default_rng(8)
    n = 600
    e = planted(rng, n, deg=16, wmax=100, c=3)
    m = len(e)
    t = orc.intermediate_size(n, m)
    T = orc.sqrt_trials(n, m, F32_95)
    cuts = [orc.sqrt_trial(n, e, t, 4, k) for k in range(T)]
    assert min(cuts) == 3
    assert all(c >= 3 for c in cuts)


def test_keep_threshold_closed_forms(orc):
    # Q_i(w) ~ (1 - 2^-i)^w * 2^64 from below; floor rounding in the squarings
    # compounds to a relative error far below 2^-40 (a wrong base or a missed
    # multiply would be off by orders of magnitude)
    from fractions import Fraction

    assert orc.keep_threshold(1, 1) == 1 << 63
    assert orc.keep_threshold(2, 1) == 3 << 62
    assert orc.keep_threshold(5, 0) == (1 << 64) - 1
    from decimal import Decimal, getcontext

    getcontext().prec = 80
    for i in (1, 2, 3, 7, 20, 40, 63):
        for w in (1, 2, 3, 5, 100, 12345, 10**6):
            q = orc.keep_threshold(i, w)
            if w <= 100:
                exact = Fraction(2**i - 1, 2**i) ** w * 2**64
                assert q <= exact and exact - q <= exact / 2**40 + 4
            else:  # (1 - 2^-i)^w = exp(w ln(1 - 2^-i)) to 80 digits
                exact = (Decimal(w) * (Decimal(1) - Decimal(2) ** -i).ln()).exp() * (Decimal(2) ** 64)
                assert Decimal(q) <= exact and exact - Decimal(q) <= exact / Decimal(2) ** 40 + 4


def test_approx_oracle_is_log_approximation(orc):
    rng = np.random.default_rng(4)
    for c in (1, 3, 8):
        e = planted(rng, 300, deg=10, wmax=100, c=c)
        a, t = orc.approx_cut(300, e, F32_95, 1)
        assert a & (a - 1) == 0 and t % 5 == 0
        assert c / 64 <= a <= 64 * c
