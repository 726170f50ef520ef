"""Pins the oracle's sparse Karger-Stein restatement (co_mincut.h:74-155):
the race key against the logarithm it approximates and the exponential race it
stands for, one contraction against scipy's connected components on the
(key, index) prefix, whole trials against exact minimum cuts, the reference's
own sparse Karger-Stein, and the dense engine where the sparse path hands over."""
import math

import numpy as np
import pytest
from graphgen import dense, planted
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

SPRS = 0x53505253


def _words(rng, k):
    hi = rng.integers(0, 2**32, size=k, dtype=np.uint64)
    lo = rng.integers(0, 2**32, size=k, dtype=np.uint64)
    return [int(h) << 32 | int(l) for h, l in zip(hi, lo)]


def test_race_key_is_the_logarithm(orc):
    # w = 1: key = L << 20 exactly, L = -log2(U / 2^53) in Q32.32 within 2^-29
    rng = np.random.default_rng(0)
    xs = _words(rng, 4000) + [0, 2**11 - 1, 2**11, 2**63, 2**64 - 2**11, 2**64 - 1]
    for x in xs:
        key = orc.race_key(x, 1)
        assert key & ((1 << 20) - 1) == 0
        L = key >> 20
        U = (x >> 11) + 1
        want = 53.0 - math.log2(U)
        assert abs(L / 2**32 - want) <= 2**-29, (x, L / 2**32, want)
    assert orc.race_key(2**64 - 1, 1) == 0  # U = 2^53: L = 0
    assert orc.race_key(0, 1) == 53 << 52  # U = 1: L = 53


def test_race_key_weight_scaling(orc):
    rng = np.random.default_rng(1)
    for x in _words(rng, 500):
        base = orc.race_key(x, 1)
        for w in (1, 2, 3, 7, 1000, 2**20 + 1, 2**40 - 1):
            assert orc.race_key(x, w) == base // w


@pytest.mark.parametrize("w1,w2", [(1, 1), (1, 3), (5, 2), (1, 40)])
def test_exponential_race(orc, w1, w2):
    # P(edge 1 first) = w1 / (w1 + w2) for independent Exp(w) keys
    rng = np.random.default_rng(w1 * 100 + w2)
    N = 20000
    a, b = _words(rng, N), _words(rng, N)
    wins = sum(orc.race_key(x, w1) < orc.race_key(y, w2) for x, y in zip(a, b))
    p = w1 / (w1 + w2)
    assert abs(wins / N - p) < 4.5 * math.sqrt(p * (1 - p) / N)


def test_triangle_contraction(orc):
    # contracting a triangle to 2 vertices removes the first edge of the race and
    # keeps the other two, in their order, as parallel edges (0, 1)
    w = [1, 2, 5]
    E = np.array([[0, 1, w[0]], [0, 2, w[1]], [1, 2, w[2]]], dtype=np.uint64)
    N = 6000
    first = np.zeros(3)
    for t in range(N):
        C = orc.sparse_contract(3, E, 2, 11, t, 2)
        assert C.shape == (2, 3)
        assert (C[:, 0] == 0).all() and (C[:, 1] == 1).all()
        kept = list(C[:, 2])
        gone = [i for i in range(3) if w[i] not in kept]
        assert len(gone) == 1 and kept == [w[i] for i in range(3) if i != gone[0]]
        first[gone[0]] += 1
    p = np.array(w) / sum(w)
    assert np.all(np.abs(first / N - p) < 4.5 * np.sqrt(p * (1 - p) / N))


def _keys(orc, m, E, seed, trial, nid):
    k = np.empty(m, dtype=np.uint64)
    for e in range(m):
        o = orc.philox4x32([trial, nid, e & 0xFFFFFFFF, e >> 32], [seed, SPRS])
        k[e] = orc.race_key(int(o[1]) << 32 | int(o[0]), int(E[e, 2]))
    return k


def _rank_labels(lab):
    # relabel components by the rank of their minimum vertex id
    first = {}
    out = np.empty(len(lab), dtype=np.int64)
    for v, l in enumerate(lab):
        out[v] = first.setdefault(int(l), len(first))
    return out


@pytest.mark.parametrize("seed", range(4))
def test_contraction_is_the_key_prefix(orc, seed):
    rng = np.random.default_rng(seed)
    n, m, tgt, nid, trial = 300, 900, 70, 5, 3
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    u, v = np.minimum(u, v), np.maximum(u, v)
    w = rng.permutation(m) + 1  # distinct weights identify every edge in the child
    E = np.stack([u, v, w], 1).astype(np.uint64)
    C = orc.sparse_contract(n, E, tgt, seed, trial, nid)
    # shortest (key, index) prefix whose union has tgt components (scipy CC)
    key = _keys(orc, m, E, seed, trial, nid)
    order = np.lexsort((np.arange(m), key))

    def comps(p):
        sel = order[:p]
        A = coo_matrix((np.ones(p), (u[sel], v[sel])), shape=(n, n))
        return connected_components(A, directed=False)

    lo, hi = 0, m
    assert comps(m)[0] <= tgt
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if comps(mid)[0] <= tgt:
            hi = mid
        else:
            lo = mid
    k, lab = comps(hi)
    assert k == tgt
    lab = _rank_labels(lab)
    want = [(min(lab[a], lab[b]), max(lab[a], lab[b]), c) for a, b, c in zip(u, v, w) if lab[a] != lab[b]]
    assert [tuple(int(x) for x in r) for r in C] == want


def test_disconnected_contraction(orc):
    E = np.array([[0, 1, 1], [2, 3, 1]], dtype=np.uint64)
    assert orc.sparse_contract(4, E, 1, 0, 0, 2) is None
    assert orc.ks_sparse_trial(4, E, 0, 0) == 0


@pytest.mark.parametrize("n,c", [(2600, 2), (3400, 3)])
def test_sparse_trials_find_planted_cut(orc, n, c):
    rng = np.random.default_rng(n)
    E = np.asarray(planted(rng, n, 14.0, 20, c), dtype=np.uint64)
    raws = [orc.ks_sparse_trial(n, E, 3, t) for t in range(6)]
    assert min(raws) == c and all(r >= c for r in raws)
    if orc.ref_available():
        assert orc.ref_ks_mincut(n, E, 0.9, 3) == c  # the reference's own sparse path


@pytest.mark.parametrize("n", [700, 2048])
def test_small_graph_hands_over_at_the_root(orc, n):
    rng = np.random.default_rng(n)
    E = np.asarray(planted(rng, n, 8.0, 9, 2), dtype=np.uint64)
    A = dense(n, E)
    for t in range(3):
        assert orc.ks_sparse_trial(n, E, 9, t) == orc.ks_trial(A, 9, t)


def test_dense_enough_graph_hands_over_at_the_root(orc):
    # m >= n^2 / 32 with 2048 < n <= 8192: the dense engine from node 1
    n = 2100
    rng = np.random.default_rng(3)
    E = np.asarray(planted(rng, n, n / 14, 5, 3), dtype=np.uint64)
    assert len(E) >= n * n // 32
    A = dense(n, E)
    assert orc.ks_sparse_trial(n, E, 4, 1) == orc.ks_trial(A, 4, 1)
