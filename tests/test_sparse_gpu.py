"""Sparse Karger-Stein on the device against the oracle: raw cut of every trial
bit-exact (race keys, halving search, relabelling, handover to the dense engine
under the node's id), disconnected inputs, the karger_stein dense switch, input
checks, and a deep recursion checked by size-independent properties."""
import numpy as np
import pytest
from graphgen import dense, planted

import cac

pytestmark = pytest.mark.gpu


def _noisy(rng, n, m, wmax):
    """Random multigraph (parallel edges, both endpoint orders) plus a path, so connected."""
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    u = np.concatenate([u, np.arange(n - 1)])
    v = np.concatenate([v, np.arange(1, n)])
    w = rng.integers(1, wmax + 1, len(u))
    return np.stack([u, v, w], 1).astype(np.uint64)


def _perturb(rng, E):
    """Planted graph with 10% of its edges repeated (parallel edges), the list
    shuffled and 30% of the endpoints swapped: the cut values move, the oracle
    sees the same list."""
    E = np.asarray(E, dtype=np.uint64)
    E = np.concatenate([E, E[rng.random(len(E)) < 0.1]])
    E = E[rng.permutation(len(E))]
    sw = rng.random(len(E)) < 0.3
    E[sw, 0], E[sw, 1] = E[sw, 1].copy(), E[sw, 0].copy()
    return E


def _canon(E):
    return np.stack([np.minimum(E[:, 0], E[:, 1]), np.maximum(E[:, 0], E[:, 1]), E[:, 2]], 1)


CASES = [
    # (name, n, builder): planted cuts below the minimum degree, so raw cuts vary by trial
    ("planted-2500", 2500, lambda rng: _perturb(rng, planted(rng, 2500, 14.0, 7, 3))),
    ("planted-heavy-3000", 3000, lambda rng: _perturb(rng, planted(rng, 3000, 16.0, 2**30, 4))),
    ("planted-4000", 4000, lambda rng: np.asarray(planted(rng, 4000, 14.0, 20, 3), dtype=np.uint64)),
    ("noisy-unit-5200", 5200, lambda rng: _noisy(rng, 5200, 3 * 5200, 1)),
]


def _canon_random(rng, n, m, wmax=None, w=None):
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    v = np.where(u == v, (v + 1) % n, v)
    ww = np.full(m, w) if w is not None else rng.integers(1, wmax + 1, m)
    return np.stack([np.minimum(u, v), np.maximum(u, v), ww], 1).astype(np.uint64)


CONTRACT_CASES = [
    # (name, n, m, tgt, weights): the halving search, the final LDS sort, ties
    ("few-candidates", 500, 800, 300, dict(wmax=9)),
    ("medium", 3000, 20000, 2123, dict(wmax=1000)),
    ("large-unit", 50000, 400000, 35357, dict(w=1)),
    # weight 2^40 - 1 leaves ~4096 key values below L = 1, so keys collide:
    # ties sorted in LDS, then a boundary key shared by more than 1024 edges
    ("ties", 3000, 2_000_000, 2000, dict(w=2**40 - 1)),
    ("ties-over-lds", 3000, 8_000_000, 2000, dict(w=2**40 - 1)),
    ("no-contraction", 400, 1000, 400, dict(wmax=5)),
]


@pytest.mark.parametrize("name,n,m,tgt,wts", CONTRACT_CASES, ids=[c[0] for c in CONTRACT_CASES])
def test_one_contraction_matches_oracle(ctx, orc, name, n, m, tgt, wts):
    rng = np.random.default_rng(m)
    E = _canon_random(rng, n, m, **wts)
    print(f"sparse contract {name}: n={n} m={m} tgt={tgt}", flush=True)
    got = ctx.sparse_contract(E, n, tgt, 13, 4, 6)
    want = orc.sparse_contract(n, E, tgt, 13, 4, 6)
    assert got.shape == want.shape and np.array_equal(got, want)
    assert (int(got[:, 1].max()) + 1 if len(got) else 0) <= tgt


def test_contraction_of_disconnected_graph(ctx, orc):
    E = np.array([[0, 1, 1], [2, 3, 1], [1, 4, 2]], dtype=np.uint64)
    with pytest.raises(cac.CacError) as ei:
        ctx.sparse_contract(E, 5, 1, 0, 0, 2)
    assert ei.value.code == cac.CAC_ENOTCONN
    assert orc.sparse_contract(5, E, 1, 0, 0, 2) is None
    got = ctx.sparse_contract(E, 5, 2, 0, 0, 2)
    assert np.array_equal(got, orc.sparse_contract(5, E, 2, 0, 0, 2))


@pytest.mark.parametrize("name,n,build", CASES, ids=[c[0] for c in CASES])
def test_sparse_trials_match_oracle(ctx, orc, name, n, build):
    rng = np.random.default_rng(n)
    E = build(rng)
    seed, first, T = 21, 2, 3
    print(f"sparse ks {name}: n={n} m={len(E)}", flush=True)
    raw, mn = ctx.ks_sparse(E, n, seed, first, T)
    want = [orc.ks_sparse_trial(n, _canon(E), seed, first + t) for t in range(T)]
    assert list(raw) == want and mn == min(want)


def test_deep_recursion_matches_oracle(ctx, orc):
    # 16384 -> 11586 -> ... -> 1452: seven sparse levels, 128 dense roots per trial
    n = 16384
    rng = np.random.default_rng(7)
    E = _perturb(rng, planted(rng, n, 14.0, 50, 2))
    print(f"sparse ks deep: n={n} m={len(E)}", flush=True)
    raw, _ = ctx.ks_sparse(E, n, 5, 0, 2)
    want = [orc.ks_sparse_trial(n, _canon(E), 5, t) for t in range(2)]
    assert list(raw) == want


@pytest.mark.parametrize("n", [600, 2048])
def test_small_graph_equals_dense_engine(ctx, n):
    rng = np.random.default_rng(n)
    E = np.asarray(planted(rng, n, 8.0, 9, 2), dtype=np.uint64)
    A = dense(n, E)
    raw_s, _ = ctx.ks_sparse(E, n, 9, 0, 4)
    raw_d, _ = ctx.ks_dense(A, 9, 0, 4)
    assert list(raw_s) == list(raw_d)


@pytest.mark.parametrize("n", [1537, 4099])
def test_handover_sums_parallel_edges(ctx, n):
    """A node handed to the dense engine at once (n <= 2048, or m >= n^2/32 at
    n <= 8192) is built band by band in LDS: n = 1537 leaves a last band of
    2 rows (bands of 5), n = 4099 has one row per band.  Repeated edges in
    both endpoint orders must add up in the matrix exactly as the dense
    engine's own input: raw trials equal ks_dense on the summed matrix."""
    rng = np.random.default_rng(n)
    if n < 2048:
        E = _perturb(rng, planted(rng, n, 8.0, 9, 2))
    else:
        E = _noisy(rng, n, n * n // 20, 7)  # m >= n^2/32: handed over by density
    A = dense(n, _canon(E))
    raw_s, _ = ctx.ks_sparse(E, n, 4, 0, 3)
    raw_d, _ = ctx.ks_dense(A, 4, 0, 3)
    assert list(raw_s) == list(raw_d)


def test_disconnected_is_zero(ctx, orc):
    rng = np.random.default_rng(3)
    a = _noisy(rng, 2500, 5000, 3)
    b = _noisy(rng, 2500, 5000, 3)
    b[:, :2] += 2500
    E = np.concatenate([a, b])
    raw, mn = ctx.ks_sparse(E, 5000, 1, 0, 2)
    assert mn == 0 and list(raw) == [0, 0]
    assert orc.ks_sparse_trial(5000, _canon(E), 1, 0) == 0


def test_karger_stein_switch(ctx, orc):
    p = 0.9
    # sparse input: T trials of the sparse engine
    n = 3000
    rng = np.random.default_rng(11)
    E = np.asarray(planted(rng, n, 14.0, 20, 2), dtype=np.uint64)
    cut, T = ctx.karger_stein(E, n, p, 4)
    assert T == orc.ks_trials(n, p)
    raw, mn = ctx.ks_sparse(E, n, 4, 0, T)
    assert cut == mn == 2
    # dense input (m >= n^2/32): the dense engine on the summed matrix
    n = 300
    E = np.asarray(planted(rng, n, 60.0, 9, 3), dtype=np.uint64)
    assert len(E) >= n * n // 32
    cut, T = ctx.karger_stein(E, n, p, 4)
    _, mn = ctx.ks_dense(dense(n, E), 4, 0, T)
    assert cut == mn == 3
    # p = 0: one trial (the reference's two-argument SequentialKargerSteinCut)
    assert ctx.karger_stein(E, n, 0.0, 4)[1] == 1


def test_rejects_bad_edges(ctx):
    E = np.array([[0, 1, 1], [1, 2, 1]], dtype=np.uint64)
    for bad, code in (([[0, 0, 1]], cac.CAC_EINVAL), ([[0, 9, 1]], cac.CAC_EINVAL), ([[0, 2, 0]], cac.CAC_ERANGE),
                      ([[0, 2, 2**40]], cac.CAC_ERANGE)):
        with pytest.raises(cac.CacError) as ei:
            ctx.ks_sparse(np.concatenate([E, np.asarray(bad, dtype=np.uint64)]), 3, 0, 0, 1)
        assert ei.value.code == code
    # the context stays usable
    raw, mn = ctx.ks_sparse(E, 3, 0, 0, 1)
    assert mn == 1


def test_large_planted_properties(ctx):
    # 2^17 vertices, ~1.1M edges: every raw cut is at least the planted cut and
    # at most the minimum degree bound; repeated calls agree
    descs, edges, c = cac.gen_planted(0, 1, 77, 1 << 17, 1 << 17, 16.0, 50, 4)
    n, m = int(descs[0]["n"]), int(descs[0]["m"])
    E = edges[:m]
    print(f"sparse ks large: n={n} m={m} planted={int(c[0])}", flush=True)
    deg = np.zeros(n, dtype=np.int64)
    np.add.at(deg, E["u"].astype(np.int64), E["w"].astype(np.int64))
    np.add.at(deg, E["v"].astype(np.int64), E["w"].astype(np.int64))
    raw, mn = ctx.ks_sparse(E, n, 3, 0, 2)
    assert all(int(c[0]) <= r <= deg.min() for r in raw)
    raw2, _ = ctx.ks_sparse(E, n, 3, 0, 2)
    assert list(raw) == list(raw2)
