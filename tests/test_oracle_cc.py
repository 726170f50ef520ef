"""Pins the CC oracle: reference fixtures (component lists stated in their
header lines), scipy's connected_components, BFS on tiny graphs and the
label invariants; plus the R-MAT generator spec against a direct restatement."""
import re
from pathlib import Path

import numpy as np
import pytest
from graphfile import read_graph

GOLD = Path(__file__).resolve().parent / "golden" / "reference_inputs"


def _fixture_partition(path):
    # "#5 connected components (0, 1, 3)  (2, 4) and the rest are singletons"
    head = open(path).readline()
    groups = [tuple(int(x) for x in g.split(",")) for g in re.findall(r"\(([\d,\s]+)\)", head)]
    count = int(re.match(r"#(\d+)", head).group(1))
    return count, groups


@pytest.mark.parametrize("name", ["cc_8.graph", "cc_13.graph"])
def test_cc_reference_fixtures(orc, name):
    n, e = read_graph(GOLD / name)
    count, groups = _fixture_partition(GOLD / name)
    lab, nc = orc.cc(e[:, :2].astype(np.uint32), n, 0)
    assert nc == count
    for g in groups:
        assert len({int(lab[v]) for v in g}) == 1 and int(lab[g[0]]) == min(g)
    inside = {v for g in groups for v in g}
    for v in range(n):
        if v not in inside:
            assert lab[v] == v
    rank, nc2 = orc.cc(e[:, :2].astype(np.uint32), n, 1)
    assert nc2 == count and sorted(set(rank.tolist())) == list(range(count))


def _bfs_labels(n, edges):
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    lab = [-1] * n
    for s in range(n):
        if lab[s] >= 0:
            continue
        stack, comp = [s], [s]
        lab[s] = s
        while stack:
            x = stack.pop()
            for y in adj[x]:
                if lab[y] < 0:
                    lab[y] = s
                    stack.append(y)
    return np.array(lab, dtype=np.uint32)


@pytest.mark.parametrize("seed", range(6))
def test_cc_bfs_small(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    m = int(rng.integers(0, 2 * n))
    e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    lab, nc = orc.cc(e, n, 0)
    assert np.array_equal(lab, _bfs_labels(n, e.tolist()))  # BFS from increasing s labels by min id
    assert nc == len(set(lab.tolist()))


@pytest.mark.parametrize("seed", range(3))
def test_cc_scipy(orc, seed):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    rng = np.random.default_rng(100 + seed)
    n = 20000
    m = 15000
    e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    lab, nc = orc.cc(e, n, 0)
    A = coo_matrix((np.ones(m), (e[:, 0], e[:, 1])), shape=(n, n))
    k, sl = connected_components(A, directed=False)
    assert nc == k
    # same partition: map scipy's label -> min vertex
    mins = np.full(k, n, dtype=np.int64)
    np.minimum.at(mins, sl, np.arange(n))
    assert np.array_equal(lab.astype(np.int64), mins[sl])
    # invariants
    assert np.all(lab <= np.arange(n))
    assert np.array_equal(lab[lab], lab)
    assert np.all(lab[e[:, 0]] == lab[e[:, 1]])


def test_cc_edge_cases(orc):
    lab, nc = orc.cc(np.zeros((0, 2), np.uint32), 5, 0)
    assert nc == 5 and lab.tolist() == [0, 1, 2, 3, 4]
    lab, nc = orc.cc(np.array([[3, 3], [1, 1]], np.uint32), 4, 0)
    assert nc == 4
    lab, nc = orc.cc(np.array([[3, 2], [2, 3], [1, 3]], np.uint32), 4, 1)
    assert nc == 2 and lab.tolist() == [0, 1, 1, 1]


def _rmat_direct(orc, scale, m, seed, th):
    """Literal restatement of the R-MAT spec in Python (draws, dedup, sort)."""
    ta, tab, tabc = th
    seen, keys, d = set(), [], 0
    while len(keys) < m:
        u = v = 0
        w = None
        for lvl in range(scale):
            if lvl % 4 == 0:
                w = orc.philox4x32([d & 0xFFFFFFFF, lvl // 4, d >> 32, 0], [seed, 0x524D4154])
            r = int(w[lvl % 4])
            bu, bv = (0, 0) if r < ta else (0, 1) if r < tab else (1, 0) if r < tabc else (1, 1)
            u, v = 2 * u + bu, 2 * v + bv
        d += 1
        if u == v:
            continue
        k = (min(u, v), max(u, v))
        if k not in seen:
            seen.add(k)
            keys.append(k)
    return np.array(sorted(keys), dtype=np.uint32), d


def test_rmat_oracle_matches_direct(orc):
    import cac

    th = cac.rmat_thresholds(0.45, 0.22, 0.22)
    assert th == (1932735283, 2877628088, 3822520893)
    e, d = orc.rmat(8, 600, 3, th)
    want, dw = _rmat_direct(orc, 8, 600, 3, th)
    assert d == dw and np.array_equal(e, want)


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("seed", [0, 1])
def test_cc_multithreaded_equals_sequential(orc, threads, seed):
    # the multithreaded CPU baseline gives orc_cc's labels and count in both
    # label forms: many small components, one giant, loops and repeats
    rng = np.random.default_rng(300 + seed)
    n = 40000
    e = np.concatenate([rng.integers(0, n, size=(30000, 2)),
                        rng.integers(0, n // 8, size=(60000, 2)),
                        np.repeat(np.arange(50)[:, None], 2, axis=1)]).astype(np.uint32)
    for mode in (0, 1):
        want, wn = orc.cc(e, n, mode)
        got, gn = orc.cc_mt(e, n, mode, threads)
        assert gn == wn
        assert np.array_equal(got, want)
