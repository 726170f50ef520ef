"""The planted-cut batch generator (cac_gen_planted, host-side): sizes, the
certificate it promises (connected halves, every weighted degree above c, c
unit cross edges), determinism, and the planted cut against exact
Stoer-Wagner (the oracle) on small graphs."""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import cac


def _graph(descs, edges, i):
    d = descs[i]
    return int(d["n"]), edges[int(d["edge_offset"]): int(d["edge_offset"] + d["m"])]


def test_sizes_and_certificate():
    descs, edges, planted = cac.gen_planted(0, 48, 3, 256, 4096)
    n = descs["n"].astype(np.int64)
    assert np.all(n % 2 == 0) and n.min() >= 256 and n.max() <= 4096
    assert np.all(descs["trials"] == 0)
    assert np.all((planted >= 1) & (planted <= 8))
    assert descs["edge_offset"][0] == 0
    assert np.all(descs["edge_offset"][1:] == descs["edge_offset"][:-1] + descs["m"][:-1])
    avg = (2 * descs["m"].sum()) / n.sum()
    assert 15.0 < avg < 17.0
    for i in range(len(descs)):
        nv, e = _graph(descs, edges, i)
        h = nv // 2
        u, v, w = e["u"].astype(np.int64), e["v"].astype(np.int64), e["w"].astype(np.int64)
        assert np.all(u < v) and np.all(v < nv)
        cross = (u < h) & (v >= h)
        assert cross.sum() == planted[i] and np.all(w[cross] == 1)
        assert np.all((w[~cross] >= 1) & (w[~cross] <= 100))
        assert len(np.unique(u * nv + v)) == len(u)  # simple graph
        deg = np.bincount(u, w, nv) + np.bincount(v, w, nv)
        assert deg.min() > planted[i]
        inner = ~cross
        g = coo_matrix((np.ones(inner.sum()), (u[inner], v[inner])), shape=(nv, nv))
        nc, lab = connected_components(g, directed=False)
        assert nc == 2 and len(set(lab[:h])) == 1 and len(set(lab[h:])) == 1


def test_deterministic_and_addressable():
    a = cac.gen_planted(0, 10, 7, 256, 2048)
    b = cac.gen_planted(0, 10, 7, 256, 2048)
    c = cac.gen_planted(6, 3, 7, 256, 2048)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    for k in range(3):
        na, ea = _graph(a[0], a[1], 6 + k)
        nc_, ec = _graph(c[0], c[1], k)
        assert na == nc_ and np.array_equal(ea, ec) and a[2][6 + k] == c[2][k]


def test_planted_cut_is_the_min_cut(orc):
    descs, edges, planted = cac.gen_planted(0, 12, 11, 256, 600)
    for i in range(len(descs)):
        nv, e = _graph(descs, edges, i)
        adj = np.zeros((nv, nv), dtype=np.int64)
        np.add.at(adj, (e["u"], e["v"]), e["w"].astype(np.int64))
        adj += adj.T
        assert orc.stoer_wagner(adj) == planted[i]


def test_size_distribution_is_log_uniform():
    ns = np.empty(4096, dtype=np.uint32)
    import ctypes

    b = ctypes.c_uint64()
    assert cac.load_library().cac_gen_planted_sizes(0, 4096, 0, 256, 65536, 16.0, 8,
                                                    ns.ctypes.data_as(cac._pu32), ctypes.byref(b)) == 0
    x = (np.log(ns.astype(np.float64)) - np.log(256)) / (np.log(65536) - np.log(256))
    hist = np.histogram(x, bins=8, range=(0, 1))[0]
    assert hist.min() > 4096 / 8 * 0.8 and hist.max() < 4096 / 8 * 1.2
    assert b.value >= 8 * ns.astype(np.int64).sum()


def test_gnp_generator():
    e = cac.gen_gnp(5000, 12.0, 3, 7, wmax=100)
    u, v = e["u"].astype(np.int64), e["v"].astype(np.int64)
    assert np.all(u < v) and np.all(v < 5000)
    assert np.all(np.diff(u * 5000 + v) > 0)  # sorted, no duplicates
    assert abs(2 * len(e) / 5000 - 12.0) < 0.5
    assert e["w"].min() >= 1 and e["w"].max() <= 100
    assert np.array_equal(e, cac.gen_gnp(5000, 12.0, 3, 7, wmax=100))
    assert not np.array_equal(e[:100], cac.gen_gnp(5000, 12.0, 3, 8, wmax=100)[:100])
    # pair frequencies: each of the n(n-1)/2 pairs with probability d/(n-1)
    counts = [len(cac.gen_gnp(400, 8.0, 11, s)) for s in range(40)]
    mean = 400 * 8.0 / 2
    assert abs(np.mean(counts) - mean) < 4 * np.sqrt(mean / 40)


def test_ws_generator():
    n, d = 2000, 8
    e = cac.gen_ws(n, d, 5)
    u, v = e["u"].astype(np.int64), e["v"].astype(np.int64)
    assert len(e) == n * d and np.all(u < v) and np.all(v < n) and np.all(e["w"] == 100)
    assert np.array_equal(e, cac.gen_ws(n, d, 5))
    # beta = 0: the ring lattice, edge (u, j) joins u and u + j mod n
    r = cac.gen_ws(n, d, 5, beta=0.0)
    src = np.repeat(np.arange(n), d)
    dst = (src + np.tile(np.arange(1, d + 1), n)) % n
    assert np.array_equal(r["u"], np.minimum(src, dst)) and np.array_equal(r["v"], np.maximum(src, dst))
    # beta = 0.3: about 30% of the lattice edges rewired to uniform endpoints
    moved = (e["u"] != r["u"]) | (e["v"] != r["v"])
    N = n * d
    assert abs(moved.mean() - 0.3 * (1 - 1 / (n - 1))) < 4.5 * np.sqrt(0.21 / N)
    k, _ = connected_components(coo_matrix((np.ones(N), (u, v)), shape=(n, n)), directed=False)
    assert k == 1


def test_hard_family_structure():
    """Hard family: halves of k weighted Hamiltonian cycles (every vertex of a
    half has the same weighted degree delta), c cross edges of total
    C = floor(0.9 min delta), the same sizes as the standard family."""
    k = 8
    descs, edges, planted = cac.gen_planted_hard(0, 24, 5, 256, 4096, cycles=k)
    std = cac.gen_planted(0, 24, 5, 256, 4096)[0]
    assert np.array_equal(descs["n"], std["n"])
    for i in range(len(descs)):
        nv, e = _graph(descs, edges, i)
        h = nv // 2
        u, v, w = e["u"].astype(np.int64), e["v"].astype(np.int64), e["w"].astype(np.int64)
        assert np.all(u < v) and np.all(v < nv)
        cross = (u < h) & (v >= h)
        c = int(cross.sum())
        assert 1 <= c <= 8 and len(e) == k * nv + c
        assert w[cross].sum() == planted[i] and w[cross].max() - w[cross].min() <= 1
        inner = ~cross
        assert np.all((w[inner] >= 1) & (w[inner] <= 100))
        deg = np.bincount(u[inner], w[inner], nv) + np.bincount(v[inner], w[inner], nv)
        da, db = deg[:h], deg[h:]
        assert np.all(da == da[0]) and np.all(db == db[0])  # regular within each half
        assert planted[i] == int(np.floor(0.9 * min(da[0], db[0])))
        # each half is k cycles: every vertex has 2k edge ends inside its half
        ends = np.bincount(u[inner], minlength=nv) + np.bincount(v[inner], minlength=nv)
        assert np.all(ends == 2 * k)
        g = coo_matrix((np.ones(inner.sum()), (u[inner], v[inner])), shape=(nv, nv))
        nc, lab = connected_components(g, directed=False)
        assert nc == 2


def test_hard_family_planted_cut_is_the_min_cut(orc):
    """C is the exact minimum cut (Stoer-Wagner, the oracle) and lies within
    10% of the minimum weighted degree."""
    descs, edges, planted = cac.gen_planted_hard(0, 10, 13, 64, 300, cycles=4)
    for i in range(len(descs)):
        nv, e = _graph(descs, edges, i)
        adj = np.zeros((nv, nv), dtype=np.int64)
        np.add.at(adj, (e["u"], e["v"]), e["w"].astype(np.int64))
        adj += adj.T
        assert orc.stoer_wagner(adj) == planted[i]
        assert planted[i] >= 0.85 * adj.sum(axis=1).min()


def test_hard_family_deterministic():
    a = cac.gen_planted_hard(0, 6, 7, 256, 2048)
    b = cac.gen_planted_hard(3, 3, 7, 256, 2048)
    for j in range(3):
        na, ea = _graph(a[0], a[1], 3 + j)
        nb, eb = _graph(b[0], b[1], j)
        assert na == nb and np.array_equal(ea, eb) and a[2][3 + j] == b[2][j]
