"""The high-concurrency sampling rule of the oracle (orc_sample_contract_rule
with orc_hc_draws(n) draws per round, WeightedIteratedSparseSampling::shrink):
its draw count is the reference's float formula (its own per-round figures);
with the same Philox draws it equals the low-concurrency rule whenever that
one finishes within its first round; it ends at exactly t canonical
super-vertices; and a pure-Python walk through the rounds agrees with it on
small graphs."""
import numpy as np

import oracle as orc

SAMP = 0x53414D50


def test_hc_draw_count():
    # (unsigned) std::pow((float) n0, 1 + 0.1f / 2): std::pow of two floats is
    # the float overload, so 1,412 / 114,104 / 1,995,261 draws per round for
    # n0 = 1000 / 65536 / 10^6 (in double the last would be 1,995,262)
    assert orc.hc_draws(1000) == 1412
    assert orc.hc_draws(65536) == 114104
    assert orc.hc_draws(10**6) == 1995261
    for n0 in (2, 3, 17, 318, 4096, 100000):
        want = float(np.float32(n0) ** np.float32(1.05))
        assert abs(orc.hc_draws(n0) - int(want)) <= 1


def _er(rng, n, m):
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    return np.stack([u[keep], v[keep], rng.integers(1, 20, keep.sum())], 1).astype(np.uint64)


def test_hc_equals_lc_within_one_round():
    rng = np.random.default_rng(3)
    same = 0
    for trial in range(40):
        n = 200
        e = _er(rng, n, 1200)
        t = int(rng.integers(60, 150))
        lab_lc, nv_lc = orc.sample_contract(n, e, t, 7, trial)
        lab_hc, nv_hc = orc.sample_contract(n, e, t, 7, trial, orc.hc_draws(n))
        # the low-concurrency rule finished in its first round when the first
        # n draws already reach t: then the high-concurrency round, whose first
        # n draws are the same words, stops at the same prefix
        lab_one, nv_one = orc.sample_contract(n, e, t, 7, trial, n)
        if nv_one == t:
            assert np.array_equal(lab_hc, lab_lc) and nv_hc == nv_lc == t
            same += 1
        assert nv_hc == t and nv_lc == t
        for lab in (lab_lc, lab_hc):
            assert np.all(lab <= np.arange(n)) and np.all(lab[lab] == lab)
            assert len(np.unique(lab)) == t
    assert same >= 10


def _py_rounds(n, e, t, seed, trial, dpr):
    lab = list(range(n))
    nv = n
    rnd = 0
    while nv > t:
        P, W = [], 0
        for u, v, w in e:
            if lab[u] != lab[v]:
                W += int(w)
            P.append(W)
        if W == 0:
            break
        par = list(range(n))

        def root(x):
            while par[x] != x:
                x = par[x]
            return x

        for d in range(dpr if dpr else nv):
            o = orc.philox4x32([trial, rnd, d, 0], [seed, SAMP])
            x = (int(o[1]) << 32) | int(o[0])
            r = 1 + ((x * W) >> 64)
            i = next(k for k, p in enumerate(P) if p >= r)
            a, b = root(lab[int(e[i][0])]), root(lab[int(e[i][1])])
            if a == b:
                continue
            par[max(a, b)] = min(a, b)
            nv -= 1
            if nv == t:
                break
        lab = [root(lab[v]) for v in range(n)]
        rnd += 1
    return lab, nv


def test_hc_rounds_match_python_walk():
    rng = np.random.default_rng(11)
    for k in range(12):
        n = int(rng.integers(8, 30))
        e = _er(rng, n, int(rng.integers(n, 4 * n)))
        t = int(rng.integers(2, max(3, n // 2)))
        for dpr in (orc.hc_draws(n), 3, 0):
            lab, nv = orc.sample_contract(n, e, t, 5, k, dpr)
            plab, pnv = _py_rounds(n, e, t, 5, k, dpr)
            assert nv == pnv and lab.tolist() == plab, (k, dpr)
