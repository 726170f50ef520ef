"""The Erdos-Renyi input generator (cac_gen_er, host-side) that makes the C1
square_root graph, against an independent numpy restatement of its rule: pair
(u < v) kept iff word 0 of Philox-4x32-10(ctr = {p_lo, p_hi, 0, 0}, key =
{seed, 'GENR'}) with p = u n + v is below threshold = p * 2^32, weight
1 + floor(wmax * word1 / 2^32).  The numpy Philox is checked against the
oracle's (which the Random123 known-answer tests pin) before it is used."""
from fractions import Fraction

import numpy as np

import cac

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def _philox(c0, c1, c2, c3, k0, k1):
    """Philox-4x32-10 on uint64 arrays holding 32-bit words (vectorised)."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) & _MASK for x in (c0, c1, c2, c3))
    for r in range(10):
        ka = np.uint64((k0 + r * _W0) & 0xFFFFFFFF)
        kb = np.uint64((k1 + r * _W1) & 0xFFFFFFFF)
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ ka, lo1, hi0 ^ c3 ^ kb, lo0
    return c0, c1, c2, c3


def _er_rule(n, p, seed, wmax):
    thr = np.uint64(int(Fraction(str(p)) * (1 << 32)))
    u, v = np.triu_indices(n, 1)
    pair = u.astype(np.uint64) * np.uint64(n) + v.astype(np.uint64)
    w0, w1, _, _ = _philox(pair & _MASK, pair >> np.uint64(32), 0, 0, seed, 0x47454E52)
    keep = w0 < thr
    w = np.uint64(1) + ((np.uint64(wmax) * w1[keep]) >> np.uint64(32))
    return u[keep].astype(np.uint64), v[keep].astype(np.uint64), w


def test_numpy_philox_equals_oracle(orc):
    rng = np.random.default_rng(3)
    for _ in range(64):
        ctr = [int(x) for x in rng.integers(0, 1 << 32, 4, dtype=np.uint64)]
        key = [int(x) for x in rng.integers(0, 1 << 32, 2, dtype=np.uint64)]
        got = [int(x[0]) for x in _philox(*[np.array([c]) for c in ctr], key[0], key[1])]
        assert got == [int(x) for x in orc.philox4x32(ctr, key)]


def test_er_c1_graph_equals_rule():
    # the C1 configuration: ER(1000, 0.05), unit weights, seed 0 (redrawn with
    # seed + 1, ... until connected)
    e, seed = cac.gen_er(1000, 0.05, 0)
    u, v, w = _er_rule(1000, 0.05, seed, 1)
    assert np.array_equal(e["u"], u) and np.array_equal(e["v"], v) and np.array_equal(e["w"], w)
    assert (seed, len(e)) == (0, 25111)


def test_er_weighted_equals_rule():
    e, seed = cac.gen_er(300, 0.1, 7, wmax=50)
    u, v, w = _er_rule(300, 0.1, seed, 50)
    assert np.array_equal(e["u"], u) and np.array_equal(e["v"], v) and np.array_equal(e["w"], w)
    assert w.min() >= 1 and w.max() <= 50 and len(np.unique(w)) > 10
