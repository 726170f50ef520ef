"""Pins the oracle's counter PRNG and uniform-integer rule.

Philox-4x32-10: the published Random123 known-answer vectors (Salmon et al.,
SC'11, kat_vectors).  Philox-4x64-10 shares the round structure and is pinned
against numpy.random.Philox, an independent implementation: a wrong round,
key schedule or word permutation fails one of these.
"""
import numpy as np
import pytest

KAT_4x32 = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT_4x32)
def test_philox4x32_kat(orc, ctr, key, want):
    assert [int(x) for x in orc.philox4x32(ctr, key)] == want


@pytest.mark.parametrize("key", [0, 1, 0xDEADBEEF, (1 << 127) + 12345])
@pytest.mark.parametrize("start", [0, 7, (1 << 64) - 1, (1 << 200) + 3])
def test_philox4x64_matches_numpy(orc, key, start):
    # numpy increments the 256-bit counter before producing each block
    g = np.random.Philox(key=key, counter=start)
    got = g.random_raw(8)
    for b in range(2):
        c = start + 1 + b
        ctr = [(c >> (64 * i)) & ((1 << 64) - 1) for i in range(4)]
        k = [key & ((1 << 64) - 1), key >> 64]
        assert [int(x) for x in orc.philox4x64(ctr, k)] == [int(x) for x in got[4 * b:4 * b + 4]]


def test_philox4x64_zero_kat(orc):
    assert [int(x) for x in orc.philox4x64([0] * 4, [0, 0])] == [
        0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]


def test_uniform1_closed_forms(orc):
    # 1 + floor(x * W / 2^64): endpoints and exact quarter points
    assert orc.uniform1(0, 1) == 1
    assert orc.uniform1((1 << 64) - 1, 1) == 1
    assert orc.uniform1(0, 1000) == 1
    assert orc.uniform1((1 << 64) - 1, 1000) == 1000
    assert orc.uniform1(1 << 63, 4) == 3
    assert orc.uniform1((1 << 62) - 1, 4) == 1
    assert orc.uniform1(1 << 62, 4) == 2
    W = (1 << 63) + 12345
    x = 0x0123456789ABCDEF
    assert orc.uniform1(x, W) == 1 + (x * W >> 64)


def test_uniform1_is_uniform(orc):
    # every value in [1, W] is hit equally often over a full sweep of the top bits
    W = 7
    counts = np.zeros(W + 1, dtype=np.int64)
    for t in range(7 * 64):
        x = (t * ((1 << 64) // (7 * 64)))
        counts[orc.uniform1(x, W)] += 1
    assert counts[0] == 0 and counts[1:].min() >= 63 and counts[1:].max() <= 65
