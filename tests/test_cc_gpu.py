"""CC parity on the device: libcac.so labels must equal the oracle's bit for bit
(min-id and dense-rank forms), on the reference fixtures, random graphs with
loops/duplicates, edge cases, the R-MAT generator and the full 2^24/2^28 R-MAT."""
import sys
from pathlib import Path

import numpy as np
import pytest
from graphfile import read_graph

import cac

GOLD = Path(__file__).resolve().parent / "golden" / "reference_inputs"
pytestmark = pytest.mark.gpu
RMAT_TH = cac.rmat_thresholds(0.45, 0.22, 0.22)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["cc_8.graph", "cc_13.graph"])
@pytest.mark.parametrize("mode", [cac.LABEL_MIN_ID, cac.LABEL_RANK])
def test_cc_fixtures(ctx, orc, name, mode):
    _say(f"cc fixture {name} mode {mode}")
    n, e = read_graph(GOLD / name)
    e = e[:, :2].astype(np.uint32)
    lab, nc = ctx.cc(e, n, mode)
    want, wn = orc.cc(e, n, mode)
    assert nc == wn == 5
    assert np.array_equal(lab, want)


@pytest.mark.parametrize("seed", range(8))
def test_cc_random(ctx, orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([1, 2, 17, 1000, 70000, 300000]))
    m = int(rng.integers(0, 3 * n + 1))
    _say(f"cc random seed {seed} n {n} m {m}")
    e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    if m > 10:  # sprinkle loops and duplicates
        e[: m // 10, 1] = e[: m // 10, 0]
        e[m // 10: m // 5] = e[m // 5: m // 5 + (m // 5 - m // 10)]
    for mode in (cac.LABEL_MIN_ID, cac.LABEL_RANK):
        lab, nc = ctx.cc(e, n, mode)
        want, wn = orc.cc(e, n, mode)
        assert nc == wn
        assert np.array_equal(lab, want)


def test_cc_large_sampled_path(ctx, orc):
    # big enough to take the sample + settle + hook path (m >= 2^18)
    rng = np.random.default_rng(7)
    n, m = 1 << 20, 3 << 20
    _say(f"cc large n {n} m {m}")
    e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    lab, nc = ctx.cc(e, n)
    want, wn = orc.cc(e, n)
    assert nc == wn and np.array_equal(lab, want)
    lab2, _ = ctx.cc(e, n)
    assert np.array_equal(lab, lab2)  # canonical labels are run-to-run deterministic


@pytest.mark.parametrize("shape", ["sorted", "islands", "one_run", "ragged_tail"])
def test_cc_sorted_runs(ctx, orc, shape):
    """The run-sampling + giant-bitmap path on u-sorted lists: long runs, no
    giant component (many equal-size islands), a single run of one u, and a
    length that leaves a partial last tile."""
    rng = np.random.default_rng(["sorted", "islands", "one_run", "ragged_tail"].index(shape) + 100)
    if shape == "sorted":
        n, m = 1 << 18, (1 << 19) + 3
        e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    elif shape == "islands":  # 2^16 islands of 8 vertices, no component dominates
        n, m = 1 << 19, 1 << 19
        isl = rng.integers(0, n // 8, size=m).astype(np.uint32)
        e = np.stack([isl * 8 + rng.integers(0, 8, size=m), isl * 8 + rng.integers(0, 8, size=m)], 1).astype(np.uint32)
    elif shape == "one_run":  # one star: every edge has u = 5
        n, m = 1 << 18, 1 << 18
        e = np.stack([np.full(m, 5), rng.integers(0, n, size=m)], 1).astype(np.uint32)
    else:
        n, m = 3000, (1 << 18) + 511
        e = rng.integers(0, n, size=(m, 2)).astype(np.uint32)
    e = e[np.lexsort((e[:, 1], e[:, 0]))]
    _say(f"cc sorted runs {shape} n {n} m {m}")
    for mode in (cac.LABEL_MIN_ID, cac.LABEL_RANK):
        lab, nc = ctx.cc(e, n, mode)
        want, wn = orc.cc(e, n, mode)
        assert nc == wn
        assert np.array_equal(lab, want)


def test_cc_edge_cases(ctx):
    _say("cc edge cases")
    lab, nc = ctx.cc(np.zeros((0, 2), np.uint32), 4)
    assert nc == 4 and lab.tolist() == [0, 1, 2, 3]
    lab, nc = ctx.cc(np.array([[2, 2]], np.uint32), 3)
    assert nc == 3 and lab.tolist() == [0, 1, 2]
    lab, nc = ctx.cc(np.array([[1, 0], [0, 1], [1, 0]], np.uint32), 2, cac.LABEL_RANK)
    assert nc == 1 and lab.tolist() == [0, 0]
    with pytest.raises(cac.CacError) as ei:
        ctx.cc(np.array([[0, 5]], np.uint32), 5)
    assert ei.value.code == cac.CAC_EINVAL


def test_rmat_generator_matches_oracle(ctx, orc):
    scale, m, seed = 14, 200000, 11
    _say(f"rmat generator scale {scale} m {m}")
    d = ctx.alloc(m * 8)
    used = ctx.gen_rmat_dev(scale, m, seed, RMAT_TH, d.ptr)
    got = d.download(np.uint32, 2 * m).reshape(m, 2)
    d.free()
    want, dw = orc.rmat(scale, m, seed, RMAT_TH)
    assert used == dw
    assert np.array_equal(got, want)


def test_cc_rmat_full_size(ctx, orc):
    """Config: R-MAT 2^24 vertices, 2^28 undirected edges, bit-exact vs the oracle."""
    scale, m = 24, 1 << 28
    n = 1 << scale
    _say(f"rmat full size scale {scale} m {m}: generate")
    de = ctx.alloc(m * 8)
    dl = ctx.alloc(n * 4)
    ctx.gen_rmat_dev(scale, m, 0, RMAT_TH, de.ptr)
    _say("rmat full size: device CC")
    nc = ctx.cc_dev(de.ptr, m, n, dl.ptr)
    lab = dl.download(np.uint32, n)
    e = de.download(np.uint32, 2 * m)
    de.free()
    dl.free()
    _say("rmat full size: oracle CC")
    want, wn = orc.cc(e, n)
    assert nc == wn
    assert np.array_equal(lab, want)
    _say(f"rmat full size: {nc} components")


@pytest.mark.parametrize("nbytes", [16, 1000, (1 << 20) + 7])
def test_copy_kernel(ctx, nbytes):
    """The byte-counter calibration copy moves exactly nbytes (vector body + tail)."""
    _say(f"copy kernel {nbytes} bytes")
    rng = np.random.default_rng(nbytes)
    src = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
    a = ctx.alloc(nbytes)
    b = ctx.alloc(nbytes + 64)
    try:
        a.upload(src)
        b.upload(np.full(nbytes + 64, 0xAB, dtype=np.uint8))
        ctx.copy_dev(b.ptr, a.ptr, nbytes)
        ctx.synchronize()
        got = b.download(np.uint8, nbytes + 64)
        assert np.array_equal(got[:nbytes], src)
        assert np.all(got[nbytes:] == 0xAB)
    finally:
        a.free()
        b.free()
