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
