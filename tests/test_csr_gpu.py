"""Connected components on a symmetric CSR (Afforest) and the CSR builder, on the
device: the CSR against a numpy construction, labels bit-exact against the
oracle (min-id and dense-rank forms) on random graphs, stars and paths (long
adjacency lists, long chains, no dominant component), the R-MAT generator, and
at the full 2^24/2^28 R-MAT size against the oracle."""
import sys

import numpy as np
import pytest

import cac

pytestmark = pytest.mark.gpu
RMAT_TH = cac.rmat_thresholds(0.45, 0.22, 0.22)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _numpy_csr(e, n):
    e = np.asarray(e, dtype=np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.searchsorted(src, np.arange(n + 1), side="left").astype(np.uint32)
    return off, dst.astype(np.uint32)


def _graph(kind, rng):
    if kind == "empty":
        return 5, np.zeros((0, 2), np.uint32)
    if kind == "single":
        return 1, np.array([[0, 0]], np.uint32)
    if kind == "random-sparse":
        n = 50000
        return n, rng.integers(0, n, size=(n // 2, 2)).astype(np.uint32)  # many small components
    if kind == "random-loops-dups":
        n = 20000
        e = rng.integers(0, n, size=(60000, 2)).astype(np.uint32)
        e[:3000, 1] = e[:3000, 0]
        e[3000:6000] = e[6000:9000]
        return n, e
    if kind == "two-stars":
        # two stars of 5000 leaves: no dominant root, long adjacency lists outside the sampled one
        n = 10002
        a = np.stack([np.zeros(5000), np.arange(2, 5002)], 1)
        b = np.stack([np.ones(5000), np.arange(5002, 10002)], 1)
        return n, np.concatenate([a, b]).astype(np.uint32)
    if kind == "path":
        n = 100000
        p = rng.permutation(n)
        return n, np.stack([p[:-1], p[1:]], 1).astype(np.uint32)
    if kind == "many-paths":
        n = 90000
        p = rng.permutation(n).reshape(30, -1)
        return n, np.concatenate([np.stack([q[:-1], q[1:]], 1) for q in p]).astype(np.uint32)
    if kind == "hooked-giant":
        # Star 1000 <- 1001..19999 is the sampled component.  Vertex 30000 (first
        # neighbours 0, 1) reaches leaf 1001 only through its third neighbour, and
        # 1001 sees 30000 only after 1000 and 1002, so root 1000 is hooked under 0
        # in the final pass: every vertex of the star needs its label redone.
        n = 30001
        star = np.stack([np.full(18999, 1000), np.arange(1001, 20000)], 1)
        extra = np.array([[1001, 1002], [1001, 1003], [30000, 0], [30000, 1], [30000, 1001]])
        return n, np.concatenate([star, extra]).astype(np.uint32)
    if kind == "random-above-hubs":
        # beyond the 2^16 hub prefixes (bulk pass, sampled root), average degree 3:
        # many vertices stay outside the sampled component and link in the final pass
        n = 150000
        return n, rng.integers(0, n, size=(225000, 2)).astype(np.uint32)
    if kind == "hooked-giant-above-hubs":
        # as hooked-giant with a bulk range: 70000 links only its first neighbour 0
        # in the bulk pass and reaches leaf 1001 in the final pass, hooking the
        # sampled root 1000 under 0
        n = 70001
        star = np.stack([np.full(58999, 1000), np.arange(1001, 60000)], 1)
        extra = np.array([[1001, 1002], [1001, 1003], [70000, 0], [70000, 1001]])
        return n, np.concatenate([star, extra]).astype(np.uint32)
    raise ValueError(kind)


KINDS = ["empty", "single", "random-sparse", "random-loops-dups", "two-stars", "path", "many-paths", "hooked-giant",
         "random-above-hubs", "hooked-giant-above-hubs"]


@pytest.mark.parametrize("kind", KINDS)
def test_csr_and_labels_match_oracle(ctx, orc, kind):
    rng = np.random.default_rng(len(kind))
    n, e = _graph(kind, rng)
    _say(f"csr cc {kind}: n {n} m {len(e)}")
    for mode in (cac.LABEL_MIN_ID, cac.LABEL_RANK):
        lab, nc, off, cols = ctx.cc_csr(e, n, mode)
        want, wn = orc.cc(e, n, mode)
        assert nc == wn
        assert np.array_equal(lab, want)
    woff, wcols = _numpy_csr(e, n)
    assert np.array_equal(off, woff) and np.array_equal(cols, wcols)


def test_rmat_scale18(ctx, orc):
    scale, m = 18, 1 << 22
    n = 1 << scale
    _say(f"csr cc rmat scale {scale} m {m}")
    de = ctx.alloc(m * 8)
    ctx.gen_rmat_dev(scale, m, 3, RMAT_TH, de.ptr)
    e = de.download(np.uint32, 2 * m).reshape(m, 2)
    de.free()
    lab, nc, off, cols = ctx.cc_csr(e, n)
    want, wn = orc.cc(e, n)
    assert nc == wn and np.array_equal(lab, want)
    woff, wcols = _numpy_csr(e, n)
    assert np.array_equal(off, woff) and np.array_equal(cols, wcols)


def test_csr_rmat_full_size(ctx, orc):
    """R-MAT 2^24 vertices / 2^28 edges: Afforest labels bit-exact vs the oracle."""
    scale, m = 24, 1 << 28
    n = 1 << scale
    _say(f"csr cc rmat full size scale {scale} m {m}: generate + build")
    de = ctx.alloc(m * 8)
    do = ctx.alloc((n + 1) * 4)
    dc = ctx.alloc(2 * m * 4)
    dl = ctx.alloc(n * 4)
    ctx.gen_rmat_dev(scale, m, 0, RMAT_TH, de.ptr)
    nnz = ctx.csr_build_dev(de.ptr, m, n, do.ptr, dc.ptr)
    assert nnz == 2 * m  # the generator drops loops and repeats
    _say("csr cc rmat full size: device CC")
    nc = ctx.cc_csr_dev(do.ptr, dc.ptr, n, dl.ptr)
    lab = dl.download(np.uint32, n)
    e = de.download(np.uint32, 2 * m)
    for b in (de, do, dc, dl):
        b.free()
    _say("csr cc rmat full size: oracle CC")
    want, wn = orc.cc(e, n)
    assert nc == wn
    assert np.array_equal(lab, want)


def test_csr_build_rejects_bad_ids(ctx):
    e = np.array([[0, 1], [2, 7]], np.uint32)
    de = ctx.alloc(e.nbytes)
    do = ctx.alloc(4 * 4)
    dc = ctx.alloc(4 * 4)
    de.upload(e)
    with pytest.raises(cac.CacError) as ei:
        ctx.csr_build_dev(de.ptr, 2, 3, do.ptr, dc.ptr)
    assert ei.value.code == cac.CAC_EINVAL
    for b in (de, do, dc):
        b.free()


@pytest.mark.parametrize("mode", [cac.LABEL_MIN_ID, cac.LABEL_RANK])
def test_csr_graph_replays_match_oracle(ctx, orc, mode):
    # cac_cc_csr_graph_dev: the captured step replayed four times per graph,
    # with eager launches and synchronisations between the replays, on two
    # graphs (the second call captures again); the labels buffer is overwritten
    # before every replay, so each replay must rewrite all of it
    rng = np.random.default_rng(40 + mode)
    for kind in ("random-above-hubs", "many-paths"):
        n, e = _graph(kind, rng)
        want, wn = orc.cc(e, n, mode)
        off, cols = _numpy_csr(e, n)
        do, dc = ctx.alloc(off.nbytes), ctx.alloc(max(cols.nbytes, 4))
        do.upload(off)
        dc.upload(cols)
        dl, de = ctx.alloc(n * 4), ctx.alloc(n * 4)
        junk = np.full(n, 0xFFFFFFFF, np.uint32)
        for rep in range(4):
            _say(f"csr graph replay {kind} mode {mode} rep {rep}")
            dl.upload(junk)
            nc = ctx.cc_csr_graph_dev(do.ptr, dc.ptr, n, dl.ptr, mode)
            assert nc == wn
            assert np.array_equal(dl.download(np.uint32, n), want)
            assert ctx.cc_csr_dev(do.ptr, dc.ptr, n, de.ptr, mode) == wn  # eager work between replays
            ctx.synchronize()
            assert np.array_equal(de.download(np.uint32, n), want)
        for b in (do, dc, dl, de):
            b.free()
