"""Batched square-root min cut (cac_mincut_batch*): every graph's cut in the
mixed batch equals its cut alone (same trials, same seed), equals the planted
cut, and equals exact Stoer-Wagner; the reference trial rule is applied when a
graph asks for it; bad input is rejected."""
import sys

import numpy as np
import pytest
from graphgen import erdos_renyi

import cac

pytestmark = pytest.mark.gpu
F32_95 = float(np.float32(0.95))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixed_batch(seed):
    descs, edges, planted = cac.gen_planted(0, 10, seed, 256, 1500)
    rng = np.random.default_rng(seed)
    extra = [erdos_renyi(rng, n, 10.0 / n, wmax=20) for n in (90, 300, 700)]
    all_d = list(descs)
    chunks = [edges]
    off = len(edges)
    for (n, e) in zip((90, 300, 700), extra):
        ee = cac.to_edges(e)
        all_d.append((off, len(ee), n, 0))
        chunks.append(ee)
        off += len(ee)
    d = np.array([tuple(x) for x in all_d], dtype=cac.GRAPH_DESC)
    d["trials"] = 3
    return d, np.concatenate(chunks), planted, extra


def test_batch_equals_alone_and_planted(ctx, orc):
    d, e, planted, extra = _mixed_batch(5)
    _say(f"batch of {len(d)} graphs, {len(e)} edges")
    cuts, trials = ctx.mincut_batch(e, d, F32_95, 17)
    assert np.all(trials == 3)
    for g in range(len(d)):
        off, m, n = int(d[g]["edge_offset"]), int(d[g]["m"]), int(d[g]["n"])
        eg = e[off: off + m]
        alone, _ = ctx.mincut_batch(eg, np.array([(0, m, n, 3)], dtype=cac.GRAPH_DESC), F32_95, 17)
        t = cac.intermediate_size(n, m)
        raw = ctx.sqrt_cut_trials(eg, n, t, 17, 0, 3)
        assert cuts[g] == alone[0] == raw.min(), g
        e3 = np.stack([eg["u"], eg["v"], eg["w"]], axis=1).astype(np.uint64)
        want = [orc.sqrt_trial(n, e3, t, 17, i) for i in range(3)]
        assert np.array_equal(raw, want), g
        if g < len(planted):
            assert cuts[g] == planted[g], (g, cuts[g], planted[g])


def test_batch_dev_and_reference_rule(ctx, orc):
    rng = np.random.default_rng(3)
    n = 120
    e = cac.to_edges(erdos_renyi(rng, n, 0.1, wmax=9))
    d = np.array([(0, len(e), n, 0)], dtype=cac.GRAPH_DESC)
    _say(f"batch reference rule n {n} m {len(e)}")
    buf = ctx.alloc(e.nbytes)
    try:
        buf.upload(e)
        cuts, trials = ctx.mincut_batch_dev(buf.ptr, d, F32_95, 4)
    finally:
        buf.free()
    assert trials[0] == cac.sqrt_trials(n, len(e), F32_95)
    single, T = ctx.mincut(e, n, F32_95, 4)
    assert T == trials[0] and cuts[0] == single
    adj = np.zeros((n, n), dtype=np.int64)
    np.add.at(adj, (e["u"], e["v"]), e["w"].astype(np.int64))
    adj += adj.T
    assert cuts[0] == orc.stoer_wagner(adj)


def test_batch_rejects_bad_input(ctx):
    e = cac.to_edges(np.array([[0, 1, 1], [1, 5, 1]], dtype=np.uint64))
    d = np.array([(0, 2, 4, 1)], dtype=cac.GRAPH_DESC)
    with pytest.raises(cac.CacError) as ei:
        ctx.mincut_batch(e, d, F32_95, 0)
    assert ei.value.code == cac.CAC_EINVAL
    d = np.array([(0, 2, 6, 1), (2, 0, 6, 1)], dtype=cac.GRAPH_DESC)
    with pytest.raises(cac.CacError) as ei:
        ctx.mincut_batch(e, d, F32_95, 0)
    assert ei.value.code == cac.CAC_EINVAL
