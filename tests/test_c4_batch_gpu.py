"""Batched square-root min cut at the C4 size and on the hard family.

* The full batch of 4096 planted-cut graphs (n log-uniform in [256, 65536],
  halves of average degree 16 with weights U[1, 100], c in [1, 8] unit cross
  edges) in one call, with the light trial rule T_g = max(2, ks_trials(t_g)):
  every cut equals its planted c; for the three largest graphs the raw trials
  of the one-graph form equal the oracle's square-root trials bit for bit and
  the batch's cut is their minimum.
* The hard family (halves of weighted Hamiltonian cycles, planted cut within
  10% of the minimum weighted degree, exact by construction) under the
  reference's trial rule (SquareRootCut::numberOfTrials): at least 95% of the
  graphs return their planted cut; every cut is at least the planted one."""
import sys

import numpy as np
import pytest

import cac

pytestmark = pytest.mark.gpu
F32_95 = float(np.float32(0.95))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _light(descs):
    t = np.array([cac.intermediate_size(int(a), int(b)) for a, b in zip(descs["n"], descs["m"])])
    return np.maximum(2, [cac.ks_trials(int(x), F32_95) for x in t]), t


def test_c4_full_batch(ctx, orc):
    _say("c4: generating 4096 planted graphs")
    descs, edges, planted = cac.gen_planted(0, 4096, 0)
    descs["trials"], t = _light(descs)
    _say(f"c4: {len(edges)} edges, {int(descs['trials'].sum())} trials")
    buf = ctx.alloc(edges.nbytes)
    try:
        buf.upload(edges)
        cuts, trials = ctx.mincut_batch_dev(buf.ptr, descs, F32_95, 0)
    finally:
        buf.free()
    assert np.array_equal(trials, descs["trials"])
    assert np.array_equal(cuts, planted.astype(np.int64))
    for g in np.argsort(descs["n"])[-3:]:
        d = descs[g]
        n, T = int(d["n"]), int(d["trials"])
        e = edges[int(d["edge_offset"]): int(d["edge_offset"] + d["m"])]
        raw = ctx.sqrt_cut_trials(e, n, int(t[g]), 0, 0, T)
        e3 = np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)
        want = [orc.sqrt_trial(n, e3, int(t[g]), 0, k) for k in range(T)]
        _say(f"  graph {g}: n {n} m {len(e)} t {t[g]} raw {raw.tolist()} oracle {want}")
        assert raw.tolist() == want
        assert cuts[g] == min(want)


def test_hard_family_reference_rule(ctx):
    descs, edges, planted = cac.gen_planted_hard(0, 24, 0, 256, 2048)
    cuts, trials = ctx.mincut_batch(edges, descs, F32_95, 0)  # trials = 0: the reference's rule
    want_T = [cac.sqrt_trials(int(a), int(b), F32_95) for a, b in zip(descs["n"], descs["m"])]
    assert trials.tolist() == want_T
    ok = cuts == planted
    _say(f"hard family, reference rule: {int(ok.sum())}/{len(ok)} planted cuts, trials {min(want_T)}-{max(want_T)}")
    assert np.all(cuts >= planted)  # a raw cut is never below the minimum cut
    assert ok.mean() >= 0.95
