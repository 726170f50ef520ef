"""The serving loop (cac_server_*): a Poisson stream of CC and min-cut queries
of different sizes, admitted by footprint and run in launch groups on several
workers, returns for every query exactly its alone result and the oracle's:
CC labels and count equal the oracle's union-find; a min cut with a trial
budget equals the minimum of the oracle's square-root trials (same seed, same
trial ids); a min cut with the reference's rule (trials = 0) equals cac_mincut
alone.  Queries with different seeds share launch groups; groups on the
workers overlap; a bad query fails alone."""
import sys

import numpy as np
import pytest

import cac

pytestmark = pytest.mark.gpu
F32_95 = float(np.float32(0.95))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _stream(rng, count, nlo=200, nhi=20000, dlo=2.0, dhi=12.0, budget=3):
    """(kind, n, edges, labels, seed, trials) per query; min-cut queries carry
    their own seeds and a budget of `budget` trials."""
    qs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(nlo), np.log(nhi))))
        d = rng.uniform(dlo, dhi)
        e = cac.gen_gnp(n, d, 9, i, wmax=100)
        kind = cac.QUERY_CC if rng.random() < 0.7 else cac.QUERY_MINCUT
        labels = np.zeros(n, dtype=np.uint32) if kind == cac.QUERY_CC else None
        qs.append((kind, n, e, labels, 100 + i, budget))
    return qs


def _e3(e):
    return np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)


def _check_cc(orc, n, e, lab, ncomp, i):
    want, nc = orc.cc(np.stack([e["u"], e["v"]], 1).astype(np.uint32), n)
    assert ncomp == nc and np.array_equal(lab, want), i


def test_served_results_equal_oracle(ctx, orc):
    rng = np.random.default_rng(42)
    qs = _stream(rng, 48)
    arrivals = np.cumsum(rng.exponential(1.0 / 400.0, size=len(qs)))
    _say(f"serve {len(qs)} queries over {arrivals[-1]:.3f} s")
    with cac.Server(0, workers=3, hbm_budget=8 << 30, max_batch=16) as srv:
        recs = [cac.Server.query(k, e, n, F32_95, sd, lab, trials=T) for (k, n, e, lab, sd, T) in qs]
        res = srv.replay(recs, arrivals)
    assert np.all(res["status"] == 0)
    for i, (k, n, e, lab, sd, Tq) in enumerate(qs):
        if k == cac.QUERY_CC:
            _check_cc(orc, n, e, lab, res[i]["ncomp"], i)
            continue
        t, T = cac.serve_mincut_plan(n, len(e), F32_95, Tq)
        assert T == Tq and res[i]["t"] == t and res[i]["trials"] == T
        want = [orc.sqrt_trial(n, _e3(e), t, sd, k) for k in range(T)]
        assert res[i]["cut"] == min(want), (i, n, res[i]["cut"], want)
        _say(f"  query {i}: n {n} m {len(e)} t {t} T {T} cut {res[i]['cut']} = oracle")
    assert np.all(res["admit_s"] >= res["submit_s"]) and np.all(res["done_s"] >= res["admit_s"])


def test_reference_rule_equals_mincut(ctx):
    """trials = 0: the reference's rule, the same t, T and cut as cac_mincut."""
    rng = np.random.default_rng(5)
    qs = []
    for i in range(6):
        n = int(rng.integers(60, 400))
        qs.append((n, cac.gen_gnp(n, float(rng.uniform(4, 10)), 3, i, wmax=50), 7 + i))
    with cac.Server(0, workers=2, hbm_budget=4 << 30, max_batch=8) as srv:
        recs = [cac.Server.query(cac.QUERY_MINCUT, e, n, F32_95, sd) for (n, e, sd) in qs]
        res = srv.replay(recs, np.zeros(len(recs)))
    for i, (n, e, sd) in enumerate(qs):
        assert res[i]["status"] == 0
        cut, T = ctx.mincut(e, n, F32_95, sd)
        assert res[i]["trials"] == T == cac.sqrt_trials(n, len(e), F32_95)
        assert res[i]["t"] == cac.intermediate_size(n, len(e))
        assert res[i]["cut"] == cut, i


def test_burst_groups_mixed_seeds_and_overlaps(ctx):
    """A burst arriving at once: queries waiting behind a running group are
    admitted together as launch groups, min-cut queries of different seeds
    share a group, and groups on the two workers overlap."""
    rng = np.random.default_rng(7)
    qs = _stream(rng, 24)
    _say("serve burst of 24 queries")
    with cac.Server(0, workers=2, hbm_budget=8 << 30, max_batch=8) as srv:
        recs = [cac.Server.query(k, e, n, F32_95, sd, lab, trials=T) for (k, n, e, lab, sd, T) in qs]
        res = srv.replay(recs, np.zeros(len(qs)))
    assert np.all(res["status"] == 0)
    mc = [i for i, q in enumerate(qs) if q[0] == cac.QUERY_MINCUT]
    for i, (k, n, e, lab, sd, T) in enumerate(qs):
        if k == cac.QUERY_CC:
            want, nc = ctx.cc(np.stack([e["u"], e["v"]], 1), n)
            assert res[i]["ncomp"] == nc and np.array_equal(lab, want), i
        else:
            t, _ = cac.serve_mincut_plan(n, len(e), F32_95, T)
            assert res[i]["cut"] == ctx.sqrt_cut_trials(e, n, t, sd, 0, T).min(), i
    assert res["batch_size"].max() > 1  # launch groups formed
    assert len(mc) >= 2 and res["batch_size"][mc].max() > 1  # min-cut queries of different seeds grouped
    order = np.argsort(res["admit_s"])
    overlap = any(res[order[j + 1]]["admit_s"] < res[order[j]]["done_s"] for j in range(len(order) - 1))
    assert overlap  # a group was admitted while another ran


def test_bad_query_fails_alone(ctx):
    good = cac.gen_gnp(300, 6.0, 1, 0, wmax=5)
    bad = good.copy()
    bad["v"][3] = 300  # id >= n
    heavy = good.copy()
    heavy["w"][0] = 1 << 41
    lab = np.zeros(300, dtype=np.uint32)
    with cac.Server(0, workers=1, hbm_budget=1 << 30, max_batch=8) as srv:
        recs = [cac.Server.query(cac.QUERY_CC, bad, 300), cac.Server.query(cac.QUERY_CC, good, 300, labels=lab),
                cac.Server.query(cac.QUERY_MINCUT, heavy, 300, F32_95, 0, trials=4),
                cac.Server.query(cac.QUERY_MINCUT, bad, 300, F32_95, 0, trials=4),
                cac.Server.query(cac.QUERY_MINCUT, good, 300, F32_95, 0, trials=4)]
        res = srv.replay(recs, np.zeros(len(recs)))
    assert res[0]["status"] == cac.CAC_EINVAL
    assert res[1]["status"] == 0
    want, nc = ctx.cc(np.stack([good["u"], good["v"]], 1), 300)
    assert res[1]["ncomp"] == nc and np.array_equal(lab, want)
    assert res[2]["status"] == cac.CAC_ERANGE
    assert res[3]["status"] == cac.CAC_EINVAL
    assert res[4]["status"] == 0
    t, T = cac.serve_mincut_plan(300, len(good), F32_95, 4)
    assert res[4]["cut"] == ctx.sqrt_cut_trials(good, 300, t, 0, 0, T).min()


def test_footprint_over_budget_is_rejected():
    e = cac.gen_gnp(2000, 20.0, 2, 0, wmax=5)
    with cac.Server(0, workers=1, hbm_budget=4096, max_batch=4) as srv:
        res = srv.replay([cac.Server.query(cac.QUERY_CC, e, 2000)], np.zeros(1))
    assert res[0]["status"] == cac.CAC_ENOMEM


def test_c5_workload_stream(ctx, orc):
    """The C5 shape: G(n, d/n) queries with n log-uniform in [10^3, 10^6] and d
    uniform in [8, 64], weights U[1, 100], 70% CC / 30% min cut with their own
    seeds and the bench's trial budget, Poisson arrivals.  CC labels and counts
    equal the oracle's; every min cut equals its alone result, and for queries
    with n <= 10^5 the alone run's first trials equal the oracle's."""
    rng = np.random.default_rng(2026)
    count = 20
    ns = np.exp(rng.uniform(np.log(1e3), np.log(1e6), size=count)).astype(np.int64)
    ds = rng.uniform(8, 64, size=count)
    kinds = np.where(rng.random(count) < 0.7, cac.QUERY_CC, cac.QUERY_MINCUT)
    kinds[:2] = cac.QUERY_MINCUT  # at least two min-cut queries
    arrivals = np.cumsum(rng.exponential(1.0 / 20.0, size=count))
    qs = []
    for i in range(count):
        n = int(ns[i])
        e = cac.gen_gnp(n, float(ds[i]), 0, i, wmax=100)
        t = min(cac.intermediate_size(n, len(e)), 4096)
        T = max(2, cac.ks_trials(t, F32_95))
        lab = np.zeros(n, dtype=np.uint32) if kinds[i] == cac.QUERY_CC else None
        qs.append((int(kinds[i]), n, e, lab, 1000 + i, T))
    _say(f"c5 stream: {count} queries, n {ns.min()}..{ns.max()}, {sum(len(q[2]) for q in qs)} edges")
    with cac.Server(0, workers=4, hbm_budget=160 << 30, max_batch=64) as srv:
        recs = [cac.Server.query(k, e, n, F32_95, sd, lab, trials=T) for (k, n, e, lab, sd, T) in qs]
        res = srv.replay(recs, arrivals)
    assert np.all(res["status"] == 0)
    for i, (k, n, e, lab, sd, Tq) in enumerate(qs):
        if k == cac.QUERY_CC:
            _check_cc(orc, n, e, lab, res[i]["ncomp"], i)
            _say(f"  cc query {i}: n {n} m {len(e)} components {res[i]['ncomp']} = oracle")
            continue
        t, T = cac.serve_mincut_plan(n, len(e), F32_95, Tq)
        raw = ctx.sqrt_cut_trials(e, n, t, sd, 0, T)
        assert res[i]["cut"] == raw.min(), i
        checked = 0
        if n <= 100000:
            e3 = _e3(e)
            for k in range(min(T, 2)):
                assert raw[k] == orc.sqrt_trial(n, e3, t, sd, k), (i, k)
                checked += 1
        _say(f"  min-cut query {i}: n {n} m {len(e)} t {t} T {T} cut {res[i]['cut']}, {checked} trials = oracle")


def test_resident_graphs_equal_uploaded(ctx, orc):
    """Graphs registered once (device-resident, checked once) and queried by
    handle give the results of the same queries bringing their edges; an
    unregistered graph is freed after its queued queries; unknown handles and
    invalid graphs are refused."""
    rng = np.random.default_rng(21)
    gs = []
    for i in range(5):
        n = int(rng.integers(300, 6000))
        gs.append((n, cac.gen_gnp(n, float(rng.uniform(3, 10)), 4, i, wmax=100)))
    with cac.Server(0, workers=2, hbm_budget=8 << 30, max_batch=8) as srv:
        hs = [srv.register_graph(e, n) for (n, e) in gs]
        qs, labs = [], []
        for h, (n, e) in zip(hs, gs):
            lab = np.zeros(n, dtype=np.uint32)
            labs.append(lab)
            qs.append(cac.Server.query(cac.QUERY_CC, None, 0, labels=lab, graph=h))
            qs.append(cac.Server.query(cac.QUERY_MINCUT, None, 0, F32_95, 30 + h, trials=3, graph=h))
        qs.append(cac.Server.query(cac.QUERY_CC, None, 0, graph=999))
        ids = [srv.submit(q, keep) for (q, keep) in qs]
        srv.unregister_graph(hs[0])  # its queued queries still run; freed after them
        got = {}
        while len(got) < len(ids):
            for r in srv.poll():
                got[int(r["id"])] = r
        res = [got[i] for i in ids]
        assert res[-1]["status"] == cac.CAC_EINVAL
        bad = gs[1][1].copy()
        bad["u"][0] = gs[1][0]
        with pytest.raises(cac.CacError) as ei:
            srv.register_graph(bad, gs[1][0])
        assert ei.value.code == cac.CAC_EINVAL
        with pytest.raises(cac.CacError):
            srv.unregister_graph(hs[0])  # already released
    for k, (h, (n, e)) in enumerate(zip(hs, gs)):
        rc, rm = res[2 * k], res[2 * k + 1]
        assert rc["status"] == 0 and rm["status"] == 0
        _check_cc(orc, n, e, labs[k], rc["ncomp"], k)
        t, T = cac.serve_mincut_plan(n, len(e), F32_95, 3)
        assert rm["cut"] == ctx.sqrt_cut_trials(e, n, t, 30 + h, 0, T).min(), k
