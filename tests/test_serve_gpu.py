"""The serving loop (cac_server_*): a Poisson stream of CC and min-cut queries
of different sizes, admitted by footprint and run in launch groups on several
workers, returns for every query exactly its alone result (CC: labels and
count of cac_cc; min cut: the minimum of its trials run alone); groups of more
than one query and overlapping queries occur; a bad query fails alone."""
import sys

import numpy as np
import pytest

import cac

pytestmark = pytest.mark.gpu
F32_95 = float(np.float32(0.95))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _stream(rng, count):
    qs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(200), np.log(20000))))
        d = rng.uniform(2.0, 12.0)
        e = cac.gen_gnp(n, d, 9, i, wmax=100)
        kind = cac.QUERY_CC if rng.random() < 0.7 else cac.QUERY_MINCUT
        labels = np.zeros(n, dtype=np.uint32) if kind == cac.QUERY_CC else None
        qs.append((kind, n, e, labels))
    return qs


def test_served_results_equal_alone(ctx):
    rng = np.random.default_rng(42)
    qs = _stream(rng, 48)
    arrivals = np.cumsum(rng.exponential(1.0 / 400.0, size=len(qs)))
    _say(f"serve {len(qs)} queries over {arrivals[-1]:.3f} s")
    with cac.Server(0, workers=3, hbm_budget=8 << 30, max_batch=16) as srv:
        recs = [cac.Server.query(k, e, n, F32_95, 5, lab) for (k, n, e, lab) in qs]
        res = srv.replay(recs, arrivals)
    assert np.all(res["status"] == 0)
    for i, (k, n, e, lab) in enumerate(qs):
        if k == cac.QUERY_CC:
            want, nc = ctx.cc(np.stack([e["u"], e["v"]], 1), n)
            assert res[i]["ncomp"] == nc and np.array_equal(lab, want), i
        else:
            t, T = cac.serve_mincut_plan(n, len(e), F32_95)
            raw = ctx.sqrt_cut_trials(e, n, t, 5, 0, T)
            assert res[i]["t"] == t and res[i]["trials"] == T
            assert res[i]["cut"] == raw.min(), i
    assert np.all(res["admit_s"] >= res["submit_s"]) and np.all(res["done_s"] >= res["admit_s"])


def test_burst_forms_groups_that_overlap(ctx):
    """A burst arriving at once: queries waiting behind a running group are
    admitted together as launch groups, and groups on the two workers overlap."""
    rng = np.random.default_rng(7)
    qs = _stream(rng, 24)
    _say("serve burst of 24 queries")
    with cac.Server(0, workers=2, hbm_budget=8 << 30, max_batch=8) as srv:
        recs = [cac.Server.query(k, e, n, F32_95, 5, lab) for (k, n, e, lab) in qs]
        res = srv.replay(recs, np.zeros(len(qs)))
    assert np.all(res["status"] == 0)
    for i, (k, n, e, lab) in enumerate(qs):
        if k == cac.QUERY_CC:
            want, nc = ctx.cc(np.stack([e["u"], e["v"]], 1), n)
            assert res[i]["ncomp"] == nc and np.array_equal(lab, want), i
    assert res["batch_size"].max() > 1  # launch groups formed
    order = np.argsort(res["admit_s"])
    overlap = any(res[order[j + 1]]["admit_s"] < res[order[j]]["done_s"] for j in range(len(order) - 1))
    assert overlap  # a group was admitted while another ran


def test_bad_query_fails_alone(ctx):
    rng = np.random.default_rng(1)
    good = cac.gen_gnp(300, 6.0, 1, 0, wmax=5)
    bad = good.copy()
    bad["v"][3] = 300  # id >= n
    heavy = good.copy()
    heavy["w"][0] = 1 << 41
    lab = np.zeros(300, dtype=np.uint32)
    with cac.Server(0, workers=1, hbm_budget=1 << 30, max_batch=8) as srv:
        recs = [cac.Server.query(cac.QUERY_CC, bad, 300), cac.Server.query(cac.QUERY_CC, good, 300, labels=lab),
                cac.Server.query(cac.QUERY_MINCUT, heavy, 300, F32_95, 0),
                cac.Server.query(cac.QUERY_MINCUT, bad, 300, F32_95, 0),
                cac.Server.query(cac.QUERY_MINCUT, good, 300, F32_95, 0)]
        res = srv.replay(recs, np.zeros(len(recs)))
    assert res[0]["status"] == cac.CAC_EINVAL
    assert res[1]["status"] == 0
    want, nc = ctx.cc(np.stack([good["u"], good["v"]], 1), 300)
    assert res[1]["ncomp"] == nc and np.array_equal(lab, want)
    assert res[2]["status"] == cac.CAC_ERANGE
    assert res[3]["status"] == cac.CAC_EINVAL
    assert res[4]["status"] == 0
    t, T = cac.serve_mincut_plan(300, len(good), F32_95)
    assert res[4]["cut"] == ctx.sqrt_cut_trials(good, 300, t, 0, 0, T).min()


def test_footprint_over_budget_is_rejected():
    rng = np.random.default_rng(2)
    e = cac.gen_gnp(2000, 20.0, 2, 0, wmax=5)
    with cac.Server(0, workers=1, hbm_budget=4096, max_batch=4) as srv:
        res = srv.replay([cac.Server.query(cac.QUERY_CC, e, 2000)], np.zeros(1))
    assert res[0]["status"] == cac.CAC_ENOMEM
