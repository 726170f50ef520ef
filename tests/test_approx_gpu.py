"""Approximate min cut on the device: the level at which the subsampled copies
first disconnect equals the oracle's exactly (same keyed draws, same fixed-point
keep threshold)."""
import numpy as np
import pytest
from graphgen import dense, erdos_renyi, planted

import cac

pytestmark = pytest.mark.gpu
F32_95 = float(np.float32(0.95))


@pytest.mark.parametrize("seed", range(4))
def test_approx_matches_oracle(ctx, orc, seed):
    rng = np.random.default_rng(seed)
    if seed % 2:
        n = 500
        e = planted(rng, n, deg=10, wmax=100, c=int(rng.integers(1, 9)))
    else:
        n = 800
        e = erdos_renyi(rng, n, 0.02, wmax=int(rng.choice([1, 20, 1000])))
    print(f"approx seed {seed} n {n} m {len(e)}", flush=True)
    got = ctx.approx_cut(e, n, F32_95, 7)
    want = orc.approx_cut(n, e, F32_95, 7)
    assert got == want
    lam = orc.stoer_wagner(dense(n, e))
    assert got[0] <= 4 * lam * np.log2(n)  # O(log n) approximation from above


def test_approx_fixtures(ctx, orc):
    from pathlib import Path

    from graphfile import read_graph

    gold = Path(__file__).resolve().parent / "golden" / "reference_inputs"
    for name in ("cycle_4.graph", "cycle_10.graph", "complete_4.graph"):
        n, e = read_graph(gold / name)
        assert ctx.approx_cut(e.astype(np.int64), n, F32_95, 0) == orc.approx_cut(n, e, F32_95, 0)
