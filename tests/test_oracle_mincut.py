"""Pins the exact min-cut oracle (Stoer-Wagner): brute force over all cuts for
n <= 14, the reference's hand-checkable fixtures, and networkx.stoer_wagner
golden values."""
import json
from pathlib import Path

import numpy as np
import pytest
from graphfile import read_graph

GOLD = Path(__file__).resolve().parent / "golden"

# Answers derived by hand from the fixture files (test_inputs/*.graph in the reference).
FIXTURE_CUTS = {
    "cycle_4.graph": 200,      # 4-cycle, weights 100
    "cycle_6.graph": 10000,    # 4-cycle plus path 3-4-5: leaf 5 hangs on one edge of 10000
    "cycle_10.graph": 10,      # 4-cycle plus a path; edge 5-6 weighs 10
    "cycle_20.graph": 100,     # edge 5-6 weighs 100
    "complete_4.graph": 85,    # {0,1} | {2,3}: 57 + 15 + 1 + 12
}


def _rand_matrix(rng, n, density, wmax):
    A = np.zeros((n, n), dtype=np.int64)
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < density:
                A[i, j] = A[j, i] = rng.integers(1, wmax + 1)
    return A


@pytest.mark.parametrize("seed", range(40))
def test_stoer_wagner_vs_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 15))
    A = _rand_matrix(rng, n, float(rng.uniform(0.2, 1.0)), int(rng.choice([1, 3, 50])))
    assert orc.stoer_wagner(A) == orc.brute_mincut(A)


@pytest.mark.parametrize("name,cut", sorted(FIXTURE_CUTS.items()))
def test_reference_fixture_cuts(orc, name, cut):
    n, e = read_graph(GOLD / "reference_inputs" / name)
    A = orc.edges_to_dense(n, e.astype(np.int64))
    assert orc.stoer_wagner(A) == cut
    assert orc.brute_mincut(A) == cut


def test_networkx_golden(orc):
    cases = json.loads((GOLD / "mincut_nx.json").read_text())
    assert len(cases) >= 20
    for c in cases:
        A = orc.edges_to_dense(c["n"], np.array(c["edges"], dtype=np.int64))
        assert orc.stoer_wagner(A) == c["mincut"]


def test_disconnected_is_zero(orc):
    A = np.zeros((4, 4), dtype=np.int64)
    A[0, 1] = A[1, 0] = 5
    A[2, 3] = A[3, 2] = 7
    assert orc.stoer_wagner(A) == 0 == orc.brute_mincut(A)
