"""Writes mincut_nx.json: exact minimum cuts computed by networkx.stoer_wagner
(an independent implementation, run in the build container only) for seeded
random weighted graphs.  The committed JSON is the fixture; this script is how
it was made:  python tests/golden/make_mincut_golden.py"""
import json
from pathlib import Path

import networkx as nx
import numpy as np


def main():
    cases = []
    rng = np.random.default_rng(20261202)
    for i in range(24):
        n = int(rng.integers(5, 60))
        p = float(rng.uniform(0.15, 0.9))
        wmax = int(rng.choice([1, 5, 100, 10**6]))
        while True:
            G = nx.gnp_random_graph(n, p, seed=int(rng.integers(1 << 30)))
            if nx.is_connected(G):
                break
        edges = []
        for u, v in G.edges():
            w = int(rng.integers(1, wmax + 1))
            G[u][v]["weight"] = w
            edges.append([int(u), int(v), w])
        cut, _ = nx.stoer_wagner(G, weight="weight")
        cases.append({"n": n, "edges": edges, "mincut": int(cut)})
    out = Path(__file__).with_name("mincut_nx.json")
    out.write_text(json.dumps(cases, separators=(",", ":")))
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
