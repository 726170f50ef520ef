"""Edge-list files in the reference format (src/input/GraphInputIterator.cpp:3-36):
one comment line, then "n m", then m lines "from to weight"."""
from __future__ import annotations

import numpy as np


def read_graph(path):
    with open(path) as f:
        f.readline()
        toks = f.read().split()
    n, m = int(toks[0]), int(toks[1])
    vals = np.array(toks[2:2 + 3 * m], dtype=np.uint64).reshape(m, 3)
    return n, vals


def write_graph(path, n, edges, comment="# generated"):
    edges = np.asarray(edges)
    with open(path, "w") as f:
        f.write(comment.rstrip("\n") + "\n")
        f.write(f"{n} {len(edges)}\n")
        for u, v, w in edges:
            f.write(f"{int(u)} {int(v)} {int(w)}\n")
