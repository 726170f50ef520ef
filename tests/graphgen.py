"""Seeded synthetic graphs for the tests (numpy; test infrastructure)."""
from __future__ import annotations

import numpy as np


def planted(rng, n: int, deg: float, wmax: int, c: int):
    """Two G(n/2, deg/(n/2-1)) halves with weights U[1, wmax], joined by c
    distinct weight-1 cross edges; redrawn until each half is connected and the
    minimum weighted degree exceeds c, so the minimum cut is exactly c."""
    h = n // 2
    for _ in range(1000):
        edges = []
        for off, size in ((0, h), (h, n - h)):
            p = min(1.0, deg / max(1, size - 1))
            iu = np.triu_indices(size, 1)
            keep = rng.random(iu[0].size) < p
            u, v = iu[0][keep] + off, iu[1][keep] + off
            w = rng.integers(1, wmax + 1, size=u.size)
            edges.append(np.stack([u, v, w], axis=1))
        pairs = set()
        while len(pairs) < c:
            pairs.add((int(rng.integers(0, h)), int(rng.integers(h, n))))
        edges.append(np.array([[a, b, 1] for a, b in sorted(pairs)], dtype=np.int64))
        e = np.concatenate(edges).astype(np.int64)
        A = dense(n, e)
        if A.sum(axis=1).min() <= c:
            continue
        if _connected(A[:h, :h]) and _connected(A[h:, h:]):
            return e
    raise RuntimeError(f"planted: no valid graph for n={n} deg={deg} c={c}")


def necklace(rng, n: int, k: int, heavy: int, deg_in: float):
    """k clusters of about n/k vertices in a ring.  Inside a cluster a path plus
    about deg_in * size / 2 random chords, all of weight `heavy`; one edge of
    weight U[10^6, 1.1 * 10^6) joins each cluster to the next.  With heavy at
    or above the minimum weighted degree, contracted heavy edges reach the
    Karger-Stein bound (so the PR pass merges), while the ring's many near-equal
    2-cuts below it make the trials' raw cuts differ."""
    g = n // k
    e = []
    for c in range(k):
        base = c * g
        size = g if c < k - 1 else n - base
        for x in range(1, size):
            e.append((base + x - 1, base + x, heavy))
        for _ in range(int(deg_in * size / 2)):
            a, b = rng.integers(0, size, 2)
            if a != b:
                e.append((base + int(a), base + int(b), heavy))
    for c in range(k):
        a = c * g + int(rng.integers(0, g))
        b = ((c + 1) % k) * g + int(rng.integers(0, g))
        e.append((a, b, int(rng.integers(10**6, 11 * 10**5))))
    return np.array(e, dtype=np.int64)


def erdos_renyi(rng, n: int, p: float, wmax: int = 1):
    """G(n, p) with weights U[1, wmax], redrawn until connected."""
    for _ in range(1000):
        iu = np.triu_indices(n, 1)
        keep = rng.random(iu[0].size) < p
        u, v = iu[0][keep], iu[1][keep]
        w = rng.integers(1, wmax + 1, size=u.size)
        e = np.stack([u, v, w], axis=1).astype(np.int64)
        if _connected(dense(n, e)):
            return e
    raise RuntimeError(f"erdos_renyi: no connected graph for n={n} p={p}")


def dense(n: int, e) -> np.ndarray:
    A = np.zeros((n, n), dtype=np.int64)
    e = np.asarray(e, dtype=np.int64)
    if e.size:
        keep = e[:, 0] != e[:, 1]
        np.add.at(A, (e[keep, 0], e[keep, 1]), e[keep, 2])
        np.add.at(A, (e[keep, 1], e[keep, 0]), e[keep, 2])
    return A


def _connected(A) -> bool:
    from scipy.sparse.csgraph import connected_components

    k, _ = connected_components(A > 0, directed=False)
    return k == 1
