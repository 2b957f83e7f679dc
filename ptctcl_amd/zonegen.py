"""Seeded synthetic DBMs for the zone-kernel microbenchmark (2^20 zones, d in {4, 8, 16, 32}).

Random matrices with v ~ U[-1000, 1000] are almost always empty once closed, so zones are built
as the canonical hull of a random point cloud instead, which is non-empty by construction, then
loosened at random entries (some to strict bounds, some to infinity) and closed again. The closure
here is numpy Floyd-Warshall over whole batches (min-plus with strictness), vectorised over zones.
"""
from __future__ import annotations

import numpy as np

INF = 1 << 30
ALU_NEG_INF = -(1 << 29)


def radd(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    inf = (a >= INF) | (b >= INF)
    v = (a >> 1) + (b >> 1)
    r = v * 2 + (a & b & 1)
    return np.where(inf, INF, r).astype(np.int32)


def close_batch(z: np.ndarray, d: int) -> np.ndarray:
    """Floyd-Warshall over a batch [n, d, d] (int32 raws)."""
    z = z.reshape(-1, d, d).copy()
    for k in range(d):
        via = radd(z[:, :, k:k + 1], z[:, k:k + 1, :])
        z = np.minimum(z, via)
    return z.reshape(-1, d * d)


def hull_zones(n: int, d: int, rng: np.random.Generator, loosen: float = 0.3, vmax: int = 1000) -> np.ndarray:
    """n canonical non-empty zones of dimension d."""
    nc = d - 1
    k = rng.integers(1, 4, size=n)
    pts = rng.integers(0, vmax + 1, size=(n, 3, nc)).astype(np.int64)
    # duplicate the first point where a zone uses fewer than 3 points
    for j in (1, 2):
        m = k <= j
        pts[m, j] = pts[m, 0]
    full = np.concatenate([np.zeros((n, 3, 1), np.int64), pts], axis=2)      # clock 0 = 0
    diff = full[:, :, :, None] - full[:, :, None, :]                         # [n, 3, d, d]
    mx = diff.max(axis=1)                                                    # [n, d, d]
    z = (mx * 2 + 1).astype(np.int64)
    # loosen: add slack, make strict, or drop to infinity
    r = rng.random((n, d, d))
    slack = rng.integers(0, 60, size=(n, d, d))
    z = np.where(r < loosen, z + 2 * slack, z)
    strict = rng.random((n, d, d)) < 0.25
    z = np.where(strict & (r < loosen), z + 1, z)          # (v,<=) -> (v+1,<): still holds every point
    drop = rng.random((n, d, d)) < 0.08
    z = np.where(drop, INF, z)
    idx = np.arange(d)
    z[:, idx, idx] = 1
    z = np.where(z >= INF, INF, z).astype(np.int32)
    return close_batch(z.reshape(n, d * d), d)


def unclose(z: np.ndarray, d: int, rng: np.random.Generator) -> np.ndarray:
    """Loosen entries of canonical zones so that closing them is real work (and stays non-empty)."""
    n = z.shape[0]
    z3 = z.reshape(n, d, d).astype(np.int64).copy()
    r = rng.random((n, d, d)) < 0.5
    z3 = np.where(r, np.minimum(z3 + 2 * rng.integers(1, 200, size=(n, d, d)), INF), z3)
    idx = np.arange(d)
    z3[:, idx, idx] = 1
    return z3.reshape(n, d * d).astype(np.int32)


def bounds(n: int, d: int, rng: np.random.Generator, neg_inf_frac: float = 0.0):
    L = rng.integers(0, 1001, size=(n, d)).astype(np.int32)
    U = rng.integers(0, 1001, size=(n, d)).astype(np.int32)
    if neg_inf_frac > 0:
        L[rng.random((n, d)) < neg_inf_frac] = ALU_NEG_INF
        U[rng.random((n, d)) < neg_inf_frac] = ALU_NEG_INF
    L[:, 0] = 0
    U[:, 0] = 0
    return L, U


def constraints(z: np.ndarray, d: int, rng: np.random.Generator) -> np.ndarray:
    """One (i, j, raw) per zone near the zone's current bound, so results are mixed empty/non-empty."""
    n = z.shape[0]
    i = rng.integers(0, d, size=n)
    j = (i + rng.integers(1, d, size=n)) % d
    cur = z.reshape(n, d, d)[np.arange(n), i, j].astype(np.int64)
    back = z.reshape(n, d, d)[np.arange(n), j, i].astype(np.int64)
    base = np.where(cur >= INF, np.where(back >= INF, 0, -(back >> 1)), cur >> 1)
    v = base + rng.integers(-300, 40, size=n)
    raw = v * 2 + rng.integers(0, 2, size=n)
    return np.stack([i, j, raw], axis=1).astype(np.int32)


def pairs_included(z: np.ndarray, d: int, rng: np.random.Generator):
    """(A, B) pairs: half B = A tightened by random constraints (so B is inside A), half independent."""
    n = z.shape[0]
    a = z
    b = hull_zones(n, d, rng)
    half = rng.random(n) < 0.5
    bt = a.reshape(n, d, d).astype(np.int64).copy()
    for _ in range(2):
        i = rng.integers(0, d, size=n)
        j = (i + rng.integers(1, d, size=n)) % d
        cur = bt[np.arange(n), i, j]
        back = bt[np.arange(n), j, i]
        # tighten towards the opposite bound but keep the zone non-empty: v in (-back, cur]
        lo = np.where(back >= INF, -1000, -(back >> 1))
        hi = np.where(cur >= INF, 1000, cur >> 1)
        v = np.where(hi > lo, rng.integers(0, 1 << 30, size=n) % np.maximum(hi - lo, 1) + lo, hi)
        new = v * 2 + 1
        bt[np.arange(n), i, j] = np.minimum(cur, new)
        bt = close_batch(bt.reshape(n, d * d).astype(np.int32), d).reshape(n, d, d).astype(np.int64)
    bt = bt.reshape(n, d * d).astype(np.int32)
    diag = bt.reshape(n, d, d)[:, np.arange(d), np.arange(d)]
    ok = (diag >= 1).all(axis=1)
    b = np.where((half & ok)[:, None], bt, b)
    return a, b


def pairs_union(z: np.ndarray, d: int, rng: np.random.Generator):
    """(A, B) pairs whose union is convex about half the time: a zone cut in two by a hyperplane
    x_i - x_j <= c / >= c (convex union), mixed with independent zones (almost never convex)."""
    n = z.shape[0]
    i = rng.integers(0, d, size=n)
    j = (i + rng.integers(1, d, size=n)) % d
    z3 = z.reshape(n, d, d).astype(np.int64)
    cur = z3[np.arange(n), i, j]
    back = z3[np.arange(n), j, i]
    lo = np.where(back >= INF, -1000, -(back >> 1))
    hi = np.where(cur >= INF, 1000, cur >> 1)
    c = lo + (rng.integers(0, 1 << 30, size=n) % np.maximum(hi - lo + 1, 1))
    a3 = z3.copy()
    b3 = z3.copy()
    a3[np.arange(n), i, j] = np.minimum(cur, c * 2 + 1)          # x_i - x_j <= c
    b3[np.arange(n), j, i] = np.minimum(back, -c * 2 + 1)        # x_j - x_i <= -c
    a = close_batch(a3.reshape(n, d * d).astype(np.int32), d)
    b = close_batch(b3.reshape(n, d * d).astype(np.int32), d)
    other = hull_zones(n, d, rng)
    pick = rng.random(n) < 0.5
    b = np.where(pick[:, None], b, other)
    # keep only non-empty operands (the operation's precondition)
    ok = (a.reshape(n, d, d)[:, np.arange(d), np.arange(d)] >= 1).all(axis=1) & \
         (b.reshape(n, d, d)[:, np.arange(d), np.arange(d)] >= 1).all(axis=1)
    b = np.where(ok[:, None], b, a)
    a = np.where(ok[:, None], a, z)
    b = np.where(ok[:, None], b, z)
    return a.astype(np.int32), b.astype(np.int32)
