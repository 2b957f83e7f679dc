"""The oracle's DBM algebra pinned against set semantics (CPU only).

A zone with integer bounds over n clocks is a union of regions, and every region contains a point
whose coordinates are multiples of 1/(n+1). So on that grid, two canonical zones are equal iff they
hold the same grid points. The canonical form itself can also be read off the grid: entry (i, j)
is (<=, g) when the largest grid value g of x_i - x_j in the zone is an integer, (<, ceil(g))
otherwise, and infinite when the difference runs to the grid's edge.

Each operation is checked against that brute force on small seeded zones:
close / constrain / up / reset (result = canonical form of the expected set), includes (set
inclusion), try_union (hull == union), a_LU (every point of Z is LU-simulated by a point of Z',
Herbreteau-Srivathsan-Walukiewicz's definition), Extra_LU+ (Z inside Extra(Z) inside a_LU(Z),
Behrmann et al. STTT 2006). Plus the reference's own known answers (tests/dbm_tests.cpp:24-75).
"""
import itertools

import numpy as np
import pytest

from oracle import binding
from ptctcl_amd import hip

INF = 1 << 30
NEG = -(1 << 29)


def le(v):
    return v * 2 + 1


def lt(v):
    return v * 2


class Grid:
    """Points x in [0, R]^n with step 1/K (K = n + 1), held as integers X = K * x."""

    def __init__(self, n, R):
        self.n, self.K, self.R = n, n + 1, R
        axis = np.arange(0, R * self.K + 1)
        pts = np.array(list(itertools.product(axis, repeat=n)), dtype=np.int64)
        self.P = np.concatenate([np.zeros((len(pts), 1), np.int64), pts], axis=1)   # x_0 = 0

    def member(self, z):
        d = self.n + 1
        z = np.asarray(z, np.int64).reshape(d, d)
        ok = np.ones(len(self.P), bool)
        for i in range(d):
            for j in range(d):
                if i == j or z[i, j] >= INF:
                    continue
                v, weak = z[i, j] >> 1, z[i, j] & 1
                diff = self.P[:, i] - self.P[:, j]
                ok &= (diff <= self.K * v) if weak else (diff < self.K * v)
        return ok

    def canon(self, mask, edge_margin):
        """Canonical DBM of the point set `mask` (None if empty)."""
        d = self.n + 1
        if not mask.any():
            return None
        P = self.P[mask]
        z = np.zeros((d, d), np.int64)
        for i in range(d):
            for j in range(d):
                if i == j:
                    z[i, j] = le(0)
                    continue
                g = (P[:, i] - P[:, j]).max()
                if g >= self.K * (self.R - edge_margin):
                    z[i, j] = INF
                elif g % self.K == 0:
                    z[i, j] = le(g // self.K)
                else:
                    z[i, j] = lt(-((-g) // self.K))
        return z.reshape(-1)


def rand_zone(rng, n, C):
    """A canonical non-empty zone: hull of 1-3 random integer points, loosened, closed."""
    d = n + 1
    k = rng.integers(1, 4)
    pts = np.concatenate([np.zeros((k, 1), np.int64), rng.integers(0, C + 1, size=(k, n))], axis=1)
    z = np.zeros((d, d), np.int64)
    for i in range(d):
        for j in range(d):
            z[i, j] = le(int((pts[:, i] - pts[:, j]).max()))
            if i != j and rng.random() < 0.3:
                z[i, j] = INF if rng.random() < 0.3 else z[i, j] + 2 * int(rng.integers(0, 2)) + int(rng.integers(0, 2))
    z[0, :] = np.minimum(z[0, :], le(0))     # clocks are non-negative
    out, flag = binding.zone_op(hip.ZOP_CLOSE, d, z.reshape(1, -1).astype(np.int32))
    assert flag[0] == 1
    return out[0].astype(np.int64)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_close_matches_grid_canonical_form(n):
    rng = np.random.default_rng(10 + n)
    C = 3
    G = Grid(n, 3 * C + 3)
    d = n + 1
    for _ in range(60):
        z = rng.integers(-C, C + 1, size=(d, d)) * 2 + rng.integers(0, 2, size=(d, d))
        z = np.where(rng.random((d, d)) < 0.25, INF, z)
        z[0, :] = np.minimum(z[0, :], le(0))     # clocks are non-negative
        np.fill_diagonal(z, le(0))
        mask = G.member(z)
        out, flag = binding.zone_op(hip.ZOP_CLOSE, d, z.reshape(1, -1).astype(np.int32))
        assert bool(flag[0]) == bool(mask.any())
        if flag[0]:
            assert np.array_equal(out[0], G.canon(mask, C + 1)), z


@pytest.mark.parametrize("n", [1, 2, 3])
def test_constrain_up_reset_match_grid(n):
    rng = np.random.default_rng(20 + n)
    C = 4
    G = Grid(n, 3 * C + 4)
    d = n + 1
    for _ in range(40):
        z = rand_zone(rng, n, C)
        base = G.member(z)
        assert base.any()
        # constrain
        a = int(rng.integers(0, d))
        b = int((a + rng.integers(1, d)) % d)
        r = le(int(rng.integers(-C, C + 1))) - int(rng.integers(0, 2))
        out, flag = binding.zone_op(hip.ZOP_CONSTRAIN, d, z.reshape(1, -1).astype(np.int32),
                                    arg=np.array([[a, b, r]], np.int32))
        cz = np.full((d, d), INF, np.int64)
        np.fill_diagonal(cz, le(0))
        cz[a, b] = r
        expect = base & G.member(cz)
        assert bool(flag[0]) == bool(expect.any())
        if flag[0]:
            assert np.array_equal(out[0], G.canon(expect, 2 * C + 1))
        # up: Y is in up(Z) iff Y - t*(1,..,1) is in Z for some grid t >= 0
        out, _ = binding.zone_op(hip.ZOP_UP, d, z.reshape(1, -1).astype(np.int32))
        side = G.R * G.K + 1
        cube = base.reshape((side,) * n)
        up = cube.copy()
        for t in range(1, side):
            sl_dst = tuple(slice(t, None) for _ in range(n))
            sl_src = tuple(slice(0, side - t) for _ in range(n))
            up[sl_dst] |= cube[sl_src]
        assert np.array_equal(out[0], G.canon(up.reshape(-1), 2 * C + 1))
        # reset(k): the projection of Z onto x_k = 0
        k = int(rng.integers(1, d))
        out, _ = binding.zone_op(hip.ZOP_RESET, d, z.reshape(1, -1).astype(np.int32), arg=np.array([k], np.int32))
        anyk = cube.any(axis=k - 1, keepdims=True)
        rs = np.zeros_like(cube)
        idx = [slice(None)] * n
        idx[k - 1] = slice(0, 1)
        rs[tuple(idx)] = anyk
        assert np.array_equal(out[0], G.canon(rs.reshape(-1), 2 * C + 1))


@pytest.mark.parametrize("n", [1, 2])
def test_includes_and_union_match_sets(n):
    rng = np.random.default_rng(30 + n)
    C = 4
    G = Grid(n, 3 * C + 4)
    d = n + 1
    seen = {True: 0, False: 0}
    for _ in range(150):
        a = rand_zone(rng, n, C)
        b = rand_zone(rng, n, C) if rng.random() < 0.5 else a.copy()
        if rng.random() < 0.6:   # make b a neighbour of a: cut a by a hyperplane
            i = int(rng.integers(0, d)); j = int((i + rng.integers(1, d)) % d)
            c = int(rng.integers(-C, C + 1))
            a2 = a.reshape(d, d).copy(); b2 = a.reshape(d, d).copy()
            a2[i, j] = min(a2[i, j], le(c)); b2[j, i] = min(b2[j, i], le(-c))
            oa, fa = binding.zone_op(hip.ZOP_CLOSE, d, a2.reshape(1, -1).astype(np.int32))
            ob, fb = binding.zone_op(hip.ZOP_CLOSE, d, b2.reshape(1, -1).astype(np.int32))
            if fa[0] and fb[0]:
                a, b = oa[0].astype(np.int64), ob[0].astype(np.int64)
        ma, mb = G.member(a), G.member(b)
        _, inc = binding.zone_op(hip.ZOP_INCLUDES, d, a.reshape(1, -1).astype(np.int32), b.reshape(1, -1).astype(np.int32))
        assert bool(inc[0]) == bool((mb & ~ma).sum() == 0)
        out, fl = binding.zone_op(hip.ZOP_UNION, d, a.reshape(1, -1).astype(np.int32), b.reshape(1, -1).astype(np.int32))
        hull = np.maximum(a, b)
        convex = np.array_equal(G.member(hull), ma | mb)
        assert bool(fl[0]) == convex
        seen[convex] += 1
        if convex:
            assert np.array_equal(out[0], hull)
    assert seen[True] > 10 and seen[False] > 10


def lu_simulated(v, Zp, G, L, U):
    """Is grid point v LU-simulated by some point of Zp? v' must satisfy, per clock x:
    v'_x = v_x, or L_x < v'_x < v_x, or U_x < v_x < v'_x (the LU preorder)."""
    n, K = G.n, G.K
    d = n + 1
    options = []
    for x in range(1, d):
        vx = v[x]
        opts = [("eq", vx)]
        lo = -1 if L[x] == NEG else L[x] * K
        if vx > lo + 1 or L[x] == NEG:
            opts.append(("between", (lo, vx)))
        if U[x] == NEG or vx > U[x] * K:
            opts.append(("above", vx))
        options.append(opts)
    for choice in itertools.product(*options):
        # DBM over scaled coordinates (integers): Zp scaled by K, plus the choice's bounds;
        # strict bounds between integers become (<= value - 1)
        z = np.array(Zp, np.int64).reshape(d, d)
        Z = np.full((d, d), INF, np.int64)
        for i in range(d):
            for j in range(d):
                if z[i, j] < INF:
                    val, weak = z[i, j] >> 1, z[i, j] & 1
                    Z[i, j] = K * val if weak else K * val - 1
        np.fill_diagonal(Z, 0)
        for x, (kind, val) in enumerate(choice, start=1):
            if kind == "eq":
                Z[x, 0] = min(Z[x, 0], val); Z[0, x] = min(Z[0, x], -val)
            elif kind == "between":
                lo, hi = val
                Z[x, 0] = min(Z[x, 0], hi - 1)
                if lo >= 0:
                    Z[0, x] = min(Z[0, x], -(lo + 1))
            else:
                Z[0, x] = min(Z[0, x], -(val + 1))
        # Floyd-Warshall on plain integers (a different representation from the oracle's)
        for k in range(d):
            Z = np.minimum(Z, np.where((Z[:, k:k + 1] >= INF) | (Z[k:k + 1, :] >= INF), INF, Z[:, k:k + 1] + Z[k:k + 1, :]))
        if (np.diag(Z) >= 0).all():
            return True
    return False


@pytest.mark.parametrize("n", [1, 2])
def test_alu_matches_lu_simulation(n):
    rng = np.random.default_rng(40 + n)
    C = 3
    G = Grid(n, 2 * C + 3)
    d = n + 1
    seen = {True: 0, False: 0}
    for _ in range(200 if n == 1 else 120):
        Z = rand_zone(rng, n, C)
        Zp = rand_zone(rng, n, C) if rng.random() < 0.6 else Z.copy()
        L = [0] + [int(v) if rng.random() > 0.2 else NEG for v in rng.integers(0, C + 1, size=n)]
        U = [0] + [int(v) if rng.random() > 0.2 else NEG for v in rng.integers(0, C + 1, size=n)]
        _, fl = binding.zone_op(hip.ZOP_ALU_LE, d, Z.reshape(1, -1).astype(np.int32), Zp.reshape(1, -1).astype(np.int32),
                                L=np.array([L], np.int32), U=np.array([U], np.int32))
        mz = G.member(Z)
        # points beyond every constant by more than one unit behave like the grid's last ones
        pts = G.P[mz]
        sim = all(lu_simulated(v, Zp, G, L, U) for v in pts)
        assert bool(fl[0]) == sim, (Z, Zp, L, U)
        seen[sim] += 1
    assert seen[True] > 5 and seen[False] > 5


@pytest.mark.parametrize("n", [1, 2])
def test_extra_lu_plus_properties(n):
    """Z is inside Extra_LU+(Z), and every point of Extra_LU+(Z) is LU-simulated by a point of Z."""
    rng = np.random.default_rng(50 + n)
    C = 3
    G = Grid(n, 2 * C + 3)
    d = n + 1
    for _ in range(80 if n == 1 else 30):
        Z = rand_zone(rng, n, C)
        L = np.array([[0] + list(rng.integers(0, C, size=n))], np.int32)
        U = np.array([[0] + list(rng.integers(0, C, size=n))], np.int32)
        out, _ = binding.zone_op(hip.ZOP_EXTRA_LU, d, Z.reshape(1, -1).astype(np.int32), L=L, U=U)
        E = out[0]
        mz, me = G.member(Z), G.member(E)
        assert not (mz & ~me).any()
        for v in G.P[me]:
            assert lu_simulated(v, Z, G, list(L[0]), list(U[0]))


def test_reference_known_answers():
    """tests/dbm_tests.cpp:15-75: zone1(lo, hi), reset/delay, inclusion, two-clock diagonal."""
    def zone1(lo, hi):
        z = np.array([[le(0), le(0)], [le(0), le(0)]], np.int32).reshape(1, -1)
        z, _ = binding.zone_op(hip.ZOP_UP, 2, z)
        if hi >= 0:
            z, _ = binding.zone_op(hip.ZOP_CONSTRAIN, 2, z, arg=np.array([[1, 0, le(hi)]], np.int32))
        z, f = binding.zone_op(hip.ZOP_CONSTRAIN, 2, z, arg=np.array([[0, 1, le(-lo)]], np.int32))
        return z, f
    z, f = zone1(0, 5)
    assert f[0] == 1 and z[0][2] == le(5) and z[0][1] == le(0)
    e = np.array([[le(0), le(0), le(0), le(0)]], np.int32)
    e, _ = binding.zone_op(hip.ZOP_UP, 2, e)
    e, _ = binding.zone_op(hip.ZOP_CONSTRAIN, 2, e, arg=np.array([[1, 0, le(5)]], np.int32))
    _, f = binding.zone_op(hip.ZOP_CONSTRAIN, 2, e, arg=np.array([[0, 1, le(-6)]], np.int32))
    assert f[0] == 0
    z, _ = zone1(2, 5)
    z, _ = binding.zone_op(hip.ZOP_RESET, 2, z, arg=np.array([1], np.int32))
    assert z[0][2] == le(0) and z[0][1] == le(0)
    z, _ = binding.zone_op(hip.ZOP_UP, 2, z)
    assert z[0][2] == INF and z[0][1] == le(0)
    a, _ = zone1(0, 5)
    b, _ = zone1(3, 5)
    assert binding.zone_op(hip.ZOP_INCLUDES, 2, a, b)[1][0] == 1
    assert binding.zone_op(hip.ZOP_INCLUDES, 2, b, a)[1][0] == 0
    # clocks x=1, y=2: delay; x<=5; y>=2; x-y<=1 non-empty; adding y-x<=-5 empties it
    z = np.full((1, 9), le(0), np.int32)
    z, _ = binding.zone_op(hip.ZOP_UP, 3, z)
    for c in ([1, 0, le(5)], [0, 2, le(-2)], [1, 2, le(1)]):
        z, f = binding.zone_op(hip.ZOP_CONSTRAIN, 3, z, arg=np.array([c], np.int32))
        assert f[0] == 1
    _, f = binding.zone_op(hip.ZOP_CONSTRAIN, 3, z, arg=np.array([[2, 1, le(-5)]], np.int32))
    assert f[0] == 0
