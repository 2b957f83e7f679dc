"""The oracle's liveness restatement pinned (CPU only).

* The backward-pass zone operations against set semantics on a point grid (the grid argument of
  test_oracle_dbm.py): down (time predecessors), free (existential projection of clocks), and
  zone difference (pieces pairwise disjoint, union = p minus y).
* Verdicts against the reference's recorded results for its fischer family (gen_fischer.py):
  `P0.req --> P0.wait` satisfied (benchmarks/results/live/hora_backward/hora_live_t1.csv:67-88),
  `P0.cs --> P0.req` and `P0.wait --> P0.cs` violated (benchmarks/queries/live_violated/fischer.q,
  README.md:31-35), `A[] not deadlock` satisfied (benchmarks/results/deadlock/hora/
  hora_deadlock_t1.csv:69-85), and the four `E[]` rows of benchmarks/queries/eg/results-4-5.txt:2-11.
* Verdicts on small automata whose answers follow from their definition (tests/live_models.py):
  timelock, forced progress, a cycle witness that no single state certifies, a partial deadlock.
* Both covering routes give every verdict; the route itself follows the seed condition of
  online_seed_key (src/liveness/refinement.cpp:176-282).
* The early witnesses (leadsto_witness, online_eg_terminal_witness) settle the query the fixpoint
  would settle.
* The construction's storing rule against the reference's recorded liveness rows
  (tests/golden/live_recorded_counts.json, written from benchmarks/results/live/hora/
  hora_live_t1.csv): the number of records expanded equals the recorded `states` on fischer,
  arbiter, phil, twolock and hddi (the online construction, online_intern, src/liveness/
  construction.cpp:480-683) and on single (the base construction, intern, :863-1003, which
  structurally_empty_d selects for its one discrete state). A dropped retirement, a union not
  taken, a cover scanned in the wrong order or a record expanded after its retirement changes
  these counts.
"""
import json
import os

import numpy as np
import pytest

from oracle import binding
from ptctcl_amd import hip, liveness as lv, models
from ptctcl_amd.network import c_le

import live_models
import live_xml_cases
from test_oracle_dbm import INF, Grid, le, rand_zone

P = lv.DPred


def member_pts(pts, z, d, K):
    z = np.asarray(z, np.int64).reshape(d, d)
    ok = np.ones(len(pts), bool)
    for i in range(d):
        for j in range(d):
            if i == j or z[i, j] >= INF:
                continue
            v, weak = z[i, j] >> 1, z[i, j] & 1
            diff = pts[:, i] - pts[:, j]
            ok &= (diff <= K * v) if weak else (diff < K * v)
    return ok


@pytest.mark.parametrize("n", [1, 2, 3])
def test_down_matches_grid(n):
    rng = np.random.default_rng(300 + n)
    C = 3
    G = Grid(n, 3 * C + 3)
    d = n + 1
    for _ in range(40):
        z = rand_zone(rng, n, C)
        out, _ = binding.zone_op(hip.ZOP_DOWN, d, z.reshape(1, -1).astype(np.int32))
        mask = np.zeros(len(G.P), bool)
        for s in range(0, (C + 3) * G.K + 1):      # p + s/K for every grid delay that can matter
            sh = G.P.copy()
            sh[:, 1:] += s
            mask |= member_pts(sh, z, d, G.K)
        assert np.array_equal(out[0], G.canon(mask, C + 1)), z


@pytest.mark.parametrize("n", [1, 2, 3])
def test_free_matches_grid(n):
    rng = np.random.default_rng(320 + n)
    C = 3
    G = Grid(n, 3 * C + 3)
    d = n + 1
    for _ in range(40):
        z = rand_zone(rng, n, C)
        mask_bits = int(rng.integers(1, 1 << n)) << 1
        out, _ = binding.zone_op(hip.ZOP_FREE, d, z.reshape(1, -1).astype(np.int32),
                                 arg=np.array([mask_bits], np.int32))
        freed = [c for c in range(1, d) if (mask_bits >> c) & 1]
        # the grid is a product (itertools order): project out the freed axes and broadcast back
        A = G.R * G.K + 1
        m = G.member(z).reshape((A,) * n)
        proj = np.any(m, axis=tuple(c - 1 for c in freed), keepdims=True)
        mask = np.broadcast_to(proj, m.shape).reshape(-1)
        assert np.array_equal(out[0], G.canon(mask, C + 1)), (z, mask_bits)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_subtract_matches_grid(n):
    rng = np.random.default_rng(340 + n)
    C = 4
    G = Grid(n, C + 3)
    d = n + 1
    ps = np.stack([rand_zone(rng, n, C) for _ in range(60)])
    ys = np.stack([rand_zone(rng, n, C) for _ in range(60)])
    pieces = binding.zone_subtract(d, ps, ys)
    for p, y, pc in zip(ps, ys, pieces):
        assert pc is not None
        want = G.member(p) & ~G.member(y)
        masks = [G.member(q) for q in pc]
        got = np.zeros(len(G.P), bool)
        for m in masks:
            assert m.any()                       # no empty piece
            assert not (got & m).any()           # pairwise disjoint
            got |= m
        assert np.array_equal(got, want)


def test_dpred_negation_is_complement():
    net = models.fischer(3)
    a, b = P.at(net, "P0", "cs"), P.at(net, "P1", "req")
    for pred in (a, a & b, a | b, ~(a & b), P.true(), P.false(), P.var(net, "id", lv.OP_GE, 2)):
        for locs in np.ndindex(4, 4, 4):
            for idv in range(4):
                assert pred.eval(locs, [idv]) != (~pred).eval(locs, [idv])


@pytest.mark.parametrize("cover", [lv.COVER_AUTO, lv.COVER_AM])
@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_fischer_leadsto_recorded_verdicts(n, cover):
    net = models.fischer(n)
    for early in (True, False):
        r, st = binding.live(net, lv.leads_to(P.at(net, "P0", "req"), P.at(net, "P0", "wait"), early, cover))
        assert r and not st.capped and st.nonempty == 0
        r, st = binding.live(net, lv.leads_to(P.at(net, "P0", "cs"), P.at(net, "P0", "req"), early, cover))
        assert not r
        r, st = binding.live(net, lv.leads_to(P.at(net, "P0", "wait"), P.at(net, "P0", "cs"), early, cover))
        assert not r
        if early:
            assert st.early == 1     # a single waiting state already deadlocks or diverges: stop witness


@pytest.mark.parametrize("n", [4, 5])
def test_fischer_deadlock_free_recorded(n):
    net = models.fischer(n)
    for early in (True, False):
        r, st = binding.live(net, lv.not_deadlock(early))
        assert r


@pytest.mark.parametrize("n", [4, 5])
def test_fischer_eg_recorded_verdicts(n):
    net = models.fischer(n)
    x0 = [c_le(1, 40)]
    rows = [(P.true(), True), (~P.at(net, "P1", "cs"), True),
            (P.at(net, "P0", "idle") | P.at(net, "P0", "wait"), False),
            (~(P.at(net, "P1", "cs") & P.at(net, "P2", "cs")), True)]
    for psi, want in rows:
        rnet, q = lv.exists_always(net, psi, x0)
        r, st = binding.live(rnet, q)
        assert r == want and not st.capped


@pytest.mark.parametrize("builder", live_models.ALL, ids=lambda b: b.__name__)
def test_small_models_known_verdicts(builder):
    net, queries = builder()
    for cover in (lv.COVER_AUTO, lv.COVER_AM):
        for q, want in queries:
            q.cover = cover
            r, st = binding.live(net, q)
            assert r == want, (net.name, q.kind, q.early, cover)


def test_route_choice():
    """fischer seeds no deadlock atom (each state with an upper invariant, `req`, can always leave
    it), so its graph is built under a_LU; the partial-deadlock automaton's first location seeds one
    (a valuation past x = 4 is stuck while the edge is enabled below it), so it takes the a_M graph."""
    net = models.fischer(5)
    r, st = binding.live(net, lv.leads_to(P.at(net, "P0", "req"), P.at(net, "P0", "wait"), early=False))
    assert r and st.route == 0 and st.seed_keys == 0
    r_am, st_am = binding.live(net, lv.leads_to(P.at(net, "P0", "req"), P.at(net, "P0", "wait"), early=False,
                                                cover=lv.COVER_AM))
    assert r_am and st_am.route == 1 and st_am.outer_nodes > st.outer_nodes
    net, _ = live_models.partial_deadlock()
    r, st = binding.live(net, lv.leads_to(P.at(net, "A", "a"), P.at(net, "A", "c"), early=False))
    assert not r and st.route == 1 and st.seed_keys == 1


def test_eg_terminal_witness():
    """E[] true on fischer: a state where every process idles lets time diverge, a maximal run
    that stays in psi, so the construction settles E[] at that level without the fixpoint; without
    the early witness the fixpoint gives the same verdict."""
    net = models.fischer(4)
    q = lv.LiveQuery(lv.EG, P.true())
    r, st = binding.live(net, q)
    assert r and st.early == 1 and st.closure == 0
    q.early = False
    r, st = binding.live(net, q)
    assert r and st.early == 0 and st.closure > 0


def test_early_witness_needs_no_fixpoint():
    net, _ = live_models.cycle()
    a, b = P.at(net, "A", "a"), P.at(net, "A", "b")
    r, st = binding.live(net, lv.leads_to(a, b))
    assert not r and st.early == 0 and st.nonempty >= 1      # cycle: only the fixpoint decides
    net, _ = live_models.timelock()
    a, b = P.at(net, "A", "a"), P.at(net, "A", "b")
    r, st = binding.live(net, lv.leads_to(a, b))
    assert not r and st.early == 1 and st.closure == 0       # stop witness at the first state


def _csma(n):
    import os
    from ptctcl_amd import uppaal
    return uppaal.load(os.path.join(os.path.dirname(__file__), "golden", "models", f"csma{n}.xml"))


@pytest.mark.parametrize("n", [4, 6])
def test_csma_broadcast_recorded_verdicts(n):
    """csma synchronises over a broadcast channel. `Bus.Transmitting --> Bus.Idle` satisfied
    (benchmarks/results/live/hora_backward/hora_live_t1.csv:36-41), `A[] not deadlock` satisfied
    (results/deadlock/hora/hora_deadlock_t1.csv:36-41), `Bus.Transmitting --> Bus.Collided`
    violated by a cycle (no early witness) and `Bus.Idle --> Bus.Collided` violated at the initial
    state (benchmarks/queries/live_violated/README.md:38-48), and the four convex `E[]` rows of
    benchmarks/queries/eg/results-4-5.txt:22-25 (csma4, all satisfied)."""
    net = _csma(n)
    T, I, Col = P.at(net, "Bus", "Transmitting"), P.at(net, "Bus", "Idle"), P.at(net, "Bus", "Collided")
    r, st = binding.live(net, lv.leads_to(T, I))
    assert r and not st.capped
    r, st = binding.live(net, lv.not_deadlock())
    assert r
    r, st = binding.live(net, lv.leads_to(T, Col))
    assert not r and st.early == 0
    r, st = binding.live(net, lv.leads_to(I, Col))
    assert not r and st.early == 1 and st.outer_nodes < 10
    if n == 4:
        y0 = net.clock_names.index("y0")
        for psi in (P.true(), ~Col, P.at(net, "S0", "Wait") | P.at(net, "S0", "Send"),
                    ~(P.at(net, "S0", "Send") & P.at(net, "S1", "Send"))):
            rnet, q = lv.exists_always(net, psi, [c_le(y0, 40)])
            r, st = binding.live(rnet, q)
            assert r and not st.capped


@pytest.mark.parametrize("case", live_xml_cases.CASES, ids=lambda c: f"{c[0]}:{c[1]}")
def test_reference_models_recorded_verdicts(case):
    """The reference's liveness rows on its own models (tests/live_xml_cases.py cites each)."""
    name, text, holds, shape, _ = case
    net = live_xml_cases.load(name)
    for early in (True, False):
        r, st = binding.live(net, live_xml_cases.leads_to(net, text, early))
        assert r == holds and not st.capped, (name, text, early)
        if early and not holds:
            live_xml_cases.check_shape(shape, st)


@pytest.mark.parametrize("case", live_xml_cases.DEADLOCK, ids=lambda c: c[0])
def test_reference_models_recorded_deadlock(case):
    name, want, _ = case
    net = live_xml_cases.load(name)
    for early in (True, False):
        r, st = binding.live(net, lv.not_deadlock(early))
        assert r == want and not st.capped, (name, early)


@pytest.mark.parametrize("case", live_xml_cases.EG, ids=lambda c: c[5])
def test_reference_models_recorded_eg(case):
    name, clock, bound, part, want, _ = case
    rnet, q = live_xml_cases.eg_query(live_xml_cases.load(name), clock, bound, part)
    r, st = binding.live(rnet, q)
    assert r == want and not st.capped


RECORDED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "live_recorded_counts.json")))


@pytest.mark.parametrize("row", [r for r in RECORDED if r["states"] <= 50_000], ids=lambda r: r["model"])
def test_recorded_construction_counts(row):
    name = row["model"]
    net = models.fischer(int(name[len("fischer"):])) if name.startswith("fischer") else live_xml_cases.load(name)
    r, st = binding.live(net, live_xml_cases.leads_to(net, row["query"]))
    assert r == row["holds"]
    assert st.expanded_nodes == row["states"], (name, st.expanded_nodes, row["states"], row["source"])
    assert st.route == 0 and st.early == 0


def test_online_construction_retires_and_joins():
    """The online rule on phil8: records retired by a newer covering record, convex unions taken,
    and fewer records expanded than stored (the retired ones that never got their turn)."""
    net = live_xml_cases.load("phil8")
    r, st = binding.live(net, live_xml_cases.leads_to(net, "P0.Eat --> P0.Dropped"))
    assert r and st.retired > 0 and st.unions > 0 and st.expanded_nodes < st.outer_nodes
