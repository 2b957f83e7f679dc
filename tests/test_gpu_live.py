"""The device liveness pass against the oracle's restatement: same verdict, same zone graph (node
keys and zones, CSR rows, edge targets and transitions, expanded flags), same closure and the same
per-node emptiness of the converged fixpoint.

Cases: the reference's fischer family with its recorded queries (`P0.req --> P0.wait` satisfied,
`P0.cs --> P0.req` and `P0.wait --> P0.cs` violated, `A[] not deadlock` satisfied, four `E[]` rows
with a clock part), the seeded token ring (binary synchronisation), and the small automata of
tests/live_models.py with their definitional verdicts. At sizes the oracle needs minutes for, the
pin is the reference's own recorded count of each row (tests/golden/live_recorded_counts.json,
written from the reference's result tables by tests/golden/make_live_recorded_counts.py): the
number of records the construction expands. Both covering routes are checked: the automatic one
(the online construction unless a discrete state seeds a deadlock atom) and the a_M graph.
"""
import copy
import json
import os

import numpy as np
import pytest

from oracle import binding
from ptctcl_amd import hip, liveness as lv, models
from ptctcl_amd.network import c_le

import live_models
import live_xml_cases

pytestmark = pytest.mark.gpu
P = lv.DPred
RECORDED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "live_recorded_counts.json")))
STATS = ("outer_nodes", "expanded_nodes", "fwd_edges", "disc_keys", "generated", "covered", "retired", "unions",
         "closure", "nonempty", "early", "route", "seed_keys")
GRAPH = ("dk", "zones", "row_off", "etgt", "ese", "ere", "expanded", "in_s", "w_nonempty")


def check(net, query, label=""):
    print("case", net.name, label, query.kind, query.early, query.cover, flush=True)
    r_o, st_o, g_o = binding.live(net, query, graph=True)
    with hip.HipContext(net, None) as ctx:
        r_h, st_h = ctx.live(query)
        g_h = ctx.live_graph()
    print("  oracle", r_o, {k: getattr(st_o, k) for k in STATS}, flush=True)
    print("  device", r_h, {k: getattr(st_h, k) for k in STATS}, flush=True)
    assert r_h == r_o
    for k in STATS:
        assert getattr(st_h, k) == getattr(st_o, k), k
    if query.kind != lv.EG:
        assert st_h.levels == st_o.levels
    for k in GRAPH:
        assert np.array_equal(g_h[k], g_o[k]), k
    return r_h, st_h


def fischer_queries(net):
    a = lambda loc: P.at(net, "P0", loc)   # noqa: E731
    return [("req-->wait", lv.leads_to(a("req"), a("wait")), True),
            ("req-->wait full", lv.leads_to(a("req"), a("wait"), early=False), True),
            ("cs-->req full", lv.leads_to(a("cs"), a("req"), early=False), False),
            ("wait-->cs", lv.leads_to(a("wait"), a("cs")), False),
            ("wait-->cs full", lv.leads_to(a("wait"), a("cs"), early=False), False),
            ("deadlock", lv.not_deadlock(), True),
            ("deadlock full", lv.not_deadlock(False), True)]


@pytest.mark.parametrize("cover", [lv.COVER_AUTO, lv.COVER_AM])
@pytest.mark.parametrize("N", [3, 4, 5, 6])
def test_fischer_recorded_queries(N, cover):
    net = models.fischer(N)
    for label, q, want in fischer_queries(net):
        q.cover = cover
        r, _ = check(net, q, label)
        assert r == want, label


@pytest.mark.parametrize("N", [4, 5])
def test_fischer_eg_recorded(N):
    net = models.fischer(N)
    rows = [(P.true(), True), (~P.at(net, "P1", "cs"), True),
            (P.at(net, "P0", "idle") | P.at(net, "P0", "wait"), False),
            (~(P.at(net, "P1", "cs") & P.at(net, "P2", "cs")), True)]
    for k, (psi, want) in enumerate(rows):
        rnet, q = lv.exists_always(net, psi, [c_le(1, 40)])
        r, _ = check(rnet, q, f"eg row {k}")
        assert r == want


@pytest.mark.parametrize("N", [4, 6])
def test_eg_early_witness(N):
    net = models.fischer(N)
    for early in (True, False):
        r, st = check(net, lv.LiveQuery(lv.EG, P.true(), early=early), f"E[] true early={early}")
        assert r and st.early == (1 if early else 0)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("N", [4, 6, 8, 10])
def test_token_ring(N, seed):
    net = models.token_ring(N, seed)
    h0, h1, i0 = P.at(net, "P0", "hold"), P.at(net, "P1", "hold"), P.at(net, "P0", "idle")
    for q in (lv.leads_to(h0, h1), lv.leads_to(h0, i0, early=False), lv.leads_to(i0, h0, early=False),
              lv.not_deadlock(False), lv.LiveQuery(lv.EG, ~h1)):
        check(net, q)


@pytest.mark.parametrize("builder", live_models.ALL, ids=lambda b: b.__name__)
def test_small_models(builder):
    net, queries = builder()
    for cover in (lv.COVER_AUTO, lv.COVER_AM):
        for q, want in queries:
            q.cover = cover
            r, _ = check(net, q)
            assert r == want


def test_rerun_at_full_capacity(monkeypatch):
    """Stop sets and fixpoint rounds first run every node with small LDS lists and rerun the nodes
    that outgrow them with the full lists. With lists of one or two zones, one constraint and one
    transition nearly every node reruns; graphs, stop sets and converged values must not change."""
    monkeypatch.setenv("HORA_HIP_GFP_CAPS", "1,1,2,1")
    monkeypatch.setenv("HORA_HIP_STOP_CAPS", "1,1,1,1")
    net = models.fischer(5)
    for label, q, want in fischer_queries(net):
        r, _ = check(net, q, label)
        assert r == want, label
    for builder in live_models.ALL:
        net, queries = builder()
        for q, want in queries:
            r, _ = check(net, q)
            assert r == want


def test_rerun_through_middle_tier(monkeypatch):
    """With small lists of one or two zones and middle lists of two or three, nodes pass from the
    small launch to the middle one (the nodes that outgrow it in its own rerun list) and from there
    to the full lists: graphs, stop sets and converged values must not change."""
    monkeypatch.setenv("HORA_HIP_GFP_CAPS", "1,1,2,1")
    monkeypatch.setenv("HORA_HIP_GFP_MID_CAPS", "2,2,3")
    monkeypatch.setenv("HORA_HIP_STOP_CAPS", "1,1,1,1")
    monkeypatch.setenv("HORA_HIP_STOP_MID_CAPS", "2,2")
    net = models.fischer(5)
    for label, q, want in fischer_queries(net):
        r, _ = check(net, q, label)
        assert r == want, label
    for builder in live_models.ALL:
        net, queries = builder()
        for q, want in queries:
            r, _ = check(net, q)
            assert r == want


def test_device_memory_tier(monkeypatch):
    """Lists of one or two zones in every LDS launch: the nodes that outgrow the full LDS lists run
    again with their federation lists in device memory (up to 4096 zones, the reference's cap).
    Graphs, stop sets and converged values must not change, and that tier must have run."""
    monkeypatch.setenv("HORA_HIP_GFP_CAPS", "1,1,2,1")
    monkeypatch.setenv("HORA_HIP_GFP_MID_CAPS", "1,1,2")
    monkeypatch.setenv("HORA_HIP_GFP_FULL_CAPS", "1,1,2")
    monkeypatch.setenv("HORA_HIP_STOP_CAPS", "1,1,1,1")
    monkeypatch.setenv("HORA_HIP_STOP_FULL_CAPS", "1,1")
    tier = 0
    net = models.fischer(5)
    for label, q, want in fischer_queries(net):
        r, st = check(net, q, label)
        assert r == want, label
        tier += st.fed_hbm_nodes
    for builder in live_models.ALL:
        net, queries = builder()
        for q, want in queries:
            r, st = check(net, q)
            assert r == want
            tier += st.fed_hbm_nodes
    assert tier > 0


def test_capped_without_device_memory_tier(monkeypatch):
    """With the device-memory tier off (HORA_HIP_FED_HBM=0) and full LDS lists of one or two zones,
    a query whose federations outgrow them stops with HORA_HIP_E_CAPPED (the reference's `capped`,
    no verdict), and leaves no graph to read."""
    monkeypatch.setenv("HORA_HIP_FED_HBM", "0")
    monkeypatch.setenv("HORA_HIP_GFP_FULL_CAPS", "1,1,2")
    monkeypatch.setenv("HORA_HIP_STOP_FULL_CAPS", "1,1")
    net = models.fischer(5)
    capped = 0
    with hip.HipContext(net, None) as ctx:
        for label, q, want in fischer_queries(net):
            try:
                r, _ = ctx.live(q)
                assert r == want, label
            except hip.Capped as exc:
                capped += 1
                assert exc.stats.capped == 1
                with pytest.raises(hip.HipError, match="hora_hip_live_graph"):
                    ctx.live_graph()
    assert capped > 0


def test_pools_grow_and_compact(monkeypatch):
    """Stop sets and fixpoint values live in pools. With pools that start at the least they can
    hold and grow by exactly what is asked, the stop-set phase runs again after growing, and rounds
    compact the value pool and run again; graphs, stop sets and converged values must not change."""
    monkeypatch.setenv("HORA_HIP_LIVE_POOL", "tight")
    net = models.fischer(5)
    for label, q, want in fischer_queries(net):
        r, _ = check(net, q, label)
        assert r == want, label
    net = models.token_ring(6, 2)
    check(net, lv.leads_to(P.at(net, "P0", "hold"), P.at(net, "P1", "hold"), early=False))
    for builder in live_models.ALL:
        net, queries = builder()
        for q, want in queries:
            r, _ = check(net, q)
            assert r == want


def test_fischer7_full_graph_parity():
    """Fischer 7 node for node against the oracle on both routes; the automatic one expands the
    reference's recorded 11,951 records (live/hora/hora_live_t1.csv:73)."""
    net = models.fischer(7)
    for cover in (lv.COVER_AUTO, lv.COVER_AM):
        q = lv.leads_to(P.at(net, "P0", "req"), P.at(net, "P0", "wait"), early=False, cover=cover)
        r, st = check(net, q, f"cover {cover}")
        assert r
        if cover == lv.COVER_AUTO:
            assert st.expanded_nodes == 11951 and st.route == 0


def test_fischer8_full_graph_parity():
    """At fischer 8 (64,534 nodes) the oracle still finishes in about a second per query: the two
    violated leads-to queries, the deadlock query and an E[] query, all without early witnesses, so
    the whole fixpoint runs, node for node against the oracle."""
    net = models.fischer(8)
    a = lambda loc: P.at(net, "P0", loc)   # noqa: E731
    for label, q, want in (("cs-->req full", lv.leads_to(a("cs"), a("req"), early=False), False),
                           ("wait-->cs full", lv.leads_to(a("wait"), a("cs"), early=False), False),
                           ("deadlock full", lv.not_deadlock(False), True),
                           ("E[] !P1.cs", lv.LiveQuery(lv.EG, ~P.at(net, "P1", "cs"), early=False), True)):
        r, _ = check(net, q, label)
        assert r == want, label


def test_partial_deadlock_takes_the_am_route():
    net, _ = live_models.partial_deadlock()
    q = lv.leads_to(P.at(net, "A", "a"), P.at(net, "A", "c"), early=False)
    r, st = check(net, q)
    assert not r and st.route == 1 and st.seed_keys >= 1


def _recorded_net(row):
    name = row["model"]
    if name.startswith("fischer"):
        return models.fischer(int(name[len("fischer"):]))
    return live_xml_cases.load(name)


@pytest.mark.parametrize("row", [r for r in RECORDED if r["states"] <= 5_000_000], ids=lambda r: r["model"])
def test_recorded_construction_counts(row):
    """The reference's recorded liveness rows at full size, on the device alone: the verdict and the
    number of records the construction expands (the `states` column; the source line is in the
    fixture). Where the fixpoint outgrows the device's federation lists (live_xml_cases.DEVICE_CAPPED)
    there is no verdict, and the construction's count is read from the capped result."""
    net = _recorded_net(row)
    q = live_xml_cases.leads_to(net, row["query"], early=True)
    print("row", row["model"], row["query"], row["states"], row["source"], flush=True)
    capped = (row["model"], row["query"]) in live_xml_cases.DEVICE_CAPPED
    with hip.HipContext(net, None) as ctx:
        if capped:
            with pytest.raises(hip.Capped) as exc:
                ctx.live(q)
            st = exc.value.stats
            assert st.capped == 1
        else:
            r, st = ctx.live(q)
            assert r == row["holds"]
    print("  device", st.expanded_nodes, st.outer_nodes, st.retired, st.unions, flush=True)
    assert st.route == 0 and st.expanded_nodes == row["states"]


@pytest.mark.parametrize("name", ["single4", "single5", "single6"])
def test_base_construction_single(name):
    """single has one possible discrete state and no partial deadlock region: the reference decides
    `p --> q` by its base construction there (structurally_empty_d). The device's graph, node for
    node and edge for edge, against the oracle's, with and without the early witness; the online
    construction still answers under HORA_COVER_ONLINE."""
    net = live_xml_cases.load(name)
    assert lv.base_route(net) == 2
    for early in (True, False):
        q = live_xml_cases.leads_to(net, "P.A --> P.A", early=early)
        r, st = check(net, q, "early" if early else "full")
        assert r and st.route == 0 and st.levels == 0
    q = live_xml_cases.leads_to(net, "P.A --> P.A")
    q.cover = lv.COVER_ONLINE
    r, st = check(net, q, "online")
    assert r


def _time_open(net):
    """The network with every location invariant dropped: no upper invariant, so time-open when it
    has no urgency (structurally_empty_d's first case)."""
    r = copy.deepcopy(net)
    for p in r.procs:
        for loc in p.locations:
            loc.invariant = []
    r._analysed = False
    r.name = net.name + "open"
    return r.finalize()


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("N", [3, 5, 7])
def test_base_construction_time_open_ring(N, seed):
    """Time-open token rings: p --> q by the base construction (FIFO, in-place enlargement, convex
    unions, re-expansion of rewritten nodes) against the oracle's, graph and verdict."""
    net = _time_open(models.token_ring(N, seed))
    assert lv.base_route(net) == 1
    h0, h1, i0 = P.at(net, "P0", "hold"), P.at(net, "P1", "hold"), P.at(net, "P0", "idle")
    for q in (lv.leads_to(h0, h1), lv.leads_to(h0, h1, early=False), lv.leads_to(i0, h0, early=False),
              lv.leads_to(P.true(), i0, early=False)):
        check(net, q)


@pytest.mark.parametrize("seed", range(1, 16))
def test_base_construction_random_time_open(seed):
    """Seeded random time-open networks, where the base construction enlarges nodes in place,
    takes convex unions and expands rewritten nodes again: graph, counts and verdict against the
    oracle's, with and without the early witness."""
    net = live_models.open_random(seed)
    assert lv.base_route(net) == 1
    for q in (lv.leads_to(P.at(net, "P0", "l1"), P.at(net, "P0", "l2"), early=False),
              lv.leads_to(P.at(net, "P0", "l1"), P.at(net, "P1", "l2")),
              lv.leads_to(P.at(net, "P2", "l0"), P.at(net, "P3", "l3"))):
        check(net, q)


def test_base_construction_buffers_grow(monkeypatch):
    """The base construction's node, edge, key-array and FIFO buffers and its key table, started
    small (HORA_HIP_LIVE_POOL=tight) so that each grows and the table is rebuilt several times
    (time-open fischer 5: 3,552 nodes, 15,200 edges; single6: 844 nodes in one key): the graph
    and verdict against the oracle's."""
    monkeypatch.setenv("HORA_HIP_LIVE_POOL", "tight")
    net = _time_open(models.fischer(5))
    a = lambda loc: P.at(net, "P0", loc)   # noqa: E731
    check(net, lv.leads_to(a("req"), a("wait"), early=False))
    net = live_xml_cases.load("single6")
    check(net, live_xml_cases.leads_to(net, "P.A --> P.A"))
    for seed in (5, 14):
        net = live_models.open_random(seed)
        check(net, lv.leads_to(P.at(net, "P0", "l1"), P.at(net, "P0", "l2"), early=False))


@pytest.mark.parametrize("N", [3, 4, 5])
def test_base_construction_time_open_fischer(N):
    """Fischer's protocol with its invariants dropped (time-open, many discrete states with several
    zones each): the base construction's graph and verdicts against the oracle's."""
    net = _time_open(models.fischer(N))
    assert lv.base_route(net) == 1
    a = lambda loc: P.at(net, "P0", loc)   # noqa: E731
    for q in (lv.leads_to(a("req"), a("wait")), lv.leads_to(a("req"), a("wait"), early=False),
              lv.leads_to(a("wait"), a("cs"), early=False), lv.leads_to(a("cs"), a("req"))):
        check(net, q)


@pytest.mark.parametrize("case", live_xml_cases.CASES, ids=lambda c: f"{c[0]}:{c[1]}")
def test_reference_models_recorded_verdicts(case):
    """The reference's liveness rows on its own models (tests/live_xml_cases.py cites each): the
    recorded verdict and witness shape, and the oracle's graph and fixpoint, node for node."""
    name, text, holds, shape, _ = case
    net = live_xml_cases.load(name)
    if (name, text) in live_xml_cases.DEVICE_CAPPED:
        with hip.HipContext(net, None) as ctx:
            with pytest.raises(hip.HipError, match="capacity"):
                ctx.live(live_xml_cases.leads_to(net, text, False))
            # a query that stopped before its fixpoint finished leaves no graph to read
            with pytest.raises(hip.HipError, match="hora_hip_live_graph"):
                ctx.live_graph()
        return
    for early in (True, False):
        r, st = check(net, live_xml_cases.leads_to(net, text, early), f"{name} {text}")
        assert r == holds
        if early and not holds:
            live_xml_cases.check_shape(shape, st)


@pytest.mark.parametrize("case", live_xml_cases.DEADLOCK, ids=lambda c: c[0])
def test_reference_models_recorded_deadlock(case):
    name, want, _ = case
    net = live_xml_cases.load(name)
    if (name, "deadlock") in live_xml_cases.DEVICE_CAPPED:
        with hip.HipContext(net, None) as ctx:
            with pytest.raises(hip.HipError, match="capacity"):
                ctx.live(lv.not_deadlock(False))
        return
    for early in (True, False):
        r, _ = check(net, lv.not_deadlock(early), f"{name} deadlock")
        assert r == want


@pytest.mark.parametrize("case", live_xml_cases.EG, ids=lambda c: c[5])
def test_reference_models_recorded_eg(case):
    name, clock, bound, part, want, _ = case
    rnet, q = live_xml_cases.eg_query(live_xml_cases.load(name), clock, bound, part)
    r, _ = check(rnet, q, f"{name} EG")
    assert r == want


def test_urgent_location_deadlock():
    """(B, x = 0) in tests/urgent_models.py's urgent-location network can neither delay nor leave:
    `A[] not deadlock` fails, found at its turn or by the scan, on the device as in the oracle."""
    import urgent_models
    net, _, _, _ = urgent_models.urgent_location()
    for early in (True, False):
        r, _ = check(net, lv.not_deadlock(early))
        assert not r
