"""Small timed automata whose liveness answers follow from their definition (test fixtures).

Each builder returns the network and a dict of queries -> expected verdict. The reasoning for each
verdict is in the docstring; the answers do not depend on any engine.
"""
import random

from ptctcl_amd import liveness as lv
from ptctcl_amd.network import Channel, Edge, Location, Network, Process, c_ge, c_gt, c_le

P = lv.DPred


def _net(locs, edges, name, nclocks=1):
    return Network(nclocks, [], [], [Process("A", locs, edges, 0)], name=name).finalize()


def timelock():
    """a (x <= 5) -> b [x >= 10]: the edge can never fire and time stops at x = 5, so every run
    ends in a at a deadlock. `a --> b` fails at the first state (stop witness), `A[] not
    deadlock` fails, `E[] true` holds (the maximal run stays in `true`), `E[] a` holds."""
    net = _net([Location("a", invariant=[c_le(1, 5)]), Location("b")], [Edge(0, 1, guard=[c_ge(1, 10)])], "timelock")
    a, b = P.at(net, "A", "a"), P.at(net, "A", "b")
    return net, [(lv.leads_to(a, b), False), (lv.leads_to(a, b, early=False), False),
                 (lv.not_deadlock(), False), (lv.not_deadlock(False), False),
                 (lv.LiveQuery(lv.EG, P.true()), True), (lv.LiveQuery(lv.EG, a), True)]


def progress():
    """a (x <= 5) -> b [x >= 3]; b has no invariant and no edge. Time in a is bounded and the edge
    is enabled on [3, 5], so every run reaches b: `a --> b` holds. b enables no action ever: `A[]
    not deadlock` fails. No run stays in a forever: `E[] a` fails; `E[] true` holds (b lets time
    diverge)."""
    net = _net([Location("a", invariant=[c_le(1, 5)]), Location("b")], [Edge(0, 1, guard=[c_ge(1, 3)])], "progress")
    a, b = P.at(net, "A", "a"), P.at(net, "A", "b")
    return net, [(lv.leads_to(a, b), True), (lv.leads_to(a, b, early=False), True),
                 (lv.not_deadlock(), False), (lv.LiveQuery(lv.EG, a), False),
                 (lv.LiveQuery(lv.EG, P.true()), True)]


def cycle():
    """a (x <= 2) with a self-loop [x >= 1] {x := 0} and a -> b [x >= 1]. From a a transition is
    always reachable by delay (no deadlock) and time cannot diverge in a, so no single state is a
    witness; but the loop keeps the run in a forever: `a --> b` fails through the cycle only. `E[]
    a` holds, `A[] not deadlock` fails (b is stuck)."""
    net = _net([Location("a", invariant=[c_le(1, 2)]), Location("b")],
               [Edge(0, 0, guard=[c_ge(1, 1)], resets=[1]), Edge(0, 1, guard=[c_ge(1, 1)])], "cycle")
    a, b = P.at(net, "A", "a"), P.at(net, "A", "b")
    return net, [(lv.leads_to(a, b), False), (lv.leads_to(a, b, early=False), False),
                 (lv.LiveQuery(lv.EG, a), True), (lv.not_deadlock(), False)]


def forced():
    """a (x <= 2) -> b [x >= 1] {x := 0}; b (x <= 2) -> a [x >= 1] {x := 0}. Both locations must
    be left in time, neither deadlocks: `a --> b` and `b --> a` hold, `A[] not deadlock` holds,
    `E[] a` fails, `E[] true` holds."""
    net = _net([Location("a", invariant=[c_le(1, 2)]), Location("b", invariant=[c_le(1, 2)])],
               [Edge(0, 1, guard=[c_ge(1, 1)], resets=[1]), Edge(1, 0, guard=[c_ge(1, 1)], resets=[1])], "forced")
    a, b = P.at(net, "A", "a"), P.at(net, "A", "b")
    return net, [(lv.leads_to(a, b), True), (lv.leads_to(b, a), True), (lv.not_deadlock(), True),
                 (lv.LiveQuery(lv.EG, a), False), (lv.LiveQuery(lv.EG, P.true()), True)]


def partial_deadlock():
    """a (x <= 10) -> b [x <= 4]: valuations of a with x > 4 can never fire the edge and cannot
    delay past 10, so they are deadlocked; a run may wait past 4. `a --> b` fails (partial stop
    set, not the whole zone), `A[] not deadlock` fails. With the guard widened to x <= 10 instead,
    see `no_partial_deadlock`."""
    net = _net([Location("a", invariant=[c_le(1, 10)]), Location("b", invariant=[c_le(1, 0)]),
                Location("c")],
               [Edge(0, 1, guard=[c_le(1, 4)], resets=[1]), Edge(1, 2, resets=[1])], "partial")
    a, c = P.at(net, "A", "a"), P.at(net, "A", "c")
    return net, [(lv.leads_to(a, c), False), (lv.leads_to(a, c, early=False), False),
                 (lv.not_deadlock(), False)]


def no_partial_deadlock():
    """a (x <= 10) -> b [x <= 10] {x := 0}; b (x <= 0) -> c {x := 0}; c (x <= 1) -> a [x >= 1]
    {x := 0}. Every state can move on: `A[] not deadlock` holds, `a --> c` holds."""
    net = _net([Location("a", invariant=[c_le(1, 10)]), Location("b", invariant=[c_le(1, 0)]),
                Location("c", invariant=[c_le(1, 1)])],
               [Edge(0, 1, guard=[c_le(1, 10)], resets=[1]), Edge(1, 2, resets=[1]),
                Edge(2, 0, guard=[c_ge(1, 1)], resets=[1])], "nopartial")
    a, c = P.at(net, "A", "a"), P.at(net, "A", "c")
    return net, [(lv.leads_to(a, c), True), (lv.not_deadlock(), True), (lv.not_deadlock(False), True)]


ALL = [timelock, progress, cycle, forced, partial_deadlock, no_partial_deadlock]


def open_random(seed, nproc=4, nloc=4, nclk=2):
    """A seeded random network with no invariant and no urgency (time-open, so `p --> q` takes the
    base construction): each process has `nloc` locations and `nclk` clocks, one or two edges out of
    each location with random lower, strict-lower and upper clock guards, random resets, and a
    third of the edges synchronising on one of two channels. No verdict is fixed here: the tests
    compare engines on it."""
    rng = random.Random(seed)
    chans = [Channel("c0"), Channel("c1")]
    procs = []
    for p in range(nproc):
        clocks = [1 + p * nclk + k for k in range(nclk)]
        edges = []
        for i in range(nloc):
            for _ in range(rng.randint(1, 2)):
                g = []
                for x in clocks:
                    r = rng.random()
                    if r < 0.3:
                        g.append(c_ge(x, rng.randint(1, 5)))
                    elif r < 0.5:
                        g.append(c_le(x, rng.randint(2, 8)))
                    elif r < 0.6:
                        g.append(c_gt(x, rng.randint(0, 3)))
                rs = [x for x in clocks if rng.random() < 0.4]
                ch, d = -1, 0
                if rng.random() < 0.3:
                    ch, d = rng.randrange(2), rng.choice((1, -1))
                edges.append(Edge(i, rng.randrange(nloc), guard=g, resets=rs, chan=ch, dir=d))
        procs.append(Process(f"P{p}", [Location(f"l{i}") for i in range(nloc)], edges, 0))
    return Network(nproc * nclk, [], chans, procs, name=f"open{seed}").finalize()
