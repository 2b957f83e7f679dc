"""Small networks with urgent locations and urgent channels whose answers follow from their
definition (test fixtures). The rule restated (include/hora/detail/semantics.hpp:118-169): time
may not pass while some process sits in an urgent location, or while a synchronisation on an
urgent channel is enabled -- a send with an integer-enabled partner, or any integer-enabled send
on an urgent broadcast channel, which fires with no receiver. Edges on urgent channels carry no
clock guard.

Each builder returns (network, target, reachable, explored): the E<> answer and the number of
symbolic states the search explores, derived in the docstring. A back end that lets time pass in
any of these states finds the target, or explores other zones.
"""
from ptctcl_amd.network import (Channel, Edge, IntVar, Location, Network, OP_EQ, Process, Target, c_ge)

X = 1   # the one clock


def _target(net, proc, loc):
    p = [q.name for q in net.procs].index(proc)
    return Target(loc_req=[(p, net.procs[p].loc_index(loc))])


def urgent_location():
    """A -> B {x := 0}, B urgent, B -> C [x >= 1]. The initial zone is x >= 0 (A lets time pass).
    B is entered with x = 0 and time is frozen there, so B -> C is never enabled: C is unreachable.
    Explored: the initial state and (B, x = 0), 2. (Without urgency B's zone would be x >= 0 and C
    reachable.) `A[] not deadlock` fails at (B, x = 0)."""
    locs = [Location("A"), Location("B", urgent=True), Location("C")]
    edges = [Edge(0, 1, resets=[X]), Edge(1, 2, guard=[c_ge(X, 1)])]
    net = Network(1, [], [], [Process("P", locs, edges, 0)], name="urgent_loc").finalize()
    return net, _target(net, "P", "C"), False, 2


def urgent_channel(v0: int):
    """P: A -u!-> B, A -> D [x >= 1]; Q: Q0 -u? [v == 1]-> Q1; u urgent, int[0,1] v = v0.
    v0 = 1: the synchronisation is enabled in the initial state, so time is frozen there (zone
    x = 0) and A -> D never fires: D is unreachable. The search explores the initial state and
    (B, Q1), 2. v0 = 0: the receiver's integer guard fails, nothing is urgent, the initial zone is
    x >= 0 and D is reached at the first level (explored: the initial state, 1, when the target is
    found)."""
    u = 0
    p = Process("P", [Location("A"), Location("B"), Location("D")],
                [Edge(0, 1, chan=u, dir=+1), Edge(0, 2, guard=[c_ge(X, 1)])], 0)
    q = Process("Q", [Location("Q0"), Location("Q1")], [Edge(0, 1, chan=u, dir=-1, int_guard=[(0, OP_EQ, 1)])], 0)
    net = Network(1, [IntVar("v", 0, 1, v0)], [Channel("u", urgent=True)], [p, q], name=f"urgent_chan{v0}").finalize()
    return net, _target(net, "P", "D"), v0 == 0, 2 if v0 == 1 else 1


def urgent_broadcast():
    """P: A -b!-> B, A -> D [x >= 1]; b an urgent broadcast channel and no process receives on it.
    A broadcast send fires with no receiver, so it is enabled in the initial state and time is
    frozen there: D is unreachable. Explored: the initial state and (B, x >= 0), 2."""
    p = Process("P", [Location("A"), Location("B"), Location("D")],
                [Edge(0, 1, chan=0, dir=+1), Edge(0, 2, guard=[c_ge(X, 1)])], 0)
    net = Network(1, [], [Channel("b", broadcast=True, urgent=True)], [p], name="urgent_bcast").finalize()
    return net, _target(net, "P", "D"), False, 2


ALL = [urgent_location, lambda: urgent_channel(1), lambda: urgent_channel(0), urgent_broadcast]
