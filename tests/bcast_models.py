"""Seeded networks that synchronise over a broadcast channel (test fixtures).

``bcast_net(N, seed)``: N processes, one clock each, a shared ``int[0,2] v``, a broadcast channel
``go`` and a binary channel ``ack``. Every process can send ``go!`` from ``ready`` (resetting its
clock), and can receive ``go?`` from ``idle`` and ``ready`` on one or two edges whose clock guards
are drawn per seed (none, a lower bound, an upper bound, or both), so a broadcast meets receivers
that must join, receivers that may stay out (their guards fail in part of the zone), and processes
with no receiving edge. Some processes enter a committed location ``hand`` on receiving and leave
it by ``ack!`` or a local step; integer guards and updates on ``v`` gate some edges.
"""
import random

from ptctcl_amd.network import (Channel, Edge, IntVar, Location, Network, OP_EQ, OP_NE, Process, c_ge, c_gt,
                                c_le, c_lt)


def bcast_net(N: int, seed: int) -> Network:
    rng = random.Random(seed)
    GO, ACK = 0, 1
    procs = []
    for p in range(N):
        x = p + 1
        cap = rng.randint(3, 6)
        locs = [Location("idle"), Location("ready", invariant=[c_le(x, cap)]), Location("done"),
                Location("hand", committed=rng.random() < 0.5)]
        lo = rng.randint(1, cap - 1)
        edges = [
            Edge(0, 1, guard=[c_ge(x, 1)], resets=[x]),                                # idle -> ready
            Edge(1, 2, guard=[c_ge(x, lo)], resets=[x], chan=GO, dir=+1,               # ready: go!
                 int_updates=[(0, rng.randint(0, 2))]),
            Edge(2, 0, guard=[c_ge(x, 2)], resets=[x]),                                # done -> idle
            Edge(3, 0, chan=ACK, dir=+1) if rng.random() < 0.5 else Edge(3, 0),        # hand -> idle
            Edge(2, 0, chan=ACK, dir=-1),                                              # done: ack?
        ]

        def guard():
            k = rng.randrange(4)
            a, b = rng.randint(1, 3), rng.randint(2, 5)
            return [[], [c_gt(x, a)], [c_lt(x, b)], [c_ge(x, a), c_le(x, a + 2)]][k]

        for src in (0, 1):
            for _ in range(rng.randint(0, 2)):                                          # go? edges
                ig = [(0, rng.choice([OP_EQ, OP_NE]), rng.randint(0, 2))] if rng.random() < 0.3 else []
                edges.append(Edge(src, rng.choice([2, 3]), guard=guard(), int_guard=ig, resets=[x], chan=GO,
                                  dir=-1))
        procs.append(Process(f"P{p}", locs, edges, 0))
    return Network(N, [IntVar("v", 0, 2, 0)], [Channel("go", broadcast=True), Channel("ack")], procs,
                   name=f"bcast{N}s{seed}").finalize()
