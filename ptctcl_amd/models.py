"""Seeded model generators for the workloads the back end is measured and checked on.

``fischer(N, k)`` is the reference's own benchmark family, built exactly as
benchmarks/scripts/generators/gen_fischer.py:13-46 writes it (recorded state counts for N = 4..15
are the golden pins). ``fischer_variant(N, k)`` is the mutual-exclusion variant the 4-process
plain-reachability workload names (guard ``x <= k`` on req->wait, no invariant, no reset on
idle->req; return edges as gen_fischer.py:36-41). ``token_ring(N, seed)`` is the seeded token ring
of the liveness workload.
"""
from __future__ import annotations

import random
from typing import Tuple

from .network import (Channel, Edge, IntVar, Location, Network, OP_EQ, OP_NE, Process, Target, c_ge, c_gt,
                      c_le)


def fischer(N: int, k: int = 2) -> Network:
    """Fischer's protocol, N flattened processes, one clock each, shared ``int[0,N] id``.

    Mirrors gen_fischer.py: idle->req [id==0] {x:=0}; req inv x<=k; req->wait {id:=pid, x:=0};
    wait->cs [x>k && id==pid]; wait->idle [x>k && id!=pid]; cs->idle {id:=0}.
    """
    procs = []
    for i in range(N):
        x = i + 1
        pid = i + 1
        locs = [Location("idle"), Location("req", invariant=[c_le(x, k)]), Location("wait"), Location("cs")]
        edges = [
            Edge(0, 1, int_guard=[(0, OP_EQ, 0)], resets=[x]),
            Edge(1, 2, int_updates=[(0, pid)], resets=[x]),
            Edge(2, 3, guard=[c_gt(x, k)], int_guard=[(0, OP_EQ, pid)]),
            Edge(2, 0, guard=[c_gt(x, k)], int_guard=[(0, OP_NE, pid)]),
            Edge(3, 0, int_updates=[(0, 0)]),
        ]
        procs.append(Process(f"P{i}", locs, edges, 0))
    return Network(N, [IntVar("id", 0, N, 0)], [], procs, name=f"fischer{N}").finalize()


def fischer_variant(N: int, k: int = 2) -> Network:
    """The mutual-exclusion variant of the plain-reachability workload.

    idle->req [id==0]; req->wait [x<=k] {id:=pid, x:=0}; wait->cs [x>k && id==pid];
    return edges as gen_fischer.py: wait->idle [x>k && id!=pid], cs->idle {id:=0}.
    """
    procs = []
    for i in range(N):
        x = i + 1
        pid = i + 1
        locs = [Location("idle"), Location("req"), Location("wait"), Location("cs")]
        edges = [
            Edge(0, 1, int_guard=[(0, OP_EQ, 0)]),
            Edge(1, 2, guard=[c_le(x, k)], int_updates=[(0, pid)], resets=[x]),
            Edge(2, 3, guard=[c_gt(x, k)], int_guard=[(0, OP_EQ, pid)]),
            Edge(2, 0, guard=[c_gt(x, k)], int_guard=[(0, OP_NE, pid)]),
            Edge(3, 0, int_updates=[(0, 0)]),
        ]
        procs.append(Process(f"P{i}", locs, edges, 0))
    return Network(N, [IntVar("id", 0, N, 0)], [], procs, name=f"fischerv{N}").finalize()


def mutex_target(net: Network, a: int = 1, b: int = 2) -> Target:
    """``Pa.cs && Pb.cs`` -- the body of ``A[] not (Pa.cs && Pb.cs)`` (benchmarks/queries/reach/fischer.q:1)."""
    return Target(loc_req=[net.resolve(f"P{a}", "cs"), net.resolve(f"P{b}", "cs")])


def token_ring(N: int, seed: int, clocks_per_proc: Tuple[int, int] = (1, 2)) -> Network:
    """Seeded token ring: process i holds the token in ``hold`` and passes it to i+1 mod N.

    Each process has 1-2 clocks (seeded). ``hold`` has invariant ``x <= c`` with c ~ U[1, 20];
    passing needs ``x >= c // 2`` (integer division) and synchronises on channel ``pass_i``
    with the next process's ``idle -> hold`` edge, which resets its clocks. A second clock,
    when present, times a ``work`` location entered from ``hold`` and left back to ``hold``
    (invariant ``y <= c2``, guard ``y >= 1``).
    """
    rng = random.Random(seed)
    nclk = [rng.randint(clocks_per_proc[0], clocks_per_proc[1]) for _ in range(N)]
    procs, chans = [], [Channel(f"pass{i}") for i in range(N)]
    base = 1
    for i in range(N):
        x = base
        y = base + 1 if nclk[i] == 2 else None
        base += nclk[i]
        c = rng.randint(1, 20)
        locs = [Location("idle"), Location("hold", invariant=[c_le(x, c)])]
        edges = [
            Edge(1, 0, guard=[c_ge(x, c // 2)], chan=i, dir=+1),                  # pass_i!
            Edge(0, 1, resets=[x] + ([y] if y else []), chan=(i - 1) % N, dir=-1),  # pass_{i-1}?
        ]
        if y is not None:
            c2 = rng.randint(1, 20)
            locs.append(Location("work", invariant=[c_le(y, c2)]))
            edges.append(Edge(1, 2, guard=[c_le(x, c - 1)] if c > 1 else [], resets=[y]))
            edges.append(Edge(2, 1, guard=[c_ge(y, 1)]))
        procs.append(Process(f"P{i}", locs, edges, 1 if i == 0 else 0))
    return Network(base - 1, [], chans, procs, name=f"ring{N}s{seed}").finalize()
