"""The batched forward search on the mutual-exclusion variant of the plain-reachability workload at
the batched search's sizes: fischer_variant(N, k=N), N = 8..12 processes, zone dimension 9..13.

E<> P1.cs && P2.cs (the body of A[] not (P1.cs && P2.cs)), full search: the device against the
sequential engine's counters, written by tests/golden/make_variant_counts.py for N = 8..10 (the
oracle needs about an hour at N = 10). The engine is order-sensitive on this family, so every
counter must match, not only the verdict; two counter identities hold as well: every stored zone is
still live, killed or absorbed by a union (inserted = stored_live + killed + unions), and every
generated successor was dropped as covered or stored, the initial zone being stored without being
generated (generated + 1 = subsumed + inserted).

A[] not deadlock, with early witnesses: the family deadlocks. In a deadlock state no process may be
in wait (one of its two edges is enabled once x > k) or in cs (its edge is unguarded); an idle
process needs id != 0 to be stuck, but id != 0 names the last process that wrote it, which sits in
wait or cs. So every process is in req with its clock past k, N transitions from the initial state:
the search stops at that record's turn in level N, having counted N + 1 levels. Device against the
oracle at N = 9..11 (seconds on the CPU) and against the oracle's fixture at N = 12.
"""
import json
import os

import pytest

from oracle import binding
from ptctcl_amd import hip, liveness as lv, models

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden")
REACH = json.load(open(os.path.join(HERE, "variant_counts.json")))
DEADLOCK = json.load(open(os.path.join(HERE, "variant_deadlock_counts.json")))
KEYS = ("explored", "generated", "stored_live", "keys", "inserted", "subsumed", "killed", "unions",
        "levels", "found", "emode_switch_level")
LIVE = ("outer_nodes", "fwd_edges", "disc_keys", "expanded_nodes", "generated", "covered", "levels",
        "retired", "unions", "early", "route")


@pytest.mark.parametrize("N", sorted(int(n) for n in REACH))
def test_variant_reach_counts(N):
    want = REACH[str(N)]
    net = models.fischer_variant(N, k=N)
    print("case variant reach", N, flush=True)
    with hip.HipContext(net, models.mutex_target(net)) as ctx:
        r, st = ctx.reach()
    print(N, r, {k: (getattr(st, k), want[k]) for k in KEYS}, flush=True)
    assert r == want["reachable"] is False
    for k in KEYS:
        assert getattr(st, k) == want[k], k
    assert st.inserted == st.stored_live + st.killed + st.unions
    assert st.generated + 1 == st.subsumed + st.inserted


def _deadlock(N):
    net = models.fischer_variant(N, k=N)
    print("case variant deadlock", N, flush=True)
    with hip.HipContext(net, None) as ctx:
        holds, st = ctx.live(lv.not_deadlock(True))
    print(N, holds, {k: getattr(st, k) for k in LIVE}, flush=True)
    assert not holds and st.early == 1 and st.levels == N + 1
    return net, st


@pytest.mark.parametrize("N", [9, 10, 11])
def test_variant_deadlock_early_witness(N):
    net, st = _deadlock(N)
    holds_o, st_o = binding.live(net, lv.not_deadlock(True))
    assert not holds_o
    for k in LIVE:
        assert getattr(st, k) == getattr(st_o, k), k


def test_variant_deadlock_early_witness_n12():
    _, st = _deadlock(12)
    want = DEADLOCK["12"]
    assert want["holds"] is False
    for k in LIVE:
        assert getattr(st, k) == want[k], k
