"""Broadcast synchronisation on the device against the oracle: verdict and every counter equal
(reachability), the same graph, stop sets and fixpoint (liveness).

bcast_net (tests/bcast_models.py) draws receivers that must join, receivers that may stay out
where their clock guards fail, processes with no receiving edge, committed locations and integer
guards; the hand network of test_oracle_bcast.py has all three options of one part; csma is the
reference's broadcast family, with its recorded liveness verdicts.
"""
import os

import pytest

from oracle import binding
from ptctcl_amd import hip, liveness as lv, uppaal
from ptctcl_amd.network import Target, c_gt, c_le, c_lt

from bcast_models import bcast_net
from test_gpu_live import check as live_check
from test_oracle_bcast import hand_net

pytestmark = pytest.mark.gpu
P = lv.DPred
KEYS = ("explored", "generated", "stored_live", "keys", "inserted", "subsumed", "killed", "unions",
        "levels", "found", "emode_switch_level")


def check(net, target):
    r_o, st_o = binding.reach(net, target, 0)
    with hip.HipContext(net, target) as ctx:
        r_h, st_h = ctx.reach()
    print(net.name, r_o, {k: (st_o[k], getattr(st_h, k)) for k in KEYS}, flush=True)
    assert r_h == r_o
    for k in KEYS:
        assert getattr(st_h, k) == st_o[k], k


# seeds whose broadcasts have at most 256 combinations in any state (the device's recipe list)
@pytest.mark.parametrize("N,seed", [(3, s) for s in range(6)] + [(4, s) for s in range(6)] +
                         [(5, s) for s in (0, 3, 4, 5)])
def test_bcast_net(N, seed):
    check(bcast_net(N, seed), None)


def test_bcast_target():
    net = bcast_net(4, 1)
    check(net, Target(loc_req=[net.resolve("P0", "hand"), net.resolve("P1", "done")]))


@pytest.mark.parametrize("guards", [[[c_lt(1, 2)], [c_gt(1, 3)]], [[], [c_gt(1, 3)]]])
@pytest.mark.parametrize("committed", [False, True])
def test_hand_network(guards, committed):
    check(hand_net(guards, committed), None)


def _live_queries(net):
    return [("ready-->done", lv.leads_to(P.at(net, "P0", "ready"), P.at(net, "P0", "done"))),
            ("idle-->hand full", lv.leads_to(P.at(net, "P0", "idle"), P.at(net, "P1", "hand"), early=False)),
            ("deadlock", lv.not_deadlock()),
            ("deadlock aM", lv.not_deadlock(cover=lv.COVER_AM)),
            ("EG !hand", lv.LiveQuery(lv.EG, ~P.at(net, "P0", "hand"), early=False))]


@pytest.mark.parametrize("N,seed", [(3, s) for s in range(6)] + [(4, s) for s in (0, 3, 5)])
def test_bcast_liveness(N, seed):
    """The liveness pass with broadcast transitions: the graph (edges carry the combination), stop
    sets (broadcast escapes) and exact predecessors (receivers' guards and resets, stay-out
    conjunctions) equal the oracle's, node for node."""
    net = bcast_net(N, seed)
    for label, q in _live_queries(net):
        live_check(net, q, label)


@pytest.mark.parametrize("n", [4, 6])
def test_csma_liveness_recorded(n):
    """csma's recorded verdicts (tests/test_oracle_live.py cites them), on the device."""
    net = uppaal.load(os.path.join(os.path.dirname(__file__), "golden", "models", f"csma{n}.xml"))
    T, I, Col = P.at(net, "Bus", "Transmitting"), P.at(net, "Bus", "Idle"), P.at(net, "Bus", "Collided")
    for label, q, want in [("T-->I", lv.leads_to(T, I), True), ("deadlock", lv.not_deadlock(), True),
                           ("T-->Col", lv.leads_to(T, Col), False), ("I-->Col", lv.leads_to(I, Col), False)]:
        r, _ = live_check(net, q, label)
        assert r == want, label
    if n == 4:
        y0 = net.clock_names.index("y0")
        for psi in (P.true(), ~Col, P.at(net, "S0", "Wait") | P.at(net, "S0", "Send"),
                    ~(P.at(net, "S0", "Send") & P.at(net, "S1", "Send"))):
            rnet, q = lv.exists_always(net, psi, [c_le(y0, 40)])
            r, _ = live_check(rnet, q, "EG")
            assert r
