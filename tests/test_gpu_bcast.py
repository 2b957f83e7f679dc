"""Broadcast synchronisation on the device against the oracle: verdict and every counter equal.

bcast_net (tests/bcast_models.py) draws receivers that must join, receivers that may stay out
where their clock guards fail, processes with no receiving edge, committed locations and integer
guards; the hand network of test_oracle_bcast.py has all three options of one part.
"""
import pytest

from oracle import binding
from ptctcl_amd import hip, liveness as lv
from ptctcl_amd.network import Target

from bcast_models import bcast_net
from test_oracle_bcast import hand_net
from ptctcl_amd.network import c_gt, c_lt

pytestmark = pytest.mark.gpu
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


def test_liveness_refuses_broadcast():
    net = bcast_net(3, 0)
    with hip.HipContext(net, None) as ctx:
        with pytest.raises(hip.OutOfScope, match="broadcast"):
            ctx.live(lv.not_deadlock())
