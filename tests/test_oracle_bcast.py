"""The oracle's broadcast synchronisation (build_broadcast_parts, semantics.hpp:275-303) on a
network small enough to derive every successor by hand (CPU only).

Sender S: a -> b on go!; receiver R: r0 -> r1 on go? [x < 2], r0 -> r2 on go? [x > 3]; T has no
receiving edge and takes no part. From (a, r0, t0) with 0 <= x <= 5 the broadcast has three
successors, in option order: R joins on its first edge (x < 2), on its second (x > 3), or stays
out where both guards fail (2 <= x <= 3); each is then delayed (no invariant), giving x >= 0,
x > 3 and x >= 2. With an unguarded receiving edge R cannot stay out; with S's source location
committed and R's not, the broadcast stays enabled (the sender is committed).
"""
import numpy as np

from oracle import binding
from ptctcl_amd.network import Channel, Edge, Location, Network, Process, c_ge, c_gt, c_lt, le, lt

INF = 1 << 30


def hand_net(r_guards, s_committed=False):
    S = Process("S", [Location("a", committed=s_committed), Location("b")], [Edge(0, 1, chan=0, dir=1)])
    loop = [Edge(k, k, guard=[c_ge(1, 10)]) for k in range(3)]   # keeps x active everywhere
    R = Process("R", [Location("r0"), Location("r1"), Location("r2")],
                [Edge(0, 1 + k, guard=g, chan=0, dir=-1) for k, g in enumerate(r_guards)] + loop)
    T = Process("T", [Location("t0")], [])
    return Network(1, [], [Channel("go", broadcast=True)], [S, R, T], name="bhand").finalize()


def successors(net):
    z = np.array([le(0), le(0), le(5), le(0)], np.int32)     # 0 <= x <= 5
    dks, zones, _, _ = binding.expand(net, None, np.array([0], np.uint64), z.reshape(1, -1))
    # key: S's location in bit 0, R's in bits 1-2
    return [((int(k) & 1, (int(k) >> 1) & 3), int(zz[1]), int(zz[2])) for k, zz in zip(dks, zones)]


def test_join_or_stay_out():
    got = successors(hand_net([[c_lt(1, 2)], [c_gt(1, 3)]]))
    assert got == [((1, 1), le(0), INF), ((1, 2), lt(-3), INF), ((1, 0), le(-2), INF)]


def test_unguarded_receiver_must_join():
    got = successors(hand_net([[], [c_gt(1, 3)]]))
    assert got == [((1, 1), le(0), INF), ((1, 2), lt(-3), INF)]


def test_committed_sender_keeps_the_broadcast():
    got = successors(hand_net([[c_lt(1, 2)], [c_gt(1, 3)]], s_committed=True))
    assert [g[0] for g in got] == [(1, 1), (1, 2), (1, 0)]
