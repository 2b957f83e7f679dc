"""The oracle's forward search against the reference's recorded state counts (CPU only)."""
import json
import os

import numpy as np
import pytest

from oracle import binding
from ptctcl_amd import models

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reach_counts.json")))


@pytest.mark.parametrize("N", [4, 5, 6, 7, 8, 9])
def test_fischer_recorded_counts_seq(N):
    net = models.fischer(N)
    reachable, st = binding.reach(net, models.mutex_target(net), 0)
    g = GOLD["fischer_reach"][str(N)]
    assert (not reachable) == g["verdict_satisfied"]
    assert st["explored"] == g["states"]


@pytest.mark.parametrize("N", [6, 8])
def test_fischer_recorded_counts_omp(N):
    # order-invariant family: the parallel engine reaches the same count at any thread count
    net = models.fischer(N)
    _, st = binding.reach(net, models.mutex_target(net), 4)
    assert st["explored"] == GOLD["fischer_reach"][str(N)]["states"]


def test_fischer_k_scaling_invariance():
    # Fischer's zone graph with guard = invariant constant k does not depend on k
    for k in (2, 5, 7):
        net = models.fischer(6, k)
        assert binding.reach(net, models.mutex_target(net), 0)[1]["explored"] == 3458


def test_reachable_target_stops_early():
    # P0.cs alone is reachable; the search stops at the first level that generates it
    net = models.fischer(5)
    from ptctcl_amd.network import Target
    t = Target(loc_req=[net.resolve("P0", "cs")])
    r, st = binding.reach(net, t, 0)
    assert r and st["found"] == 1
    assert st["explored"] < 977


def test_python_analyses_match_oracle():
    # the product's static analyses (ptctcl_amd/network.py) equal the oracle's restatement
    for net in (models.fischer(5), models.fischer_variant(4), models.token_ring(6, 3)):
        L, U, LL, UL, AM = binding.analyses(net)
        enc = net.encode()
        assert np.array_equal(L, enc.arrays["lbound"]) and np.array_equal(U, enc.arrays["ubound"])
        assert np.array_equal(LL, enc.arrays["lbound_loc"]) and np.array_equal(UL, enc.arrays["ubound_loc"])
        assert np.array_equal(AM, enc.arrays["active_mask"])


def test_token_ring_runs():
    net = models.token_ring(6, 1)
    r, st = binding.reach(net, None, 0)
    assert not r and st["explored"] > 0


@pytest.mark.parametrize("k", range(4))
def test_urgent_semantics_known_answers(k):
    """Urgent locations and urgent (binary, broadcast) channels: the answers and explored counts
    derived in tests/urgent_models.py (include/hora/detail/semantics.hpp:118-169)."""
    import urgent_models
    net, target, reachable, explored = urgent_models.ALL[k]()
    r, st = binding.reach(net, target, 0)
    assert r == reachable and st["explored"] == explored


def test_urgency_is_what_blocks_the_target():
    """The same networks with the urgency taken away: time passes and every target is reached, so the
    known answers above hold only under the urgent rule."""
    import urgent_models
    for k in (0, 1, 3):
        net, target, reachable, _ = urgent_models.ALL[k]()
        assert not reachable
        for p in net.procs:
            for loc in p.locations:
                loc.urgent = False
        for ch in net.channels:
            ch.urgent = False
        net._analysed = False
        net.finalize()
        r, _ = binding.reach(net, target, 0)
        assert r, net.name
