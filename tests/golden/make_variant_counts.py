"""Writes the sequential engine's counters on the mutual-exclusion variant of the plain-reachability
workload at the batched search's sizes, guard constant k = N (fischer_variant(N, N)):

  variant_counts.json           E<> P1.cs && P2.cs (A[] not (P1.cs && P2.cs)), full search, N = 8..10
  variant_deadlock_counts.json  A[] not deadlock with early witnesses, N = 12

The sequential search takes about an hour at N = 10 and the deadlock search half a minute at N = 12,
too long for a test; the device tests compare against these counters.
Run: python tests/golden/make_variant_counts.py [reach] [deadlock]"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import binding  # noqa: E402
from ptctcl_amd import liveness as lv  # noqa: E402
from ptctcl_amd import models  # noqa: E402

KEYS = ("explored", "generated", "stored_live", "keys", "inserted", "subsumed", "killed", "unions", "levels",
        "found", "emode_switch_level")
LIVE_KEYS = ("outer_nodes", "fwd_edges", "disc_keys", "expanded_nodes", "generated", "covered", "levels",
             "retired", "unions", "early", "route")


def reach():
    out = {}
    for n in (8, 9, 10):
        net = models.fischer_variant(n, n)
        t0 = time.time()
        reachable, st = binding.reach(net, models.mutex_target(net), 0)
        out[str(n)] = dict({k: st[k] for k in KEYS}, reachable=reachable, oracle_seconds=round(time.time() - t0, 1))
        print(n, out[str(n)], flush=True)
        with open(os.path.join(HERE, "variant_counts.json"), "w") as f:
            json.dump(out, f, indent=1)


def deadlock():
    out = {}
    for n in (12,):
        net = models.fischer_variant(n, n)
        t0 = time.time()
        holds, st = binding.live(net, lv.not_deadlock(True))
        out[str(n)] = dict({k: getattr(st, k) for k in LIVE_KEYS}, holds=holds,
                           oracle_seconds=round(time.time() - t0, 1))
        print(n, out[str(n)], flush=True)
    with open(os.path.join(HERE, "variant_deadlock_counts.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    which = sys.argv[1:] or ["reach", "deadlock"]
    if "deadlock" in which:
        deadlock()
    if "reach" in which:
        reach()
