"""The device search on the reference's own benchmark models, read from their UPPAAL XML.

* Every counter equal to the oracle's sequential engine (the device search is bit-exact with it):
  arbiter, phil, twolock, fischer, and the families whose recorded counts were taken under the
  G-simulation (grc, single, hddi; hddi10 has 31 clocks).
* Past the sizes the oracle finishes in seconds, the reference's recorded counts are the pin
  (tests/golden/xml_reach_counts.json, each entry citing its CSV line).
* csma synchronises over a broadcast channel (its collision detection): counters equal the
  oracle's, and the recorded count at csma10.
"""
import json
import os

import pytest

from oracle import binding
from ptctcl_amd import hip, uppaal

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)
GOLD = json.load(open(os.path.join(HERE, "golden", "xml_reach_counts.json")))["models"]
KEYS = ("explored", "generated", "stored_live", "keys", "inserted", "subsumed", "killed", "unions",
        "levels", "found", "emode_switch_level")


def load(name):
    net = uppaal.load(os.path.join(HERE, "golden", "models", name + ".xml"))
    inv, tgt = uppaal.safety_target(net, GOLD[name]["query"])
    return net, tgt


@pytest.mark.parametrize("name", ["arbiter4", "arbiter8", "arbiter12", "phil4", "phil6", "phil8", "phil10",
                                  "twolock4", "twolock6", "twolock8", "fischer8", "grc4", "grc5", "single4",
                                  "single5", "single6", "hddi2", "hddi10", "csma4", "csma6", "csma8", "viking4",
                                  "viking6", "viking10"])
def test_counters_equal_oracle(name):
    net, tgt = load(name)
    r_o, st_o = binding.reach(net, tgt, 0)
    with hip.HipContext(net, tgt) as ctx:
        r_h, st_h = ctx.reach()
    print(name, r_o, {k: (st_o[k], getattr(st_h, k)) for k in KEYS}, flush=True)
    assert r_h == r_o and (not r_h) == GOLD[name]["verdict_satisfied"]
    for k in KEYS:
        assert getattr(st_h, k) == st_o[k], k


@pytest.mark.parametrize("name", ["arbiter16", "arbiter18", "twolock10", "twolock12", "twolock13", "fischer12",
                                  "csma10", "csma14", "csma16", "viking12", "viking15", "viking17", "viking20"])
def test_recorded_counts_at_size(name):
    net, tgt = load(name)
    with hip.HipContext(net, tgt) as ctx:
        r, st = ctx.reach()
    print(name, r, st.explored, GOLD[name]["states"], flush=True)
    assert not r and st.explored == GOLD[name]["states"], GOLD[name]["cite"]
