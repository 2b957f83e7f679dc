"""The device search against the oracle's sequential engine: verdict and every counter equal.

fischer: the reference's family (recorded counts are the golden pins); fischer_variant: the
plain-reachability variant, order-sensitive from N = 7 (the parallel CPU engine's counts differ
there, the sequential and device ones must not); token_ring: binary synchronisation.
"""
import json
import os

import pytest

from oracle import binding
from ptctcl_amd import hip, models
from ptctcl_amd.network import Target

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reach_counts.json")))
# every counter with a meaning in the sequential engine; chain_reads counts work, which the
# device's parallel cover pre-filter changes by design
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
    return r_h, st_h


@pytest.mark.parametrize("N", [4, 5, 6, 7, 8])
def test_fischer_counts(N):
    net = models.fischer(N)
    r, st = check(net, models.mutex_target(net))
    assert not r and st.explored == GOLD["fischer_reach"][str(N)]["states"]


@pytest.mark.parametrize("N", [10, 11, 12, 13, 14, 15])
def test_fischer_recorded_counts_at_size(N):
    """Past what the oracle finishes in seconds, the reference's recorded counts are the pin (its
    16-thread runs take 1.6 s to 1,358 s for these sizes). Fischer 15 stores 157,592,231 zones of
    dimension 16 as dense int32 (about 161 GB of zones, 214 GiB of device memory in all): its
    levels run in chunks and the zone arena grows by mapping memory, never by copying."""
    net = models.fischer(N)
    with hip.HipContext(net, models.mutex_target(net)) as ctx:
        r, st = ctx.reach()
    print(f"fischer{N}", r, st.explored, st.stored_live, flush=True)
    assert not r and st.explored == GOLD["fischer_reach"][str(N)]["states"]


@pytest.mark.parametrize("N", [6, 8])
def test_chunked_levels(N, monkeypatch):
    """A level split into chunks of at most 2,000 successor slots, expanded and merged one after
    another, is the sequential fold in the same order: every counter as the oracle's, on the
    order-sensitive variant (zones an earlier chunk stored and a later one retires leave the next
    frontier)."""
    monkeypatch.setenv("HORA_HIP_CHUNK_SLOTS", "2000")
    net = models.fischer_variant(N, k=N)
    check(net, models.mutex_target(net))


@pytest.mark.parametrize("N", [4, 5, 6, 7])
def test_variant_order_sensitive(N):
    net = models.fischer_variant(N, k=N)
    check(net, models.mutex_target(net))


def test_variant_config0():
    net = models.fischer_variant(4, k=2)
    r, st = check(net, models.mutex_target(net))
    assert not r


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_token_ring(seed):
    net = models.token_ring(6, seed)
    check(net, None)


def test_reachable_target_early_stop():
    net = models.fischer(6)
    r, st = check(net, Target(loc_req=[net.resolve("P2", "cs")]))
    assert r and st.found == 1


def test_clock_target():
    net = models.fischer(5)
    # P0 in wait with x0 > 3 (x0 is clock 1)
    t = Target(loc_req=[net.resolve("P0", "wait")], clock_cond=[(0, 1, -3 * 2)])
    check(net, t)


def test_adaptive_switch_level():
    """The model whose switch level follows from the 4x rule by hand (tests/extra_cases.py): level 3,
    on the device as in the oracle, with every counter equal."""
    import extra_cases
    net, target, level = extra_cases.switch_model()
    r, st = check(net, target)
    assert not r and st.emode_switch_level == level


@pytest.mark.parametrize("k", range(4))
def test_urgent_semantics(k):
    """Urgent locations and urgent channels (tests/urgent_models.py): the derived answer and
    explored count, on the device as in the oracle, every counter equal."""
    import urgent_models
    net, target, reachable, explored = urgent_models.ALL[k]()
    r, st = check(net, target)
    assert r == reachable and st.explored == explored


@pytest.mark.parametrize("make", [lambda: models.fischer(8), lambda: models.fischer_variant(7, k=7)])
@pytest.mark.parametrize("at", ["0", "100000", None])
def test_narrow_storage(make, at, monkeypatch):
    """Stored zones switch from int32 to int16 (the reference's narrowing rule for small constants,
    engine.cpp:1596-1611) once the int32 arena would pass HORA_HIP_NARROW_AT bytes: from the start
    (0), in place during the search (100,000 bytes: a few hundred zones), or never at these sizes
    (the 8 GiB default). Every counter equals the oracle's in each case; the variant exercises the
    heavy keys' pair pre-pass and the ordered merge reading narrow zones."""
    if at is not None:
        monkeypatch.setenv("HORA_HIP_NARROW_AT", at)
    net = make()
    _, st = check(net, models.mutex_target(net))
    assert st.zone_entry_bytes == (4 if at is None else 2)


def test_narrow_storage_off(monkeypatch):
    monkeypatch.setenv("HORA_HIP_NARROW_AT", "0")
    monkeypatch.setenv("HORA_HIP_NARROW", "0")
    net = models.fischer(7)
    _, st = check(net, models.mutex_target(net))
    assert st.zone_entry_bytes == 4


@pytest.mark.parametrize("at", ["0", "20000"])
def test_narrow_overflow_searches_again_wide(at, monkeypatch):
    """Constants of 20,000 do not fit 16 bits (raw bound 40,001). Unforced, such a model stays
    int32. Forced narrow, the first such entry is found where it is narrowed (a zone stored narrow,
    or the in-place rewrite of the arena), and the search runs again with int32 zones: every
    counter as the oracle's, on that search and on the context's next one."""
    net = models.fischer(5, k=20000)
    target = models.mutex_target(net)
    monkeypatch.setenv("HORA_HIP_NARROW_AT", at)
    _, st = check(net, target)
    assert st.zone_entry_bytes == 4
    monkeypatch.setenv("HORA_HIP_NARROW", "1")
    with hip.HipContext(net, target) as ctx:
        r1, st1 = ctx.reach()
        r2, st2 = ctx.reach()
    r_o, st_o = binding.reach(net, target, 0)
    for r, s in ((r1, st1), (r2, st2)):
        assert r == r_o and s.zone_entry_bytes == 4
        for k in KEYS:
            assert getattr(s, k) == st_o[k], k
