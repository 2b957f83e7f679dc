"""The reference's liveness queries on its benchmark models, with the verdicts it records (test
fixtures; the models are the copies in tests/golden/models/).

Satisfied leads-to rows: benchmarks/results/live/hora_backward/hora_live_t1.csv (line per row).
`A[] not deadlock`: benchmarks/results/deadlock/hora/hora_deadlock_t1.csv (line per row).
`E[]` with a clock part: benchmarks/queries/eg/results-4-5.txt (line per row; hora and verifyta
agree on every one).
Violated rows: benchmarks/queries/live_violated/*.q, "deliberately falsified variants"
(benchmarks/queries/README.md:12); the witness shapes are the ones
benchmarks/queries/live_violated/README.md:29-48 names: a stop witness found during the
construction, a cycle no single state certifies (no early witness), or the initial state.
"""
import os

from ptctcl_amd import liveness as lv, uppaal
from ptctcl_amd.network import c_le

HERE = os.path.dirname(__file__)
P = lv.DPred
EARLY, CYCLE, INITIAL = "early", "cycle", "initial"

# (model, query, holds, witness shape for violated rows, citation)
CASES = [
    ("arbiter4", "C0.cs --> C0.idle", True, None, "hora_live_t1.csv:2"),
    ("arbiter8", "C0.cs --> C0.idle", True, None, "hora_live_t1.csv:10"),
    ("arbiter4", "C0.cs --> C0.want", False, CYCLE, "live_violated/arbiter.q, README.md:40"),
    ("arbiter4", "C0.want --> C0.cs", False, None, "live_violated/arbiter.q"),
    ("grc4", "T0.in --> T0.far", True, None, "hora_live_t1.csv:166"),
    ("grc4", "T0.near --> T0.far", False, EARLY, "live_violated/grc.q, README.md:34"),
    ("grc4", "T0.far --> T0.in", False, INITIAL, "live_violated/grc.q, README.md:43"),
    ("hddi2", "ST1.station_z_sync --> ST1.station_z_idle", True, None, "hora_live_t1.csv:158"),
    ("hddi10", "ST1.station_z_sync --> ST1.station_z_idle", True, None, "hora_live_t1.csv:160"),
    ("phil4", "P0.Eat --> P0.Dropped", True, None, "hora_live_t1.csv:90"),
    ("phil6", "P0.Eat --> P0.Dropped", True, None, "hora_live_t1.csv:94"),
    ("phil4", "P0.Eat --> P0.Think", False, CYCLE, "live_violated/phil.q, README.md:41"),
    ("single4", "P.A --> P.A", True, None, "hora_live_t1.csv:149"),
    ("twolock4", "P0.reqB --> P0.waitB", True, None, "hora_live_t1.csv:126"),
    ("twolock6", "P0.reqB --> P0.waitB", True, None, "hora_live_t1.csv:130"),
    ("twolock4", "P0.waitB --> P0.reqB", False, EARLY, "live_violated/twolock.q, README.md:33"),
    ("twolock4", "P0.reqB --> P0.cs", False, None, "live_violated/twolock.q"),
]


def load(name):
    return uppaal.load(os.path.join(HERE, "golden", "models", name + ".xml"))


def leads_to(net, text, early=True):
    a, b = [t.strip() for t in text.split("-->")]
    return lv.leads_to(P.at(net, *a.split(".")), P.at(net, *b.split(".")), early)


def check_shape(shape, st):
    """The early-witness behaviour the reference records for a violated row."""
    if shape == CYCLE:
        assert st.early == 0
    elif shape == EARLY:
        assert st.early == 1
    elif shape == INITIAL:
        assert st.early == 1 and st.outer_nodes < 10


# (model, deadlock-free, citation)
DEADLOCK = [
    ("arbiter4", True, "hora_deadlock_t1.csv:2"),
    ("grc4", False, "hora_deadlock_t1.csv:146"),
    ("hddi2", True, "hora_deadlock_t1.csv:138"),
    ("hddi10", True, "hora_deadlock_t1.csv:140"),
    ("phil4", False, "hora_deadlock_t1.csv:155"),
    ("single4", True, "hora_deadlock_t1.csv:92"),
    ("twolock4", True, "hora_deadlock_t1.csv:101"),
    ("csma4", True, "hora_deadlock_t1.csv:36"),
]

# (model, clock, bound, discrete part of psi, holds, citation): E[] psi with psi = part && clock <= bound
EG = [
    ("arbiter4", "x0", 40, lambda n: P.true(), True, "results-4-5.txt:12"),
    ("arbiter4", "x0", 40, lambda n: ~P.at(n, "C1", "cs"), True, "results-4-5.txt:13"),
    ("arbiter4", "x0", 40, lambda n: P.at(n, "C0", "idle") | P.at(n, "C0", "want"), False, "results-4-5.txt:14"),
    ("arbiter4", "x0", 40, lambda n: ~(P.at(n, "C1", "cs") & P.at(n, "C2", "cs")), True, "results-4-5.txt:15"),
    ("phil4", "c0", 40, lambda n: P.true(), True, "results-4-5.txt:32"),
    ("phil4", "c0", 40, lambda n: ~P.at(n, "P1", "Eat"), True, "results-4-5.txt:33"),
    ("phil4", "c0", 40, lambda n: ~(P.at(n, "P0", "Eat") & P.at(n, "P1", "Eat")), True, "results-4-5.txt:34"),
    ("grc4", "x0", 15, lambda n: P.true(), True, "results-4-5.txt:40"),
    ("grc4", "x0", 15, lambda n: ~P.at(n, "T0", "in"), True, "results-4-5.txt:41"),
    ("grc4", "x0", 15, lambda n: P.at(n, "T0", "far") | P.at(n, "T0", "near"), True, "results-4-5.txt:42"),
    ("grc4", "x0", 15, lambda n: ~(P.at(n, "T0", "in") & P.at(n, "T1", "in")), True, "results-4-5.txt:43"),
    ("twolock4", "x0", 30, lambda n: P.true(), True, "results-4-5.txt:50"),
    ("twolock4", "x0", 30, lambda n: ~P.at(n, "P1", "cs"), True, "results-4-5.txt:51"),
    ("twolock4", "x0", 30, lambda n: P.at(n, "P0", "idle") | P.at(n, "P0", "waitA"), False, "results-4-5.txt:52"),
]


def eg_query(net, clock, bound, part):
    """(network restricted by clock <= bound, E[] query)."""
    return lv.exists_always(net, part(net), [c_le(net.clock_names.index(clock), bound)])


# Rows whose fixpoint or stop sets outgrow the device's federation lists, where the device reports
# HORA_HIP_E_CAPPED, the reference's `capped` (no verdict): none since the liveness kernels' lists
# go to device memory (up to the reference's 4096 zones per federation) when they outgrow LDS.
# hddi10 (31 clocks: LDS lists of 8 to 11 zones) and phil12 P0.Eat --> P0.Dropped (candidate
# unions of 120 zones of dimension 13) take that tier.
DEVICE_CAPPED: set = set()
