"""The reference's liveness queries on its benchmark models, with the verdicts it records (test
fixtures; the models are the copies in tests/golden/models/).

Satisfied rows: benchmarks/results/live/hora_backward/hora_live_t1.csv (line per row below).
Violated rows: benchmarks/queries/live_violated/*.q, "deliberately falsified variants"
(benchmarks/queries/README.md:12); the witness shapes are the ones
benchmarks/queries/live_violated/README.md:29-48 names: a stop witness found during the
construction, a cycle no single state certifies (no early witness), or the initial state.
"""
import os

from ptctcl_amd import liveness as lv, uppaal

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
