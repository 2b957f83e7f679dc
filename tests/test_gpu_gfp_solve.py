"""hora_hip_gfp_solve: the device's backward pass over a zone graph the caller supplies (the call
site of LivenessEngine::solve_greatest_fixpoint, src/liveness/fixpoint.cpp:48, and its readouts).
The graph is the oracle's own construction, moved to device memory; the verdict and every node's
converged emptiness must equal the oracle's fixpoint over the same graph, and, on the device's own
graph, the result of hora_hip_live."""
import numpy as np
import pytest

# torch first: the library then binds to the HIP runtime torch loaded, and the device arrays torch
# allocates are that runtime's
torch = pytest.importorskip("torch")

from oracle import binding  # noqa: E402
from ptctcl_amd import hip, liveness as lv, models  # noqa: E402

import live_models  # noqa: E402
import live_xml_cases  # noqa: E402

pytestmark = pytest.mark.gpu
P = lv.DPred


def cases():
    out = []
    for n in (4, 5, 6):
        net = models.fischer(n)
        a = lambda loc, net=net: P.at(net, "P0", loc)   # noqa: E731
        out += [(net, lv.leads_to(a("req"), a("wait"), early=False)),
                (net, lv.leads_to(a("cs"), a("req"), early=False)),
                (net, lv.not_deadlock(False)),
                (net, lv.LiveQuery(lv.EG, ~P.at(net, "P1", "cs"), early=False))]
    net = models.token_ring(6, 2)
    out.append((net, lv.leads_to(P.at(net, "P0", "hold"), P.at(net, "P1", "hold"), early=False)))
    for b in live_models.ALL:
        net, qs = b()
        for q, _ in qs:
            q.early = False
            out.append((net, q))
    for name, text in (("phil4", "P0.Eat --> P0.Think"), ("arbiter4", "C0.cs --> C0.want"),
                       ("twolock4", "P0.reqB --> P0.cs")):
        net = live_xml_cases.load(name)
        out.append((net, live_xml_cases.leads_to(net, text, early=False)))
    return out


@pytest.mark.parametrize("k", range(len(cases())))
def test_solve_on_oracle_graph(k):
    net, q = cases()[k]
    print("case", k, net.name, q.kind, flush=True)
    r_o, st_o, g_o = binding.live(net, q, graph=True)
    with hip.HipContext(net, None) as ctx:
        r_h, st_h, wne = ctx.gfp_solve(g_o, q)
    print("  oracle", r_o, st_o.closure, st_o.nonempty, " device", r_h, st_h.closure, st_h.nonempty, flush=True)
    assert r_h == r_o
    assert st_h.closure == st_o.closure and st_h.nonempty == st_o.nonempty
    assert np.array_equal(wne, g_o["w_nonempty"])


def test_solve_on_device_graph():
    """The device's own graph handed back through the boundary decides as hora_hip_live did."""
    net = models.fischer(6)
    q = lv.leads_to(P.at(net, "P0", "cs"), P.at(net, "P0", "req"), early=False)
    with hip.HipContext(net, None) as ctx:
        r1, st1 = ctx.live(q)
        g = ctx.live_graph()
        r2, st2, wne = ctx.gfp_solve(g, q)
    assert r1 == r2 and not r1
    assert st2.closure == st1.closure and st2.nonempty == st1.nonempty
    assert np.array_equal(wne, g["w_nonempty"])


def test_bad_graph_is_refused():
    net = models.fischer(4)
    with hip.HipContext(net, None) as ctx:
        g = {k: np.zeros(0, np.int64) for k in ("dk", "zones", "etgt", "ese", "ere", "expanded")}
        g["row_off"] = np.zeros(3, np.int64)      # two nodes, but no key, zone or flag arrays
        with pytest.raises(hip.HipError, match="graph arrays"):
            ctx.gfp_solve(g, lv.not_deadlock(False))
