"""hora_hip_gfp_solve_noalloc against hora_hip_gfp_solve through the C ABI: two calls in a row on
the same GPU, the second with a smaller graph, each equal to the allocating entry point's result
(verdict, fixpoint counters and every node's converged emptiness); a graph beyond the static
scratch is refused with its reason."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import binding  # noqa: E402
from ptctcl_amd import hip, liveness as lv, models  # noqa: E402

pytestmark = pytest.mark.gpu
P = lv.DPred
COUNTERS = ("closure", "nonempty", "iters", "pre_calls", "products", "fed_max")


def both(net, q, g):
    with hip.HipContext(net, None) as ctx:
        r_a, st_a, w_a = ctx.gfp_solve(g, q)
        r_n, st_n, w_n = ctx.gfp_solve(g, q, entry="hora_hip_gfp_solve_noalloc")
    print(net.name, "alloc", r_a, [getattr(st_a, k) for k in COUNTERS],
          "noalloc", r_n, [getattr(st_n, k) for k in COUNTERS], flush=True)
    assert r_n == r_a
    for k in COUNTERS:
        assert getattr(st_n, k) == getattr(st_a, k), k
    assert np.array_equal(w_n, w_a)
    return r_n


def test_two_calls_second_smaller():
    big = models.fischer(6)
    q_big = lv.leads_to(P.at(big, "P0", "cs"), P.at(big, "P0", "req"), early=False)
    _, _, g_big = binding.live(big, q_big, graph=True)
    small = models.fischer(4)
    q_small = lv.leads_to(P.at(small, "P0", "req"), P.at(small, "P0", "wait"), early=False)
    _, _, g_small = binding.live(small, q_small, graph=True)
    assert g_small["row_off"].shape[0] < g_big["row_off"].shape[0]
    assert both(big, q_big, g_big) is False
    assert both(small, q_small, g_small) is True


def test_graph_beyond_scratch_is_refused():
    net = models.fischer(4)
    with hip.HipContext(net, None) as ctx:
        g = {k: np.zeros(1, np.int64) for k in ("dk", "zones", "etgt", "ese", "ere", "expanded")}
        g["row_off"] = np.zeros(2, np.int64)
        enc = lv.not_deadlock(False).encode()
        import ctypes as C
        st = lv.hora_live_stats()
        r = C.c_int32(0)
        t = {k: torch.from_numpy(v).cuda() for k, v in g.items()}
        gs = lv.hora_gfp_graph(50_000_000, 0, *(t[k].data_ptr() for k in ("dk", "zones", "row_off", "etgt", "ese",
                                                                         "ere", "expanded")))
        rc = hip.lib().hora_hip_gfp_solve_noalloc(ctx.h, C.byref(gs), C.byref(enc.struct), C.byref(r), C.byref(st),
                                                  None)
        assert rc == hip.E_MEMORY
        assert "beyond the static scratch" in hip.lib().hora_hip_last_error(ctx.h).decode()
