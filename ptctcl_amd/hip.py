"""Host API of the HIP back end, over the C ABI of include/hora_hip.h (ctypes).

Mirrors the reference's entry points for this path:
  * ``reachable_net(net, target, backend, stats)`` -- include/hora/reachability.hpp:61-62, with
    ``Backend.Hip`` beside the reference's ``Seq``/``Omp``/``Cuda``/``Metal`` (include/hora/reach.hpp:21);
  * ``ReachStats`` -- include/hora/reach.hpp:10-19;
  * the factory's out-of-scope answer (``make_net_gpu_*`` returning null,
    include/hora/gpu/network.hpp:121-128) surfaces as :class:`OutOfScope`.

Only the HIP back end lives here. The library must be built (``python -m ptctcl_amd.build``);
every call fails loudly if it is not, and nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import liveness as lv
from .network import NetEnc, Network, Target, TargetEnc, hora_reach_stats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhora_hip.so")

E_OK, E_ARG, E_SCOPE, E_DEVICE, E_MEMORY, E_INTERNAL, E_CAPPED = 0, -1, -2, -3, -4, -5, -6

(ZOP_CLOSE, ZOP_UP, ZOP_RESET, ZOP_CONSTRAIN, ZOP_EXTRA_LU, ZOP_INCLUDES, ZOP_ALU_LE, ZOP_UNION, ZOP_EMPTY,
 ZOP_DOWN, ZOP_FREE) = range(11)

EXPORTS = ("hora_hip_abi_version", "hora_hip_build_info", "hora_hip_open", "hora_hip_close",
           "hora_hip_last_error", "hora_hip_reach", "hora_hip_reach_profile", "hora_hip_zone_op",
           "hora_hip_live", "hora_hip_live_graph", "hora_hip_gfp_solve", "hora_hip_gfp_solve_noalloc",
           "hora_hip_server_open", "hora_hip_server_submit_reach",
           "hora_hip_server_submit_live", "hora_hip_server_wait", "hora_hip_server_now", "hora_hip_server_close")


class Backend(enum.Enum):
    Seq = 0
    Omp = 1
    Cuda = 2
    Metal = 3
    Hip = 4


class hora_reach_profile(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("expand_ms", "probe_ms", "merge_ms", "other_ms",
                                          "expand_bytes", "probe_bytes", "merge_bytes")] + \
               [(k, C.c_int64) for k in ("expand_launches", "probe_launches", "merge_launches")]


class HipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hora_hip error {code}: {msg}")
        self.code = code


class OutOfScope(HipError):
    """The network is outside the device path's scope; the reference's caller falls back to its CPU engine."""


class Capped(HipError):
    """A federation outgrew the device's capacity: no verdict (``GfpStats::capped``). ``stats`` holds
    the query's counters up to that point (the whole construction when the fixpoint capped)."""

    def __init__(self, code: int, msg: str, stats):
        super().__init__(code, msg)
        self.stats = stats


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run python -m ptctcl_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.hora_hip_abi_version.restype = C.c_int
        L.hora_hip_build_info.restype = C.c_char_p
        L.hora_hip_open.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.hora_hip_close.argtypes = [C.c_void_p]
        L.hora_hip_close.restype = None
        L.hora_hip_last_error.argtypes = [C.c_void_p]
        L.hora_hip_last_error.restype = C.c_char_p
        L.hora_hip_reach.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(hora_reach_stats)]
        L.hora_hip_reach_profile.argtypes = [C.c_void_p, C.POINTER(hora_reach_profile)]
        L.hora_hip_zone_op.argtypes = [C.c_int32, C.c_int32, C.c_int64] + [C.c_void_p] * 8
        L.hora_hip_live.argtypes = [C.c_void_p, C.POINTER(lv.hora_live_query), C.POINTER(C.c_int32),
                                    C.POINTER(lv.hora_live_stats)]
        L.hora_hip_live_graph.argtypes = [C.c_void_p, C.POINTER(lv.hora_live_graph)]
        L.hora_hip_gfp_solve.argtypes = [C.c_void_p, C.POINTER(lv.hora_gfp_graph), C.POINTER(lv.hora_live_query),
                                         C.POINTER(C.c_int32), C.POINTER(lv.hora_live_stats), C.c_void_p]
        L.hora_hip_gfp_solve_noalloc.argtypes = L.hora_hip_gfp_solve.argtypes
        _lib = L
    return _lib


@dataclass
class ReachStats:
    explored: int = 0
    generated: int = 0
    expanded: int = 0
    stored_live: int = 0
    keys: int = 0
    inserted: int = 0
    subsumed: int = 0
    killed: int = 0
    unions: int = 0
    chain_reads: int = 0
    levels: int = 0
    found: int = 0
    emode_switch_level: int = -1
    zone_entry_bytes: int = 4
    seconds: float = 0.0

    @classmethod
    def from_c(cls, s: hora_reach_stats) -> "ReachStats":
        return cls(**s.as_dict())


class HipContext:
    """One device context for one network and query (``hora_hip_open`` .. ``hora_hip_close``)."""

    def __init__(self, net: Network, target: Optional[Target]):
        self.enc = net.encode()
        self.tenc = TargetEnc(target)
        h = C.c_void_p()
        rc = lib().hora_hip_open(C.addressof(self.enc.struct), C.addressof(self.tenc.struct), C.byref(h))
        self.h = h
        if rc != E_OK:
            msg = lib().hora_hip_last_error(h).decode() if h else ""
            if h:
                lib().hora_hip_close(h)
                self.h = None
            raise (OutOfScope if rc == E_SCOPE else HipError)(rc, msg)

    def reach(self):
        r = C.c_int32(0)
        st = hora_reach_stats()
        rc = lib().hora_hip_reach(self.h, C.byref(r), C.byref(st))
        if rc != E_OK:
            raise HipError(rc, lib().hora_hip_last_error(self.h).decode())
        return bool(r.value), ReachStats.from_c(st)

    def live(self, query: "lv.LiveQuery"):
        """Decide one liveness query (``hora_hip_live``): returns (holds, LiveStats)."""
        enc = query.encode()
        r = C.c_int32(0)
        st = lv.hora_live_stats()
        rc = lib().hora_hip_live(self.h, C.byref(enc.struct), C.byref(r), C.byref(st))
        if rc == E_CAPPED:
            raise Capped(rc, lib().hora_hip_last_error(self.h).decode(), lv.LiveStats.from_c(st))
        if rc != E_OK:
            raise (OutOfScope if rc == E_SCOPE else HipError)(rc, lib().hora_hip_last_error(self.h).decode())
        return bool(r.value), lv.LiveStats.from_c(st)

    def gfp_solve(self, graph: dict, query: "lv.LiveQuery", entry: str = "hora_hip_gfp_solve"):
        """The backward pass over a zone graph (``hora_hip_gfp_solve``): `graph` holds the arrays of
        :func:`liveness.graph_arrays` (dk, zones, row_off, etgt, ese, ere, expanded) on the host; they
        are copied to device memory, and the result is (holds, LiveStats, per-node non-empty flags)."""
        import torch
        dev = torch.device("cuda")
        host = {k: np.ascontiguousarray(graph[k]) for k in ("dk", "zones", "row_off", "etgt", "ese", "ere", "expanded")}
        host["dk"] = host["dk"].view(np.int64)   # the same 64-bit keys, in a type torch moves
        t = {k: torch.from_numpy(v).to(dev) for k, v in host.items()}
        n, e = int(graph["row_off"].shape[0]) - 1, int(graph["etgt"].shape[0])
        wne = torch.zeros(max(n, 1), dtype=torch.int8, device=dev)
        g = lv.hora_gfp_graph(n, e, *(t[k].data_ptr() for k in ("dk", "zones", "row_off", "etgt", "ese", "ere",
                                                              "expanded")))
        enc = query.encode()
        r = C.c_int32(0)
        st = lv.hora_live_stats()
        torch.cuda.synchronize()
        rc = getattr(lib(), entry)(self.h, C.byref(g), C.byref(enc.struct), C.byref(r), C.byref(st), wne.data_ptr())
        if rc == E_CAPPED:
            raise Capped(rc, lib().hora_hip_last_error(self.h).decode(), lv.LiveStats.from_c(st))
        if rc != E_OK:
            raise (OutOfScope if rc == E_SCOPE else HipError)(rc, lib().hora_hip_last_error(self.h).decode())
        return bool(r.value), lv.LiveStats.from_c(st), wne[:n].cpu().numpy()

    def live_graph(self) -> dict:
        """The last liveness query's zone graph and fixpoint emptiness, copied to host arrays."""
        g = lv.hora_live_graph()
        lib().hora_hip_live_graph(self.h, C.byref(g))
        n, e, d = g.nodes, g.edges, g.d
        arrs = lv.graph_arrays(n, e, d)
        lv.fill_graph_struct(g, arrs)
        rc = lib().hora_hip_live_graph(self.h, C.byref(g))
        if rc != E_OK:
            raise HipError(rc, "hora_hip_live_graph")
        return lv.trim_graph(arrs, n, e, d)

    def profile(self) -> dict:
        """Per-kernel device time (ms) and algorithmic bytes of the last search."""
        p = hora_reach_profile()
        lib().hora_hip_reach_profile(self.h, C.byref(p))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def close(self):
        if self.h:
            lib().hora_hip_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reachable_net(net: Network, target: Target, backend: Backend = Backend.Hip,
                  stats: Optional[ReachStats] = None) -> bool:
    """``E<> target`` on the device. Only ``Backend.Hip`` is served here."""
    if backend is not Backend.Hip:
        raise ValueError("this package implements Backend.Hip only; seq/omp are the reference's CPU engines")
    with HipContext(net, target) as ctx:
        r, st = ctx.reach()
    if stats is not None:
        stats.__dict__.update(st.__dict__)
    return r


def leads_to_otf(net: Network, p: "lv.DPred", q: "lv.DPred", gs: Optional["lv.LiveStats"] = None,
                 early: bool = True) -> bool:
    """``p --> q`` on the device (``leads_to_otf``, include/hora/liveness.hpp:96-120)."""
    with HipContext(net, None) as ctx:
        r, st = ctx.live(lv.leads_to(p, q, early))
    if gs is not None:
        gs.__dict__.update(st.__dict__)
    return r


def exists_invariant_run_otf(net: Network, psi: "lv.DPred", clocks=(), gs: Optional["lv.LiveStats"] = None) -> bool:
    """``E[] psi`` on the device (``exists_invariant_run_otf``, include/hora/liveness.hpp:121-140)."""
    rnet, query = lv.exists_always(net, psi, clocks)
    with HipContext(rnet, None) as ctx:
        r, st = ctx.live(query)
    if gs is not None:
        gs.__dict__.update(st.__dict__)
    return r


def deadlock_free(net: Network, gs: Optional["lv.LiveStats"] = None, early: bool = True) -> bool:
    """``A[] not deadlock`` on the device (the negation of ``deadlock_reachable``,
    include/hora/liveness.hpp:144-170)."""
    with HipContext(net, None) as ctx:
        r, st = ctx.live(lv.not_deadlock(early))
    if gs is not None:
        gs.__dict__.update(st.__dict__)
    return r


def zone_op(op: int, d: int, a, b=None, arg=None, L=None, U=None, out=None, flag=None, stream=None) -> None:
    """Batched zone operation on device tensors (torch, int32, contiguous), asynchronous on `stream`."""
    def ptr(t):
        return None if t is None else t.data_ptr()
    n = a.shape[0]
    rc = lib().hora_hip_zone_op(op, d, n, ptr(a), ptr(b), ptr(arg), ptr(L), ptr(U), ptr(out), ptr(flag),
                                None if stream is None else stream)
    if rc != E_OK:
        raise HipError(rc, "hora_hip_zone_op")
