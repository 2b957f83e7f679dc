"""ctypes binding of oracle/liboracle.so (built by oracle/Makefile). Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from ptctcl_amd import liveness as lv
from ptctcl_amd.network import NetEnc, TargetEnc, hora_reach_stats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.oracle_reach.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(hora_reach_stats)]
        _lib.oracle_last_error.restype = C.c_char_p
        _lib.oracle_zone_op.argtypes = [C.c_int32, C.c_int32, C.c_int64] + [C.c_void_p] * 7
        _lib.oracle_expand.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_int64)]
        _lib.oracle_analyses.argtypes = [C.c_void_p] * 6
        _lib.oracle_reach_bounded.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                              C.POINTER(C.c_int32), C.POINTER(hora_reach_stats),
                                              C.POINTER(C.c_int32)]
        _lib.oracle_live.argtypes = [C.c_void_p, C.POINTER(lv.hora_live_query), C.POINTER(C.c_int32),
                                     C.POINTER(lv.hora_live_stats), C.POINTER(C.c_void_p)]
        _lib.oracle_live_graph.argtypes = [C.c_void_p, C.POINTER(lv.hora_live_graph)]
        _lib.oracle_live_free.argtypes = [C.c_void_p]
        _lib.oracle_live_free.restype = None
        _lib.oracle_live_last_error.restype = C.c_char_p
        _lib.oracle_zone_subtract.argtypes = [C.c_int32, C.c_int64] + [C.c_void_p] * 2 + [C.c_int32] + [C.c_void_p] * 2
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def reach(net, target, nthreads: int = 0):
    """Run the oracle search; nthreads == 0 is `seq`, > 0 is `omp` with that many threads."""
    enc = net.encode() if not isinstance(net, NetEnc) else net
    te = TargetEnc(target)
    r = C.c_int32(0)
    st = hora_reach_stats()
    rc = lib().oracle_reach(C.addressof(enc.struct), C.addressof(te.struct), nthreads, C.byref(r), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_reach: {rc} {lib().oracle_last_error().decode()}")
    return bool(r.value), st.as_dict()


def analyses(net):
    enc = net.encode()
    d = net.dim
    tl = int(enc.struct.total_locs)
    L = np.zeros(d, np.int32); U = np.zeros(d, np.int32)
    LL = np.zeros(tl * d, np.int32); UL = np.zeros(tl * d, np.int32); AM = np.zeros(tl, np.uint64)
    rc = lib().oracle_analyses(C.addressof(enc.struct), _p(L), _p(U), _p(LL), _p(UL), _p(AM))
    assert rc == 0
    return L, U, LL, UL, AM


def zone_op(op: int, d: int, a, b=None, arg=None, L=None, U=None):
    n = a.shape[0]
    out = np.zeros_like(a)
    flag = np.zeros(n, np.int32)
    a = np.ascontiguousarray(a, np.int32)
    b = None if b is None else np.ascontiguousarray(b, np.int32)
    arg = None if arg is None else np.ascontiguousarray(arg, np.int32)
    L = None if L is None else np.ascontiguousarray(L, np.int32)
    U = None if U is None else np.ascontiguousarray(U, np.int32)
    rc = lib().oracle_zone_op(op, d, n, _p(a), _p(b), _p(arg), _p(L), _p(U), _p(out), _p(flag))
    assert rc == 0, rc
    return out, flag


def expand(net, target, dks, zones, lu_plus=False, cap=None):
    enc = net.encode() if not isinstance(net, NetEnc) else net
    te = TargetEnc(target)
    d2 = net.dim * net.dim if not isinstance(net, NetEnc) else (net.struct.nclocks + 1) ** 2
    n = len(dks)
    cap = cap or max(1, n * 64)
    dks = np.ascontiguousarray(dks, np.uint64)
    zones = np.ascontiguousarray(zones, np.int32).reshape(-1)
    odk = np.zeros(cap, np.uint64); oz = np.zeros(cap * d2, np.int32)
    osrc = np.zeros(cap, np.int32); otg = np.zeros(cap, np.int32)
    on = C.c_int64(0)
    rc = lib().oracle_expand(C.addressof(enc.struct), C.addressof(te.struct), 1 if lu_plus else 0, n, _p(dks),
                             _p(zones), cap, _p(odk), _p(oz), _p(osrc), _p(otg), C.byref(on))
    if rc != 0:
        raise RuntimeError(f"oracle_expand: {rc} {lib().oracle_last_error().decode()}")
    k = on.value
    return odk[:k], oz[:k * d2].reshape(k, d2), osrc[:k], otg[:k]


def reach_bounded(net, target, nthreads: int, max_seconds: float):
    """Run the oracle for at most ~max_seconds (checked per BFS level). Returns (complete, stats)."""
    enc = net.encode()
    te = TargetEnc(target)
    r = C.c_int32(0)
    done = C.c_int32(0)
    st = hora_reach_stats()
    rc = lib().oracle_reach_bounded(C.addressof(enc.struct), C.addressof(te.struct), nthreads, max_seconds,
                                    C.byref(r), C.byref(st), C.byref(done))
    if rc != 0:
        raise RuntimeError(f"oracle_reach_bounded: {rc}")
    return bool(done.value), st.as_dict()


def live(net, query, graph: bool = False):
    """One liveness query on the CPU restatement: (holds, LiveStats[, graph arrays])."""
    enc = net.encode()
    qe = query.encode()
    r = C.c_int32(0)
    st = lv.hora_live_stats()
    h = C.c_void_p()
    rc = lib().oracle_live(C.addressof(enc.struct), C.byref(qe.struct), C.byref(r), C.byref(st), C.byref(h))
    try:
        if rc != 0:
            raise RuntimeError(f"oracle_live: {rc} {lib().oracle_live_last_error().decode()}")
        stats = lv.LiveStats.from_c(st)
        if not graph:
            return bool(r.value), stats
        g = lv.hora_live_graph()
        lib().oracle_live_graph(h, C.byref(g))
        n, e, d = g.nodes, g.edges, g.d
        arrs = lv.graph_arrays(n, e, d)
        lv.fill_graph_struct(g, arrs)
        lib().oracle_live_graph(h, C.byref(g))
        return bool(r.value), stats, lv.trim_graph(arrs, n, e, d)
    finally:
        if h:
            lib().oracle_live_free(h)


def zone_subtract(d: int, p, y, cap: int = 256):
    """p \\ y per pair: list of piece arrays (None where more than cap pieces)."""
    p = np.ascontiguousarray(p, np.int32)
    y = np.ascontiguousarray(y, np.int32)
    n = p.shape[0]
    out = np.zeros((n, cap, d * d), np.int32)
    cnt = np.zeros(n, np.int32)
    lib().oracle_zone_subtract(d, n, _p(p), _p(y), cap, _p(out), _p(cnt))
    return [None if cnt[k] < 0 else out[k, : cnt[k]] for k in range(n)]
