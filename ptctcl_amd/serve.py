"""Serving: verification jobs arriving over time, continuously batched onto one GPU.

:class:`Server` wraps the C ABI's server (``hora_hip_server_*``): ``slots`` device slots, each a HIP
stream with its own device-resident passed list and frontier; jobs are admitted in arrival order
when a slot is free and the HBM footprint budget allows, and admitted jobs overlap on the device.

:func:`job_mix` is the seeded serving workload: jobs mixing plain reachability on the
mutual-exclusion variant (k = 2 and k = N), that family's deadlock query, and the token-ring
liveness query, with sizes drawn per job, and :func:`poisson_arrivals` the arrival schedule.
"""
from __future__ import annotations

import ctypes as C
import random
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import hip, liveness as lv, models
from .network import Network, Target, TargetEnc, hora_reach_stats


class hora_serve_job_result(C.Structure):
    _fields_ = [("job", C.c_int64), ("kind", C.c_int32), ("status", C.c_int32), ("verdict", C.c_int32),
                ("slot", C.c_int32), ("reach", hora_reach_stats), ("live", lv.hora_live_stats),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        L.hora_hip_server_open.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]
        L.hora_hip_server_submit_reach.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                                   C.POINTER(C.c_int64)]
        L.hora_hip_server_submit_live.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(lv.hora_live_query), C.c_int64,
                                                  C.POINTER(C.c_int64)]
        L.hora_hip_server_wait.argtypes = [C.c_void_p, C.POINTER(hora_serve_job_result), C.c_int32,
                                           C.POINTER(C.c_int32), C.c_int32]
        L.hora_hip_server_now.argtypes = [C.c_void_p]
        L.hora_hip_server_now.restype = C.c_double
        L.hora_hip_server_close.argtypes = [C.c_void_p]
        L.hora_hip_server_close.restype = None
        _bound = True
    return L


@dataclass
class Result:
    job: int
    kind: int
    status: int
    verdict: bool
    slot: int
    reach: Optional[hip.ReachStats]
    live: Optional[lv.LiveStats]
    t_submit: float
    t_start: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class Server:
    def __init__(self, slots: int, hbm_budget_bytes: int = 0):
        h = C.c_void_p()
        rc = _lib().hora_hip_server_open(slots, hbm_budget_bytes, C.byref(h))
        if rc != hip.E_OK:
            raise hip.HipError(rc, "hora_hip_server_open")
        self.h = h

    def submit_reach(self, net: Network, target: Optional[Target], est_bytes: int = 0) -> int:
        enc, te = net.encode(), TargetEnc(target)
        jid = C.c_int64()
        rc = _lib().hora_hip_server_submit_reach(self.h, C.addressof(enc.struct), C.addressof(te.struct), est_bytes,
                                                 C.byref(jid))
        if rc != hip.E_OK:
            raise hip.HipError(rc, "hora_hip_server_submit_reach")
        return jid.value

    def submit_live(self, net: Network, query: lv.LiveQuery, est_bytes: int = 0) -> int:
        enc, qe = net.encode(), query.encode()
        jid = C.c_int64()
        rc = _lib().hora_hip_server_submit_live(self.h, C.addressof(enc.struct), C.byref(qe.struct), est_bytes,
                                                C.byref(jid))
        if rc != hip.E_OK:
            raise hip.HipError(rc, "hora_hip_server_submit_live")
        return jid.value

    def wait(self, max_results: int = 256, timeout_ms: int = 100) -> List[Result]:
        buf = (hora_serve_job_result * max_results)()
        n = C.c_int32(0)
        _lib().hora_hip_server_wait(self.h, buf, max_results, C.byref(n), timeout_ms)
        out = []
        for r in buf[: n.value]:
            out.append(Result(r.job, r.kind, r.status, bool(r.verdict), r.slot,
                              hip.ReachStats.from_c(r.reach) if r.kind == 0 else None,
                              lv.LiveStats.from_c(r.live) if r.kind == 1 else None,
                              r.t_submit, r.t_start, r.t_done))
        return out

    def now(self) -> float:
        return _lib().hora_hip_server_now(self.h)

    def close(self):
        if self.h:
            _lib().hora_hip_server_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---- the serving workload -----------------------------------------------------------------------

@dataclass
class JobSpec:
    label: str
    kind: int                  # 0 reachability, 1 liveness
    net: Network
    target: Optional[Target] = None
    query: Optional[lv.LiveQuery] = None

    def submit(self, server: Server) -> int:
        if self.kind == 0:
            return server.submit_reach(self.net, self.target)
        return server.submit_live(self.net, self.query)

    def run_alone(self, ctx: "hip.HipContext"):
        if self.kind == 0:
            return ctx.reach()
        return ctx.live(self.query)


FULL_SEARCH_MAX_N = 8


def job_mix(n: int, seed: int, full_max: int = FULL_SEARCH_MAX_N) -> List[JobSpec]:
    """n seeded jobs, four types in equal proportion, each with its own N:

    * plain reachability, E<> P1.cs && P2.cs, on the mutual-exclusion variant with k = 2 (the
      4-process CPU case's family), N ~ U[3, full_max];
    * the same query on the same family with k = N (the batched forward search family),
      N ~ U[3, full_max];
    * A[] not deadlock on that family with k = N, early witnesses, N ~ U[3, 10];
    * liveness P0.hold --> P1.hold on the seeded token ring, N ~ U[3, 10].

    The two full searches stop at N = 8 by default: alone, one takes 0.2 s at N = 8, about 4 s at
    N = 9 and about 16 s at N = 10, so at U[3, 10] they would hold a 512-job mix on the GPU for
    many minutes, almost all of it in the few largest jobs.
    """
    rng = random.Random(seed)
    jobs: List[JobSpec] = []
    for k in range(n):
        t = rng.randrange(4)
        if t == 0:
            N = rng.randint(3, full_max)
            net = models.fischer_variant(N, k=2)
            jobs.append(JobSpec(f"variant{N}-k2", 0, net, target=models.mutex_target(net)))
        elif t == 1:
            N = rng.randint(3, full_max)
            net = models.fischer_variant(N, k=N)
            jobs.append(JobSpec(f"variant{N}", 0, net, target=models.mutex_target(net)))
        elif t == 2:
            N = rng.randint(3, 10)
            net = models.fischer_variant(N, k=N)
            jobs.append(JobSpec(f"deadlock-variant{N}", 1, net, query=lv.not_deadlock()))
        else:
            N = rng.randint(3, 10)
            net = models.token_ring(N, rng.randint(1, 1 << 30))
            q = lv.leads_to(lv.DPred.at(net, "P0", "hold"), lv.DPred.at(net, "P1", "hold"))
            jobs.append(JobSpec(f"ring{N}", 1, net, query=q))
    return jobs


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival offsets (seconds) of a Poisson process of the given rate (jobs per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def run_schedule(server: Server, jobs: Sequence[JobSpec], arrivals: Sequence[float]) -> List[Result]:
    """Submit each job at its arrival offset (from now), collect every result."""
    ids = {}
    results: List[Result] = []
    t0 = time.perf_counter()
    for spec, at in zip(jobs, arrivals):
        delay = at - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids[spec.submit(server)] = spec
        results.extend(server.wait(256, 0))
    while len(results) < len(jobs):
        results.extend(server.wait(256, 200))
    return results


def latency_summary(results: Sequence[Result]) -> dict:
    lat = np.array(sorted(r.latency for r in results))
    span = max(r.t_done for r in results) - min(r.t_submit for r in results)
    return {"p50_ms": round(float(np.percentile(lat, 50)) * 1000, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1000, 3),
            "mean_ms": round(float(lat.mean()) * 1000, 3),
            "verdicts_per_s": round(len(results) / span, 2) if span > 0 else None,
            "span_s": round(span, 3)}
