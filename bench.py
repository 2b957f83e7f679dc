#!/usr/bin/env python3
"""Benchmark of the MI355X back end: symbolic states/sec of a full forward zone-graph search.

One step = one complete device-resident search (hora_hip_reach) of the workload, with the network
tables already resident in HBM. The default workload is the batched forward search on the
mutual-exclusion family at its largest configured size: Fischer, 12 processes, one clock each
(zone dimension 13), guard constant k = 12, query A[] not (P1.cs && P2.cs), i.e. a full search of
E<> P1.cs && P2.cs. Its explored count (4,758,404) is the reference's recorded one
(benchmarks/results/reach/hora/hora_t16.csv:87).

Prints one JSON line. `roofline` is the dominant kernel's algorithmic HBM bytes per launch over its
average launch duration (HIP events on the library's own stream). `cpu_baseline` is the oracle's
OpenMP engine -- a restatement of the reference's OMP search -- on this host's cores, on a bounded
run of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def workload(name: str):
    from ptctcl_amd import models
    if name.startswith("fischerv"):        # fischerv<N>: the plain-reachability variant, k = N
        n = int(name[len("fischerv"):])
        net = models.fischer_variant(n, k=n)
        return net, models.mutex_target(net), {
            "workload": f"fischer-variant mutual exclusion N={n}, k={n}, A[] not (P1.cs && P2.cs)",
            "processes": n, "zone_dim": n + 1, "k": n}
    if name.startswith("fischer"):          # fischer<N>: the reference's family, k = N
        n = int(name[len("fischer"):])
        net = models.fischer(n, k=n)
        return net, models.mutex_target(net), {
            "workload": f"fischer mutual exclusion N={n}, k={n}, A[] not (P1.cs && P2.cs)",
            "processes": n, "zone_dim": n + 1, "k": n}
    raise SystemExit(f"unknown workload {name}")


def physical_cores() -> int:
    try:
        cores = set()
        phys = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                cores.add((phys, line.split(":")[1].strip()))
        if cores:
            return len(cores)
    except OSError:
        pass
    return os.cpu_count() or 1


def cpu_baseline(name: str, seconds: float) -> dict:
    from oracle import binding
    net, target, _ = workload(name)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or physical_cores()
    done, st = binding.reach_bounded(net, target, threads, seconds)
    return {
        "value": st["explored"] / st["seconds"] if st["seconds"] > 0 else 0.0,
        "unit": "states/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{'full search' if done else 'first %d levels (time-bounded)' % st['levels']} of the same "
                   f"workload: {st['explored']} states in {st['seconds']:.2f} s, OpenMP engine "
                   f"(level-synchronous, stripe-locked passed set, dynamic chunk 8)"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fischer12")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this back end runs on one GPU; multi-GPU runs are not part of this build "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        return 2

    import torch
    from ptctcl_amd import hip

    net, target, cfg = workload(a.workload)
    ctx = hip.HipContext(net, target)
    for _ in range(a.warmup):
        ctx.reach()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = None
    prof = None
    for _ in range(a.steps):
        reachable, st = ctx.reach()
        p = ctx.profile()
        prof = p if prof is None else {k: prof[k] + p[k] for k in prof}
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / a.steps
    value = st.explored / (ms / 1000.0)

    # dominant kernel by device time; a launch is one BFS level's launch of that kernel
    fams = {"k_expand": "expand", "k_probe": "probe", "k_merge": "merge"}
    kern = max(fams, key=lambda k: prof[fams[k] + "_ms"])
    f = fams[kern]
    launches = max(1, int(prof[f + "_launches"]))
    kms, kbytes = prof[f + "_ms"], prof[f + "_bytes"]
    achieved = (kbytes / launches) / ((kms / launches) / 1000.0) / 1e9 if kms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": kern,
                "avg_launch_ms": round(kms / launches, 4), "bytes_per_launch": round(kbytes / launches)}
    line = {
        "metric": "symbolic states/sec (full forward zone-graph search)",
        "value": round(value, 1),
        "unit": "states/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generated model; no external data)",
        "config": dict(cfg, explored=st.explored, generated=st.generated, verdict_A_not=not reachable,
                       levels=st.levels, stored_live=st.stored_live, keys=st.keys),
        "roofline": roofline,
        "kernel_ms_per_step": {k: round(prof[k + "_ms"] / a.steps, 3) for k in ("expand", "probe", "merge", "other")},
    }
    ctx.close()
    if not a.no_cpu:
        cb = cpu_baseline(a.workload, a.cpu_seconds)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(value / cb["value"], 2) if cb["value"] > 0 else None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
