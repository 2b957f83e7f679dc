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

# The CPU baseline's OpenMP threads are pinned one per physical core, as the reference's own OMP
# runs are described (OMP_PROC_BIND=close, OMP_PLACES=cores); the OpenMP runtime reads these when it
# starts, so they are set before anything that loads it. A caller's own settings win.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# k_expand is bound by instruction issue: one wave64 VALU instruction per CU per clock (four 16-lane
# SIMDs, four clocks each), 256 CUs at 2.4 GHz. Its VALU instructions per generated successor on a
# workload are the rocprofv3 SQ_INSTS_VALU count over one search divided by the successors, from
# profiles/k_expand_fischer12_sq_counters.md.
VALU_ISSUE_PEAK = 614.4   # G instructions/s
K_EXPAND_VALU_PER_SUCCESSOR = {"fischer12": 683}


def live_workload(name: str):
    """live-fischer<N>: P0.req --> P0.wait on the reference's fischer family (k = 2), the liveness
    query of benchmarks/queries/live/fischer.q, decided by the full graph and fixpoint."""
    from ptctcl_amd import liveness as lv, models
    n = int(name[len("live-fischer"):])
    net = models.fischer(n)
    q = lv.leads_to(lv.DPred.at(net, "P0", "req"), lv.DPred.at(net, "P0", "wait"))
    return net, q, {"workload": f"fischer N={n}, k=2, P0.req --> P0.wait (forward graph + greatest fixpoint)",
                    "processes": n, "zone_dim": n + 1, "k": 2}


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


def measured_traffic(workload_name: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload_name`, from the FETCH_SIZE / WRITE_SIZE counter
    runs summarised in profiles/<kernel>_<workload>_traffic.md (scaled by the copy calibration
    recorded there), or None when no such measurement exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_name}_traffic.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch {kernel} = "):
                return int(line.split("=")[1])
    except OSError:
        pass
    return None


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
    """The oracle's OpenMP engine on this host: one warm-up run, then the best of three, each bounded
    to a quarter of `seconds` (a full search of the default workload takes about 5 s)."""
    from oracle import binding
    net, target, _ = workload(name)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or physical_cores()
    per_run = seconds / 4.0
    binding.reach_bounded(net, target, threads, per_run)
    best = None
    for _ in range(3):
        done, st = binding.reach_bounded(net, target, threads, per_run)
        rate = st["explored"] / st["seconds"] if st["seconds"] > 0 else 0.0
        if best is None or rate > best[0]:
            best = (rate, done, st)
    rate, done, st = best
    return {
        "value": rate,
        "unit": "states/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{'full search' if done else 'first %d levels (time-bounded)' % st['levels']} of the same "
                   f"workload, best of 3 after one warm-up: {st['explored']} states in {st['seconds']:.2f} s, "
                   f"OpenMP engine (level-synchronous, stripe-locked passed set, dynamic chunk 8), "
                   f"{threads} threads, OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                   f"OMP_PLACES={os.environ.get('OMP_PLACES')}"),
    }


def live_cpu_baseline(name: str, seconds: float) -> dict:
    """The oracle's single-threaded liveness restatement on the largest instance of the family it
    finishes within `seconds` (sizes tried upward from 5)."""
    from oracle import binding
    n_top = int(name[len("live-fischer"):])
    best = None
    for n in range(5, n_top + 1):
        net, q, _ = live_workload(f"live-fischer{n}")
        t0 = time.perf_counter()
        _, st = binding.live(net, q)
        dt = time.perf_counter() - t0
        best = (n, st.expanded_nodes, dt)
        if dt * 12 > seconds:      # the next size costs roughly ten times more
            break
    n, nodes, dt = best
    return {"value": round(nodes / dt, 1), "unit": "states/s", "cores": 1, "kind": "port",
            "sample": (f"oracle liveness restatement, fischer N={n} (same query and construction): "
                       f"{nodes} states expanded in {dt:.2f} s, one thread")}


def live_main(a) -> int:
    import torch
    from ptctcl_amd import hip
    net, q, cfg = live_workload(a.workload)
    ctx = hip.HipContext(net, None)
    for _ in range(a.warmup):
        ctx.live(q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        holds, st = ctx.live(q)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / a.steps
    ctx.close()
    d2 = (net.nclocks + 1) ** 2
    # construction phase, algorithmic bytes (int32 entries): frontier zones read, successors written
    # and read back, stored zones read by the cover tests, new nodes written
    b_build = 4.0 * d2 * (st.expanded_nodes + 2 * st.generated + st.chain_reads + st.outer_nodes)
    achieved = b_build / st.build_seconds / 1e9 if st.build_seconds > 0 else 0.0
    line = {
        "metric": "symbolic states/sec (liveness: states expanded by the construction per second, construction + fixpoint)",
        "value": round(st.expanded_nodes / (ms / 1000.0), 1),
        "unit": "states/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generated model; no external data)",
        "config": dict(cfg, holds=holds, states_expanded=st.expanded_nodes, records=st.outer_nodes,
                       retired=st.retired, unions=st.unions, fwd_edges=st.fwd_edges, closure=st.closure,
                       nonempty=st.nonempty, fixpoint_rounds=st.iters, pre_calls=st.pre_calls,
                       build_ms=round(st.build_seconds * 1000, 2), fixpoint_ms=round(st.gfp_seconds * 1000, 2)),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "construction phase (k_expand, k_gprobe, k_omerge, k_opost)"},
    }
    if not a.no_cpu:
        cb = live_cpu_baseline(a.workload, a.cpu_seconds)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(line["value"] / cb["value"], 2) if cb["value"] > 0 else None
    print(json.dumps(line), flush=True)
    return 0


def serve_main(a) -> int:
    """Serving: 512 seeded mixed jobs (ptctcl_amd.serve.job_mix) arriving as a Poisson process.
    The arrival rate is set from the measured one-job-at-a-time capacity (all jobs queued at once on
    one slot): 0.8 of it. The same schedule then runs on one slot (plain one-job-at-a-time GPU) and
    on `--slots` slots (continuous batching); latency is arrival -> verdict."""
    from ptctcl_amd import serve
    import numpy as np
    n = 512
    jobs = serve.job_mix(n, seed=2026)
    # warm-up: every slot allocates its buffers once
    with serve.Server(slots=a.slots) as srv:
        for spec in jobs[:64]:
            spec.submit(srv)
        got = 0
        while got < 64:
            got += len(srv.wait(256, 500))
    with serve.Server(slots=1) as srv:
        t0 = time.perf_counter()
        for spec in jobs:
            spec.submit(srv)
        res = []
        while len(res) < n:
            res.extend(srv.wait(256, 500))
        closed = time.perf_counter() - t0
    capacity = n / closed
    rate = 0.8 * capacity
    arrivals = serve.poisson_arrivals(n, rate, seed=2027)
    summary = {}
    verdicts = {}
    for slots in (1, a.slots):
        with serve.Server(slots=slots) as srv:
            res = serve.run_schedule(srv, jobs, arrivals)
        summary[slots] = serve.latency_summary(res)
        verdicts[slots] = sorted((r.job - min(x.job for x in res), r.verdict, r.status) for r in res)
    same = [v[1:] for v in verdicts[1]] == [v[1:] for v in verdicts[a.slots]]
    served, single = summary[a.slots], summary[1]
    line = {
        "metric": "served verdicts/sec (512 mixed jobs, Poisson arrivals at 0.8x one-at-a-time capacity)",
        "value": served["verdicts_per_s"],
        "unit": "verdicts/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(served["span_s"] * 1000, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded job mix; no external data)",
        "config": {"workload": "serving: 512 jobs, types mutual-exclusion variant E<> P1.cs && P2.cs with k=2 "
                               "and with k=N at N~U[3,8], variant k=N A[] not deadlock N~U[3,10], token-ring "
                               "P0.hold --> P1.hold N~U[3,10]",
                   "slots": a.slots, "arrival_rate_per_s": round(rate, 2),
                   "one_at_a_time_capacity_per_s": round(capacity, 2),
                   "served": served, "one_at_a_time": single,
                   "p99_ratio_served_over_single": round(served["p99_ms"] / single["p99_ms"], 4),
                   "verdicts_identical": same},
        "roofline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def zoneops_main(a) -> int:
    """Zone-kernel microbenchmark: each HDBM operation on 2^20 zones per dimension d in
    {4, 8, 16, 32}, device-resident int32 zones. The zones are 2^16 seeded canonical zones
    (ptctcl_amd.zonegen: point-cloud hulls, loosened; 2^16 distinct for d <= 8, 2^14 for d = 16,
    2^13 for d = 32, the host-side generation being cubic in d) tiled to 2^20; L, U ~ U[0, 1000]; inclusion
    pairs half built as B inside A. Algorithmic bytes per zone-op: 8*d*d for the unary operations
    (zone read + written), 8*d*d + 4 for the pair tests (two zones read, one flag written), 4*d + 4
    for the emptiness test (diagonal read)."""
    import numpy as np
    import torch
    from ptctcl_amd import hip, zonegen as zg
    n = 1 << 20
    rows, tot_ops, tot_s = [], 0, 0.0
    for d in (4, 8, 16, 32):
        pool = {4: 1 << 16, 8: 1 << 16, 16: 1 << 14, 32: 1 << 13}[d]
        rng = np.random.default_rng(1000 + d)
        zp = zg.hull_zones(pool, d, rng)
        up = zg.unclose(zp, d, rng)
        _, bp = zg.pairs_included(zp, d, rng)
        Lp, Up = zg.bounds(pool, d, rng)
        cons = zg.constraints(zp, d, rng)
        rep = n // pool
        dev = lambda x: torch.from_numpy(np.ascontiguousarray(np.tile(x, (rep,) + (1,) * (x.ndim - 1)))).to("cuda:0")
        Z, UZ, B, L, U = dev(zp), dev(up), dev(bp), dev(Lp), dev(Up)
        C = dev(cons).reshape(-1)
        Rk = torch.randint(1, d, (n,), dtype=torch.int32, device="cuda:0")
        Fm = (torch.randint(1, 1 << min(d - 1, 30), (n,), dtype=torch.int64, device="cuda:0") << 1).to(torch.int32)
        out = torch.empty_like(Z)
        flag = torch.empty(n, dtype=torch.int32, device="cuda:0")
        z2 = 4 * d * d
        ops = [("close", hip.ZOP_CLOSE, UZ, None, None, 2 * z2), ("up", hip.ZOP_UP, Z, None, None, 2 * z2),
               ("reset", hip.ZOP_RESET, Z, None, Rk, 2 * z2), ("constrain", hip.ZOP_CONSTRAIN, Z, None, C, 2 * z2),
               ("extra_lu", hip.ZOP_EXTRA_LU, Z, None, None, 2 * z2), ("includes", hip.ZOP_INCLUDES, Z, B, None, 2 * z2 + 4),
               ("alu_le", hip.ZOP_ALU_LE, B, Z, None, 2 * z2 + 4), ("union", hip.ZOP_UNION, Z, B, None, 2 * z2 + 4),
               ("empty", hip.ZOP_EMPTY, Z, None, None, 4 * d + 4), ("down", hip.ZOP_DOWN, Z, None, None, 2 * z2),
               ("free", hip.ZOP_FREE, Z, None, Fm, 2 * z2)]
        for name, op, A, Bx, arg, bytes_per in ops:
            def go():
                hip.zone_op(op, d, A, Bx, arg, L, U, out, flag)
            go()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 5
            e0.record()
            for _ in range(reps):
                go()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            gbs = bytes_per * n / (ms / 1000.0) / 1e9
            rows.append({"d": d, "op": name, "ms": round(ms, 4), "zone_ops_per_s": round(n / (ms / 1000.0), 1),
                         "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
            tot_ops += n
            tot_s += ms / 1000.0
        del Z, UZ, B, L, U, C, Rk, Fm, out, flag
        torch.cuda.empty_cache()
    dom = max(rows, key=lambda r: r["ms"])
    stream = [r for r in rows if r["op"] in ("up", "reset", "down", "free", "includes")]
    best = max(stream, key=lambda r: r["GB_s"])
    line = {
        "metric": "zone operations/sec (HDBM ops on 2^20 device-resident zones, d in {4,8,16,32})",
        "value": round(tot_ops / tot_s, 1), "unit": "zone-ops/s", "n_gpus": 1, "steps": 5, "warmup": 1,
        "ms_per_step": round(tot_s * 1000, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic (seeded canonical zones, 2^13-2^16 distinct tiled to 2^20)",
        "config": {"workload": "zone-kernel microbenchmark", "zones_per_op": n, "ops": rows,
                   "best_streaming_op": best},
        "roofline": {"bound": "hbm", "achieved": dom["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dom["GB_s"] / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": f"{dom['op']} d={dom['d']} (the slowest op)"},
    }
    print(json.dumps(line), flush=True)
    return 0


def calibrate_main(a) -> int:
    """Reference copies for the HBM traffic counters: `--steps` device-to-device copies of 1 GiB.
    Run under `rocprofv3 --pmc FETCH_SIZE` (then `WRITE_SIZE`); the counter per copy dispatch over
    the bytes copied is the factor that turns the counter into bytes for kernels measured the same
    way. Prints no benchmark line."""
    import torch
    n = (1 << 30) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(a.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.steps} x 1024 MiB ({n * 4} bytes each)", flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fischer12")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--slots", type=int, default=8, help="serving workload: device slots")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this back end runs on one GPU; multi-GPU runs are not part of this build "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        return 2

    if a.workload.startswith("live-"):
        return live_main(a)
    if a.workload == "serve":
        return serve_main(a)
    if a.workload == "zoneops":
        return zoneops_main(a)
    if a.workload == "calibrate":
        return calibrate_main(a)

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
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(a.workload, kern), "kernel": kern,
                "avg_launch_ms": round(kms / launches, 4), "bytes_per_launch": round(kbytes / launches)}
    issue = None
    if kern == "k_expand" and a.workload in K_EXPAND_VALU_PER_SUCCESSOR:
        per = K_EXPAND_VALU_PER_SUCCESSOR[a.workload]
        rate = per * (st.generated / (launches / a.steps)) / ((kms / launches) / 1000.0) / 1e9
        issue = {"bound": "valu_issue", "achieved": round(rate, 1), "peak": VALU_ISSUE_PEAK, "unit": "G instr/s",
                 "frac": round(rate / VALU_ISSUE_PEAK, 4), "kernel": kern, "instr_per_successor": per,
                 "source": "profiles/k_expand_fischer12_sq_counters.md"}
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
                       levels=st.levels, stored_live=st.stored_live, keys=st.keys,
                       stored_zone_bytes=st.zone_entry_bytes * (cfg["zone_dim"] ** 2 if st.zone_entry_bytes == 4
                                                                else (cfg["zone_dim"] ** 2 + 1) // 2 * 2)),
        "roofline": roofline,
        "roofline_issue": issue,
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
