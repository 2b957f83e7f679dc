"""The serving loop: each job's result is the same alone and in the mix; admission honours the HBM
footprint budget; admitted jobs overlap on the device."""
import pytest

from ptctcl_amd import hip, serve

pytestmark = pytest.mark.gpu
REACH = ("explored", "generated", "stored_live", "keys", "inserted", "subsumed", "killed", "unions", "levels",
         "found")
LIVE = ("outer_nodes", "fwd_edges", "disc_keys", "covered", "closure", "nonempty", "early")


def alone(spec):
    with hip.HipContext(spec.net, spec.target if spec.kind == 0 else None) as ctx:
        return spec.run_alone(ctx)


def collect(server, n):
    out = []
    while len(out) < n:
        out.extend(server.wait(256, 500))
    return out


def test_results_identical_alone_and_in_mix():
    jobs = serve.job_mix(512, seed=7)
    ref = [alone(s) for s in jobs]
    with serve.Server(slots=6) as srv:
        ids = [s.submit(srv) for s in jobs]
        res = {r.job: r for r in collect(srv, len(jobs))}
    slots = set()
    for spec, jid, (verdict, st) in zip(jobs, ids, ref):
        r = res[jid]
        print(spec.label, jid, r.slot, r.verdict, verdict, flush=True)
        assert r.status == hip.E_OK
        assert r.verdict == verdict, spec.label
        got = r.reach if spec.kind == 0 else r.live
        for k in (REACH if spec.kind == 0 else LIVE):
            assert getattr(got, k) == getattr(st, k), (spec.label, k)
        slots.add(r.slot)
    assert len(slots) > 1


def test_admission_respects_footprint_budget():
    jobs = serve.job_mix(12, seed=3)
    with serve.Server(slots=4, hbm_budget_bytes=1) as srv:     # below any job's footprint
        for s in jobs:
            s.submit(srv)
        res = sorted(collect(srv, len(jobs)), key=lambda r: r.t_start)
    for a, b in zip(res, res[1:]):
        assert b.t_start >= a.t_done       # one job at a time: nothing else fits beside a running job


def test_admitted_jobs_overlap():
    from ptctcl_amd import models
    with serve.Server(slots=4) as srv:
        for n in (9, 9, 10, 10):
            net = models.fischer(n, k=n)
            srv.submit_reach(net, models.mutex_target(net))
        res = collect(srv, 4)
    res.sort(key=lambda r: r.t_start)
    assert any(b.t_start < a.t_done for a, b in zip(res, res[1:]))
