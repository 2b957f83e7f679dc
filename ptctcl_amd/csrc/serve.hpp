// serve.hpp — the serving loop of the HIP back end: verification jobs of different networks arrive
// while others run, are admitted as device slots and HBM free up, and run concurrently on the GPU.
// Part of the hora_hip.hip translation unit.
//
// Each slot is one native worker thread owning one context: its own HIP stream, its own
// device-resident passed list, frontier and successor buffers, kept across jobs (a slot rebinds its
// context to the next job's network; buffers only grow). Jobs run level by level on their slot's
// stream, so the levels of different jobs overlap on the device: a small job's launches fill the
// CUs a large job's level leaves idle, and a short job admitted behind a long one does not wait for
// it (continuous batching at BFS-level granularity). Admission is FIFO in arrival order, gated by a
// free slot and by the HBM footprint: the declared footprints of the running jobs plus the next
// one stay within the server's budget (a job is always admitted when nothing else runs).
#pragma once

#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

namespace {

// Owned copy of a job's network, target and liveness query (the caller's arrays may go away).
struct JobData {
  std::deque<std::vector<int32_t>> i32;
  std::vector<uint8_t> u8[2];
  std::vector<uint64_t> u64v;
  hora_net_enc net{};
  hora_target target{};
  hora_live_query query{};
  const int32_t* keep(const int32_t* p, long n) {
    i32.emplace_back();
    std::vector<int32_t>& v = i32.back();
    v.assign(p && n > 0 ? p : nullptr, p && n > 0 ? p + n : nullptr);
    if (v.empty()) v.push_back(0);
    return v.data();
  }
  void copy_net(const hora_net_enc* e) {
    net = *e;
    const int np = e->nprocs, ni = e->nintvars, nch = e->nchans, ne = e->nedges, tl = e->total_locs;
    const int d = e->nclocks + 1;
    net.int_lo = keep(e->int_lo, ni);
    net.int_hi = keep(e->int_hi, ni);
    net.int_init = keep(e->int_init, ni);
    u8[0].assign(e->chan_bcast, e->chan_bcast + (e->chan_bcast ? nch : 0));
    u8[1].assign(e->chan_urgent, e->chan_urgent + (e->chan_urgent ? nch : 0));
    u8[0].push_back(0);
    u8[1].push_back(0);
    net.chan_bcast = u8[0].data();
    net.chan_urgent = u8[1].data();
    net.locbase = keep(e->locbase, np + 1);
    net.loc_init = keep(e->loc_init, np);
    net.loc_flags = keep(e->loc_flags, tl);
    net.inv_off = keep(e->inv_off, tl + 1);
    net.inv_data = keep(e->inv_data, 3L * e->inv_off[tl]);
    net.edge_proc = keep(e->edge_proc, ne);
    net.edge_src = keep(e->edge_src, ne);
    net.edge_tgt = keep(e->edge_tgt, ne);
    net.edge_chan = keep(e->edge_chan, ne);
    net.edge_dir = keep(e->edge_dir, ne);
    net.cg_off = keep(e->cg_off, ne + 1);
    net.cg_data = keep(e->cg_data, 3L * e->cg_off[ne]);
    net.rs_off = keep(e->rs_off, ne + 1);
    net.rs_clk = keep(e->rs_clk, e->rs_off[ne]);
    net.ig_off = keep(e->ig_off, ne + 1);
    net.ig_data = keep(e->ig_data, 3L * e->ig_off[ne]);
    net.iu_off = keep(e->iu_off, ne + 1);
    net.iu_data = keep(e->iu_data, 2L * e->iu_off[ne]);
    net.lbound = keep(e->lbound, d);
    net.ubound = keep(e->ubound, d);
    net.lbound_loc = keep(e->lbound_loc, (long)tl * d);
    net.ubound_loc = keep(e->ubound_loc, (long)tl * d);
    u64v.assign(e->active_mask, e->active_mask + tl);
    u64v.push_back(0);
    net.active_mask = u64v.data();
  }
  void copy_target(const hora_target* t) {
    if (!t) { target = hora_target{}; target.never = 1; return; }
    target = *t;
    target.loc_req = keep(t->loc_req, 2L * t->nloc_req);
    target.int_cond = keep(t->int_cond, 3L * t->nint_cond);
    target.clock_cond = keep(t->clock_cond, 3L * t->nclock_cond);
  }
  void copy_pred(const hora_dpred& src, hora_dpred& dst) {
    dst.nterms = src.nterms;
    dst.term_off = keep(src.term_off, src.nterms > 0 ? src.nterms + 1 : 0);
    const long na = src.nterms > 0 ? src.term_off[src.nterms] : 0;
    dst.atoms = keep(src.atoms, 4 * na);
  }
  void copy_query(const hora_live_query* q) {
    query = *q;
    copy_pred(q->p, query.p);
    copy_pred(q->q, query.q);
    if (q->orig_inv_off) {
      query.orig_inv_off = keep(q->orig_inv_off, net.total_locs + 1);
      query.orig_inv_data = keep(q->orig_inv_data, 3L * q->orig_inv_off[net.total_locs]);
    }
  }
};

struct Job {
  int64_t id = 0;
  int kind = 0;               // 0 reachability, 1 liveness
  int64_t est_bytes = 0;
  JobData data;
  hora_serve_job_result res{};
};

}  // namespace

struct hora_hip_server {
  int slots = 1;
  int64_t budget = 0, running_bytes = 0;
  int running = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> pending;
  std::vector<hora_serve_job_result> done;
  std::vector<std::thread> workers;
  std::vector<hora_hip_ctx*> ctxs;
  int64_t next_id = 1;
  bool stop = false;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void run_job(hora_hip_ctx* c, Job* j) {
  hora_serve_job_result& r = j->res;
  try {
    StreamScope scope(c->stream);   // null before the slot's first job: then a device-wide wait
    open_ctx(&j->data.net, j->kind == 0 ? &j->data.target : nullptr, c);
    tl_stream = c->stream;
    int32_t v = 0;
    if (j->kind == 0) r.status = run_search(c, &v, &r.reach);
    else r.status = run_live(c, &j->data.query, &v, &r.live);
    r.verdict = v;
  } catch (const ScopeFail&) {
    r.status = HORA_HIP_E_SCOPE;
  } catch (const HipFail&) {
    r.status = HORA_HIP_E_DEVICE;
  } catch (const std::exception&) {
    r.status = HORA_HIP_E_INTERNAL;
  }
}

void worker(hora_hip_server* S, int slot) {
  hora_hip_ctx* c = S->ctxs[slot];
  for (;;) {
    Job* j = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->m);
      S->cv_work.wait(lk, [&] {
        if (S->stop) return true;
        if (S->pending.empty()) return false;
        const Job* f = S->pending.front();
        return S->running == 0 || S->running_bytes + f->est_bytes <= S->budget;
      });
      if (S->stop && S->pending.empty()) return;
      if (S->pending.empty()) continue;
      j = S->pending.front();
      S->pending.pop_front();
      ++S->running;
      S->running_bytes += j->est_bytes;
    }
    j->res.slot = slot;
    j->res.t_start = S->now();
    run_job(c, j);
    j->res.t_done = S->now();
    {
      std::lock_guard<std::mutex> lk(S->m);
      --S->running;
      S->running_bytes -= j->est_bytes;
      S->done.push_back(j->res);
    }
    delete j;
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

int64_t default_footprint(const hora_net_enc* e) {
  const int64_t d2 = (int64_t)(e->nclocks + 1) * (e->nclocks + 1);
  return ((int64_t)64 << 20) + ((4 * d2) << 21);   // tables and hash slots, plus two million zones
}

int submit(hora_hip_server* S, Job* j, int64_t* id) {
  {
    std::lock_guard<std::mutex> lk(S->m);
    j->id = S->next_id++;
    j->res.job = j->id;
    j->res.kind = j->kind;
    j->res.t_submit = S->now();
    if (id) *id = j->id;       // before the job becomes visible to the workers
    S->pending.push_back(j);
  }
  S->cv_work.notify_all();
  return HORA_HIP_OK;
}

}  // namespace
