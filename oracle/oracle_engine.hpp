// oracle_engine.hpp — shared part of the oracle: network tables, static analyses and the
// successor semantics, used by the reachability search (oracle_engine.cpp) and the liveness
// restatement (oracle_live.cpp).
//
// TEST INFRASTRUCTURE ONLY: this is the oracle. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may build, load or call it; the HIP back end never does.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/hora_hip.h"
#include "oracle_dbm.hpp"

namespace orc {

struct Cons { int i, j; raw_t b; };
struct IntAtom { int var, op, val; };
struct IntAsgn { int var, val; };
struct Edge {
  int proc, src, tgt, chan, dir;
  std::vector<Cons> guard;
  std::vector<int> resets;
  std::vector<IntAtom> ig;
  std::vector<IntAsgn> iu;
};

struct Net {
  int nclocks = 0, d = 0, np = 0, ni = 0, nch = 0, ne = 0, tl = 0;
  std::vector<int> int_lo, int_hi, int_init;
  std::vector<std::uint8_t> bcast, uch;
  bool has_urgent_chan = false;
  std::vector<int> locbase, loc_init, loc_flags;
  std::vector<std::vector<Cons>> inv;   // per global location
  std::vector<Edge> edges;
  std::vector<std::vector<int>> edges_of;  // per global location, global edge ids ascending
  // analyses, recomputed here from the raw tables
  std::vector<std::int32_t> lbound, ubound, lloc, uloc;
  std::vector<std::uint64_t> amask;
  bool has_diag = false;
  int gl(int p, int l) const { return locbase[p] + l; }
};

inline void load_net(const hora_net_enc* e, Net& n) {
  n.nclocks = e->nclocks; n.d = e->nclocks + 1; n.np = e->nprocs; n.ni = e->nintvars;
  n.nch = e->nchans; n.ne = e->nedges; n.tl = e->total_locs;
  n.int_lo.assign(e->int_lo, e->int_lo + n.ni);
  n.int_hi.assign(e->int_hi, e->int_hi + n.ni);
  n.int_init.assign(e->int_init, e->int_init + n.ni);
  n.bcast.assign(n.nch, 0); n.uch.assign(n.nch, 0);
  for (int c = 0; c < n.nch; ++c) {
    n.bcast[c] = e->chan_bcast ? e->chan_bcast[c] : 0;
    n.uch[c] = e->chan_urgent ? e->chan_urgent[c] : 0;
    if (n.uch[c]) n.has_urgent_chan = true;
  }
  n.locbase.assign(e->locbase, e->locbase + n.np + 1);
  n.loc_init.assign(e->loc_init, e->loc_init + n.np);
  n.loc_flags.assign(e->loc_flags, e->loc_flags + n.tl);
  n.inv.assign(n.tl, {});
  for (int L = 0; L < n.tl; ++L)
    for (int c = e->inv_off[L]; c < e->inv_off[L + 1]; ++c)
      n.inv[L].push_back({e->inv_data[3 * c], e->inv_data[3 * c + 1], e->inv_data[3 * c + 2]});
  n.edges.resize(n.ne);
  n.edges_of.assign(n.tl, {});
  for (int g = 0; g < n.ne; ++g) {
    Edge& ed = n.edges[g];
    ed.proc = e->edge_proc[g]; ed.src = e->edge_src[g]; ed.tgt = e->edge_tgt[g];
    ed.chan = e->edge_chan[g]; ed.dir = e->edge_dir[g];
    for (int c = e->cg_off[g]; c < e->cg_off[g + 1]; ++c)
      ed.guard.push_back({e->cg_data[3 * c], e->cg_data[3 * c + 1], e->cg_data[3 * c + 2]});
    for (int r = e->rs_off[g]; r < e->rs_off[g + 1]; ++r) ed.resets.push_back(e->rs_clk[r]);
    for (int c = e->ig_off[g]; c < e->ig_off[g + 1]; ++c)
      ed.ig.push_back({e->ig_data[3 * c], e->ig_data[3 * c + 1], e->ig_data[3 * c + 2]});
    for (int c = e->iu_off[g]; c < e->iu_off[g + 1]; ++c)
      ed.iu.push_back({e->iu_data[2 * c], e->iu_data[2 * c + 1]});
    n.edges_of[n.gl(ed.proc, ed.src)].push_back(g);
  }
  auto diag = [](const std::vector<Cons>& cs) {
    for (const auto& c : cs) if (c.i > 0 && c.j > 0) return true;
    return false;
  };
  for (const auto& v : n.inv) if (diag(v)) n.has_diag = true;
  for (const auto& ed : n.edges) if (diag(ed.guard)) n.has_diag = true;
}

// ---- static analyses (src/model/network.cpp:59-245) ------------------------------------------
inline void compute_analyses(Net& n) {
  const int d = n.d;
  n.lbound.assign(d, 0); n.ubound.assign(d, 0);
  auto scan_glob = [&](const std::vector<Cons>& cs) {
    for (const auto& c : cs) {
      const int v = bnd(c.b);
      if (c.i > 0 && c.j == 0) n.ubound[c.i] = std::max(n.ubound[c.i], v);
      else if (c.i == 0 && c.j > 0) n.lbound[c.j] = std::max(n.lbound[c.j], -v);
      else if (c.i > 0 && c.j > 0) {
        const int a = std::abs(v);
        n.ubound[c.i] = std::max(n.ubound[c.i], a); n.ubound[c.j] = std::max(n.ubound[c.j], a);
        n.lbound[c.i] = std::max(n.lbound[c.i], a); n.lbound[c.j] = std::max(n.lbound[c.j], a);
      }
    }
  };
  for (const auto& v : n.inv) scan_glob(v);
  for (const auto& ed : n.edges) scan_glob(ed.guard);
  n.lbound[0] = n.ubound[0] = 0;

  // per-location bounds: backward least fixpoint within each process
  n.lloc.assign(static_cast<std::size_t>(n.tl) * d, kAluNegInf);
  n.uloc.assign(static_cast<std::size_t>(n.tl) * d, kAluNegInf);
  auto scan_loc = [&](const std::vector<Cons>& cs, std::vector<int>& Lc, std::vector<int>& Uc) {
    for (const auto& c : cs) {
      const int v = bnd(c.b);
      if (c.i > 0 && c.j == 0) Uc[c.i] = std::max(Uc[c.i], v);
      else if (c.i == 0 && c.j > 0) Lc[c.j] = std::max(Lc[c.j], -v);
      else if (c.i > 0 && c.j > 0) {
        const int a = std::abs(v);
        Uc[c.i] = std::max(Uc[c.i], a); Uc[c.j] = std::max(Uc[c.j], a);
        Lc[c.i] = std::max(Lc[c.i], a); Lc[c.j] = std::max(Lc[c.j], a);
      }
    }
  };
  for (int p = 0; p < n.np; ++p) {
    const int nl = n.locbase[p + 1] - n.locbase[p];
    for (bool changed = true; changed;) {
      changed = false;
      for (int l = 0; l < nl; ++l) {
        std::vector<int> Lc(d, kAluNegInf), Uc(d, kAluNegInf);
        const int L = n.gl(p, l);
        scan_loc(n.inv[L], Lc, Uc);
        for (int g : n.edges_of[L]) {
          const Edge& ed = n.edges[g];
          scan_loc(ed.guard, Lc, Uc);
          std::vector<char> rst(d, 0);
          for (int r : ed.resets) rst[r] = 1;
          const int T = n.gl(p, ed.tgt);
          for (int c = 1; c < d; ++c)
            if (!rst[c]) {
              Lc[c] = std::max(Lc[c], n.lloc[static_cast<std::size_t>(T) * d + c]);
              Uc[c] = std::max(Uc[c], n.uloc[static_cast<std::size_t>(T) * d + c]);
            }
        }
        for (int c = 1; c < d; ++c) {
          auto& lo = n.lloc[static_cast<std::size_t>(L) * d + c];
          auto& up = n.uloc[static_cast<std::size_t>(L) * d + c];
          if (Lc[c] > lo) { lo = Lc[c]; changed = true; }
          if (Uc[c] > up) { up = Uc[c]; changed = true; }
        }
      }
    }
  }

  // active clocks: backward least fixpoint within each process
  std::vector<std::vector<char>> act(n.tl, std::vector<char>(d, 0));
  for (int p = 0; p < n.np; ++p) {
    const int nl = n.locbase[p + 1] - n.locbase[p];
    for (bool changed = true; changed;) {
      changed = false;
      for (int l = 0; l < nl; ++l) {
        const int L = n.gl(p, l);
        std::vector<char> A(d, 0);
        for (const auto& c : n.inv[L]) { if (c.i > 0) A[c.i] = 1; if (c.j > 0) A[c.j] = 1; }
        for (int g : n.edges_of[L]) {
          const Edge& ed = n.edges[g];
          for (const auto& c : ed.guard) { if (c.i > 0) A[c.i] = 1; if (c.j > 0) A[c.j] = 1; }
          std::vector<char> rst(d, 0);
          for (int r : ed.resets) rst[r] = 1;
          const int T = n.gl(p, ed.tgt);
          for (int c = 1; c < d; ++c) if (act[T][c] && !rst[c]) A[c] = 1;
        }
        for (int c = 1; c < d; ++c)
          if (A[c] && !act[L][c]) { act[L][c] = 1; changed = true; }
      }
    }
  }
  n.amask.assign(n.tl, 0);
  for (int L = 0; L < n.tl; ++L)
    for (int c = 1; c < d; ++c)
      if (act[L][c]) n.amask[L] |= std::uint64_t(1) << c;
}

// ---- query ----------------------------------------------------------------------------------
struct Target {
  bool never = false;
  std::vector<std::pair<int, int>> loc_req;
  std::vector<IntAtom> int_cond;
  std::vector<Cons> clock_cond;
};
inline void load_target(const hora_target* t, Target& q) {
  if (!t) return;
  q.never = t->never != 0;
  for (int k = 0; k < t->nloc_req; ++k) q.loc_req.push_back({t->loc_req[2 * k], t->loc_req[2 * k + 1]});
  for (int k = 0; k < t->nint_cond; ++k)
    q.int_cond.push_back({t->int_cond[3 * k], t->int_cond[3 * k + 1], t->int_cond[3 * k + 2]});
  for (int k = 0; k < t->nclock_cond; ++k)
    q.clock_cond.push_back({t->clock_cond[3 * k], t->clock_cond[3 * k + 1], t->clock_cond[3 * k + 2]});
}

// ---- semantics (include/hora/detail/semantics.hpp) ------------------------------------------
inline bool apply_all(raw_t* z, int d, const std::vector<Cons>& cs) {
  for (const auto& c : cs)
    if (!dbm_constrain(z, d, c.i, c.j, c.b)) return false;
  return true;
}
inline bool int_ok(const std::vector<IntAtom>& g, const int* ints) {
  for (const auto& a : g) {
    const int x = ints[a.var];
    bool r = false;
    switch (a.op) {
      case HORA_OP_EQ: r = x == a.val; break;
      case HORA_OP_NE: r = x != a.val; break;
      case HORA_OP_LT: r = x < a.val; break;
      case HORA_OP_LE: r = x <= a.val; break;
      case HORA_OP_GT: r = x > a.val; break;
      case HORA_OP_GE: r = x >= a.val; break;
    }
    if (!r) return false;
  }
  return true;
}
inline Cons negate(const Cons& c) {
  const int v = bnd(c.b);
  return Cons{c.j, c.i, is_strict(c.b) ? le(-v) : lt(-v)};
}

struct Engine;

struct Engine {
  Net n;
  Target q;
  std::vector<std::int32_t> L, U, Lq, Uq;
  std::uint64_t qbits = 0;
  std::vector<int> shift;
  std::vector<std::uint64_t> fmask;
  int key_bits = 0;

  explicit Engine(const hora_net_enc* e, const hora_target* t) {
    load_net(e, n);
    compute_analyses(n);
    load_target(t, q);
    const int d = n.d;
    L = n.lbound; U = n.ubound;
    for (const auto& c : q.clock_cond) {   // extend_bounds (engine.cpp:739-752)
      const int v = bnd(c.b);
      if (c.i > 0 && c.j == 0) { if (v > U[c.i]) U[c.i] = v; }
      else if (c.i == 0 && c.j > 0) { if (-v > L[c.j]) L[c.j] = -v; }
      else if (c.i > 0 && c.j > 0) {
        const int a = std::abs(v);
        U[c.i] = std::max(U[c.i], a); U[c.j] = std::max(U[c.j], a);
        L[c.i] = std::max(L[c.i], a); L[c.j] = std::max(L[c.j], a);
      }
    }
    L[0] = U[0] = 0;
    Lq.assign(d, kAluNegInf); Uq.assign(d, kAluNegInf);
    for (const auto& c : q.clock_cond) {
      if (c.i > 0) qbits |= std::uint64_t(1) << c.i;
      if (c.j > 0) qbits |= std::uint64_t(1) << c.j;
      const int v = bnd(c.b);
      if (c.i > 0 && c.j == 0) Uq[c.i] = std::max(Uq[c.i], v);
      else if (c.i == 0 && c.j > 0) Lq[c.j] = std::max(Lq[c.j], -v);
      else if (c.i > 0 && c.j > 0) {
        const int a = std::abs(v);
        Uq[c.i] = std::max(Uq[c.i], a); Uq[c.j] = std::max(Uq[c.j], a);
        Lq[c.i] = std::max(Lq[c.i], a); Lq[c.j] = std::max(Lq[c.j], a);
      }
    }
    // key codec: one field per process, then per integer variable (engine.cpp:981-1001)
    auto fbits = [](long nv) { int b = 0; while ((1LL << b) < nv) ++b; return b; };
    int off = 0;
    for (int p = 0; p < n.np; ++p) {
      const int w = fbits(n.locbase[p + 1] - n.locbase[p]);
      shift.push_back(off); fmask.push_back((std::uint64_t(1) << w) - 1); off += w;
    }
    for (int v = 0; v < n.ni; ++v) {
      const int w = fbits(static_cast<long>(n.int_hi[v]) - n.int_lo[v] + 1);
      shift.push_back(off); fmask.push_back((std::uint64_t(1) << w) - 1); off += w;
    }
    key_bits = off;
  }

  std::uint64_t encode(const int* locs, const int* ints) const {
    std::uint64_t k = 0;
    for (int p = 0; p < n.np; ++p) k |= std::uint64_t(static_cast<unsigned>(locs[p])) << shift[p];
    for (int v = 0; v < n.ni; ++v)
      k |= std::uint64_t(static_cast<unsigned>(ints[v] - n.int_lo[v])) << shift[n.np + v];
    return k;
  }
  void decode(std::uint64_t k, int* locs, int* ints) const {
    for (int p = 0; p < n.np; ++p) locs[p] = static_cast<int>((k >> shift[p]) & fmask[p]);
    for (int v = 0; v < n.ni; ++v)
      ints[v] = static_cast<int>((k >> shift[n.np + v]) & fmask[n.np + v]) + n.int_lo[v];
  }

  // semantics.hpp:118-139
  bool urgent_sync_enabled(const int* locs, const int* ints) const {
    if (!n.has_urgent_chan) return false;
    for (int i = 0; i < n.np; ++i)
      for (int ei : n.edges_of[n.gl(i, locs[i])]) {
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || !n.uch[se.chan]) continue;
        if (!int_ok(se.ig, ints)) continue;
        if (n.bcast[se.chan]) return true;
        for (int j = 0; j < n.np; ++j) {
          if (j == i) continue;
          for (int ej : n.edges_of[n.gl(j, locs[j])]) {
            const Edge& re = n.edges[ej];
            if (re.chan != se.chan || re.dir != -1) continue;
            if (int_ok(re.ig, ints)) return true;
          }
        }
      }
    return false;
  }
  // semantics.hpp:160-169
  void frozen_flags(const int* locs, const int* ints, bool* committed, bool* urgent) const {
    *committed = *urgent = false;
    for (int p = 0; p < n.np; ++p) {
      const int f = n.loc_flags[n.gl(p, locs[p])];
      if (f & 1) *committed = true;
      if (f & 2) *urgent = true;
    }
    if (!*urgent && urgent_sync_enabled(locs, ints)) *urgent = true;
  }
  bool loc_committed(int p, int l) const { return (n.loc_flags[n.gl(p, l)] & 1) != 0; }
  // semantics.hpp:186-215 (bitmask path)
  void reset_inactive(const int* locs, raw_t* z) const {
    std::uint64_t a = qbits;
    for (int p = 0; p < n.np; ++p) a |= n.amask[n.gl(p, locs[p])];
    std::uint64_t dead = ~a & (((std::uint64_t(1) << n.nclocks) - 1) << 1);
    while (dead) {
      const int c = __builtin_ctzll(dead);
      dead &= dead - 1;
      dbm_reset(z, n.d, c);
    }
  }
  // all_invariants + apply_all (semantics.hpp:82-86)
  bool apply_inv(const int* locs, raw_t* z) const {
    for (int p = 0; p < n.np; ++p)
      if (!apply_all(z, n.d, n.inv[n.gl(p, locs[p])])) return false;
    return true;
  }
  // close_succ_zone without the avoid-set (semantics.hpp:522-568)
  bool close_succ_zone(const int* locs, const int* ints, raw_t* z, bool lu_plus) const {
    if (!apply_inv(locs, z)) return false;
    bool c, u;
    frozen_flags(locs, ints, &c, &u);
    if (!c && !u) { dbm_up(z, n.d); apply_inv(locs, z); }
    reset_inactive(locs, z);
    if (lu_plus) dbm_extra_lu_plus(z, n.d, L.data(), U.data());
    return true;
  }
  // reachable_net's match (engine.cpp:2275-2287)
  bool is_target(const int* locs, const int* ints, const raw_t* z) const {
    if (q.never) return false;
    for (const auto& lr : q.loc_req) if (locs[lr.first] != lr.second) return false;
    if (!int_ok(q.int_cond, ints)) return false;
    std::vector<raw_t> w(z, z + n.d * n.d);
    return apply_all(w.data(), n.d, q.clock_cond);
  }
  // state_bounds (engine.cpp:1458-1497), a_LU part
  void state_bounds(const int* locs, std::int32_t* L2, std::int32_t* U2) const {
    const int d = n.d;
    for (int c = 0; c < d; ++c) { L2[c] = kAluNegInf; U2[c] = kAluNegInf; }
    for (int p = 0; p < n.np; ++p) {
      const std::size_t base = static_cast<std::size_t>(n.gl(p, locs[p])) * d;
      for (int c = 1; c < d; ++c) {
        L2[c] = std::max(L2[c], n.lloc[base + c]);
        U2[c] = std::max(U2[c], n.uloc[base + c]);
      }
    }
    for (int c = 1; c < d; ++c) { L2[c] = std::max(L2[c], Lq[c]); U2[c] = std::max(U2[c], Uq[c]); }
    L2[0] = U2[0] = 0;
  }

  // Successor generation in gen_succ_reuse order (engine.cpp:541-718). `emit(locs, ints, zone)`
  // is called once per non-empty successor, in order; returning false stops the enumeration.
  template <class Emit>
  void successors(const int* slocs, const int* sints, const raw_t* szone, bool lu_plus,
                  Emit&& emit) const {
    const int d = n.d, d2 = d * d;
    std::vector<raw_t> z(d2);
    std::vector<int> nl(n.np), niv(n.ni);
    bool src_c, src_u;
    frozen_flags(slocs, sints, &src_c, &src_u);
    auto reset_disc = [&]() {
      std::copy(slocs, slocs + n.np, nl.begin());
      std::copy(sints, sints + n.ni, niv.begin());
    };
    // local edges
    for (int p = 0; p < n.np; ++p) {
      if (src_c && !loc_committed(p, slocs[p])) continue;
      for (int ei : n.edges_of[n.gl(p, slocs[p])]) {
        const Edge& e = n.edges[ei];
        if (e.chan >= 0) continue;
        if (!int_ok(e.ig, sints)) continue;
        reset_disc();
        nl[p] = e.tgt;
        for (const auto& a : e.iu) niv[a.var] = a.val;
        std::copy(szone, szone + d2, z.begin());
        if (!apply_all(z.data(), d, e.guard)) continue;
        for (int r : e.resets) dbm_reset(z.data(), d, r);
        if (!close_succ_zone(nl.data(), niv.data(), z.data(), lu_plus)) continue;
        if (dbm_empty(z.data(), d)) continue;
        if (!emit(nl.data(), niv.data(), z.data())) return;
      }
    }
    // binary synchronisation
    for (int i = 0; i < n.np; ++i)
      for (int ei : n.edges_of[n.gl(i, slocs[i])]) {
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || n.bcast[se.chan]) continue;
        if (!int_ok(se.ig, sints)) continue;
        const bool i_comm = loc_committed(i, slocs[i]);
        for (int j = 0; j < n.np; ++j) {
          if (j == i) continue;
          for (int ej : n.edges_of[n.gl(j, slocs[j])]) {
            const Edge& re = n.edges[ej];
            if (re.chan != se.chan || re.dir != -1) continue;
            if (src_c && !(i_comm || loc_committed(j, slocs[j]))) continue;
            if (!int_ok(re.ig, sints)) continue;
            reset_disc();
            nl[i] = se.tgt; nl[j] = re.tgt;
            for (const auto& a : se.iu) niv[a.var] = a.val;
            for (const auto& a : re.iu) niv[a.var] = a.val;
            std::copy(szone, szone + d2, z.begin());
            if (!apply_all(z.data(), d, se.guard)) continue;
            if (!apply_all(z.data(), d, re.guard)) continue;
            for (int r : se.resets) dbm_reset(z.data(), d, r);
            for (int r : re.resets) dbm_reset(z.data(), d, r);
            if (!close_succ_zone(nl.data(), niv.data(), z.data(), lu_plus)) continue;
            if (dbm_empty(z.data(), d)) continue;
            if (!emit(nl.data(), niv.data(), z.data())) return;
          }
        }
      }
    // broadcast synchronisation (build_broadcast_parts, semantics.hpp:275-303)
    for (int i = 0; i < n.np; ++i)
      for (int ei : n.edges_of[n.gl(i, slocs[i])]) {
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || !n.bcast[se.chan]) continue;
        if (!int_ok(se.ig, sints)) continue;
        struct Opt { int edge; std::vector<Cons> extra; };
        std::vector<std::pair<int, std::vector<Opt>>> parts;
        for (int j = 0; j < n.np; ++j) {
          if (j == i) continue;
          std::vector<Opt> opts;
          std::vector<const std::vector<Cons>*> guards;
          for (int ej : n.edges_of[n.gl(j, slocs[j])]) {
            const Edge& re = n.edges[ej];
            if (re.chan == se.chan && re.dir == -1 && int_ok(re.ig, sints)) {
              opts.push_back({ej, {}});
              guards.push_back(&re.guard);
            }
          }
          if (opts.empty()) continue;
          std::vector<std::vector<Cons>> dnf{{}};
          for (const auto* g : guards) {
            if (g->empty()) { dnf.clear(); break; }
            std::vector<std::vector<Cons>> next;
            for (const auto& acc : dnf)
              for (const auto& c : *g) { auto cp = acc; cp.push_back(negate(c)); next.push_back(cp); }
            dnf.swap(next);
          }
          for (auto& dd : dnf) opts.push_back({-1, dd});
          parts.push_back({j, opts});
        }
        std::vector<std::size_t> pick(parts.size(), 0);
        for (;;) {
          bool part_c = loc_committed(i, slocs[i]);
          for (std::size_t t = 0; t < parts.size(); ++t)
            if (parts[t].second[pick[t]].edge >= 0 && loc_committed(parts[t].first, slocs[parts[t].first]))
              part_c = true;
          if (!src_c || part_c) {
            std::copy(szone, szone + d2, z.begin());
            bool ok = apply_all(z.data(), d, se.guard);
            for (std::size_t t = 0; ok && t < parts.size(); ++t) {
              const Opt& o = parts[t].second[pick[t]];
              ok = o.edge >= 0 ? apply_all(z.data(), d, n.edges[o.edge].guard)
                               : apply_all(z.data(), d, o.extra);
            }
            if (ok) {
              reset_disc();
              nl[i] = se.tgt;
              for (const auto& a : se.iu) niv[a.var] = a.val;
              for (int r : se.resets) dbm_reset(z.data(), d, r);
              for (std::size_t t = 0; t < parts.size(); ++t) {
                const Opt& o = parts[t].second[pick[t]];
                if (o.edge < 0) continue;
                const Edge& re = n.edges[o.edge];
                nl[parts[t].first] = re.tgt;
                for (const auto& a : re.iu) niv[a.var] = a.val;
                for (int r : re.resets) dbm_reset(z.data(), d, r);
              }
              if (close_succ_zone(nl.data(), niv.data(), z.data(), lu_plus) && !dbm_empty(z.data(), d))
                if (!emit(nl.data(), niv.data(), z.data())) return;
            }
          }
          if (parts.empty()) break;
          std::size_t k = 0;
          for (; k < parts.size(); ++k) {
            if (++pick[k] < parts[k].second.size()) break;
            pick[k] = 0;
          }
          if (k == parts.size()) break;
        }
      }
  }
};

}  // namespace orc
