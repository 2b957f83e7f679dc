// oracle_engine.cpp — CPU restatement of HORA's network semantics and forward zone-graph search,
// in a sequential ("seq") and an OpenMP ("omp") form, behind a small C API for ctypes.
//
// TEST INFRASTRUCTURE ONLY: this is the oracle. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may build, load or call it; the HIP back end never does.
//
// What is restated, and from where (paths relative to the HORA source tree):
//   static analyses  Network::compute_bounds / compute_loc_bounds / compute_active
//                    src/model/network.cpp:59-245
//   semantics        apply_all, int_ok, all_invariants, urgent_sync_enabled, frozen_flags,
//                    reset_inactive, build_broadcast_parts, close_succ_zone
//                    include/hora/detail/semantics.hpp:42-568
//   successors       gen_succ_reuse (local, then binary, then broadcast), engine.cpp:541-718
//   key codec        DiscKeyCodecT, engine.cpp:967-1026; key_mix, include/hora/detail/disc_key.hpp:50
//   bounds           extend_bounds engine.cpp:739-752, state_bounds engine.cpp:1458-1497
//   passed set       NetPassedT::add / probe_kill, include/hora/detail/passed_set.hpp:239-444,
//                    kMergeScan = 32 include/hora/detail/engine_config.hpp:52
//   search loop      reach_engine CPU path, engine.cpp:1996-2202 (adaptive emode switch 2178-2184)
//   target test      reachable_net's match, engine.cpp:2275-2287
//
// Canonical next-frontier order (the one deliberate choice; engine.cpp:1254-1300 iterates an
// std::unordered_set, whose order is implementation-defined): the zones stored during a level
// that are still live at its end, in the order they were stored, i.e. by the (frontier index,
// successor index) of the add() call that stored them. Within a level, insertion order is the
// reference's own (i, sk) order (engine.cpp:2149-2163). Families the reference records as
// order-invariant (include/hora/config.hpp:150-152: fischer, arbiter, twolock, viking) give the
// recorded counts under any order; tests pin them.
#include "oracle_engine.hpp"

namespace orc {

// ---- passed set: one antichain per discrete key (passed_set.hpp:287-444) --------------------
constexpr std::size_t kMergeScan = 32;

struct ZoneArena {  // chunked so growth never moves a zone
  int d2 = 0;
  static constexpr std::size_t kChunk = 1 << 14;
  std::vector<std::unique_ptr<raw_t[]>> chunks;
  std::size_t n = 0;
  raw_t* at(std::size_t i) { return chunks[i / kChunk].get() + (i % kChunk) * d2; }
  const raw_t* at(std::size_t i) const { return chunks[i / kChunk].get() + (i % kChunk) * d2; }
  std::size_t push(const raw_t* z) {
    if (n % kChunk == 0 && n / kChunk == chunks.size())
      chunks.emplace_back(new raw_t[kChunk * static_cast<std::size_t>(d2)]);
    std::memcpy(at(n), z, sizeof(raw_t) * d2);
    return n++;
  }
};

struct Chain {
  std::vector<std::uint32_t> nodes;  // stored-zone ids, oldest first
  std::size_t ndead = 0;
};

struct Stripe {
  std::unordered_map<std::uint64_t, Chain> map;
  ZoneArena zones;
  std::vector<int> level;
  std::vector<char> dead;
  std::size_t live = 0;
};

struct FoldCounters { std::uint64_t subsumed = 0, killed = 0, unions = 0, reads = 0, inserted = 0; };

// NetPassedT::add, dense-zone form. Returns the stored id or -1 when the zone is covered.
static long passed_add(Stripe& st, std::uint64_t dk, raw_t* s, int d, int level,
                       const std::int32_t* Ls, const std::int32_t* Us, FoldCounters& fc) {
  auto it = st.map.find(dk);
  if (it == st.map.end()) {
    const std::size_t id = st.zones.push(s);
    st.level.push_back(level); st.dead.push_back(0); ++st.live;
    Chain c; c.nodes.push_back(static_cast<std::uint32_t>(id));
    st.map.emplace(dk, std::move(c));
    ++fc.inserted;
    return static_cast<long>(id);
  }
  Chain& ch = it->second;
  // live zones newest first
  std::vector<std::uint32_t> live;
  live.reserve(ch.nodes.size());
  for (std::size_t k = ch.nodes.size(); k-- > 0;)
    if (!st.dead[ch.nodes[k]]) live.push_back(ch.nodes[k]);
  // (1)+(2) covered by a stored zone -> drop (plain inclusion implies simulation)
  for (std::uint32_t id : live) {
    ++fc.reads;
    if (dbm_alu_le(s, st.zones.at(id), d, Ls, Us)) { ++fc.subsumed; return -1; }
  }
  auto retire = [&](std::uint32_t id) { st.dead[id] = 1; --st.live; ++ch.ndead; };
  // (3) retire stored zones covered by s, keeping the survivors' order
  for (std::size_t k = 0; k < live.size();) {
    ++fc.reads;
    if (dbm_alu_le(st.zones.at(live[k]), s, d, Ls, Us)) {
      retire(live[k]); ++fc.killed;
      live.erase(live.begin() + static_cast<long>(k));
    } else {
      ++k;
    }
  }
  // (4) convex-union merge over the kMergeScan newest live zones, restarting after each merge
  for (bool merged = true; merged;) {
    merged = false;
    for (std::size_t k = 0; k < std::min(live.size(), kMergeScan); ++k) {
      ++fc.reads;
      if (dbm_try_union(s, st.zones.at(live[k]), d)) {
        retire(live[k]); ++fc.unions;
        live.erase(live.begin() + static_cast<long>(k));
        merged = true;
        break;
      }
    }
  }
  // (5) store s as the newest member
  const std::size_t id = st.zones.push(s);
  st.level.push_back(level); st.dead.push_back(0); ++st.live;
  ch.nodes.push_back(static_cast<std::uint32_t>(id));
  ++fc.inserted;
  if (ch.ndead * 2 > ch.nodes.size() + 8) {   // drop tombstones, order kept
    std::vector<std::uint32_t> keep;
    for (std::uint32_t idn : ch.nodes) if (!st.dead[idn]) keep.push_back(idn);
    ch.nodes.swap(keep); ch.ndead = 0;
  }
  return static_cast<long>(id);
}

// ---- spin reader/writer lock for the OMP engine (include/hora/detail/spin_rw.hpp:27-76) -----
struct SpinRW {
  std::atomic<int> s{0};  // -1 writer, >0 readers
  void lock_shared() {
    for (;;) {
      int v = s.load(std::memory_order_relaxed);
      if (v >= 0 && s.compare_exchange_weak(v, v + 1, std::memory_order_acquire)) return;
    }
  }
  void unlock_shared() { s.fetch_sub(1, std::memory_order_release); }
  void lock() {
    for (;;) {
      int v = 0;
      if (s.compare_exchange_weak(v, -1, std::memory_order_acquire)) return;
    }
  }
  void unlock() { s.store(0, std::memory_order_release); }
};

inline std::uint64_t key_mix(std::uint64_t dk) { return dk * 0x9e3779b97f4a7c15ULL; }

struct FrontierEntry { std::uint64_t dk; std::uint32_t stripe; std::uint32_t id; };

// `max_seconds` > 0 bounds the run: like the reference's cancel flag (engine.cpp:2054) it is polled
// once per level, and the counters then describe the levels completed so far.
static int run_reach(Engine& E, int nthreads, int32_t* reachable, hora_reach_stats* st,
                     double max_seconds = 0) {
  const auto t0 = std::chrono::steady_clock::now();
  const int d = E.n.d, d2 = d * d, np = E.n.np, ni = E.n.ni;
  std::memset(st, 0, sizeof(*st));
  st->emode_switch_level = -1;
  *reachable = 0;
  const bool parallel = nthreads > 0;
  const int nth = parallel ? nthreads : 1;
  int sbits = 0;
  if (parallel) { sbits = 10; while ((1 << sbits) < std::max(16384, nth * 64)) ++sbits; }
  const std::size_t nstripes = std::size_t(1) << sbits;
  auto stripe_of = [&](std::uint64_t dk) -> std::size_t {
    return sbits ? static_cast<std::size_t>(key_mix(dk) >> (64 - sbits)) : 0;
  };
  std::vector<Stripe> stripes(nstripes);
  for (auto& s : stripes) s.zones.d2 = d2;
  std::vector<SpinRW> locks(parallel ? nstripes : 0);
  bool lu_plus = false;

  // initial state (initial_symbolic_state, engine.cpp:1322-1333)
  std::vector<int> l0(np), i0(ni);
  for (int p = 0; p < np; ++p) l0[p] = E.n.loc_init[p];
  for (int v = 0; v < ni; ++v) i0[v] = E.n.int_init[v];
  std::vector<raw_t> z0(d2);
  dbm_zero(z0.data(), d);
  if (!E.close_succ_zone(l0.data(), i0.data(), z0.data(), lu_plus)) {
    st->seconds = 0; return 0;
  }
  if (E.is_target(l0.data(), i0.data(), z0.data())) {
    *reachable = 1; st->found = 1; return 0;
  }
  FoldCounters fc0;
  std::vector<std::int32_t> Ls(d), Us(d);
  const std::uint64_t dk0 = E.encode(l0.data(), i0.data());
  E.state_bounds(l0.data(), Ls.data(), Us.data());
  const std::size_t s0 = stripe_of(dk0);
  const long id0 = passed_add(stripes[s0], dk0, z0.data(), d, 0, Ls.data(), Us.data(), fc0);
  std::vector<FrontierEntry> frontier{{dk0, static_cast<std::uint32_t>(s0), static_cast<std::uint32_t>(id0)}};

  int level = 0;
  std::atomic<bool> found{false};
  std::vector<FoldCounters> tfc(nth);
  tfc[0] = fc0;
  while (!frontier.empty()) {
    if (max_seconds > 0 &&
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > max_seconds)
      break;
    st->explored += frontier.size();
    ++level;
    std::vector<std::vector<FrontierEntry>> stored(nth);
    std::vector<std::uint64_t> tgen(nth, 0), texp(nth, 0);
    // frontier zones copied out first: the stripes' arenas grow during the level
    std::vector<raw_t> fz(frontier.size() * static_cast<std::size_t>(d2));
    for (std::size_t i = 0; i < frontier.size(); ++i)
      std::memcpy(fz.data() + i * d2, stripes[frontier[i].stripe].zones.at(frontier[i].id), sizeof(raw_t) * d2);
    if (!parallel) {
      std::vector<int> locs(np), ints(ni);
      std::vector<raw_t> sz(d2);
      std::uint64_t gen = 0;
      for (std::size_t i = 0; i < frontier.size() && !found; ++i) {
        E.decode(frontier[i].dk, locs.data(), ints.data());
        std::memcpy(sz.data(), fz.data() + i * d2, sizeof(raw_t) * d2);
        ++texp[0];
        E.successors(locs.data(), ints.data(), sz.data(), lu_plus,
                     [&](const int* nl, const int* nv, raw_t* z) {
                       ++gen;
                       if (E.is_target(nl, nv, z)) { found = true; return false; }
                       const std::uint64_t dk = E.encode(nl, nv);
                       E.state_bounds(nl, Ls.data(), Us.data());
                       const long id = passed_add(stripes[0], dk, z, d, level, Ls.data(), Us.data(), tfc[0]);
                       if (id >= 0) stored[0].push_back({dk, 0, static_cast<std::uint32_t>(id)});
                       return true;
                     });
      }
      tgen[0] = gen;
    } else {
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
      {
        const int tid = omp_get_thread_num();
        std::vector<int> locs(np), ints(ni);
        std::vector<raw_t> sz(d2);
        std::vector<std::int32_t> L2(d), U2(d);
        std::uint64_t gen = 0, ex = 0;
#pragma omp for schedule(dynamic, 8) nowait
        for (long i = 0; i < static_cast<long>(frontier.size()); ++i) {
          if (found.load(std::memory_order_relaxed)) continue;
          E.decode(frontier[i].dk, locs.data(), ints.data());
          std::memcpy(sz.data(), fz.data() + i * d2, sizeof(raw_t) * d2);
          ++ex;
          E.successors(locs.data(), ints.data(), sz.data(), lu_plus,
                       [&](const int* nl, const int* nv, raw_t* z) {
                         ++gen;
                         if (E.is_target(nl, nv, z)) { found.store(true); return false; }
                         const std::uint64_t dk = E.encode(nl, nv);
                         E.state_bounds(nl, L2.data(), U2.data());
                         const std::size_t si = stripe_of(dk);
                         Stripe& sp = stripes[si];
                         {  // shared-lock probe (passed_set.hpp:239-266), then exclusive add
                           locks[si].lock_shared();
                           bool covered = false;
                           auto itc = sp.map.find(dk);
                           if (itc != sp.map.end())
                             for (std::size_t k = itc->second.nodes.size(); k-- > 0 && !covered;) {
                               const std::uint32_t idn = itc->second.nodes[k];
                               if (!sp.dead[idn] && dbm_alu_le(z, sp.zones.at(idn), d, L2.data(), U2.data()))
                                 covered = true;
                             }
                           locks[si].unlock_shared();
                           if (covered) { ++tfc[tid].subsumed; return true; }
                         }
                         locks[si].lock();
                         const long id = passed_add(sp, dk, z, d, level, L2.data(), U2.data(), tfc[tid]);
                         locks[si].unlock();
                         if (id >= 0) stored[tid].push_back({dk, static_cast<std::uint32_t>(si), static_cast<std::uint32_t>(id)});
                         return true;
                       });
        }
        tgen[tid] = gen; texp[tid] = ex;
      }
#endif
    }
    for (int t = 0; t < nth; ++t) { st->generated += tgen[t]; st->expanded += texp[t]; }
    if (found) { *reachable = 1; st->found = 1; break; }
    // next frontier: live zones stored this level, in storing order
    std::vector<FrontierEntry> next;
    for (int t = 0; t < nth; ++t)
      for (const auto& fe : stored[t])
        if (!stripes[fe.stripe].dead[fe.id]) next.push_back(fe);
    frontier.swap(next);
    if (!lu_plus) {   // adaptive Extra_LU+ switch (engine.cpp:2178-2184)
      std::size_t nlive = 0, ndisc = 0;
      for (const auto& sp : stripes) { nlive += sp.live; ndisc += sp.map.size(); }
      if (ndisc && static_cast<double>(nlive) >= 4.0 * static_cast<double>(ndisc)) {
        lu_plus = true;
        st->emode_switch_level = level;
      }
    }
  }
  st->levels = level;
  for (const auto& sp : stripes) { st->stored_live += sp.live; st->keys += sp.map.size(); }
  for (const auto& f : tfc) {
    st->subsumed += f.subsumed; st->killed += f.killed; st->unions += f.unions;
    st->chain_reads += f.reads; st->inserted += f.inserted;
  }
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

static thread_local std::string g_err;

}  // namespace orc

extern "C" {

const char* oracle_last_error(void) { return orc::g_err.c_str(); }

// Network analyses as the oracle computes them (src/model/network.cpp:59-245).
int oracle_analyses(const hora_net_enc* e, int32_t* lbound, int32_t* ubound, int32_t* lbound_loc,
                    int32_t* ubound_loc, uint64_t* active_mask) {
  try {
    orc::Net n;
    orc::load_net(e, n);
    orc::compute_analyses(n);
    std::copy(n.lbound.begin(), n.lbound.end(), lbound);
    std::copy(n.ubound.begin(), n.ubound.end(), ubound);
    std::copy(n.lloc.begin(), n.lloc.end(), lbound_loc);
    std::copy(n.uloc.begin(), n.uloc.end(), ubound_loc);
    std::copy(n.amask.begin(), n.amask.end(), active_mask);
    return 0;
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

// E<> target search. nthreads == 0: the sequential engine; > 0: the OpenMP engine with that many
// threads (level-synchronous, stripe-locked shared passed set, dynamic schedule chunk 8).
int oracle_reach(const hora_net_enc* e, const hora_target* t, int32_t nthreads, int32_t* reachable,
                 hora_reach_stats* st) {
  try {
    orc::Engine E(e, t);
    if (E.n.has_diag) { orc::g_err = "diagonal constraints are not covered"; return HORA_HIP_E_SCOPE; }
    if (E.key_bits > 63) { orc::g_err = "discrete key wider than 63 bits"; return HORA_HIP_E_SCOPE; }
    return orc::run_reach(E, nthreads, reachable, st);
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

// Time-bounded run for the CPU baseline sample: stops at the first level boundary after
// `max_seconds`; *complete is set to 1 if the search finished within the bound.
int oracle_reach_bounded(const hora_net_enc* e, const hora_target* t, int32_t nthreads, double max_seconds,
                         int32_t* reachable, hora_reach_stats* st, int32_t* complete) {
  try {
    orc::Engine E(e, t);
    if (E.n.has_diag || E.key_bits > 63) return HORA_HIP_E_SCOPE;
    const int rc = orc::run_reach(E, nthreads, reachable, st, max_seconds);
    *complete = (st->seconds <= max_seconds || max_seconds <= 0) ? 1 : 0;
    return rc;
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

// Successors of explicit states, in generation order: for parity tests of the expansion alone.
// in: n states as packed keys + zones. out: up to cap successors (dk, zone, source index).
int oracle_expand(const hora_net_enc* e, const hora_target* t, int32_t lu_plus, int64_t n,
                  const uint64_t* dks, const int32_t* zones, int64_t cap, uint64_t* out_dk,
                  int32_t* out_zones, int32_t* out_src, int32_t* out_target, int64_t* out_n) {
  try {
    orc::Engine E(e, t);
    const int d = E.n.d, d2 = d * d;
    std::vector<int> locs(E.n.np), ints(E.n.ni);
    int64_t k = 0;
    bool overflow = false;
    for (int64_t i = 0; i < n && !overflow; ++i) {
      E.decode(dks[i], locs.data(), ints.data());
      std::vector<orc::raw_t> sz(zones + i * d2, zones + (i + 1) * d2);
      E.successors(locs.data(), ints.data(), sz.data(), lu_plus != 0,
                   [&](const int* nl, const int* nv, orc::raw_t* z) {
                     if (k >= cap) { overflow = true; return false; }
                     out_dk[k] = E.encode(nl, nv);
                     std::memcpy(out_zones + k * d2, z, sizeof(int32_t) * d2);
                     out_src[k] = static_cast<int32_t>(i);
                     out_target[k] = E.is_target(nl, nv, z) ? 1 : 0;
                     ++k;
                     return true;
                   });
    }
    *out_n = k;
    if (overflow) { orc::g_err = "output capacity exceeded"; return HORA_HIP_E_MEMORY; }
    return 0;
  } catch (const std::exception& ex) {
    orc::g_err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

// Batched zone operations on host memory, same contract as hora_hip_zone_op.
int oracle_zone_op(int32_t op, int32_t d, int64_t n, const int32_t* a, const int32_t* b,
                   const int32_t* arg, const int32_t* L, const int32_t* U, int32_t* out,
                   int32_t* out_flag) {
  using namespace orc;
  const int d2 = d * d;
  std::vector<raw_t> z(d2);
  for (int64_t k = 0; k < n; ++k) {
    const raw_t* A = a + k * d2;
    const raw_t* B = b ? b + k * d2 : nullptr;
    std::copy(A, A + d2, z.begin());
    int flag = 0;
    switch (op) {
      case HORA_ZOP_CLOSE: flag = dbm_close(z.data(), d) ? 1 : 0; break;
      case HORA_ZOP_UP: dbm_up(z.data(), d); flag = 1; break;
      case HORA_ZOP_RESET: dbm_reset(z.data(), d, arg[k]); flag = 1; break;
      case HORA_ZOP_CONSTRAIN:
        flag = dbm_constrain(z.data(), d, arg[3 * k], arg[3 * k + 1], arg[3 * k + 2]) ? 1 : 0;
        break;
      case HORA_ZOP_EXTRA_LU: dbm_extra_lu_plus(z.data(), d, L + k * d, U + k * d); flag = 1; break;
      case HORA_ZOP_INCLUDES: flag = dbm_includes(A, B, d) ? 1 : 0; break;
      case HORA_ZOP_ALU_LE: flag = dbm_alu_le(A, B, d, L + k * d, U + k * d) ? 1 : 0; break;
      case HORA_ZOP_UNION: flag = dbm_try_union(z.data(), B, d) ? 1 : 0; break;
      case HORA_ZOP_EMPTY: flag = dbm_empty(A, d) ? 1 : 0; break;
      case HORA_ZOP_DOWN: dbm_down(z.data(), d); flag = 1; break;
      case HORA_ZOP_FREE: dbm_free_many(z.data(), d, static_cast<std::uint32_t>(arg[k])); flag = 1; break;
      default: return HORA_HIP_E_ARG;
    }
    if (out) std::copy(z.begin(), z.end(), out + k * d2);
    if (out_flag) out_flag[k] = flag;
  }
  return 0;
}

}  // extern "C"
