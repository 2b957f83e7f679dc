// oracle_live.cpp — CPU restatement of HORA's liveness pass: forward zone graph with recorded
// transitions, stop sets, exact predecessor, and the backward greatest fixpoint
// W = W /\ (NoSucc \/ Pre(W)), for `p --> q`, `E[] psi` and `A[] not deadlock`.
//
// TEST INFRASTRUCTURE ONLY: this is the oracle. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may build, load or call it; the HIP back end never does.
//
// What is restated, and from where (paths relative to the HORA source tree):
//   transitions      tagged_successors, src/liveness/operators.cpp:55-206 (local, then binary
//                    synchronisation, then broadcast with the receivers' options of
//                    build_broadcast_parts, semantics.hpp:275-303; the order of gen_succ_reuse).
//                    An edge records (sender or local edge, receiver edge, or -2 - c for
//                    combination c of a broadcast). No extrapolation on this lane
//                    (emode = ExNone, src/liveness/engine.cpp:256-276).
//   graph            three constructions, chosen as decide_leads_to / the deadlock and E[] entry
//                    points choose them (src/liveness/api.cpp:357-447):
//                    * online (the default): LivenessEngine::online_base_forward with online_expand
//                      and online_intern (src/liveness/refinement.cpp:935-1065, :610-781;
//                      construction.cpp:480-683) for a map D without atoms: a successor goes to the
//                      newest active record of its key that a_LU-covers it, else it is joined with
//                      the active records whose union with it is one zone and appended, retiring
//                      the active records it covers; a record retired before its turn is not
//                      expanded; edges to unexpanded records follow the retirements
//                      (online_finalize_graph, refinement.cpp:1071-1160). `build_online`.
//                    * base, for p --> q where D is provably a_LU (structurally_empty_d,
//                      src/liveness/internal.cpp:59-103): LivenessEngine::intern
//                      (construction.cpp:863-1003): oldest cover, else in-place enlargement of the
//                      oldest covered node, else convex union, else append; a node rewritten by
//                      another node's successor is expanded again. `build_base`.
//                    * the a_M graph, when a reachable discrete state can seed a deadlock atom
//                      (online_seed_key, refinement.cpp:176-282) or when asked for: a successor goes
//                      to the first (oldest) node of its key that covers it under a_M (per-location
//                      bounds raised to M = max(L, U) on every clock, a bisimulation, so deadlocks
//                      are preserved; construction.cpp:899-913), else it becomes a new node. `build`.
//                    The reference refines a_LU at the seeded keys only (revocable covers, deferred
//                    recipes); that refinement is not restated, so a seeded network gets the a_M
//                    graph's larger node count and the same verdict.
//   order            a FIFO of records, run level by level: the frontier of a level is the records
//                    the previous level created that are still active, in creation order; successors
//                    are interned in (frontier index, transition index) order, so record ids are
//                    creation order (the reference's FIFO over the same records, api.cpp:37-117).
//   stop set         stop_set, construction.cpp:1200-1238; time_diverges :1097-1162;
//                    orig_time_open :1164-1191; state_has_deadlock,
//                    src/reachability/deadlock.cpp:74-268.
//   predecessor      exact_predecessor, src/liveness/operators.cpp:257-306 (tag resets
//                    src/liveness/fixpoint.cpp:638-647, dead mask :599-633).
//   fixpoint         solve_greatest_fixpoint's relax, src/liveness/fixpoint.cpp:750-864, run to
//                    convergence by a worklist; the limit is the unique greatest fixpoint, so any
//                    fair schedule reaches the same sets (the device runs Jacobi rounds).
//   readouts         certificate_in / accepting_closure (fixpoint.cpp:979-1101); eg_closure /
//                    eg_verdict (src/liveness/api.cpp:169-188, 290-337); early witness
//                    leadsto_witness (construction.cpp:786-813), online_eg_terminal_witness
//                    (:748-767) and the deadlock scan, checked after each level.
//
// Federations are lists of zones; only the sets they denote matter to the verdict and to the
// per-node emptiness the tests compare (zone difference splits along the subtrahend's facets in
// row-major order, oracle_dbm.hpp).
#include <array>
#include <chrono>
#include <deque>

#include "oracle_engine.hpp"

namespace orc {
namespace live {

static thread_local std::string g_err;

struct DPred {
  std::vector<std::vector<std::array<int, 4>>> terms;
  void load(const hora_dpred& p) {
    terms.clear();
    for (int t = 0; t < p.nterms; ++t) {
      std::vector<std::array<int, 4>> term;
      for (int a = p.term_off[t]; a < p.term_off[t + 1]; ++a)
        term.push_back({p.atoms[4 * a], p.atoms[4 * a + 1], p.atoms[4 * a + 2], p.atoms[4 * a + 3]});
      terms.push_back(term);
    }
  }
  bool eval(const int* locs, const int* ints) const {
    for (const auto& term : terms) {
      bool all = true;
      for (const auto& a : term) {
        const int v = a[0] == 0 ? locs[a[1]] : ints[a[1]];
        if (!int_ok({IntAtom{0, a[2], a[3]}}, &v)) { all = false; break; }
      }
      if (all) return true;
    }
    return false;
  }
};

using Fed = std::vector<raw_t>;   // zones back to back, d*d each

struct Live {
  Engine& E;
  const Net& n;
  int d, d2, kind;
  bool early_on, restricted = false;
  std::vector<std::vector<Cons>> orig_inv;   // unrestricted invariants (E[] with a clock part)
  DPred p, q;
  // graph
  std::vector<std::uint64_t> dk;
  std::vector<raw_t> zones;
  std::vector<std::int64_t> row_off;
  std::vector<std::int32_t> etgt, ese, ere;
  std::vector<std::int8_t> expanded, in_s, wne;
  std::unordered_map<std::uint64_t, std::vector<int>> chains;
  std::vector<std::int64_t> ecnt;
  hora_live_stats st{};
  std::size_t fedcap = 4096;   // fixpoint.cpp's federation cap

  Live(Engine& e, const hora_live_query* qq) : E(e), n(e.n), d(e.n.d), d2(e.n.d * e.n.d) {
    kind = qq->kind;
    early_on = qq->early != 0;
    p.load(qq->p);
    q.load(qq->q);
    if (qq->orig_inv_off) {
      restricted = true;
      orig_inv.assign(n.tl, {});
      for (int L = 0; L < n.tl; ++L)
        for (int c = qq->orig_inv_off[L]; c < qq->orig_inv_off[L + 1]; ++c)
          orig_inv[L].push_back({qq->orig_inv_data[3 * c], qq->orig_inv_data[3 * c + 1], qq->orig_inv_data[3 * c + 2]});
    }
  }

  const raw_t* Z(int v) const { return &zones[static_cast<std::size_t>(v) * d2]; }
  void state(int v, int* locs, int* ints) const { E.decode(dk[v], locs, ints); }
  bool psi(int v) const {
    int locs[64], ints[64];
    state(v, locs, ints);
    return p.eval(locs, ints);
  }
  bool is_p(int v) const { int l[64], i[64]; state(v, l, i); return p.eval(l, i); }
  bool is_q(int v) const { int l[64], i[64]; state(v, l, i); return q.eval(l, i); }
  bool frozen(const int* locs, const int* ints) const {
    bool c, u;
    E.frozen_flags(locs, ints, &c, &u);
    return c || u;
  }
  std::uint64_t dead_mask(const int* locs) const {
    std::uint64_t a = 0;
    for (int pr = 0; pr < n.np; ++pr) a |= n.amask[n.gl(pr, locs[pr])];
    return ~a & (((std::uint64_t(1) << n.nclocks) - 1) << 1);
  }
  const std::vector<Cons>& inv_of(const std::vector<std::vector<Cons>>& table, int pr, int l) const {
    return table[n.gl(pr, l)];
  }
  bool apply_inv_table(const std::vector<std::vector<Cons>>& table, const int* locs, raw_t* z) const {
    for (int pr = 0; pr < n.np; ++pr)
      if (!apply_all(z, d, table[n.gl(pr, locs[pr])])) return false;
    return true;
  }
  bool has_upper(const std::vector<std::vector<Cons>>& table, const int* locs) const {
    for (int pr = 0; pr < n.np; ++pr)
      for (const auto& c : table[n.gl(pr, locs[pr])])
        if (c.i > 0 && c.j == 0 && c.b < kInf) return true;
    return false;
  }

  // a_M bounds of a location vector: the a_LU state bounds with L and U raised to their maximum
  void mbounds(const int* locs, std::int32_t* M) const {
    std::int32_t L2[64], U2[64];
    E.state_bounds(locs, L2, U2);
    for (int c = 0; c < d; ++c) M[c] = std::max(L2[c], U2[c]);
  }

  // Broadcast parts of sender edge se from the discrete state (sl, si) (build_broadcast_parts,
  // semantics.hpp:275-303): per other process with an integer-enabled receiving edge, in process
  // order, its options: each receiving edge, then each way of staying out where all of their
  // clock guards fail (one negated atom per guard, the first guard the most significant).
  struct BOpt { int edge; std::vector<Cons> extra; };
  using BParts = std::vector<std::pair<int, std::vector<BOpt>>>;
  BParts bparts(const int* sl, const int* si, int se) const {
    BParts parts;
    const int i = n.edges[se].proc, ch = n.edges[se].chan;
    for (int j = 0; j < n.np; ++j) {
      if (j == i) continue;
      std::vector<BOpt> opts;
      std::vector<const std::vector<Cons>*> guards;
      for (int ej : n.edges_of[n.gl(j, sl[j])]) {
        const Edge& re = n.edges[ej];
        if (re.chan == ch && re.dir == -1 && int_ok(re.ig, si)) {
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
    return parts;
  }
  static long bcombos(const BParts& parts) {
    long c = 1;
    for (const auto& p : parts) c *= static_cast<long>(p.second.size());
    return c;
  }
  // the option of each part in combination c, the first part the fastest digit
  static std::vector<const BOpt*> bpick(const BParts& parts, long c) {
    std::vector<const BOpt*> pick;
    for (const auto& p : parts) {
      const long k = static_cast<long>(p.second.size());
      pick.push_back(&p.second[c % k]);
      c /= k;
    }
    return pick;
  }

  // successors with their transition (se, re): tagged_successors on the a_LU lane
  template <class F>
  void succ_rec(const int* sl, const int* si, const raw_t* sz, F&& emit) const {
    std::vector<raw_t> z(d2);
    std::vector<int> nl(n.np), niv(n.ni);
    const bool src_c = [&] { bool c, u; E.frozen_flags(sl, si, &c, &u); return c; }();
    auto fire = [&](int se, int re) {
      std::copy(sl, sl + n.np, nl.begin());
      std::copy(si, si + n.ni, niv.begin());
      const Edge& a = n.edges[se];
      nl[a.proc] = a.tgt;
      for (const auto& u : a.iu) niv[u.var] = u.val;
      if (re >= 0) {
        const Edge& b = n.edges[re];
        nl[b.proc] = b.tgt;
        for (const auto& u : b.iu) niv[u.var] = u.val;
      }
      std::copy(sz, sz + d2, z.begin());
      if (!apply_all(z.data(), d, a.guard)) return;
      if (re >= 0 && !apply_all(z.data(), d, n.edges[re].guard)) return;
      for (int r : a.resets) dbm_reset(z.data(), d, r);
      if (re >= 0) for (int r : n.edges[re].resets) dbm_reset(z.data(), d, r);
      if (!E.close_succ_zone(nl.data(), niv.data(), z.data(), false)) return;
      if (dbm_empty(z.data(), d)) return;
      emit(nl.data(), niv.data(), z.data(), se, re);
    };
    for (int pr = 0; pr < n.np; ++pr) {
      if (src_c && !E.loc_committed(pr, sl[pr])) continue;
      for (int ei : n.edges_of[n.gl(pr, sl[pr])]) {
        const Edge& e = n.edges[ei];
        if (e.chan >= 0 || !int_ok(e.ig, si)) continue;
        fire(ei, -1);
      }
    }
    for (int i = 0; i < n.np; ++i)
      for (int ei : n.edges_of[n.gl(i, sl[i])]) {
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || n.bcast[se.chan]) continue;
        if (!int_ok(se.ig, si)) continue;
        const bool i_comm = E.loc_committed(i, sl[i]);
        for (int j = 0; j < n.np; ++j) {
          if (j == i) continue;
          for (int ej : n.edges_of[n.gl(j, sl[j])]) {
            const Edge& re = n.edges[ej];
            if (re.chan != se.chan || re.dir != -1) continue;
            if (src_c && !(i_comm || E.loc_committed(j, sl[j]))) continue;
            if (!int_ok(re.ig, si)) continue;
            fire(ei, ej);
          }
        }
      }
    for (int i = 0; i < n.np; ++i)
      for (int ei : n.edges_of[n.gl(i, sl[i])]) {
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || !n.bcast[se.chan]) continue;
        if (!int_ok(se.ig, si)) continue;
        const BParts parts = bparts(sl, si, ei);
        const long combos = bcombos(parts);
        for (long c = 0; c < combos; ++c) {
          const auto pick = bpick(parts, c);
          bool part_c = E.loc_committed(i, sl[i]);
          for (std::size_t t = 0; t < parts.size(); ++t)
            if (pick[t]->edge >= 0 && E.loc_committed(parts[t].first, sl[parts[t].first])) part_c = true;
          if (src_c && !part_c) continue;
          std::copy(sl, sl + n.np, nl.begin());
          std::copy(si, si + n.ni, niv.begin());
          nl[i] = se.tgt;
          for (const auto& u : se.iu) niv[u.var] = u.val;
          std::copy(sz, sz + d2, z.begin());
          bool ok = apply_all(z.data(), d, se.guard);
          for (std::size_t t = 0; ok && t < parts.size(); ++t)
            ok = pick[t]->edge >= 0 ? apply_all(z.data(), d, n.edges[pick[t]->edge].guard)
                                    : apply_all(z.data(), d, pick[t]->extra);
          if (!ok) continue;
          for (int r : se.resets) dbm_reset(z.data(), d, r);
          for (std::size_t t = 0; t < parts.size(); ++t) {
            if (pick[t]->edge < 0) continue;
            const Edge& re = n.edges[pick[t]->edge];
            nl[parts[t].first] = re.tgt;
            for (const auto& u : re.iu) niv[u.var] = u.val;
            for (int r : re.resets) dbm_reset(z.data(), d, r);
          }
          if (!E.close_succ_zone(nl.data(), niv.data(), z.data(), false)) continue;
          if (dbm_empty(z.data(), d)) continue;
          emit(nl.data(), niv.data(), z.data(), ei, static_cast<int>(-2 - c));
        }
      }
  }

  int new_node(std::uint64_t key, const raw_t* z) {
    const int id = static_cast<int>(dk.size());
    dk.push_back(key);
    zones.insert(zones.end(), z, z + d2);
    expanded.push_back(0);
    ecnt.push_back(0);
    chains[key].push_back(id);
    return id;
  }

  bool cover_am = true;   // covering relation of the graph being built: a_M, else a_LU

  // first stored node of the same discrete state that covers z, else a new node
  int intern(std::uint64_t key, const int* locs, const raw_t* z, bool* created) {
    *created = false;
    auto it = chains.find(key);
    if (it != chains.end()) {
      std::int32_t L2[64], U2[64];
      if (cover_am) {
        mbounds(locs, L2);
        std::copy(L2, L2 + d, U2);
      } else {
        E.state_bounds(locs, L2, U2);
      }
      for (int m : it->second) {
        ++st.chain_reads;
        if (dbm_alu_le(z, Z(m), d, L2, U2)) { ++st.covered; return m; }
      }
    }
    *created = true;
    return new_node(key, z);
  }

  // ---- stop set ------------------------------------------------------------------------------
  bool time_diverges(int v) const {
    if (n.nclocks == 0) return false;
    int locs[64], ints[64];
    state(v, locs, ints);
    if (has_upper(n.inv, locs)) return false;
    if (frozen(locs, ints)) return false;
    std::vector<raw_t> z(Z(v), Z(v) + d2);
    dbm_up(z.data(), d);
    if (!E.apply_inv(locs, z.data())) return false;
    for (int i = 1; i < d; ++i)
      if (z[i * d] != kInf) return false;
    return true;
  }
  bool orig_time_open(int v) const {
    if (n.nclocks == 0) return false;
    int locs[64], ints[64];
    state(v, locs, ints);
    if (has_upper(orig_inv, locs)) return false;
    if (frozen(locs, ints)) return false;
    std::vector<raw_t> z(d2, kInf);
    for (int j = 0; j < d; ++j) z[j] = kLe0;
    for (int i = 0; i < d; ++i) z[i * d + i] = kLe0;
    if (!apply_inv_table(orig_inv, locs, z.data())) return false;
    for (int i = 1; i < d; ++i)
      if (z[i * d] != kInf) return false;
    return true;
  }

  // keep `live` inclusion-minimal (fed_add)
  void fed_add(Fed& live, const raw_t* g) const {
    const std::size_t m = live.size() / d2;
    for (std::size_t k = 0; k < m; ++k)
      if (dbm_includes(&live[k * d2], g, d)) return;
    Fed kept;
    for (std::size_t k = 0; k < m; ++k)
      if (!dbm_includes(g, &live[k * d2], d)) kept.insert(kept.end(), &live[k * d2], &live[k * d2] + d2);
    kept.insert(kept.end(), g, g + d2);
    live.swap(kept);
  }

  // z minus every zone of `sub`, as a federation
  Fed fed_subtract(const raw_t* z, const Fed& sub) const {
    Fed cur(z, z + d2);
    for (std::size_t c = 0; c < sub.size() / d2 && !cur.empty(); ++c) {
      Fed nxt;
      for (std::size_t k = 0; k < cur.size() / d2; ++k) dbm_subtract(&cur[k * d2], &sub[c * d2], d, nxt);
      cur.swap(nxt);
    }
    return cur;
  }
  bool fed_covers(const Fed& F, const raw_t* z) const { return fed_subtract(z, F).empty(); }

  bool state_has_deadlock(int v, const std::vector<std::vector<Cons>>& inv, Fed* stuck) const {
    int locs[64], ints[64];
    state(v, locs, ints);
    return state_has_deadlock(locs, ints, Z(v), inv, stuck, nullptr);
  }

  // state_has_deadlock (deadlock.cpp:74-268) of the state (locs, ints, zone) judged with invariant
  // table `inv`; `escapes` receives the enabling regions after the time predecessor, when asked
  bool state_has_deadlock(const int* locs, const int* ints, const raw_t* zone,
                          const std::vector<std::vector<Cons>>& inv, Fed* stuck, Fed* escapes) const {
    const bool src_c = [&] { bool c, u; E.frozen_flags(locs, ints, &c, &u); return c; }();
    const bool fz = frozen(locs, ints);
    if (escapes) escapes->clear();
    std::vector<raw_t> Zd(zone, zone + d2);
    if (!fz) dbm_up(Zd.data(), d);
    if (!apply_inv_table(inv, locs, Zd.data())) return false;
    Fed live;
    bool covered = false;
    std::vector<int> tl(n.np);
    // the moving edges of one escape: their resets, and the target location vector
    auto add_src = [&](std::vector<raw_t>& g, const std::vector<int>& moves) {
      std::uint64_t rm = 0;
      std::copy(locs, locs + n.np, tl.begin());
      for (int ed : moves) {
        for (int r : n.edges[ed].resets) rm |= std::uint64_t(1) << r;
        tl[n.edges[ed].proc] = n.edges[ed].tgt;
      }
      for (int pr = 0; pr < n.np; ++pr)
        for (const auto& c : inv[n.gl(pr, tl[pr])]) {
          const int i = ((rm >> c.i) & 1) ? 0 : c.i;
          const int j = ((rm >> c.j) & 1) ? 0 : c.j;
          if (i == j) { if (c.b < kLe0) return; continue; }
          if (!dbm_constrain(g.data(), d, i, j, c.b)) return;
        }
      if (std::memcmp(g.data(), Zd.data(), sizeof(raw_t) * d2) == 0) { covered = true; return; }
      fed_add(live, g.data());
    };
    std::vector<raw_t> g(d2);
    for (int pr = 0; pr < n.np && !covered; ++pr) {
      if (src_c && !E.loc_committed(pr, locs[pr])) continue;
      for (int ei : n.edges_of[n.gl(pr, locs[pr])]) {
        const Edge& e = n.edges[ei];
        if (e.chan >= 0 || !int_ok(e.ig, ints)) continue;
        std::copy(Zd.begin(), Zd.end(), g.begin());
        if (!apply_all(g.data(), d, e.guard)) continue;
        add_src(g, {ei});
        if (covered) break;
      }
    }
    for (int i = 0; i < n.np && !covered; ++i)
      for (int ei : n.edges_of[n.gl(i, locs[i])]) {
        if (covered) break;
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || n.bcast[se.chan]) continue;
        if (!int_ok(se.ig, ints)) continue;
        const bool i_comm = E.loc_committed(i, locs[i]);
        for (int j = 0; j < n.np && !covered; ++j) {
          if (j == i) continue;
          for (int ej : n.edges_of[n.gl(j, locs[j])]) {
            const Edge& re = n.edges[ej];
            if (re.chan != se.chan || re.dir != -1) continue;
            if (src_c && !(i_comm || E.loc_committed(j, locs[j]))) continue;
            if (!int_ok(re.ig, ints)) continue;
            std::copy(Zd.begin(), Zd.end(), g.begin());
            if (!apply_all(g.data(), d, se.guard) || !apply_all(g.data(), d, re.guard)) continue;
            add_src(g, {ei, ej});
            if (covered) break;
          }
        }
      }
    // broadcast escapes (deadlock.cpp:201-235): every combination of the receivers' options
    for (int i = 0; i < n.np && !covered; ++i)
      for (int ei : n.edges_of[n.gl(i, locs[i])]) {
        if (covered) break;
        const Edge& se = n.edges[ei];
        if (se.chan < 0 || se.dir != 1 || !n.bcast[se.chan]) continue;
        if (!int_ok(se.ig, ints)) continue;
        const BParts parts = bparts(locs, ints, ei);
        const long combos = bcombos(parts);
        for (long c = 0; c < combos && !covered; ++c) {
          const auto pick = bpick(parts, c);
          bool part_c = E.loc_committed(i, locs[i]);
          for (std::size_t t = 0; t < parts.size(); ++t)
            if (pick[t]->edge >= 0 && E.loc_committed(parts[t].first, locs[parts[t].first])) part_c = true;
          if (src_c && !part_c) continue;
          std::copy(Zd.begin(), Zd.end(), g.begin());
          bool ok = apply_all(g.data(), d, se.guard);
          std::vector<int> moves{ei};
          for (std::size_t t = 0; ok && t < parts.size(); ++t) {
            if (pick[t]->edge < 0) { ok = apply_all(g.data(), d, pick[t]->extra); continue; }
            ok = apply_all(g.data(), d, n.edges[pick[t]->edge].guard);
            moves.push_back(pick[t]->edge);
          }
          if (ok) add_src(g, moves);
        }
      }
    if (stuck) stuck->clear();
    if (covered) return false;
    const std::size_t m = live.size() / d2;
    if (!fz)
      for (std::size_t k = 0; k < m; ++k) dbm_down(&live[k * d2], d);
    if (escapes) *escapes = live;
    for (std::size_t k = 0; k < m; ++k)
      if (dbm_includes(&live[k * d2], Zd.data(), d)) return false;
    Fed dd = fed_subtract(Zd.data(), live);
    const bool any = !dd.empty();
    if (stuck) stuck->swap(dd);
    return any;
  }

  Fed stop_set(int v) const { return stop_set_at(v, row_off[v + 1] == row_off[v]); }

  // stop_set of node v, told whether v has a successor (leadsto_witness, construction.cpp:786-813,
  // takes it from the successors just generated)
  Fed stop_set_at(int v, bool no_successor) const {
    Fed f;
    if (!restricted) {
      if (no_successor || time_diverges(v)) { f.assign(Z(v), Z(v) + d2); return f; }
      state_has_deadlock(v, n.inv, &f);
      return f;
    }
    if (time_diverges(v)) { f.assign(Z(v), Z(v) + d2); return f; }
    if (orig_time_open(v)) return f;
    state_has_deadlock(v, orig_inv, &f);
    return f;
  }

  // ---- route: can a discrete state seed a deadlock atom? (online_seed_key,
  // src/liveness/refinement.cpp:176-282) -----------------------------------------------------------
  // D refines a_LU only at discrete states whose deadlock region is partial: a deadlocked valuation
  // exists and some escape meets the invariant. Judged on the universe zone in the unrestricted
  // network; a non-frozen state without an upper invariant, or whose invariant lets every clock
  // grow, cannot seed one except for `A[] not deadlock`.
  bool key_seeds(std::uint64_t key) const {
    int locs[64], ints[64];
    E.decode(key, locs, ints);
    const auto& inv = restricted ? orig_inv : n.inv;
    const bool fz = frozen(locs, ints);
    const bool dl = kind == HORA_LIVE_DEADLOCK;
    if (!fz && !dl && !has_upper(inv, locs)) return false;
    std::vector<raw_t> u(d2, kInf);
    for (int j = 0; j < d; ++j) u[j] = kLe0;
    for (int i = 0; i < d; ++i) u[i * d + i] = kLe0;
    std::vector<raw_t> iz(u);
    if (!apply_inv_table(inv, locs, iz.data())) return false;
    if (!fz && !dl) {
      bool open = true;
      for (int i = 1; i < d; ++i)
        if (iz[i * d] != kInf) open = false;
      if (open) return false;
    }
    Fed stuck, esc;
    if (!state_has_deadlock(locs, ints, u.data(), inv, &stuck, &esc)) return false;
    std::vector<raw_t> t(d2);
    for (std::size_t k = 0; k < esc.size() / d2; ++k)
      if (dbm_intersect(t.data(), &esc[k * d2], iz.data(), d)) return true;
    return false;                              // Dead == Inv: no distinguishing atom
  }

  // ---- exact predecessor (operators.cpp:257-306) ----------------------------------------------
  bool pre(int v, std::int64_t e, const raw_t* X, raw_t* Y) const {
    const int t = etgt[e], se = ese[e], re = ere[e];
    int sl[64], si[64], tl[64], ti[64];
    state(v, sl, si);
    state(t, tl, ti);
    std::copy(X, X + d2, Y);
    dbm_free_many(Y, d, dead_mask(tl));
    if (!E.apply_inv(tl, Y)) return false;
    if (!frozen(tl, ti)) {
      dbm_down(Y, d);
      if (!E.apply_inv(tl, Y)) return false;
    }
    // the participating edges, and a broadcast's stay-out constraints (operators.cpp:283-300)
    std::vector<int> parts{se};
    std::vector<Cons> abstain;
    if (re >= 0) parts.push_back(re);
    if (re <= -2) {
      const BParts bp = bparts(sl, si, se);
      const auto pick = bpick(bp, -2L - re);
      for (const BOpt* o : pick) {
        if (o->edge >= 0) parts.push_back(o->edge);
        else abstain.insert(abstain.end(), o->extra.begin(), o->extra.end());
      }
    }
    std::uint64_t rm = 0;
    for (int ed : parts)
      for (int r : n.edges[ed].resets) {
        if (!dbm_constrain(Y, d, r, 0, kLe0)) return false;
        if (!dbm_constrain(Y, d, 0, r, kLe0)) return false;
        rm |= std::uint64_t(1) << r;
      }
    dbm_free_many(Y, d, rm);
    for (int ed : parts)
      if (!apply_all(Y, d, n.edges[ed].guard)) return false;
    if (!apply_all(Y, d, abstain)) return false;
    if (!frozen(sl, si)) dbm_down(Y, d);
    if (!E.apply_inv(sl, Y)) return false;
    return true;
  }

  // drop members contained in another (for equal members keep the first)
  void compact(Fed& F) {
    const std::size_t m = F.size() / d2;
    std::vector<char> drop(m, 0);
    for (std::size_t k = 0; k < m; ++k)
      for (std::size_t j = 0; j < m && !drop[k]; ++j) {
        if (j == k) continue;
        if (!dbm_includes(&F[j * d2], &F[k * d2], d)) continue;
        if (!dbm_includes(&F[k * d2], &F[j * d2], d) || j < k) drop[k] = 1;
      }
    Fed out;
    for (std::size_t k = 0; k < m; ++k)
      if (!drop[k]) out.insert(out.end(), &F[k * d2], &F[k * d2] + d2);
    F.swap(out);
    if (F.size() / d2 > fedcap) st.capped = 1;
  }

  // ---- construction -----------------------------------------------------------------------------
  // Returns 1/0 when an early witness settled the query (value = holds), -1 otherwise.
  int build() {
    std::vector<int> l0(n.np), i0(n.ni);
    for (int pr = 0; pr < n.np; ++pr) l0[pr] = n.loc_init[pr];
    for (int v = 0; v < n.ni; ++v) i0[v] = n.int_init[v];
    std::vector<raw_t> z0(d2);
    dbm_zero(z0.data(), d);
    if (!E.close_succ_zone(l0.data(), i0.data(), z0.data(), false) || dbm_empty(z0.data(), d)) return -1;
    new_node(E.encode(l0.data(), i0.data()), z0.data());
    std::vector<int> frontier;
    if (kind != HORA_LIVE_EG || psi(0)) frontier.push_back(0);
    std::vector<int> sl(n.np), si(n.ni);
    while (!frontier.empty()) {
      ++st.levels;
      std::vector<int> next;
      for (int v : frontier) {
        expanded[v] = 1;
        ++st.expanded_nodes;
        state(v, sl.data(), si.data());
        const std::vector<raw_t> zv(Z(v), Z(v) + d2);
        succ_rec(sl.data(), si.data(), zv.data(), [&](const int* nl, const int* niv, const raw_t* z, int se, int re) {
          ++st.generated;
          bool created = false;
          const int t = intern(E.encode(nl, niv), nl, z, &created);
          if (created && (kind != HORA_LIVE_EG || psi(t))) next.push_back(t);
          etgt.push_back(t);
          ese.push_back(se);
          ere.push_back(re);
          ++ecnt[v];
        });
      }
      build_row_off();
      if (early_on && kind == HORA_LIVE_LEADSTO) {
        for (int v : frontier)
          if (is_p(v) && !is_q(v) && !stop_set(v).empty()) { st.early = 1; return 0; }
      }
      if (early_on && kind == HORA_LIVE_DEADLOCK) {
        for (int v : frontier)
          if (state_has_deadlock(v, n.inv, nullptr)) { st.early = 1; return 0; }
      }
      if (early_on && kind == HORA_LIVE_EG) {   // online_eg_terminal_witness, construction.cpp:748-767
        for (int v : frontier)
          if (!stop_set(v).empty()) { st.early = 1; return 1; }   // a maximal psi-run ends at v
      }
      frontier.swap(next);
    }
    return -1;
  }
  void build_row_off() {
    row_off.assign(dk.size() + 1, 0);
    for (std::size_t v = 0; v < dk.size(); ++v) row_off[v + 1] = row_off[v] + ecnt[v];
  }

  // ---- the online construction (a_LU lane, no refinement atom) -------------------------------
  // LivenessEngine::online_base_forward / online_drain_forward_work (refinement.cpp:935-1065),
  // online_expand (:610-781) and online_intern (construction.cpp:480-683) with the map D empty,
  // which is what structurally_empty_d / the seed test establish for the networks this lane runs:
  //   * a FIFO worklist of active records in activation (= creation) order; a record retired
  //     before its turn is skipped (refinement.cpp:1013-1014);
  //   * each successor of the record being expanded, in tagged_successors order, goes to the
  //     NEWEST active record of its key that a_LU-covers it (the active list is prepended on
  //     activation, construction.cpp:385-396), whether or not that record includes it
  //     (construction.cpp:515-535: inclusion, or a revocable cover with no atom to revoke it);
  //   * otherwise the successor is joined with every active record but the root whose union with
  //     it is one zone, in newest-first passes repeated until a pass joins nothing
  //     (construction.cpp:607-640), and the result is appended as a new active record
  //     (online_append, :309-327, the count `outer_nodes`); every other active record but the
  //     root that it a_LU-covers is retired (online_compact_key, :426-477).
  // Not restated: the in-place enlargement the reference applies under A[] not deadlock only
  // (construction.cpp:590-605: an active unexpanded record that the successor includes takes the
  // successor's zone). It changes which records are stored, not whether a deadlocked valuation is
  // reachable; the recorded deadlock counts come from a route this restatement does not reproduce.
  // A retired record keeps its zone and, if it was expanded, its edges; an edge to a record that
  // was never expanded is redirected along the retirements to the active record that covers it
  // (online_finalize_graph, refinement.cpp:1071-1160, online_resolve_cover, construction.cpp:359-378).
  std::vector<char> active;
  std::vector<int> retired_by;
  std::vector<int> depth;
  std::unordered_map<std::uint64_t, std::vector<int>> act;   // key -> its active records, oldest first
  bool base_route = false;

  void key_bounds(std::uint64_t key, std::int32_t* L2, std::int32_t* U2) const {
    int locs[64], ints[64];
    E.decode(key, locs, ints);
    E.state_bounds(locs, L2, U2);
  }

  void retire_covered(int v) {   // online_compact_key
    std::int32_t L2[64], U2[64];
    key_bounds(dk[v], L2, U2);
    auto& A = act[dk[v]];
    std::vector<int> keep;
    keep.reserve(A.size());
    for (int j : A) {
      if (j != 0 && j != v && dbm_alu_le(Z(j), Z(v), d, L2, U2)) {
        active[j] = 0;
        retired_by[j] = v;
        ++st.retired;
        continue;
      }
      keep.push_back(j);
    }
    A.swap(keep);
  }

  int online_intern(std::uint64_t key, std::vector<raw_t>& s, int dep, std::deque<int>& work) {
    std::int32_t L2[64], U2[64];
    key_bounds(key, L2, U2);
    {
      auto& A = act[key];
      for (auto it = A.rbegin(); it != A.rend(); ++it) {
        ++st.chain_reads;
        if (dbm_alu_le(s.data(), Z(*it), d, L2, U2)) { ++st.covered; return *it; }
      }
      for (bool grew = true; grew;) {
        grew = false;
        for (auto it = A.rbegin(); it != A.rend(); ++it) {
          const int j = *it;
          if (j == 0 || dbm_includes(s.data(), Z(j), d)) continue;
          if (dbm_try_union(s.data(), Z(j), d)) { grew = true; ++st.unions; }
        }
      }
    }
    const int id = new_node(key, s.data());
    active.push_back(1);
    retired_by.push_back(-1);
    depth.push_back(dep);
    act[key].push_back(id);
    work.push_back(id);
    retire_covered(id);
    return id;
  }

  // ---- the base construction (networks where D is provably a_LU) ------------------------------
  // LivenessEngine::base_forward (api.cpp:37-117), expand (construction.cpp:1034-1096) and intern
  // (:863-1003), which decide_leads_to takes when structurally_empty_d holds (api.cpp:392):
  // a FIFO worklist of unexpanded nodes; each successor goes to the OLDEST node of its key that
  // a_LU-covers it, else the oldest node it a_LU-covers is overwritten with it in place, else the
  // oldest node whose union with it is one zone absorbs it, else it is appended. A node rewritten
  // by a successor of another node loses its edges and is expanded again.
  std::vector<std::vector<std::array<int, 3>>> adj;

  int base_intern(std::uint64_t key, std::vector<raw_t>& s, bool* grew) {
    std::int32_t L2[64], U2[64];
    key_bounds(key, L2, U2);
    *grew = false;
    auto it = chains.find(key);
    if (it != chains.end()) {
      const std::vector<int>& b = it->second;
      for (int j : b) {
        ++st.chain_reads;
        if (dbm_alu_le(s.data(), Z(j), d, L2, U2)) { ++st.covered; return j; }
      }
      for (int j : b)
        if (dbm_alu_le(Z(j), s.data(), d, L2, U2)) {
          std::copy(s.begin(), s.end(), zones.begin() + static_cast<std::ptrdiff_t>(j) * d2);
          ++st.enlarged;
          *grew = true;
          return j;
        }
      std::vector<raw_t> t(d2);
      for (int j : b) {
        std::copy(Z(j), Z(j) + d2, t.begin());
        if (dbm_try_union(t.data(), s.data(), d)) {
          std::copy(t.begin(), t.end(), zones.begin() + static_cast<std::ptrdiff_t>(j) * d2);
          ++st.unions;
          *grew = true;
          return j;
        }
      }
    }
    const int id = new_node(key, s.data());
    adj.emplace_back();
    return id;
  }

  int build_base() {
    std::vector<int> l0(n.np), i0(n.ni);
    for (int pr = 0; pr < n.np; ++pr) l0[pr] = n.loc_init[pr];
    for (int v = 0; v < n.ni; ++v) i0[v] = n.int_init[v];
    std::vector<raw_t> z0(d2);
    dbm_zero(z0.data(), d);
    if (!E.close_succ_zone(l0.data(), i0.data(), z0.data(), false) || dbm_empty(z0.data(), d)) return -1;
    adj.clear();
    new_node(E.encode(l0.data(), i0.data()), z0.data());
    adj.emplace_back();
    std::deque<int> work{0};
    std::vector<char> inq{1};
    std::vector<int> sl(n.np), si(n.ni);
    struct Succ { std::uint64_t key; std::vector<raw_t> z; int se, re; };
    std::vector<Succ> succ;
    int result = -1;
    while (!work.empty()) {
      const int u = work.front();
      work.pop_front();
      inq[u] = 0;
      if (!expanded[u]) {
        expanded[u] = 1;
        ++st.expanded_nodes;
        state(u, sl.data(), si.data());
        succ.clear();
        {
          const std::vector<raw_t> zu(Z(u), Z(u) + d2);
          succ_rec(sl.data(), si.data(), zu.data(), [&](const int* nl, const int* niv, const raw_t* z, int se, int re) {
            succ.push_back({E.encode(nl, niv), std::vector<raw_t>(z, z + d2), se, re});
          });
        }
        if (early_on) {
          const bool none = succ.empty();
          bool hit = false;
          if (kind == HORA_LIVE_DEADLOCK) hit = state_has_deadlock(u, n.inv, nullptr);
          else if (kind == HORA_LIVE_LEADSTO) hit = is_p(u) && !is_q(u) && !stop_set_at(u, none).empty();
          if (hit) {
            st.early = 1;
            result = 0;
            break;
          }
        }
        for (auto& s : succ) {
          ++st.generated;
          bool grew = false;
          const int v = base_intern(s.key, s.z, &grew);
          adj[u].push_back({v, s.se, s.re});
          if (grew) {
            if (v != u) {
              expanded[v] = 0;
              adj[v].clear();
            } else {
              expanded[u] = 0;
            }
          }
        }
      }
      inq.resize(dk.size(), 0);
      for (const auto& e : adj[u])
        if (!expanded[e[0]] && !inq[e[0]]) { inq[e[0]] = 1; work.push_back(e[0]); }
    }
    // the edges in node order
    etgt.clear(); ese.clear(); ere.clear();
    for (std::size_t v = 0; v < dk.size(); ++v) {
      ecnt[v] = static_cast<std::int64_t>(adj[v].size());
      for (const auto& e : adj[v]) { etgt.push_back(e[0]); ese.push_back(e[1]); ere.push_back(e[2]); }
    }
    return result;
  }

  // structurally_empty_d (src/liveness/internal.cpp:59-103): no urgency and no upper invariant, or
  // a single possible discrete state whose universe seed has no partial deadlock region
  bool structurally_empty() const {
    bool time_open = !n.has_urgent_chan, one_key = true;
    for (int L = 0; L < n.tl; ++L) {
      if (n.loc_flags[L] & 3) time_open = false;
      for (const auto& c : n.inv[L])
        if (c.i > 0 && c.j == 0 && c.b < kInf) time_open = false;
    }
    for (int pr = 0; pr < n.np; ++pr)
      if (n.locbase[pr + 1] - n.locbase[pr] != 1) one_key = false;
    for (int v = 0; v < n.ni; ++v)
      if (n.int_lo[v] != n.int_hi[v]) one_key = false;
    if (time_open) return true;
    if (!one_key) return false;
    std::vector<int> l0(n.np), i0(n.ni);
    for (int pr = 0; pr < n.np; ++pr) l0[pr] = n.loc_init[pr];
    for (int v = 0; v < n.ni; ++v) i0[v] = n.int_init[v];
    std::vector<raw_t> u(d2, kInf);
    for (int j = 0; j < d; ++j) u[j] = kLe0;
    for (int i = 0; i < d; ++i) u[i * d + i] = kLe0;
    std::vector<raw_t> iz(u);
    if (!apply_inv_table(n.inv, l0.data(), iz.data())) return true;
    Fed stuck, esc;
    if (!state_has_deadlock(l0.data(), i0.data(), u.data(), n.inv, &stuck, &esc)) return true;
    if (!frozen(l0.data(), i0.data())) {
      bool open = true;
      for (int i = 1; i < d; ++i)
        if (iz[i * d] != kInf) open = false;
      if (open) return true;
    }
    std::vector<raw_t> t(d2);
    for (std::size_t k = 0; k < esc.size() / d2; ++k)
      if (dbm_intersect(t.data(), &esc[k * d2], iz.data(), d)) return false;
    return true;
  }

  // Returns 1/0 when an early witness settled the query (value = holds), -1 otherwise.
  int build_online() {
    std::vector<int> l0(n.np), i0(n.ni);
    for (int pr = 0; pr < n.np; ++pr) l0[pr] = n.loc_init[pr];
    for (int v = 0; v < n.ni; ++v) i0[v] = n.int_init[v];
    std::vector<raw_t> z0(d2);
    dbm_zero(z0.data(), d);
    if (!E.close_succ_zone(l0.data(), i0.data(), z0.data(), false) || dbm_empty(z0.data(), d)) return -1;
    std::deque<int> work;
    active.clear(); retired_by.clear(); depth.clear(); act.clear();
    {
      const std::uint64_t k0 = E.encode(l0.data(), i0.data());
      new_node(k0, z0.data());
      active.push_back(1);
      retired_by.push_back(-1);
      depth.push_back(0);
      act[k0].push_back(0);
      work.push_back(0);
    }
    struct Succ { std::uint64_t key; std::vector<raw_t> z; int se, re; };
    std::vector<Succ> succ;
    std::vector<int> sl(n.np), si(n.ni);
    int result = -1, max_depth = -1;
    while (!work.empty()) {
      const int u = work.front();
      work.pop_front();
      if (!active[u] || expanded[u]) continue;
      if (kind == HORA_LIVE_EG && !psi(u)) continue;   // outside psi: stored, never expanded
      expanded[u] = 1;
      ++st.expanded_nodes;
      max_depth = std::max(max_depth, depth[u]);
      state(u, sl.data(), si.data());
      succ.clear();
      {
        const std::vector<raw_t> zu(Z(u), Z(u) + d2);
        succ_rec(sl.data(), si.data(), zu.data(), [&](const int* nl, const int* niv, const raw_t* z, int se, int re) {
          succ.push_back({E.encode(nl, niv), std::vector<raw_t>(z, z + d2), se, re});
        });
      }
      // early witnesses at the record being expanded (online_expand, refinement.cpp:634-654)
      if (early_on) {
        const bool none = succ.empty();
        bool hit = false;
        if (kind == HORA_LIVE_DEADLOCK) hit = state_has_deadlock(u, n.inv, nullptr);
        else if (kind == HORA_LIVE_LEADSTO) hit = is_p(u) && !is_q(u) && !stop_set_at(u, none).empty();
        else hit = !stop_set_at(u, none).empty();
        if (hit) {
          st.early = 1;
          result = kind == HORA_LIVE_EG ? 1 : 0;
          break;
        }
      }
      for (auto& s : succ) {
        ++st.generated;
        const int t = online_intern(s.key, s.z, depth[u] + 1, work);
        etgt.push_back(t);
        ese.push_back(s.se);
        ere.push_back(s.re);
        ++ecnt[u];
      }
    }
    st.levels = max_depth + 1;
    // edges to records never expanded go to the active record that covers them
    for (auto& t : etgt) {
      int r = t;
      while (!expanded[r] && retired_by[r] >= 0) r = retired_by[r];
      t = r;
    }
    return result;
  }

  // ---- backward pass ------------------------------------------------------------------------------
  std::vector<int> closure() {
    const int N = static_cast<int>(dk.size());
    in_s.assign(N, 0);
    std::vector<int> stk, out;
    auto nq = [&](int v) { return kind == HORA_LIVE_EG ? psi(v) : !is_q(v); };
    if (kind == HORA_LIVE_LEADSTO) {
      for (int v = 0; v < N; ++v)
        if (expanded[v] && nq(v) && is_p(v)) { in_s[v] = 1; stk.push_back(v); }
    } else if (N > 0 && expanded[0] && nq(0)) {
      in_s[0] = 1;
      stk.push_back(0);
    }
    while (!stk.empty()) {
      const int v = stk.back();
      stk.pop_back();
      out.push_back(v);
      for (std::int64_t e = row_off[v]; e < row_off[v + 1]; ++e) {
        const int t = etgt[e];
        if (!in_s[t] && expanded[t] && nq(t)) { in_s[t] = 1; stk.push_back(t); }
      }
    }
    std::sort(out.begin(), out.end());
    return out;
  }

  std::vector<Fed> W;
  std::vector<Fed> stops;
  std::vector<char> stop_ready;

  const Fed& stop_of(int v) {
    if (!stop_ready[v]) { stops[v] = stop_set(v); stop_ready[v] = 1; }
    return stops[v];
  }

  // relax (fixpoint.cpp:750-864): W[v] := W[v] /\ (NoSucc(v) \/ U_e Pre_e(W[t])). True if changed.
  bool relax(int v) {
    Fed& wv = W[v];
    const std::size_t nmem = wv.size() / d2;
    if (nmem == 0) return false;
    std::vector<char> cov(nmem, 0);
    std::size_t ncov = 0;
    Fed cand;
    std::vector<raw_t> pz(d2);
    for (std::int64_t e = row_off[v]; e < row_off[v + 1]; ++e) {
      const int t = etgt[e];
      if (!in_s[t]) continue;
      const Fed& wt = W[t];
      for (std::size_t y = 0; y < wt.size() / d2; ++y) {
        ++st.pre_calls;
        if (!pre(v, e, &wt[y * d2], pz.data())) continue;
        for (std::size_t k = 0; k < nmem; ++k)
          if (!cov[k] && dbm_includes(pz.data(), &wv[k * d2], d)) { cov[k] = 1; ++ncov; }
        if (ncov == nmem) return false;
        cand.insert(cand.end(), pz.begin(), pz.end());
      }
    }
    compact(cand);
    auto escapes = [&] {
      for (std::size_t k = 0; k < nmem; ++k) {
        bool whole = false;
        for (std::size_t c = 0; c < cand.size() / d2 && !whole; ++c)
          if (dbm_includes(&cand[c * d2], &wv[k * d2], d)) whole = true;
        if (!whole && !fed_covers(cand, &wv[k * d2])) return true;
      }
      return false;
    };
    if (!escapes()) return false;
    const Fed& ns = stop_of(v);
    if (!ns.empty()) {
      cand.insert(cand.end(), ns.begin(), ns.end());
      compact(cand);
      if (!escapes()) return false;
    }
    Fed nw;
    std::vector<raw_t> t(d2);
    for (std::size_t k = 0; k < nmem; ++k) {
      const raw_t* z = &wv[k * d2];
      bool whole = false;
      for (std::size_t c = 0; c < cand.size() / d2 && !whole; ++c)
        if (dbm_includes(&cand[c * d2], z, d)) whole = true;
      if (whole) { nw.insert(nw.end(), z, z + d2); continue; }
      for (std::size_t c = 0; c < cand.size() / d2; ++c)
        if (dbm_intersect(t.data(), z, &cand[c * d2], d)) { ++st.products; nw.insert(nw.end(), t.begin(), t.end()); }
    }
    compact(nw);
    wv.swap(nw);
    return true;
  }

  void solve(const std::vector<int>& S) {
    const int N = static_cast<int>(dk.size());
    W.assign(N, {});
    stops.assign(N, {});
    stop_ready.assign(N, 0);
    std::vector<std::vector<int>> preds(N);
    for (int v : S) {
      W[v].assign(Z(v), Z(v) + d2);
      for (std::int64_t e = row_off[v]; e < row_off[v + 1]; ++e)
        if (in_s[etgt[e]]) preds[etgt[e]].push_back(v);
    }
    std::deque<int> work(S.begin(), S.end());
    std::vector<char> inq(N, 0);
    for (int v : S) inq[v] = 1;
    while (!work.empty() && !st.capped) {
      const int v = work.front();
      work.pop_front();
      inq[v] = 0;
      if (!relax(v)) continue;
      ++st.iters;
      for (int u : preds[v])
        if (!inq[u]) { inq[u] = 1; work.push_back(u); }
    }
    wne.assign(N, 0);
    for (int v : S) {
      const std::int64_t m = static_cast<std::int64_t>(W[v].size() / d2);
      if (m) { wne[v] = 1; ++st.nonempty; st.fed_max = std::max(st.fed_max, m); }
    }
  }

  void reset_graph() {
    dk.clear(); zones.clear(); row_off.clear(); etgt.clear(); ese.clear(); ere.clear();
    expanded.clear(); in_s.clear(); wne.clear(); chains.clear(); ecnt.clear();
    const std::int32_t seeds = st.seed_keys;
    st = hora_live_stats{};
    st.seed_keys = seeds;
  }

  int run(int32_t* holds, int cover) {
    const auto t0 = std::chrono::steady_clock::now();
    cover_am = cover == HORA_COVER_AM;
    base_route = cover == HORA_COVER_AUTO && kind == HORA_LIVE_LEADSTO && structurally_empty();
    int er = cover_am ? build() : base_route ? build_base() : build_online();
    if (er < 0 && !cover_am) {
      // the online graph stands unless one of its discrete states can seed a deadlock atom
      std::int32_t seeds = 0;
      for (const auto& kv : chains) seeds += key_seeds(kv.first) ? 1 : 0;
      st.seed_keys = seeds;
      if (seeds > 0) {
        reset_graph();
        cover_am = true;
        er = build();
      }
    }
    st.route = cover_am ? 1 : 0;
    build_row_off();
    in_s.assign(dk.size(), 0);
    wne.assign(dk.size(), 0);
    const auto t1 = std::chrono::steady_clock::now();
    st.outer_nodes = static_cast<std::int64_t>(dk.size());
    st.fwd_edges = static_cast<std::int64_t>(etgt.size());
    st.disc_keys = static_cast<std::int64_t>(chains.size());
    if (er >= 0) {
      *holds = er;
    } else if (kind == HORA_LIVE_DEADLOCK) {
      bool dl = false;
      for (std::size_t v = 0; v < dk.size() && !dl; ++v)
        if (expanded[v] && state_has_deadlock(static_cast<int>(v), n.inv, nullptr)) dl = true;
      *holds = dl ? 0 : 1;
    } else {
      const std::vector<int> S = closure();
      st.closure = static_cast<std::int64_t>(S.size());
      solve(S);
      if (kind == HORA_LIVE_LEADSTO) {
        bool violated = false;
        for (int v : S)
          if (wne[v] && is_p(v)) { violated = true; break; }
        *holds = violated ? 0 : 1;
      } else {
        *holds = (!dk.empty() && in_s[0] && wne[0]) ? 1 : 0;
      }
    }
    const auto t2 = std::chrono::steady_clock::now();
    st.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    st.gfp_seconds = std::chrono::duration<double>(t2 - t1).count();
    st.seconds = std::chrono::duration<double>(t2 - t0).count();
    return st.capped ? HORA_HIP_E_CAPPED : 0;
  }
};

struct Held {
  std::unique_ptr<Engine> E;
  std::unique_ptr<Live> L;
};

}  // namespace live
}  // namespace orc

extern "C" {

const char* oracle_live_last_error(void) { return orc::live::g_err.c_str(); }

// One liveness query on the CPU. *graph (optional) receives a handle for oracle_live_graph /
// oracle_live_free.
int oracle_live(const hora_net_enc* e, const hora_live_query* q, int32_t* holds, hora_live_stats* st,
                void** graph) {
  using namespace orc::live;
  try {
    if (!e || !q || !holds) return HORA_HIP_E_ARG;
    auto h = std::make_unique<Held>();
    hora_target never{};
    never.never = 1;
    h->E = std::make_unique<orc::Engine>(e, &never);
    if (h->E->n.has_diag) { g_err = "diagonal constraints"; return HORA_HIP_E_SCOPE; }
    h->L = std::make_unique<Live>(*h->E, q);
    const int rc = h->L->run(holds, q->cover);
    if (st) *st = h->L->st;
    if (graph) *graph = h.release();
    return rc;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

int oracle_live_graph(void* handle, hora_live_graph* g) {
  auto* h = static_cast<orc::live::Held*>(handle);
  if (!h || !g) return HORA_HIP_E_ARG;
  const auto& L = *h->L;
  g->nodes = static_cast<int64_t>(L.dk.size());
  g->edges = static_cast<int64_t>(L.etgt.size());
  g->d = L.d;
  auto cp = [](auto* dst, const auto& v) { if (dst) std::copy(v.begin(), v.end(), dst); };
  cp(g->dk, L.dk);
  cp(g->zones, L.zones);
  cp(g->row_off, L.row_off);
  cp(g->etgt, L.etgt);
  cp(g->ese, L.ese);
  cp(g->ere, L.ere);
  cp(g->expanded, L.expanded);
  cp(g->in_s, L.in_s);
  cp(g->w_nonempty, L.wne);
  return 0;
}

void oracle_live_free(void* handle) { delete static_cast<orc::live::Held*>(handle); }

// p \ y for n pairs of zones: out receives at most `cap` pieces per pair (cap*d*d int32 each),
// counts[k] the number of pieces (or -1 if more than cap).
int oracle_zone_subtract(int32_t d, int64_t n, const int32_t* p, const int32_t* y, int32_t cap, int32_t* out,
                         int32_t* counts) {
  const int d2 = d * d;
  for (int64_t k = 0; k < n; ++k) {
    std::vector<orc::raw_t> pieces;
    orc::dbm_subtract(p + k * d2, y + k * d2, d, pieces);
    const int m = static_cast<int>(pieces.size() / d2);
    counts[k] = m > cap ? -1 : m;
    if (m <= cap) std::copy(pieces.begin(), pieces.end(), out + k * static_cast<int64_t>(cap) * d2);
  }
  return 0;
}

}  // extern "C"
