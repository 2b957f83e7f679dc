// oracle_dbm.hpp — CPU restatement of the DBM zone algebra HORA takes from HDBM.
//
// TEST INFRASTRUCTURE ONLY. This file is part of the oracle: the CPU checker the HIP back end is
// compared against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// build, load or call it. The product path never routes through it.
//
// HDBM itself is absent from the reference snapshot (an un-vendored submodule, HDBM/ is empty), so
// each operation below restates the in-tree hand-written ports of it, cited per function, and the
// textbook definitions where no port exists:
//   Bengtsson & Yi, "Timed Automata: Semantics, Algorithms and Tools" (2004): close, up, reset,
//     constrain, inclusion;
//   Behrmann, Bouyer, Larsen, Pelanek, STTT 8(3) 2006: Extra_LU+;
//   Herbreteau, Srivathsan, Walukiewicz, LICS 2012: the a_LU inclusion test.
//
// Pinning (tests/test_oracle_dbm.py): canonical-form uniqueness makes close / up / reset /
// constrain / includes / try_union answers mathematically determined, and each is checked against
// a region-grid brute force on small dimensions plus the reference's own known answers
// (tests/dbm_tests.cpp:24-75). a_LU is checked against its semantic definition by brute force on
// tiny grids. Extra_LU+ is checked against the STTT 2006 case definition on a point grid
// (parity pinned against the published definition only: HDBM's exact edge cases are not in the
// snapshot).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

using raw_t = std::int32_t;
constexpr raw_t kInf = 1 << 30;
constexpr raw_t kLe0 = 1;
constexpr std::int32_t kAluNegInf = -(1 << 29);

// Encoding (src/metal/netengine_metal.mm:76-85).
inline raw_t le(int v) { return v * 2 + 1; }
inline raw_t lt(int v) { return v * 2; }
inline int bnd(raw_t r) { return r >> 1; }
inline bool is_strict(raw_t r) { return (r & 1) == 0; }
// Saturating bound addition: infinity absorbs, weak iff both weak (netengine_metal.mm:86-91).
inline raw_t radd(raw_t a, raw_t b) {
  if (a >= kInf || b >= kInf) return kInf;
  const int v = bnd(a) + bnd(b);
  return (!is_strict(a) && !is_strict(b)) ? le(v) : lt(v);
}

// The zero zone: every clock equal to 0 (HDBM Dbm(n); tests/dbm_tests.cpp:15-22).
inline void dbm_zero(raw_t* z, int d) {
  for (int k = 0; k < d * d; ++k) z[k] = kLe0;
}

// Any negative diagonal entry (netengine_metal.mm:101-104).
inline bool dbm_empty(const raw_t* z, int d) {
  for (int i = 0; i < d; ++i)
    if (z[i * d + i] < kLe0) return true;
  return false;
}

// Delay / up: drop every upper bound (netengine_metal.mm:105).
inline void dbm_up(raw_t* z, int d) {
  for (int i = 1; i < d; ++i) z[i * d] = kInf;
}

// x_k := 0 (netengine_metal.mm:106-109).
inline void dbm_reset(raw_t* z, int d, int k) {
  for (int j = 0; j < d; ++j) {
    z[k * d + j] = z[j];
    z[j * d + k] = z[j * d];
  }
  z[k * d + k] = kLe0;
}

// Floyd-Warshall closure (netengine_metal.mm:225-233). Returns false if the result is empty.
inline bool dbm_close(raw_t* z, int d) {
  for (int k = 0; k < d; ++k)
    for (int i = 0; i < d; ++i) {
      const raw_t ik = z[i * d + k];
      if (ik >= kInf) continue;
      for (int j = 0; j < d; ++j) {
        const raw_t s = radd(ik, z[k * d + j]);
        if (s < z[i * d + j]) z[i * d + j] = s;
      }
    }
  return !dbm_empty(z, d);
}

// Tighten x_a - x_b to r and restore canonical form incrementally (netengine_metal.mm:110-124).
// Returns false iff the result is empty. The zone must be canonical and non-empty on entry.
inline bool dbm_constrain(raw_t* z, int d, int a, int b, raw_t r) {
  if (r >= z[a * d + b]) return !dbm_empty(z, d);
  z[a * d + b] = r;
  for (int i = 0; i < d; ++i) {
    const raw_t via = radd(z[i * d + a], r);
    if (via < z[i * d + b]) z[i * d + b] = via;
  }
  for (int i = 0; i < d; ++i) {
    const raw_t ib = z[i * d + b];
    if (ib >= kInf) continue;
    for (int k = 0; k < d; ++k) {
      const raw_t via = radd(ib, z[b * d + k]);
      if (via < z[i * d + k]) z[i * d + k] = via;
    }
  }
  return !dbm_empty(z, d);
}

// Does a include b? Entrywise b <= a (netengine_metal.mm:235-238).
inline bool dbm_includes(const raw_t* a, const raw_t* b, int d) {
  for (int k = 0; k < d * d; ++k)
    if (b[k] > a[k]) return false;
  return true;
}

// Z <=_aLU Z' (netengine_metal.mm:240-252; Herbreteau et al. LICS 2012 Thm. 3): Z is NOT
// simulated iff some x, y (x != y) have Z[0][x] >= (<=, -U_x), Z'[y][x] < Z[y][x] and
// Z'[y][x] + (<, -L_y) < Z[0][x]. Clocks whose bound is kAluNegInf take no part.
inline bool dbm_alu_le(const raw_t* Z, const raw_t* Zp, int d, const std::int32_t* L,
                       const std::int32_t* U) {
  for (int x = 0; x < d; ++x) {
    if (U[x] == kAluNegInf) continue;
    if (Z[x] < le(-U[x])) continue;
    for (int y = 0; y < d; ++y) {
      if (x == y || L[y] == kAluNegInf) continue;
      if (Zp[y * d + x] < Z[y * d + x] && radd(Zp[y * d + x], lt(-L[y])) < Z[x]) return false;
    }
  }
  return true;
}

// Extra_LU+ then close (src/metal/engine_metal.mm:80-93; Behrmann et al. STTT 2006, Def. 6):
// with row 0 taken before any change, for i != j and z[i][j] finite:
//   z[i][j] > L_i           -> inf
//   -z[0][i] > L_i          -> inf
//   i != 0, -z[0][j] > U_j  -> inf
//   i == 0, -z[0][j] > U_j  -> (<, -U_j)
inline void dbm_extra_lu_plus(raw_t* z, int d, const std::int32_t* L, const std::int32_t* U) {
  std::vector<raw_t> row0(z, z + d);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      if (i == j) continue;
      const raw_t e = z[i * d + j];
      if (e >= kInf) continue;
      if (bnd(e) > L[i]) { z[i * d + j] = kInf; continue; }
      if (-bnd(row0[i]) > L[i]) { z[i * d + j] = kInf; continue; }
      if (i != 0 && -bnd(row0[j]) > U[j]) { z[i * d + j] = kInf; continue; }
      if (e < lt(-U[j])) z[i * d + j] = lt(-U[j]);
    }
  dbm_close(z, d);
}

// Is a U b exactly convex? If so a := hull and return true (netengine_metal.mm:310-335;
// src/cuda/netengine_cuda.cu:121-143). The hull H is the entrywise max (canonical when a and b
// are). a U b == H iff H \ a is inside b and H \ b inside a; H \ z is the union over z's facets
// (i, j) with z[i][j] < H[i][j] of H conjoined with the facet's negation. This restatement tests
// every such facet; the reference prunes to the minimal-graph facets first, which removes only
// redundant tests and never changes the answer (a redundant facet's crossing region lies inside
// the union of the irredundant ones').
inline bool dbm_try_union(raw_t* a, const raw_t* b, int d) {
  const int d2 = d * d;
  std::vector<raw_t> H(d2), R(d2);
  for (int k = 0; k < d2; ++k) H[k] = std::max(a[k], b[k]);
  for (int pass = 0; pass < 2; ++pass) {
    const raw_t* z = pass == 0 ? a : b;
    const raw_t* other = pass == 0 ? b : a;
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) {
        if (i == j) continue;
        const raw_t zij = z[i * d + j];
        if (zij >= H[i * d + j]) continue;
        std::copy(H.begin(), H.end(), R.begin());
        const int v = bnd(zij);
        const raw_t neg = is_strict(zij) ? le(-v) : lt(-v);
        if (!dbm_constrain(R.data(), d, j, i, neg)) continue;
        if (!dbm_includes(other, R.data(), d)) return false;
      }
  }
  std::copy(H.begin(), H.end(), a);
  return true;
}

// ---- operations of the backward (liveness) pass -------------------------------------------------
// Textbook definitions (Bengtsson & Yi 2004, section 4; the UPPAAL DBM library's down / freeClock),
// which HDBM's Dbm::down and dbm_free_many restate; the reference calls them from
// src/liveness/operators.cpp:257-306 and src/reachability/deadlock.cpp:74-268.

// Time predecessors: every lower bound x_j >= ... relaxed to what the other clocks force,
// z[0][j] = min(<=0, min_i z[i][j]). Canonical in, canonical out.
inline void dbm_down(raw_t* z, int d) {
  for (int j = 1; j < d; ++j) {
    raw_t v = kLe0;
    for (int i = 1; i < d; ++i) v = std::min(v, z[i * d + j]);
    z[j] = v;
  }
}

// Forget the clocks of `mask` (bit c = clock c): row c unbounded, column c takes column 0, i.e.
// only x_c >= 0 remains. Freeing several clocks at once equals freeing them one by one.
inline void dbm_free_many(raw_t* z, int d, std::uint64_t mask) {
  if (!mask) return;
  std::vector<raw_t> t(z, z + d * d);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      raw_t v;
      if (i == j) v = kLe0;
      else if ((mask >> i) & 1) v = kInf;
      else if ((mask >> j) & 1) v = t[i * d];
      else v = t[i * d + j];
      z[i * d + j] = v;
    }
}

// Definitely disjoint: a two-clock cycle through both zones is negative. A necessary condition
// for an empty intersection of canonical zones, used only to skip work.
inline bool dbm_disjoint_quick(const raw_t* a, const raw_t* b, int d) {
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      if (radd(a[i * d + j], b[j * d + i]) < kLe0) return true;
  return false;
}

// out := a /\ b (entrywise min, then closure). Returns false iff the intersection is empty.
inline bool dbm_intersect(raw_t* out, const raw_t* a, const raw_t* b, int d) {
  if (dbm_disjoint_quick(a, b, d)) return false;
  for (int k = 0; k < d * d; ++k) out[k] = std::min(a[k], b[k]);
  return dbm_close(out, d);
}

// p \ y as disjoint zones appended to `out` (flat, d*d each): walk y's facets (i, j) in row-major
// order; where y is tighter than the remainder r, split off r /\ not(x_i - x_j <| y_ij) and keep
// r /\ (x_i - x_j <| y_ij). When p and y are disjoint the answer is p itself.
inline void dbm_subtract(const raw_t* p, const raw_t* y, int d, std::vector<raw_t>& out) {
  const int d2 = d * d;
  std::vector<raw_t> r(p, p + d2), t(d2);
  if (!dbm_intersect(t.data(), p, y, d)) { out.insert(out.end(), p, p + d2); return; }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      if (i == j) continue;
      const raw_t yij = y[i * d + j];
      if (yij >= r[i * d + j]) continue;
      std::copy(r.begin(), r.end(), t.begin());
      const int v = bnd(yij);
      const raw_t neg = is_strict(yij) ? le(-v) : lt(-v);
      if (dbm_constrain(t.data(), d, j, i, neg)) out.insert(out.end(), t.begin(), t.end());
      if (!dbm_constrain(r.data(), d, i, j, yij)) return;
    }
}

}  // namespace orc
