// zone_dev.hpp — DBM zone algebra for gfx950, one zone per wavefront, staged in LDS.
//
// Each function is called by all 64 lanes of a one-wavefront workgroup with the zone in LDS
// (`__shared__`); lanes own entries e = lane, lane+64, ... of the d*d row-major matrix. The
// results are the canonical DBMs HDBM's operations produce (raw encoding and semantics as
// restated in src/metal/netengine_metal.mm:76-335 and src/metal/engine_metal.mm:80-93):
// a non-empty zone has exactly one canonical form, so a data-parallel schedule that reaches it
// is bit-identical to the sequential one.
//
// Every function begins with a workgroup barrier, so callers may write the zone just before.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hh {

constexpr int kInf = 1 << 30;
constexpr int kLe0 = 1;
constexpr int kAluNegInf = -(1 << 29);
constexpr int kWave = 64;

// Lane of the calling thread in its wavefront. Zone operations run on one wavefront; a workgroup
// may hold several wavefronts, each on a zone of its own.
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Row of entry e of a d x d matrix, e < d*d, d <= 32, by a multiply and a shift: with
// M = ceil(2^16 / d) = (2^16 + r) / d, r < d, floor(e*M / 2^16) = floor(e / d) since e*r < 2^16.
// (A division by a run-time d costs a dozen vector instructions per entry; even the one division
// that forms M costs twenty instructions where it is inlined, so M comes from a table.)
static __device__ __constant__ unsigned kEdivM[33] = {
    0,    65536, 32768, 21846, 16384, 13108, 10923, 9363, 8192, 7282, 6554, 5958, 5462, 5042, 4682, 4370, 4096,
    3856, 3641,  3450,  3277,  3121,  2979,  2850,  2731, 2622, 2521, 2428, 2341, 2260, 2185, 2115, 2048};
__device__ __forceinline__ int ediv(int e, int d) { return (int)(((unsigned)e * kEdivM[d]) >> 16); }

__device__ __forceinline__ int r_le(int v) { return v * 2 + 1; }
__device__ __forceinline__ int r_lt(int v) { return v * 2; }
__device__ __forceinline__ int r_bnd(int r) { return r >> 1; }
__device__ __forceinline__ bool r_strict(int r) { return (r & 1) == 0; }
// Saturating bound addition: infinity absorbs; the sum is weak iff both operands are weak. With
// a = 2x + sa, b = 2y + sb the result 2(x + y) + (sa & sb) is a + b - (sa | sb).
__device__ __forceinline__ int radd(int a, int b) {
  if (max(a, b) >= kInf) return kInf;
  return a + b - ((a | b) & 1);
}
// The raw bound of the complement of x_i - x_j <| v, stated as a bound on x_j - x_i.
__device__ __forceinline__ int r_neg(int r) {
  const int v = r >> 1;
  return r_strict(r) ? r_le(-v) : r_lt(-v);
}

// Ordering point for LDS inside a one-wavefront workgroup. The wavefront's LDS instructions
// execute in issue order, so a lane reading what another lane wrote needs only that the compiler
// keeps the accesses in program order and that earlier LDS operations have completed
// (lgkmcnt(0)). A workgroup barrier (__syncthreads) would also wait for every outstanding global
// load and store (vmcnt(0) counts both on CDNA), stalling each zone operation on HBM latency.
//
// Built with HORA_FED_HBM (the liveness kernels' device-memory tier, whose federation lists live in
// global memory), the ordering point is a workgroup-scope fence instead: it also waits for the
// wavefront's vector memory operations, so a lane reads what another lane stored to global memory.
__device__ __forceinline__ void wsync() {
#ifdef HORA_FED_HBM
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#else
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0); vmcnt and expcnt left at their maxima
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
#endif
}

// Full workgroup barrier with memory ordering: for global memory written by one lane and read
// by another, and for multi-wavefront workgroups.
__device__ __forceinline__ void bsync() { __syncthreads(); }

// Ordering point for global memory between the lanes of one wavefront, without a workgroup
// barrier (for a wavefront whose workgroup holds other wavefronts that do not take the same path):
// a workgroup-scope fence, which also waits for the wavefront's outstanding memory operations.
__device__ __forceinline__ void gsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void z_copy(int* __restrict__ dst, const int* __restrict__ src, int d2) {
  for (int e = lane_id(); e < d2; e += kWave) dst[e] = src[e];
}

// Narrow (int16) stored zones. Every entry of a stored zone of a model whose clock constants are
// small is a raw bound that fits 16 bits, or infinity (the narrowing rule of the reference's GPU
// passed set, include/hora/detail/gpu_zone_intern.hpp:36-47, src/reachability/engine.cpp:1596-1611).
// Infinity is stored as kInf16. A narrow zone takes nstride(d2) entries, an even number, so every
// narrow zone starts on a 4-byte boundary and its last entry shares a dword only with padding.
// Whoever narrows a zone checks each entry (fits16) instead of trusting the constants.
constexpr int kInf16 = 0x7fff;
__host__ __device__ __forceinline__ int nstride(int d2) { return (d2 + 1) & ~1; }
__device__ __forceinline__ bool fits16(int r) { return r == kInf || (r >= -32768 && r < kInf16); }
__device__ __forceinline__ short narrow16(int r) { return (short)(r == kInf ? kInf16 : r); }
__device__ __forceinline__ int widen16(int h) { return h == kInf16 ? kInf : h; }

// A narrow stored zone as an operand: entries widened on load.
struct Z16 {
  const short* p;
  __device__ __forceinline__ int operator[](int e) const { return widen16(p[e]); }
};

// One pass z[i][j] = min(z[i][j], z[i][p] + z[p][j]) for a pivot p whose row and column the pass
// leaves unchanged (z[p][p] = (<=, 0)), so lanes may read and write in any order. Branch-free: a
// lane reads its three values, then writes the minimum (the same value where nothing changes).
__device__ __forceinline__ void z_pivot_pass(int* z, int d, int p) {
  const int l = lane_id(), d2 = d * d;
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    const int ip = z[i * d + p], pj = z[p * d + j], ij = z[e];
    z[e] = min(ij, radd(ip, pj));
  }
}

__device__ __forceinline__ bool z_empty(const int* z, int d) {
  wsync();
  const int l = lane_id();
  const bool neg = l < d && z[l * d + l] < kLe0;
  return __any(neg);
}

// Delay: every upper bound x_i - 0 dropped.
__device__ __forceinline__ void z_up(int* z, int d) {
  wsync();
  const int l = lane_id();
  if (l >= 1 && l < d) z[l * d] = kInf;
}

// x_k := 0: row k takes row 0, column k takes column 0, z[k][k] = (<=, 0).
__device__ __forceinline__ void z_reset(int* z, int d, int k) {
  wsync();
  const int l = lane_id();
  int r0 = 0, c0 = 0;
  if (l < d) { r0 = z[l]; c0 = z[l * d]; }
  wsync();
  if (l < d) {
    z[k * d + l] = (l == k) ? kLe0 : r0;
    z[l * d + k] = (l == k) ? kLe0 : c0;
  }
}

// Tighten x_a - x_b to r and re-close incrementally. Returns false iff the result is empty.
// The input must be canonical and non-empty. The new system is empty exactly when the cycle
// a -> b -> a closes negative, z[b][a] + r < (<=, 0); otherwise the closure of one added edge is
// z[i][k] = min(z[i][k], z[i][b] + z[b][k]) after the column z[i][b] = min(z[i][b], z[i][a] + r).
// On an empty result the zone's contents are left unspecified; callers discard it.
__device__ __forceinline__ bool z_constrain(int* z, int d, int a, int b, int r) {
  wsync();
  const int zab = z[a * d + b];
  if (r >= zab) return true;
  if (radd(z[b * d + a], r) < kLe0) return false;
  const int l = lane_id();
  int col = 0;
  if (l < d) col = radd(z[l * d + a], r);
  wsync();
  if (l < d && col < z[l * d + b]) z[l * d + b] = col;
  wsync();
  z_pivot_pass(z, d, b);
  return true;
}

// Floyd-Warshall closure, one pivot per step. Within a step, row k and column k do not change on
// a zone without negative cycles, so the parallel update equals the sequential one.
__device__ __forceinline__ void z_close(int* z, int d) {
  for (int k = 0; k < d; ++k) {
    wsync();
    z_pivot_pass(z, d, k);
  }
}

// Extra_LU+ (Behrmann, Bouyer, Larsen, Pelanek, STTT 2006) with row 0 read before any change,
// then closure. `row0` is d ints of LDS scratch.
__device__ __forceinline__ void z_extra_lu_plus(int* z, int d, const int* L, const int* U, int* row0) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  if (l < d) row0[l] = z[l];
  wsync();
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    if (i == j) continue;
    const int v = z[e];
    if (v >= kInf) continue;
    int nv = v;
    if (r_bnd(v) > L[i]) nv = kInf;
    else if (-r_bnd(row0[i]) > L[i]) nv = kInf;
    else if (i != 0 && -r_bnd(row0[j]) > U[j]) nv = kInf;
    else if (v < r_lt(-U[j])) nv = r_lt(-U[j]);
    z[e] = nv;
  }
  z_close(z, d);
}

// Z <=_aLU Zp (Herbreteau, Srivathsan, Walukiewicz, LICS 2012): Z is not simulated iff some pair
// x != y has U_x, L_y finite, Z[0][x] >= (<=,-U_x), Zp[y][x] < Z[y][x] and
// Zp[y][x] + (<,-L_y) < Z[0][x]. Lanes take the pairs (y, x) = entries of the matrix, so both
// operands are read with unit stride; either operand may be in LDS or global memory (ZA, ZB: an
// int pointer or a narrow zone, Z16).
template <class ZA, class ZB>
__device__ __forceinline__ bool z_alu_le(ZA Z, ZB Zp, int d, const int* L, const int* U) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  bool wit = false;
  for (int e = l; e < d2; e += kWave) {
    const int y = ediv(e, d), x = e - y * d;
    if (x == y) continue;
    const int ux = U[x], ly = L[y];
    if (ux == kAluNegInf || ly == kAluNegInf) continue;
    const int z0x = Z[x];
    if (z0x < r_le(-ux)) continue;
    const int zp = Zp[e];
    if (zp < Z[e] && radd(zp, r_lt(-ly)) < z0x) wit = true;
  }
  return !__any(wit);
}

// Entrywise b <= a.
__device__ __forceinline__ bool z_includes(const int* a, const int* b, int d) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  bool bad = false;
  for (int e = l; e < d2; e += kWave)
    if (b[e] > a[e]) bad = true;
  return !__any(bad);
}

// Is a U b convex? If so a := hull (entrywise max) and return true.
//
// H \ a is the union over facets (i, j) of a with a_ij < H_ij (i.e. a_ij < b_ij) of
// R = H /\ (x_j - x_i <| neg a_ij); R is never empty for such a facet, and its canonical form is
// R[p][q] = min(H_pq, H_pj + neg + H_iq). R is inside b iff that is <= b_pq for every entry,
// which holds trivially where H_pq = b_pq; so only the entries with a_pq > b_pq need the test.
// Symmetrically for H \ b. With FA = {a < b} and PA = {a > b} the whole test is the pairs
// FA x PA twice, checked here across lanes. `fa`/`pa` are d*d ints of LDS scratch each;
// `cnt` is 2 ints of LDS scratch. Both a and b must be canonical, non-empty and in LDS.
__device__ __forceinline__ bool z_try_union(int* a, const int* b, int d, int* fa, int* pa, int* cnt) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  {
    // A gap between the two zones' projections on some x_i - x_j (canonical: a's upper bound
    // a[i][j], b's lower bound -b[j][i]) makes the union non-convex, as the projection of a convex
    // set is an interval: the answer is "no" without the facet pairs.
    bool gap = false;
    for (int e = l; e < d2; e += kWave) {
      const int i = ediv(e, d), j = e - i * d;
      const int ra = a[e], rb = b[j * d + i];
      if (i != j && ra < kInf && rb < kInf) {
        const int s = r_bnd(ra) + r_bnd(rb);
        if (s < 0 || (s == 0 && r_strict(ra) && r_strict(rb))) gap = true;
      }
    }
    if (__any(gap)) return false;
  }
  if (l == 0) { cnt[0] = 0; cnt[1] = 0; }
  wsync();
  // ballot compaction of the differing entries, in index order
  int nf = 0, np = 0;
  for (int base = 0; base < d2; base += kWave) {
    const int e = base + l;
    const int av = e < d2 ? a[e] : 0, bv = e < d2 ? b[e] : 0;
    const bool f = e < d2 && av < bv, p = e < d2 && av > bv;
    const uint64_t mf = __ballot(f), mp = __ballot(p);
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    if (f) fa[nf + __popcll(mf & below)] = e;
    if (p) pa[np + __popcll(mp & below)] = e;
    nf += __popcll(mf);
    np += __popcll(mp);
  }
  wsync();
  if (nf > 0 && np > 0) {
    // a wavefront of pairs at a time; the first crossing point found ends the test for all lanes
    const int pairs = nf * np;
    for (int base = 0; base < pairs; base += kWave) {
      // pass 1: facet of a from FA, entry from PA, compared against b
      const int t = base + l;
      bool bad = false;
      if (t < pairs) {
        const int f = fa[t / np], e = pa[t % np];
        const int i = ediv(f, d), j = f - i * d, p = ediv(e, d), q = e - p * d;
        const int neg = r_neg(a[f]);
        const int hpj = max(a[p * d + j], b[p * d + j]);
        const int hiq = max(a[i * d + q], b[i * d + q]);
        bad = radd(radd(hpj, neg), hiq) > b[e];
      }
      if (__any(bad)) return false;
    }
    for (int base = 0; base < pairs; base += kWave) {
      // pass 2: facet of b from PA, entry from FA, compared against a
      const int t = base + l;
      bool bad = false;
      if (t < pairs) {
        const int f = pa[t / nf], e = fa[t % nf];
        const int i = ediv(f, d), j = f - i * d, p = ediv(e, d), q = e - p * d;
        const int neg = r_neg(b[f]);
        const int hpj = max(a[p * d + j], b[p * d + j]);
        const int hiq = max(a[i * d + q], b[i * d + q]);
        bad = radd(radd(hpj, neg), hiq) > a[e];
      }
      if (__any(bad)) return false;
    }
  }
  wsync();
  for (int e = l; e < d2; e += kWave) a[e] = max(a[e], b[e]);
  return true;
}

// ---- backward-pass operations (textbook down / free, Bengtsson & Yi 2004 section 4) -----------

// Time predecessors: z[0][j] = min(<=0, min_{i>=1} z[i][j]). Canonical in, canonical out.
__device__ __forceinline__ void z_down(int* z, int d) {
  wsync();
  const int l = lane_id();
  int v = kLe0;
  if (l >= 1 && l < d)
    for (int i = 1; i < d; ++i) v = min(v, z[i * d + l]);
  wsync();
  if (l >= 1 && l < d) z[l] = v;
}

// Forget the clocks of `mask`: row c unbounded, column c takes column 0. `tmp` is d*d LDS scratch.
__device__ __forceinline__ void z_free_many(int* z, int d, uint64_t mask, int* tmp) {
  if (!mask) return;
  wsync();
  const int l = lane_id(), d2 = d * d;
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    int v;
    if (i == j) v = kLe0;
    else if ((mask >> i) & 1) v = kInf;
    else if ((mask >> j) & 1) v = z[i * d];
    else v = z[e];
    tmp[e] = v;
  }
  wsync();
  for (int e = l; e < d2; e += kWave) z[e] = tmp[e];
}

// Z /\ (x_c = 0 for every clock c of `mask`), then those clocks freed: the exact predecessor's
// "resets to zero, then free" in two passes. The equalities merge M = mask + {0} into one node
// with zero-weight edges, so on a canonical Z a shortest path enters M once:
// z'[i][j] = min(z[i][j], min_{a in M} z[i][a] + min_{b in M} z[b][j]), and the conjunction is
// empty iff some z[a][b] (a != b in M) admits no zero difference. The result is the canonical
// zone the clock-by-clock constrain and free produce. `tmp` is 2*d ints of LDS scratch.
__device__ __forceinline__ bool z_zero_free(int* z, int d, uint64_t mask, int* tmp) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  const uint64_t M = mask | 1ull;
  bool bad = false;
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    if (i != j && ((M >> i) & 1) && ((M >> j) & 1) && z[e] < kLe0) bad = true;
  }
  if (__any(bad)) return false;
  if (l < d) {
    int rmin = kInf, cmin = kInf;
    for (uint64_t m = M; m; m &= m - 1) {
      const int a = __builtin_ctzll(m);
      rmin = min(rmin, z[l * d + a]);
      cmin = min(cmin, z[a * d + l]);
    }
    tmp[l] = rmin;
    tmp[d + l] = cmin;
  }
  wsync();
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    int v;
    if (i == j) v = kLe0;
    else if ((mask >> i) & 1) v = kInf;
    else if ((mask >> j) & 1) v = min(z[i * d], radd(tmp[i], tmp[d]));
    else v = min(z[e], radd(tmp[i], tmp[d + j]));
    z[e] = v;
  }
  return true;
}

// A two-clock cycle through both zones is negative: the zones are certainly disjoint.
__device__ __forceinline__ bool z_disjoint_quick(const int* a, const int* b, int d) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  bool neg = false;
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    if (radd(a[e], b[j * d + i]) < kLe0) neg = true;
  }
  return __any(neg);
}

// out := a /\ b (entrywise min, closure). Returns false iff empty; out is then unspecified.
__device__ __forceinline__ bool z_intersect(int* out, const int* a, const int* b, int d) {
  if (z_disjoint_quick(a, b, d)) return false;
  const int l = lane_id(), d2 = d * d;
  for (int e = l; e < d2; e += kWave) out[e] = min(a[e], b[e]);
  z_close(out, d);
  return !z_empty(out, d);
}

// Equal entrywise.
__device__ __forceinline__ bool z_equal(const int* a, const int* b, int d) {
  wsync();
  const int l = lane_id(), d2 = d * d;
  bool diff = false;
  for (int e = l; e < d2; e += kWave)
    if (a[e] != b[e]) diff = true;
  return !__any(diff);
}

// p \ y as disjoint zones appended to the list `out` (stride `zs` ints, `*n` entries, at most
// `cap`): y's facets (i, j) in row-major order; where y is tighter than the remainder r, split off
// r /\ not(x_i - x_j <| y_ij) and keep r /\ (x_i - x_j <| y_ij). The facets are taken a wavefront
// of entries at a time: one ballot names those tighter than the remainder (the remainder only
// shrinks, so no other entry of the chunk can cut it later), and each is checked again when its
// turn comes. `r`, `t` are d*d LDS scratch. Returns false if the list would exceed `cap`.
__device__ __forceinline__ bool z_subtract(const int* p, const int* y, int d, int* out, int zs, int* n, int cap,
                                           int* r, int* t) {
  const int l = lane_id(), d2 = d * d;
  if (!z_intersect(t, p, y, d)) {
    if (*n >= cap) return false;
    wsync();
    for (int e = l; e < d2; e += kWave) out[*n * zs + e] = p[e];
    ++*n;
    return true;
  }
  wsync();
  for (int e = l; e < d2; e += kWave) r[e] = p[e];
  for (int base = 0; base < d2; base += kWave) {
    wsync();
    const int e0 = base + l;
    uint64_t cut = __ballot(e0 < d2 && y[e0] < r[e0]);   // diagonal entries never cut: both hold <= 0 there
    for (; cut; cut &= cut - 1) {
      const int f = base + __builtin_ctzll(cut);
      const int i = ediv(f, d), j = f - i * d;
      wsync();
      const int yij = y[f];
      if (yij >= r[f]) continue;   // implied by the facets taken since
      wsync();
      for (int e = l; e < d2; e += kWave) t[e] = r[e];
      if (z_constrain(t, d, j, i, r_neg(yij))) {
        if (*n >= cap) return false;
        wsync();
        for (int e = l; e < d2; e += kWave) out[*n * zs + e] = t[e];
        ++*n;
      }
      if (!z_constrain(r, d, i, j, yij)) return true;
    }
  }
  return true;
}

}  // namespace hh
