// hora_hip.hip — the MI355X back end of HORA's forward zone-graph search, behind the C ABI of
// include/hora_hip.h.
//
// One BFS level runs as a short pipeline of hand-written gfx950 kernels over device-resident
// state; the host only sizes launches and reads a few counters between kernels:
//
//   k_count    one wavefront per frontier state: count the state's transitions ("recipes") in
//              the reference's enumeration order (gen_succ_reuse, src/reachability/engine.cpp:
//              541-718)
//   scan       recipe offsets, so every successor has a fixed slot in (state, recipe) order
//   k_expand   one wavefront per frontier state, zones staged in LDS: guards, resets, target
//              invariant, delay unless frozen, invariant, inactive-clock reset, Extra_LU+ when on,
//              emptiness, target test (close_succ_zone, include/hora/detail/semantics.hpp:522-568)
//   scan       compaction of the non-empty successors, keeping (state, recipe) order
//   k_find     discrete key -> slot of the device-resident passed list (open addressing over a
//              64-bit CAS)
//   k_probe    cover pre-filter of each successor against its key's previous-level zones
//   k_ticket   a ticket in the key's per-level work list for each successor the probe kept
//   k_merge    one wavefront per touched discrete key: the key's new zones in (state, recipe)
//              order, each through NetPassedT::add's folds (include/hora/detail/passed_set.hpp:
//              287-444): a_LU cover test, retirement of covered stored zones, convex-union merge
//              over the 32 newest live zones with restart, append
//   scan       next frontier: zones stored this level that are still live, in storing order
//
// add() never lets two discrete keys interact, so running keys in parallel while each key
// replays its own successors in order is exactly the sequential insertion order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hora_hip.h"
#include "zone_dev.hpp"

using u32 = uint32_t;
using u64 = uint64_t;
using namespace hh;

namespace {

constexpr u64 kEmptyKey = ~0ull;
constexpr int kMergeScan = 32;       // include/hora/detail/engine_config.hpp:52
constexpr int kSortLds = 1024;       // bucket work lists up to this length are sorted in LDS
constexpr u32 kDead = 0x80000000u;   // tombstone bit on a chain entry
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kStageMax = 8192;  // table blob (ints) staged in LDS by the walking kernels when it fits
constexpr int kLightMax = 8;   // working-chain size up to which a key takes the light merge launch

struct DevNet {
  int d, d2, np, ni, nclocks, lu_plus;
  int znar;               // the search stores its zones narrow (int16); 0 outside the search
  const int* locbase;
  const int* loc_flags;
  const int* inv_off;
  const int* inv_data;
  const int* le_off;      // [total_locs+1] edges leaving each global location, ascending
  const int* le_edges;
  const int* edge_proc;
  const int* edge_tgt;
  const int* edge_chan;
  const int* edge_dir;
  const int* cg_off;
  const int* cg_data;
  const int* rs_off;
  const int* rs_clk;
  const int* ig_off;
  const int* ig_data;
  const int* iu_off;
  const int* iu_data;
  const int* chan_bcast;  // one int per channel
  const int* chan_urgent;
  int has_urgent_chan;
  int has_bcast;          // some channel is a broadcast channel
  const int* Lx;          // extrapolation bounds [d] (global bounds widened by the query)
  const int* Ux;
  const int* lloc;        // per-location a_LU bounds [total_locs*d]
  const int* uloc;
  const int* Lq;          // query floors [d]
  const int* Uq;
  const u64* amask;
  u64 qbits;
  const int* kshift;      // key field offsets [np+ni]
  const u64* kmask;       // key field masks
  const int* int_lo;
  int never, nlr, nic, ncc;
  int rcap, icap;         // bounds on one state's transitions and one location vector's invariant
  const int* lr;
  const int* ic;
  const int* cc;
  // every table above except lloc/uloc lives in one blob, so that a workgroup can stage all of
  // them in LDS with one copy and rebase the pointers
  const int* blob;
  int blob_ints;
  int stage;              // 1: kernels that walk the tables stage the blob in LDS
  int inv_upper;          // every invariant constraint is an upper bound x_a - x_0 <| r
};

// Rebase every blob pointer of N onto an LDS copy of the blob.
__device__ __forceinline__ DevNet rebase(const DevNet& N, const int* lds) {
  DevNet L = N;
  const char* from = (const char*)N.blob;
  const char* to = (const char*)lds;
  auto rb = [&](auto p) { return (decltype(p))(to + ((const char*)p - from)); };
  L.locbase = rb(N.locbase); L.loc_flags = rb(N.loc_flags); L.inv_off = rb(N.inv_off); L.inv_data = rb(N.inv_data);
  L.le_off = rb(N.le_off); L.le_edges = rb(N.le_edges); L.edge_proc = rb(N.edge_proc); L.edge_tgt = rb(N.edge_tgt);
  L.edge_chan = rb(N.edge_chan); L.edge_dir = rb(N.edge_dir); L.cg_off = rb(N.cg_off); L.cg_data = rb(N.cg_data);
  L.rs_off = rb(N.rs_off); L.rs_clk = rb(N.rs_clk); L.ig_off = rb(N.ig_off); L.ig_data = rb(N.ig_data);
  L.iu_off = rb(N.iu_off); L.iu_data = rb(N.iu_data); L.chan_bcast = rb(N.chan_bcast); L.chan_urgent = rb(N.chan_urgent);
  L.Lx = rb(N.Lx); L.Ux = rb(N.Ux); L.Lq = rb(N.Lq); L.Uq = rb(N.Uq); L.amask = rb(N.amask); L.kshift = rb(N.kshift);
  L.kmask = rb(N.kmask); L.int_lo = rb(N.int_lo); L.lr = rb(N.lr); L.ic = rb(N.ic); L.cc = rb(N.cc);
  L.blob = lds;
  return L;
}

// Stage the table blob in LDS (all lanes copy) when the context asked for it.
__device__ __forceinline__ DevNet stage_tables(const DevNet& N, int* lds) {
  if (!N.stage) return N;
  for (int k = threadIdx.x; k < N.blob_ints; k += blockDim.x) lds[k] = N.blob[k];
  __syncthreads();
  return rebase(N, lds);
}

// Statistics are striped over cache lines by workgroup so that hundreds of thousands of
// workgroups per launch do not serialise on one address; the host sums the stripes.
constexpr int kStripes = 64;
enum { C_SUB, C_KILL, C_UNI, C_READS, C_INS, C_KEYS, C_LIVE, C_PREADS, C_PSUB, C_FULL, C_WOVF, C_NSTAT };
struct Counters {
  unsigned long long st[kStripes][16];
  int err;
  int first_target;
  int ntouched;
  int nlight, nheavy, npair;
  int pad[2];
};

__device__ __forceinline__ void stat_add(Counters* c, int k, unsigned long long v) {
  if (v) atomicAdd(&c->st[blockIdx.x % kStripes][k], v);
}

// Wavefront-aggregated append: lanes with `pred` get consecutive indices from `*counter`, with one
// atomic per wavefront. Every lane of the wavefront must call it.
// Workgroup-aggregated append (blockDim.x a multiple of 64, at most 1024): one atomic per workgroup.
// Every thread of the workgroup must call it; the order across workgroups is that of the atomics.
__device__ __forceinline__ int block_append(int* counter, bool pred) {
  __shared__ int wcnt[16], wbase;
  const u64 b = __ballot(pred);
  const int wid = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63), nw = (int)(blockDim.x >> 6);
  __syncthreads();   // a previous call's readers are done with wcnt and wbase
  if (lane == 0) wcnt[wid] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < nw; ++w) tot += wcnt[w];
    wbase = tot ? atomicAdd(counter, tot) : 0;
  }
  __syncthreads();
  int base = wbase;
  for (int w = 0; w < wid; ++w) base += wcnt[w];
  return pred ? base + __popcll(b & ((1ull << lane) - 1)) : -1;
}

__device__ __forceinline__ int wave_append(int* counter, bool pred) {
  const u64 b = __ballot(pred);
  if (!b) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(b);
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(b));
  base = __shfl(base, leader);
  return pred ? base + __popcll(b & ((1ull << lane) - 1)) : -1;
}

// Wavefront-aggregated statistic: one atomic per wavefront. Every lane must call it.
__device__ __forceinline__ void wave_stat(Counters* c, int k, unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) stat_add(c, k, v);
}

enum { ERR_NONE = 0, ERR_HASH_FULL = 1, ERR_ZONES_FULL = 2, ERR_NODES_FULL = 3, ERR_RECIPES = 4, ERR_NARROW = 5 };
constexpr int kInvCap = 128;      // invariant constraints of one location vector held in LDS

// ---------------------------------------------------------------- discrete state helpers

// A value every lane of the wavefront holds (the discrete state being expanded). Forcing it into a
// scalar register (readfirstlane) was measured 5x slower on the expansion: the LDS read then has
// to complete before each table lookup is issued, which serialises the loop. Left as a marker.
__device__ __forceinline__ int uni(int v) { return v; }

__device__ __forceinline__ bool atom_holds(int v, int op, int k) {
  switch (op) {
    case HORA_OP_EQ: return v == k;
    case HORA_OP_NE: return v != k;
    case HORA_OP_LT: return v < k;
    case HORA_OP_LE: return v <= k;
    case HORA_OP_GT: return v > k;
    default: return v >= k;
  }
}

__device__ __forceinline__ bool int_ok(const DevNet& N, int e, const int* ints) {
  for (int c = N.ig_off[e]; c < N.ig_off[e + 1]; ++c)
    if (!atom_holds(uni(ints[N.ig_data[3 * c]]), N.ig_data[3 * c + 1], N.ig_data[3 * c + 2])) return false;
  return true;
}

__device__ __forceinline__ bool committed_at(const DevNet& N, int p, int l) {
  return (N.loc_flags[N.locbase[p] + uni(l)] & 1) != 0;
}

// Lanes decode the packed key (DiscKeyCodecT::decode, engine.cpp:1018-1025) into LDS.
__device__ __forceinline__ void decode_key(const DevNet& N, u64 dk, int* locs, int* ints) {
  const int l = lane_id();
  if (l < N.np) locs[l] = (int)((dk >> N.kshift[l]) & N.kmask[l]);
  if (l < N.ni) ints[l] = (int)((dk >> N.kshift[N.np + l]) & N.kmask[N.np + l]) + N.int_lo[l];
}

__device__ __forceinline__ u64 encode_key(const DevNet& N, const int* locs, const int* ints) {
  u64 k = 0;
  for (int p = 0; p < N.np; ++p) k |= (u64)(unsigned)locs[p] << N.kshift[p];
  for (int v = 0; v < N.ni; ++v) k |= (u64)(unsigned)(ints[v] - N.int_lo[v]) << N.kshift[N.np + v];
  return k;
}

// urgent_sync_enabled (semantics.hpp:118-139)
__device__ __forceinline__ bool urgent_sync_enabled(const DevNet& N, const int* locs, const int* ints) {
  if (!N.has_urgent_chan) return false;
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + uni(locs[i]);
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || !N.chan_urgent[ch]) continue;
      if (!int_ok(N, se, ints)) continue;
      if (N.chan_bcast[ch]) return true;
      for (int j = 0; j < N.np; ++j) {
        if (j == i) continue;
        const int gj = N.locbase[j] + uni(locs[j]);
        for (int t2 = N.le_off[gj]; t2 < N.le_off[gj + 1]; ++t2) {
          const int re = N.le_edges[t2];
          if (N.edge_chan[re] != ch || N.edge_dir[re] != -1) continue;
          if (int_ok(N, re, ints)) return true;
        }
      }
    }
  }
  return false;
}

// frozen_flags (semantics.hpp:160-169): bit0 committed, bit1 urgent
__device__ __forceinline__ int frozen_flags(const DevNet& N, const int* locs, const int* ints) {
  int f = 0;
  for (int p = 0; p < N.np; ++p) f |= N.loc_flags[N.locbase[p] + uni(locs[p])] & 3;
  if (!(f & 2) && urgent_sync_enabled(N, locs, ints)) f |= 2;
  return f;
}

// Conjoin the combined invariant of `locs` (all_invariants + apply_all, semantics.hpp:82-86).
__device__ __forceinline__ bool apply_inv(const DevNet& N, const int* locs, int* z) {
  for (int p = 0; p < N.np; ++p) {
    const int g = N.locbase[p] + uni(locs[p]);
    for (int c = N.inv_off[g]; c < N.inv_off[g + 1]; ++c)
      if (!z_constrain(z, N.d, N.inv_data[3 * c], N.inv_data[3 * c + 1], N.inv_data[3 * c + 2])) return false;
  }
  return true;
}

__device__ __forceinline__ bool apply_edge_guard(const DevNet& N, int e, int* z) {
  for (int c = N.cg_off[e]; c < N.cg_off[e + 1]; ++c)
    if (!z_constrain(z, N.d, N.cg_data[3 * c], N.cg_data[3 * c + 1], N.cg_data[3 * c + 2])) return false;
  return true;
}

__device__ __forceinline__ void apply_edge_resets(const DevNet& N, int e, int* z) {
  for (int r = N.rs_off[e]; r < N.rs_off[e + 1]; ++r) z_reset(z, N.d, N.rs_clk[r]);
}

// close_succ_zone (semantics.hpp:522-568) followed by the caller's emptiness test.
__device__ __forceinline__ bool close_succ(const DevNet& N, const int* locs, const int* ints, int* z, int* row0) {
  if (!apply_inv(N, locs, z)) return false;
  const int fz = frozen_flags(N, locs, ints);
  if (!fz) {
    z_up(z, N.d);
    apply_inv(N, locs, z);
  }
  u64 a = N.qbits;
  for (int p = 0; p < N.np; ++p) a |= N.amask[N.locbase[p] + uni(locs[p])];
  u64 dead = ~a & (((1ull << N.nclocks) - 1) << 1);
  while (dead) {
    const int c = __builtin_ctzll(dead);
    dead &= dead - 1;
    z_reset(z, N.d, c);
  }
  if (N.lu_plus) z_extra_lu_plus(z, N.d, N.Lx, N.Ux, row0);
  return !z_empty(z, N.d);
}

// reachable_net's match (engine.cpp:2275-2287); `tmp` is d*d LDS scratch.
__device__ __forceinline__ bool is_target(const DevNet& N, const int* locs, const int* ints, const int* z, int* tmp) {
  if (N.never) return false;
  for (int k = 0; k < N.nlr; ++k)
    if (locs[N.lr[2 * k]] != N.lr[2 * k + 1]) return false;
  for (int c = 0; c < N.nic; ++c)
    if (!atom_holds(ints[N.ic[3 * c]], N.ic[3 * c + 1], N.ic[3 * c + 2])) return false;
  if (N.ncc == 0) return true;
  wsync();
  z_copy(tmp, z, N.d2);
  for (int k = 0; k < N.ncc; ++k)
    if (!z_constrain(tmp, N.d, N.cc[3 * k], N.cc[3 * k + 1], N.cc[3 * k + 2])) return false;
  return true;
}

// ---------------------------------------------------------------- kernels: expansion

struct ExpandOut {
  int* valid;      // [R] 1 = non-empty successor
  u64* dk;         // [R] packed discrete key
  int* zone;       // [R*d2]
  int* tgt;        // [R] 1 = matches the target
  int* se;         // [R] transition of the slot (sender or local edge), or null: not recorded
  int* re;         // [R] receiver edge or -1
};

// Wavefront-wide helpers. One lane per process (np <= 64): the per-process table lookups of a
// location vector are issued together instead of as a chain of dependent loads.

__device__ __forceinline__ u64 wave_or64(u64 v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int wave_excl_scan(int v, int* total) {
  const int l = lane_id();
  int x = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (l >= o) x += y;
  }
  *total = __shfl(x, kWave - 1);
  return x - v;
}

// Packed key of a location vector and valuation (DiscKeyCodecT::encode, engine.cpp:1006-1013).
__device__ __forceinline__ u64 wave_encode_key(const DevNet& N, const int* locs, const int* ints) {
  const int l = lane_id();
  u64 k = 0;
  if (l < N.np) k |= (u64)(unsigned)locs[l] << N.kshift[l];
  if (l < N.ni) k |= (u64)(unsigned)(ints[l] - N.int_lo[l]) << N.kshift[N.np + l];
  return wave_or64(k);
}

// The target side of close_succ_zone for a location vector: frozen flags (semantics.hpp:160-169),
// the active-clock mask union (reset_inactive, semantics.hpp:195-205) and the combined invariant
// gathered into `inv` as (i, j, raw) triples (all_invariants, semantics.hpp:82-86). The list is
// written only when it fits `cap` triples; `ninv` is its length either way.
__device__ __forceinline__ void wave_loc_info(const DevNet& N, const int* locs, const int* ints, int* frozen,
                                              u64* amask, int* inv, int* ninv, int cap) {
  const int l = lane_id();
  int f = 0, ia = 0, ib = 0;
  u64 am = 0;
  if (l < N.np) {
    const int g = N.locbase[l] + locs[l];
    f = N.loc_flags[g] & 3;
    am = N.amask[g];
    ia = N.inv_off[g];
    ib = N.inv_off[g + 1];
  }
  int fr = (__ballot(f & 1) ? 1 : 0) | (__ballot(f & 2) ? 2 : 0);
  if (!(fr & 2) && N.has_urgent_chan && urgent_sync_enabled(N, locs, ints)) fr |= 2;
  *frozen = fr;
  *amask = wave_or64(am);
  int total = 0;
  const int off = wave_excl_scan(ib - ia, &total);
  if (total <= cap)
    for (int c = ia; c < ib; ++c) {
      const int o = 3 * (off + c - ia);
      inv[o] = N.inv_data[3 * c];
      inv[o + 1] = N.inv_data[3 * c + 1];
      inv[o + 2] = N.inv_data[3 * c + 2];
    }
  *ninv = total;
}

// Conjoin a list of constraints. Constraints that do not tighten the zone now never will (entries
// only decrease), so a parallel check selects the ones that need the incremental closure; the
// final zone does not depend on the order, and emptiness is detected either way.
__device__ __forceinline__ bool z_apply_list(int* z, int d, const int* cs, int n) {
  for (int base = 0; base < n; base += kWave) {
    wsync();
    const int k = base + lane_id();
    bool cand = false;
    if (k < n) cand = cs[3 * k + 2] < z[cs[3 * k] * d + cs[3 * k + 1]];
    for (u64 b = __ballot(cand); b; b &= b - 1) {
      const int c = base + __builtin_ctzll(b);
      if (!z_constrain(z, d, cs[3 * c], cs[3 * c + 1], cs[3 * c + 2])) return false;
    }
  }
  return true;
}

// Reset every clock in `mask` at once. Resetting x_k copies row 0 into row k and column 0 into
// column k; for a set K that is z[i][j] := z[i'][j'] with i' = 0 if i is in K (else i), likewise
// j' -- the composition of the single resets in any order. In place, in one pass: every entry
// read has both indices outside K (0 is never in K) and every entry written has one inside, so no
// lane reads an entry any lane writes, whatever the order. (`tmp` is unused.)
__device__ __forceinline__ void z_reset_many(int* z, int d, u64 mask, int* tmp) {
  if (!mask) return;
  wsync();
  const int l = lane_id(), d2 = d * d;
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    const bool ri = (mask >> i) & 1, rj = (mask >> j) & 1;
    if (ri || rj) z[e] = z[(ri ? 0 : i) * d + (rj ? 0 : j)];
  }
}

__device__ __forceinline__ u64 edge_reset_mask(const DevNet& N, int e) {
  u64 m = 0;
  for (int r = N.rs_off[e]; r < N.rs_off[e + 1]; ++r) m |= 1ull << N.rs_clk[r];
  return m;
}

// One part of a broadcast (build_broadcast_parts, semantics.hpp:275-303): the receiving edges on
// channel `ch` that the integer guards enable at global location gj, and the options they give the
// process: receive on one of them (nrecv), or, when each carries a clock guard, stay out in a zone
// where all of them fail -- one negated atom per guard, nnone = the product of the guard sizes.
struct BPart {
  int nrecv, nnone;
};

constexpr int kRecipeCap = 256;    // transitions of one discrete state held in LDS

__device__ __forceinline__ BPart bcast_part(const DevNet& N, int gj, int ch, const int* ints) {
  BPart b{0, 1};
  for (int t = N.le_off[gj]; t < N.le_off[gj + 1]; ++t) {
    const int re = N.le_edges[t];
    if (N.edge_chan[re] != ch || N.edge_dir[re] != -1 || !int_ok(N, re, ints)) continue;
    ++b.nrecv;
    b.nnone = min(b.nnone * (N.cg_off[re + 1] - N.cg_off[re]), kRecipeCap + 1);
  }
  if (b.nrecv == 0) b.nnone = 0;
  return b;
}

// The k-th enabled receiving edge of a part.
__device__ __forceinline__ int bcast_recv_edge(const DevNet& N, int gj, int ch, const int* ints, int k) {
  for (int t = N.le_off[gj]; t < N.le_off[gj + 1]; ++t) {
    const int re = N.le_edges[t];
    if (N.edge_chan[re] != ch || N.edge_dir[re] != -1 || !int_ok(N, re, ints)) continue;
    if (k == 0) return re;
    --k;
  }
  return -1;
}

// Option `combo` of broadcast sender edge se, part by part in process order: the options of a part
// are its receiving edges in edge order, then its stay-out conjunctions; the first part is the
// fastest digit of `combo` (the odometer of semantics.hpp:275-303). fn(j, gj, re, q) gets the
// receiving edge, or re = -1 and the index q of the stay-out conjunction.
template <class F>
__device__ __forceinline__ void bcast_walk(const DevNet& N, const int* locs, const int* ints, int se, int combo,
                                           F&& fn) {
  const int i = N.edge_proc[se], ch = N.edge_chan[se];
  for (int j = 0; j < N.np; ++j) {
    if (j == i) continue;
    const int gj = N.locbase[j] + locs[j];
    const BPart b = bcast_part(N, gj, ch, ints);
    if (b.nrecv == 0) continue;
    const int nopt = b.nrecv + b.nnone;
    const int pick = combo % nopt;
    combo /= nopt;
    fn(j, gj, pick < b.nrecv ? bcast_recv_edge(N, gj, ch, ints, pick) : -1, pick - b.nrecv);
  }
}

// Stay-out conjunction q of the part at gj: one negated atom of each receiving edge's clock guard,
// the first edge's atom the most significant digit of q.
__device__ __forceinline__ bool apply_bcast_none(const DevNet& N, int gj, int ch, const int* ints, int q, int* z) {
  for (int t = N.le_off[gj + 1] - 1; t >= N.le_off[gj]; --t) {
    const int re = N.le_edges[t];
    if (N.edge_chan[re] != ch || N.edge_dir[re] != -1 || !int_ok(N, re, ints)) continue;
    const int a = N.cg_off[re], sz = N.cg_off[re + 1] - a;
    const int c = a + q % sz;
    q /= sz;
    if (!z_constrain(z, N.d, N.cg_data[3 * c + 1], N.cg_data[3 * c], r_neg(N.cg_data[3 * c + 2]))) return false;
  }
  return true;
}

// Transition recipes of a discrete state, in gen_succ_reuse's order (engine.cpp:541-718), into
// `rec` as (sender edge, receiver edge or -1) pairs: local edges in process order -- each lane
// enumerates its own process, a prefix sum places them -- then binary synchronisations (sender
// process and edge, then receiver process and edge), then broadcasts (sender process and edge,
// then the combinations of bcast_walk in order), written (sender edge, -2 - combination).
// Returns the number of recipes, or -1 if more than `cap`.
template <bool BC = true>
__device__ __forceinline__ int build_recipes(const DevNet& N, const int* locs, const int* ints, bool src_c,
                                             int* rec, int cap) {
  const int l = lane_id();
  int cnt = 0, g = 0;
  bool comm = false;
  if (l < N.np) {
    g = N.locbase[l] + locs[l];
    comm = (N.loc_flags[g] & 1) != 0;
    if (!src_c || comm)
      for (int t = N.le_off[g]; t < N.le_off[g + 1]; ++t) {
        const int e = N.le_edges[t];
        if (N.edge_chan[e] < 0 && int_ok(N, e, ints)) ++cnt;
      }
  }
  int total = 0;
  int off = wave_excl_scan(cnt, &total);
  if (total > cap) return -1;
  if (l < N.np && cnt > 0)
    for (int t = N.le_off[g]; t < N.le_off[g + 1]; ++t) {
      const int e = N.le_edges[t];
      if (N.edge_chan[e] < 0 && int_ok(N, e, ints)) {
        rec[2 * off] = e;
        rec[2 * off + 1] = -1;
        ++off;
      }
    }
  // binary synchronisations: per sender edge (process order, then edge order), the receivers of
  // every other process at once, one lane per receiving process, placed by a prefix sum (the
  // receiver process, then edge order)
  const u64 has_send = __ballot(l < N.np && N.le_off[g + 1] > N.le_off[g]);
  if (!has_send) return total;
  int n = total;
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + locs[i];
    const bool i_comm = committed_at(N, i, locs[i]);
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || N.chan_bcast[ch]) continue;
      if (!int_ok(N, se, ints)) continue;
      // lane l: the receiving edges of process l (g is its location)
      int rc = 0;
      const bool part = l < N.np && l != i && (!src_c || i_comm || comm);
      if (part)
        for (int t2 = N.le_off[g]; t2 < N.le_off[g + 1]; ++t2) {
          const int re = N.le_edges[t2];
          if (N.edge_chan[re] == ch && N.edge_dir[re] == -1 && int_ok(N, re, ints)) ++rc;
        }
      int tot = 0;
      int o = wave_excl_scan(rc, &tot);
      if (n + tot > cap) return -1;
      if (rc > 0)
        for (int t2 = N.le_off[g]; t2 < N.le_off[g + 1]; ++t2) {
          const int re = N.le_edges[t2];
          if (N.edge_chan[re] != ch || N.edge_dir[re] != -1 || !int_ok(N, re, ints)) continue;
          rec[2 * (n + o)] = se;
          rec[2 * (n + o) + 1] = re;
          ++o;
        }
      n += tot;
    }
  }
  if (!BC || !N.has_bcast) return n;
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + locs[i];
    const bool i_comm = committed_at(N, i, locs[i]);
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || !N.chan_bcast[ch]) continue;
      if (!int_ok(N, se, ints)) continue;
      int nopt = 1;   // lane j: the options of process j (1 when it takes no part)
      if (l < N.np && l != i) {
        const BPart b = bcast_part(N, N.locbase[l] + locs[l], ch, ints);
        if (b.nrecv > 0) nopt = b.nrecv + b.nnone;
      }
      int combos = 1;
      for (int j = 0; j < N.np; ++j) combos = min(combos * __builtin_amdgcn_readlane(nopt, j), cap + 1);
      if (n + combos > cap) return -1;
      for (int c = 0; c < combos; ++c) {
        // from a committed state, the sender or some receiving process must be committed
        bool keep = !src_c || i_comm;
        if (!keep)
          bcast_walk(N, locs, ints, se, c, [&](int j, int, int re, int) {
            if (re >= 0 && committed_at(N, j, locs[j])) keep = true;
          });
        if (!keep) continue;
        if (l == 0) { rec[2 * n] = se; rec[2 * n + 1] = -2 - c; }
        ++n;
      }
    }
  }
  return n;
}

// The number of recipes build_recipes lists for the state with key dk (-1 where it would fail for
// `cap`), one thread per state: the location and variable fields are read from the key as needed.
// Broadcast combinations are counted by product, the committed rule by its complement: from a
// committed state a combination whose sender is not committed survives unless every committed
// process that takes part stays out.
__device__ __forceinline__ int count_recipes(const DevNet& N, u64 dk, int cap) {
  auto loc = [&](int p) { return (int)((dk >> N.kshift[p]) & N.kmask[p]); };
  auto iok = [&](int e) {
    for (int c = N.ig_off[e]; c < N.ig_off[e + 1]; ++c) {
      const int v = N.ig_data[3 * c];
      const int x = (int)((dk >> N.kshift[N.np + v]) & N.kmask[N.np + v]) + N.int_lo[v];
      if (!atom_holds(x, N.ig_data[3 * c + 1], N.ig_data[3 * c + 2])) return false;
    }
    return true;
  };
  auto comm = [&](int p) { return (N.loc_flags[N.locbase[p] + loc(p)] & 1) != 0; };
  bool src_c = false;
  for (int p = 0; p < N.np; ++p) src_c |= comm(p);
  long long n = 0;
  for (int p = 0; p < N.np; ++p) {
    if (src_c && !comm(p)) continue;
    const int g = N.locbase[p] + loc(p);
    for (int t = N.le_off[g]; t < N.le_off[g + 1]; ++t) {
      const int e = N.le_edges[t];
      if (N.edge_chan[e] < 0 && iok(e)) ++n;
    }
  }
  if (n > cap) return -1;
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + loc(i);
    const bool i_comm = comm(i);
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || N.chan_bcast[ch] || !iok(se)) continue;
      for (int j = 0; j < N.np; ++j) {
        if (j == i) continue;
        if (src_c && !(i_comm || comm(j))) continue;
        const int gj = N.locbase[j] + loc(j);
        for (int t2 = N.le_off[gj]; t2 < N.le_off[gj + 1]; ++t2) {
          const int re = N.le_edges[t2];
          if (N.edge_chan[re] == ch && N.edge_dir[re] == -1 && iok(re)) ++n;
        }
      }
    }
  }
  if (n > cap) return -1;
  if (!N.has_bcast) return (int)n;
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + loc(i);
    const bool i_comm = comm(i);
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || !N.chan_bcast[ch] || !iok(se)) continue;
      long long combos = 1, all_out = 1;   // all_out: combinations where every committed part stays out
      for (int j = 0; j < N.np; ++j) {
        if (j == i) continue;
        const int gj = N.locbase[j] + loc(j);
        int nrecv = 0, nnone = 1;
        for (int t2 = N.le_off[gj]; t2 < N.le_off[gj + 1]; ++t2) {
          const int re = N.le_edges[t2];
          if (N.edge_chan[re] != ch || N.edge_dir[re] != -1 || !iok(re)) continue;
          ++nrecv;
          nnone = min(nnone * (N.cg_off[re + 1] - N.cg_off[re]), kRecipeCap + 1);
        }
        if (nrecv == 0) continue;
        const long long nopt = nrecv + nnone;
        combos = min(combos * nopt, (long long)cap + 1);
        all_out = min(all_out * (comm(j) ? (long long)nnone : nopt), (long long)cap + 1);
      }
      if (n + combos > cap) return -1;   // the limit build_recipes applies before filtering
      n += (src_c && !i_comm) ? combos - all_out : combos;
    }
  }
  return (int)n;
}

__global__ __launch_bounds__(256) void k_count(const DevNet* __restrict__ Np, const u64* __restrict__ fr_dk, int F,
                                               int* __restrict__ cnt, Counters* ctr) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= F) return;
  const int r = count_recipes(*Np, fr_dk[i], Np->rcap);
  cnt[i] = r < 0 ? 0 : r;
  if (r < 0) atomicExch(&ctr->err, ERR_RECIPES);
}

// LDS layout of k_expand: 3 zones + row-0 scratch + 2 discrete states + recipes (rcap) + the source
// state's invariant triples with their process (icap) + per-process flags and active-clock masks.
__host__ __device__ constexpr int expand_lds_ints(int d2, int rcap, int icap) {
  return 3 * d2 + HORA_MAX_DIM + 2 * (HORA_MAX_PROCS + HORA_MAX_INTS) + 2 * rcap + 4 * icap +
         3 * HORA_MAX_PROCS + 2;
}


// z_apply_list over the source state's invariant triples, skipping those of the processes in
// `moved`: `tag[k]` is the process of triple k.
__device__ __forceinline__ bool z_apply_list_skip(int* z, int d, const int* cs, const int* tag, int n, u64 moved) {
  for (int base = 0; base < n; base += kWave) {
    wsync();
    const int k = base + lane_id();
    bool cand = false;
    if (k < n && !((moved >> tag[k]) & 1)) cand = cs[3 * k + 2] < z[cs[3 * k] * d + cs[3 * k + 1]];
    for (u64 b = __ballot(cand); b; b &= b - 1) {
      const int c = base + __builtin_ctzll(b);
      if (!z_constrain(z, d, cs[3 * c], cs[3 * c + 1], cs[3 * c + 2])) return false;
    }
  }
  return true;
}

// The invariant of global location g (wave-uniform), conjoined.
__device__ __forceinline__ bool apply_loc_inv(const DevNet& N, int g, int* z) {
  for (int c = N.inv_off[g]; c < N.inv_off[g + 1]; ++c)
    if (!z_constrain(z, N.d, N.inv_data[3 * c], N.inv_data[3 * c + 1], N.inv_data[3 * c + 2])) return false;
  return true;
}

// The combined invariant of a successor: the source state's, minus the old locations of the
// processes in `moved`, plus the new locations g1, g2 (global locations; -1 for none). The
// receivers of a broadcast add theirs after this.
__device__ __forceinline__ bool apply_succ_inv(const DevNet& N, int* z, const int* sinv, const int* stag, int ns,
                                               u64 moved, int g1, int g2) {
  if (!z_apply_list_skip(z, N.d, sinv, stag, ns, moved)) return false;
  if (g1 >= 0 && !apply_loc_inv(N, g1, z)) return false;
  if (g2 >= 0 && !apply_loc_inv(N, g2, z)) return false;
  return true;
}

// After z_up no clock has a finite upper bound. Conjoining upper bounds x_a <| r, triples (a, 0, r)
// handed out by `each`, then needs no incremental closure per triple: a shortest path through the
// new edges a -> 0 uses one of them (a second would close a cycle through 0, never negative in a
// non-empty zone), so the closed result is
//   z'[i][0] = min(z[i][0], min_a z[i][a] + r_a),  z'[i][j] = min(z[i][j], z'[i][0] + z[0][j]),
// and it is empty iff min_a z[0][a] + r_a < (<=, 0). A non-empty zone has one canonical form, so
// this is the zone the triples give one constraint at a time. Returns false iff empty.
template <class Each>
__device__ __forceinline__ bool z_upper_after_up(int* z, int d, Each each) {
  wsync();
  const int l = lane_id();
  int m = kInf;
  if (l < d) each([&](int a, int r) { m = min(m, radd(z[l * d + a], r)); });
  if (__any(l == 0 && m < kLe0)) return false;
  const bool tight = l >= 1 && l < d && m < z[l * d];
  if (!__any(tight)) return true;
  wsync();
  if (tight) z[l * d] = m;
  wsync();
  const int d2 = d * d;
  for (int e = l; e < d2; e += kWave) {
    const int i = ediv(e, d), j = e - i * d;
    if (i == 0 || j == 0) continue;
    const int i0 = z[i * d];
    if (i0 >= kInf) continue;
    const int v = radd(i0, z[j]);
    if (v < z[e]) z[e] = v;
  }
  return true;
}

__device__ __forceinline__ int bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ u64 bcast64(u64 v, int lane) {
  const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)v, lane);
  const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(v >> 32), lane);
  return ((u64)hi << 32) | lo;
}

// Successors of frontier states [0, F) into slots rec_off[i] + r, r < rec_cnt[i]. With `init` set, the frontier
// is the single initial state: its zone is the zero zone and its one recipe is "no transition".
// A workgroup holds blockDim.x / 64 wavefronts, one frontier state each, with LDS regions of their own.
//
// The discrete side of each successor (target key, frozen flags, active-clock mask, which processes
// moved and where, the target test's discrete part, the reset mask) is computed for up to 64
// transitions at once, one lane per transition, from the source state's per-process tables held in
// LDS; the zone side then runs transition by transition over the whole wavefront, reading those
// values with readlane. The combined invariant of a successor is the source state's triples
// without the moved processes', plus the moved processes' new locations'.
// Eight wavefronts per SIMD (64 VGPRs, no spills to memory) measured ~5 % faster than the seven
// the compiler picks by itself. BC: the network has broadcast channels; that instance also walks
// the parts of broadcast recipes (more registers, launched only for such networks).
template <bool BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_expand(const DevNet* __restrict__ Np, const u64* __restrict__ fr_dk,
                                                const u32* __restrict__ fr_zid, const int* __restrict__ zarena,
                                                int F, const long long* __restrict__ rec_off,
                                                const int* __restrict__ rec_cnt, ExpandOut out,
                                                Counters* ctr, int init) {
  extern __shared__ int smem[];
  const DevNet& N = *Np;
  const int wid = (int)(threadIdx.x >> 6);
  int* sm = smem + wid * expand_lds_ints(N.d2, N.rcap, N.icap);
  int* src = sm;
  int* z = src + N.d2;
  int* tmp = z + N.d2;
  int* row0 = tmp + N.d2;
  int* locs = row0 + HORA_MAX_DIM;
  int* ints = locs + HORA_MAX_PROCS;
  int* nl = ints + HORA_MAX_INTS;
  int* niv = nl + HORA_MAX_PROCS;
  int* rec = niv + HORA_MAX_INTS;
  int* sinv = rec + 2 * N.rcap;                     // 3 * icap
  int* stag = sinv + 3 * N.icap;                    // icap
  int* sflags = stag + N.icap;                      // HORA_MAX_PROCS
  u64* samask = (u64*)(((uintptr_t)(sflags + HORA_MAX_PROCS) + 7) & ~(uintptr_t)7);   // HORA_MAX_PROCS
  const int i = (int)blockIdx.x * (int)(blockDim.x >> 6) + wid;
  if (i >= F) return;
  const int l = lane_id(), d = N.d, d2 = N.d2;
  const u64 dk = fr_dk[i];
  decode_key(N, dk, locs, ints);
  if (init) {
    for (int e = l; e < d2; e += kWave) src[e] = kLe0;
  } else if (N.znar) {
    const short* zs = (const short*)zarena + (size_t)fr_zid[i] * nstride(d2);
    for (int e = l; e < d2; e += kWave) src[e] = widen16(zs[e]);
  } else {
    const int* zs = zarena + (size_t)fr_zid[i] * d2;
    for (int e = l; e < d2; e += kWave) src[e] = zs[e];
  }
  wsync();
  // source state: per-process flags, active clocks and invariant triples (lane p = process p)
  int fz_src = 0, ns = 0;
  {
    int f = 0, ia = 0, ib = 0;
    u64 am = 0;
    if (l < N.np) {
      const int g = N.locbase[l] + locs[l];
      f = N.loc_flags[g] & 3;
      am = N.amask[g];
      ia = N.inv_off[g];
      ib = N.inv_off[g + 1];
      sflags[l] = f;
      samask[l] = am;
    }
    fz_src = (__ballot(f & 1) ? 1 : 0) | (__ballot(f & 2) ? 2 : 0);
    if (!(fz_src & 2) && N.has_urgent_chan && urgent_sync_enabled(N, locs, ints)) fz_src |= 2;
    const int o0 = wave_excl_scan(ib - ia, &ns);
    if (ns > N.icap) {
      if (l == 0) atomicExch(&ctr->err, ERR_RECIPES);
      return;
    }
    for (int c = ia; c < ib; ++c) {
      const int o = o0 + c - ia;
      sinv[3 * o] = N.inv_data[3 * c];
      sinv[3 * o + 1] = N.inv_data[3 * c + 1];
      sinv[3 * o + 2] = N.inv_data[3 * c + 2];
      stag[o] = l;
    }
  }
  const u64 clocks = ((1ull << N.nclocks) - 1) << 1;
  // the invariant of the successor of transition (se, re); a broadcast's receivers add theirs
  auto succ_inv = [&](int se, int re, u64 moved, int g1, int g2) __attribute__((always_inline)) -> bool {
    if (!apply_succ_inv(N, z, sinv, stag, ns, moved, g1, g2)) return false;
    bool ok = true;
    if (BC && re <= -2)
      bcast_walk(N, locs, ints, se, -2 - re, [&](int j, int, int rj, int) {
        if (ok && rj >= 0) ok = apply_loc_inv(N, N.locbase[j] + N.edge_tgt[rj], z);
      });
    return ok;
  };
  // the zone side of one successor; `fz`, `am` describe the target location vector
  auto finish = [&](int slot, int se, int re, u64 moved, int g1, int g2, int fz, u64 am, u64 rm, u64 key,
                    int tdisc) __attribute__((always_inline)) {
    wsync();
    bool ok = true;
    const bool guarded = se >= 0 && ((BC && re <= -2) || N.cg_off[se + 1] > N.cg_off[se] ||
                                     (re >= 0 && N.cg_off[re + 1] > N.cg_off[re]));
    if (se >= 0 && !guarded) {
      // no clock guard: the copy and every reset in one pass (z_reset_many's gather)
      for (int e = l; e < d2; e += kWave) {
        const int i = ediv(e, d), j = e - i * d;
        z[e] = src[(((rm >> i) & 1) ? 0 : i) * d + (((rm >> j) & 1) ? 0 : j)];
      }
    } else {
      z_copy(z, src, d2);
    }
    if (guarded) {
      // guards (sender, then the receiver or each broadcast part), then every reset at once
      ok = apply_edge_guard(N, se, z);
      if (ok && re >= 0) ok = apply_edge_guard(N, re, z);
      if (BC && ok && re <= -2)
        bcast_walk(N, locs, ints, se, -2 - re, [&](int, int gj, int rj, int q) {
          if (ok) ok = rj >= 0 ? apply_edge_guard(N, rj, z) : apply_bcast_none(N, gj, N.edge_chan[se], ints, q, z);
        });
      if (ok) z_reset_many(z, d, rm, tmp);
    }
    // close_succ_zone (semantics.hpp:522-568): invariant, delay unless frozen, invariant,
    // inactive clocks reset, Extra_LU+ when on. A canonical zone that survives the invariant stays
    // non-empty through the delay, the second invariant, the resets and the extrapolation.
    if (ok) ok = succ_inv(se, re, moved, g1, g2);
    if (ok) {
      if (!fz) {
        // the invariant holds before the delay, so the zone stays non-empty after it
        z_up(z, d);
        if (N.inv_upper && !(BC && re <= -2)) {
          z_upper_after_up(z, d, [&](auto&& f) __attribute__((always_inline)) {
            for (int k = 0; k < ns; ++k)
              if (!((moved >> stag[k]) & 1)) f(sinv[3 * k], sinv[3 * k + 2]);
            for (int g : {g1, g2})
              if (g >= 0)
                for (int c = N.inv_off[g]; c < N.inv_off[g + 1]; ++c) f(N.inv_data[3 * c], N.inv_data[3 * c + 2]);
          });
        } else {
          succ_inv(se, re, moved, g1, g2);
        }
      }
      z_reset_many(z, d, ~(am | N.qbits) & clocks, tmp);
      if (N.lu_plus) z_extra_lu_plus(z, d, N.Lx, N.Ux, row0);
    }
    bool tg = ok && tdisc && !N.never;
    if (tg && N.ncc > 0) {
      wsync();
      z_copy(tmp, z, d2);
      for (int k = 0; k < N.ncc && tg; ++k) tg = z_constrain(tmp, d, N.cc[3 * k], N.cc[3 * k + 1], N.cc[3 * k + 2]);
    }
    wsync();
    if (l == 0) {
      out.valid[slot] = ok ? 1 : 0;
      out.tgt[slot] = tg ? 1 : 0;
      out.dk[slot] = ok ? key : 0;
      if (tg) atomicMin(&ctr->first_target, slot);
      if (out.se) {
        out.se[slot] = se;
        out.re[slot] = re;
      }
    }
    if (ok) {
      int* dst = out.zone + (size_t)slot * d2;
      for (int e = l; e < d2; e += kWave) dst[e] = z[e];
    }
  };
  // the discrete side of transition (se, re) from the source state, in the calling lane
  auto describe = [&](int se, int re, u64* moved, int* g1, int* g2, int* fz, u64* am, u64* rm, u64* key,
                      int* tdisc) __attribute__((always_inline)) {
    u64 k = dk, mv = 0, r = 0;
    int fl = 0;
    u64 a = 0;
    // process q moves along edge e: its location field, integer updates (in order: sender, then
    // the receiver or the broadcast parts in process order), flags, active clocks and resets
    auto move = [&](int e) __attribute__((always_inline)) {
      const int q = N.edge_proc[e], t = N.edge_tgt[e], h = N.locbase[q] + t;
      k = (k & ~(N.kmask[q] << N.kshift[q])) | ((u64)(unsigned)t << N.kshift[q]);
      for (int c = N.iu_off[e]; c < N.iu_off[e + 1]; ++c) {
        const int v = N.iu_data[2 * c], val = N.iu_data[2 * c + 1];
        const int f = N.np + v;
        k = (k & ~(N.kmask[f] << N.kshift[f])) | ((u64)(unsigned)(val - N.int_lo[v]) << N.kshift[f]);
      }
      fl |= N.loc_flags[h] & 3;
      a |= N.amask[h];
      for (int c = N.rs_off[e]; c < N.rs_off[e + 1]; ++c) r |= 1ull << N.rs_clk[c];
      mv |= 1ull << q;
      return h;
    };
    const int h1 = move(se);
    const int h2 = re >= 0 ? move(re) : -1;
    if (BC && re <= -2) bcast_walk(N, locs, ints, se, -2 - re, [&](int, int, int rj, int) { if (rj >= 0) move(rj); });
    for (int p = 0; p < N.np; ++p)
      if (!((mv >> p) & 1)) {
        fl |= sflags[p];
        a |= samask[p];
      }
    // discrete part of the target test (reachable_net's match, engine.cpp:2275-2287)
    int td = 1;
    for (int t = 0; t < N.nlr && td; ++t) {
      const int p = N.lr[2 * t], want = N.lr[2 * t + 1];
      const int lp = (int)((k >> N.kshift[p]) & N.kmask[p]);
      td = lp == want;
    }
    for (int t = 0; t < N.nic && td; ++t) {
      const int v = N.ic[3 * t];
      const int x = (int)((k >> N.kshift[N.np + v]) & N.kmask[N.np + v]) + N.int_lo[v];
      td = atom_holds(x, N.ic[3 * t + 1], N.ic[3 * t + 2]);
    }
    *moved = mv; *g1 = h1; *g2 = h2; *fz = fl; *am = a; *rm = r; *key = k; *tdisc = td;
  };
  if (init) {
    // no transition: the initial location vector, its invariant, delay unless frozen, ...
    int td = 1;
    for (int t = 0; t < N.nlr && td; ++t) td = locs[N.lr[2 * t]] == N.lr[2 * t + 1];
    for (int t = 0; t < N.nic && td; ++t) td = atom_holds(ints[N.ic[3 * t]], N.ic[3 * t + 1], N.ic[3 * t + 2]);
    u64 am0 = 0;
    for (int p = 0; p < N.np; ++p) am0 |= samask[p];
    finish(0, -1, -1, 0ull, -1, -1, fz_src, am0, 0ull, dk, td);
    return;
  }
  // the slots reserved for this state bound what is written: a state filtered out after counting
  // (rec_cnt[i] = 0) writes nothing
  const int nr = min(build_recipes<BC>(N, locs, ints, (fz_src & 1) != 0, rec, N.rcap), rec_cnt[i]);
  wsync();
  const int base = (int)rec_off[i];
  for (int b0 = 0; b0 < nr; b0 += kWave) {
    int g1 = -1, g2 = -1, fz = 0, td = 0;
    u64 moved = 0, am = 0, rm = 0, key = 0;
    if (b0 + l < nr) describe(rec[2 * (b0 + l)], rec[2 * (b0 + l) + 1], &moved, &g1, &g2, &fz, &am, &rm, &key, &td);
    const int nb = min(kWave, nr - b0);
    for (int r = 0; r < nb; ++r) {
      const int se = rec[2 * (b0 + r)], re = rec[2 * (b0 + r) + 1];
      int fzr = bcast(fz, r);
      if (!(fzr & 2) && N.has_urgent_chan) {
        // urgent synchronisations read the whole target state: materialise it
        wsync();
        if (l < N.np) nl[l] = locs[l];
        if (l < N.ni) niv[l] = ints[l];
        wsync();
        if (l == 0) {
          auto step = [&](int e) {
            nl[N.edge_proc[e]] = N.edge_tgt[e];
            for (int c = N.iu_off[e]; c < N.iu_off[e + 1]; ++c) niv[N.iu_data[2 * c]] = N.iu_data[2 * c + 1];
          };
          step(se);
          if (re >= 0) step(re);
          if (BC && re <= -2) bcast_walk(N, locs, ints, se, -2 - re, [&](int, int, int rj, int) { if (rj >= 0) step(rj); });
        }
        wsync();
        if (urgent_sync_enabled(N, nl, niv)) fzr |= 2;
      }
      finish(base + b0 + r, se, re, bcast64(moved, r), bcast(g1, r), bcast(g2, r), fzr, bcast64(am, r),
             bcast64(rm, r), bcast64(key, r), bcast(td, r));
    }
  }
}

// ---------------------------------------------------------------- scans

// Exclusive scan, three phases: per-block totals, a scan of the totals, per-block downsweep.
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int* __restrict__ in, long n,
                                                            long long* __restrict__ bsum, const int* __restrict__ np) {
  __shared__ long long s[kScanBlock];
  if (np) n = min(n, (long)*np);   // n an upper bound, the length in device memory
  const long base = (long)blockIdx.x * kScanBlock * kScanItems;
  long long acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const long idx = base + (long)k * kScanBlock + threadIdx.x;
    if (idx < n) acc += in[idx];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kScanBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(long long* __restrict__ bsum, long nb,
                                                          long long* __restrict__ total) {
  // one block: sequential chunks of kScanBlock, Hillis-Steele within a chunk
  __shared__ long long s[kScanBlock];
  long long carry = 0;
  for (long base = 0; base < nb; base += kScanBlock) {
    const long idx = base + threadIdx.x;
    const long long v = idx < nb ? bsum[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
      const long long t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < nb) bsum[idx] = carry + s[threadIdx.x] - v;  // exclusive
    carry += s[kScanBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const int* __restrict__ in, long n,
                                                          const long long* __restrict__ bsum,
                                                          long long* __restrict__ out, const int* __restrict__ np) {
  __shared__ long long s[kScanBlock];
  if (np) n = min(n, (long)*np);
  const long base = (long)blockIdx.x * kScanBlock * kScanItems;
  long long carry = bsum[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const long idx = base + (long)k * kScanBlock + threadIdx.x;
    const long long v = idx < n ? in[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
      const long long t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < n) out[idx] = carry + s[threadIdx.x] - v;
    carry += s[kScanBlock - 1];
    __syncthreads();
  }
}

// The whole scan in one workgroup, for n <= kScanBlock * kScanItems: one launch instead of three.
__global__ __launch_bounds__(kScanBlock) void k_scan_small(const int* __restrict__ in, long n,
                                                           long long* __restrict__ out,
                                                           long long* __restrict__ total, const int* __restrict__ np) {
  __shared__ long long s[kScanBlock];
  if (np) n = min(n, (long)*np);
  long long carry = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const long idx = (long)k * kScanBlock + threadIdx.x;
    const long long v = idx < n ? in[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
      const long long t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (idx < n) out[idx] = carry + s[threadIdx.x] - v;
    carry += s[kScanBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void k_offsets_to_int(const long long* __restrict__ in, long n, int* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int)in[i];
}

// compaction of the non-empty successor slots: sj[pos[r]] = r where valid[r]
// (with `ctr`, also clears the touched-key count for the level's tickets)
__global__ void k_compact_valid(const int* __restrict__ valid, const long long* __restrict__ pos, long R,
                                int* __restrict__ sj, Counters* ctr) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0 && ctr) ctr->ntouched = 0;
  if (r < R && valid[r]) sj[pos[r]] = (int)r;
}

// ---------------------------------------------------------------- passed list: hashing

struct Passed {
  u64* key;        // [H] packed discrete key or kEmptyKey
  u64* off;        // [H] chain segment offset into nodes
  u32* len;        // [H] live chain length
  u32* ncnt;       // [H] successors queued here this level
  u64* noff;       // [H] offset of this level's work list
  u64* nseg;       // [H] chain segment reserved for this level
  u32* nodes;      // chain segments (zone ids, oldest first)
  int* zarena;     // stored zones: d2 int32 entries each, or nstride(d2) int16 entries when nar
  u64 hmask;
  int hshift;
  int nar;         // stored zones are narrow (int16)
  u64 zcap, ncap;
};

// A stored zone as an operand: int32 entries, or narrow ones widened on load.
template <bool NAR> struct StoredZone;
template <> struct StoredZone<false> {
  using T = const int*;
  static __device__ __forceinline__ T at(const Passed& P, u64 zid, int d2) { return P.zarena + zid * (u64)d2; }
};
template <> struct StoredZone<true> {
  using T = Z16;
  static __device__ __forceinline__ T at(const Passed& P, u64 zid, int d2) {
    return Z16{(const short*)P.zarena + zid * (u64)nstride(d2)};
  }
};
// Entry e of stored zone zid, whichever the storage.
__device__ __forceinline__ int stored_entry(const Passed& P, u64 zid, int e, int d2) {
  return P.nar ? widen16(((const short*)P.zarena)[zid * (u64)nstride(d2) + e]) : P.zarena[zid * (u64)d2 + e];
}

__device__ __forceinline__ u64 key_mix(u64 dk) { return dk * 0x9e3779b97f4a7c15ULL; }

__device__ __forceinline__ long find_or_insert(Passed P, u64 dk, bool* inserted) {
  u64 h = key_mix(dk) >> P.hshift;
  for (u64 probe = 0; probe <= P.hmask; ++probe) {
    const u64 cur = __hip_atomic_load(&P.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == dk) return (long)h;
    if (cur == kEmptyKey) {
      const u64 prev = atomicCAS((unsigned long long*)&P.key[h], (unsigned long long)kEmptyKey,
                                 (unsigned long long)dk);
      if (prev == kEmptyKey) { *inserted = true; return (long)h; }
      if (prev == dk) return (long)h;
    }
    h = (h + 1) & P.hmask;
  }
  return -1;
}

// successor -> passed-list slot of its discrete key (inserting new keys)
// (with `survived`, also clears the successor's survivor id and frontier flag)
__global__ void k_find(Passed P, const int* __restrict__ sj, const u64* __restrict__ sdk, long G,
                       long* __restrict__ gbk, Counters* ctr, int* __restrict__ survived, int* __restrict__ flag) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool ins = false;
  if (g < G && survived) {
    survived[g] = -1;
    flag[g] = 0;
  }
  if (g < G) {
    const long h = find_or_insert(P, sdk[sj[g]], &ins);
    if (h < 0) atomicExch(&ctr->err, ERR_HASH_FULL);
    gbk[g] = h;
  }
  wave_stat(ctr, C_KEYS, ins ? 1ull : 0ull);
}

// Cover pre-filter. A successor covered (a_LU) by a zone its key held before this level is
// dropped by the sequential add() whatever else the level does: a stored zone is only retired by
// a zone that covers it (a kill) or that contains it (a convex-union hull), simulation is
// transitive and inclusion implies it, so at the successor's turn the chain still holds a cover
// (the argument of NetPassedT::probe_kill, include/hora/detail/passed_set.hpp:233-234). Such
// successors are decided here, in parallel; the ordered merge sees only the rest.
//
// Four 16-lane groups per wavefront, one successor per group, lanes over matrix entries; the
// group walks the key's previous-level chain oldest first (whether some zone covers the successor
// does not depend on the order; on the fischer variant's long chains the oldest zones are the
// likelier covers: 30 % fewer zones read at N = 10).
//
// The successor's side of the a_LU test is fixed for the whole walk, so each lane reduces its
// entries (y, x) once: with c = (<, -L_y) strict and M[y][x] < S[y][x] <= inf, M[y][x] + c < S[0][x]
// reads (M[y][x] & ~1) < S[0][x] - c. An entry the test skips gets S[y][x] = INT_MIN, which no
// stored entry is below. Up to kProbeRegs entries per lane (d*d <= 192) stay in registers, and a
// chain step is one load, one mask and two compares per entry. The per-successor bounds take one
// load per (process, clock) pair, issued together.
constexpr int kGroup = 16;
constexpr int kProbeRegs = 12;

template <bool NAR>
__global__ __launch_bounds__(64) void k_probe(const DevNet* __restrict__ Np, Passed P, const long* __restrict__ gbk,
                                              const int* __restrict__ sj, const u64* __restrict__ sdk,
                                              const int* __restrict__ szone, long G, int* __restrict__ covered,
                                              Counters* ctr) {
  __shared__ int Lg[4][HORA_MAX_DIM], Ug[4][HORA_MAX_DIM], gl[4][HORA_MAX_PROCS];
  const DevNet& N = *Np;
  const int lane = threadIdx.x, q = lane / kGroup, t = lane % kGroup;
  const long g = (long)blockIdx.x * 4 + q;
  const bool active = g < G;
  const int d = N.d, d2 = N.d2, np = N.np;
  long h = active ? gbk[g] : -1;
  const u32 len = (h >= 0) ? P.len[h] : 0u;
  const bool work = active && len > 0;
  // per-successor a_LU bounds from its locations (state_bounds, engine.cpp:1487-1497): the global
  // location rows first, then each clock's maximum over the processes
  if (work) {
    const u64 dk = sdk[sj[g]];
    for (int p = t; p < np; p += kGroup) gl[q][p] = (N.locbase[p] + (int)((dk >> N.kshift[p]) & N.kmask[p])) * d;
  }
  __syncthreads();
  if (work) {
    for (int c = t; c < d; c += kGroup) {
      int lb = 0, ub = 0;
      if (c > 0) {
        lb = max(kAluNegInf, N.Lq[c]);
        ub = max(kAluNegInf, N.Uq[c]);
        for (int p0 = 0; p0 < np; p0 += 8) {
          int lv[8], uv[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            lv[k] = uv[k] = kAluNegInf;
            if (p0 + k < np) {
              const int b = gl[q][p0 + k] + c;
              lv[k] = N.lloc[b];
              uv[k] = N.uloc[b];
            }
          }
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            lb = max(lb, lv[k]);
            ub = max(ub, uv[k]);
          }
        }
      }
      Lg[q][c] = lb;
      Ug[q][c] = ub;
    }
  }
  __syncthreads();
  const int* S = active ? szone + (size_t)sj[active ? g : 0] * d2 : szone;
  const u64 gmask = 0xffffull << (q * kGroup);
  bool cov = false;
  int reads = 0;
  const int* L = Lg[q];
  const int* U = Ug[q];
  // the successor's terms of entry (y, x): the bound M must stay below, and S[y][x] (INT_MIN: skip)
  auto terms = [&](int e, int* se, int* th) __attribute__((always_inline)) {
    *se = INT_MIN;
    *th = 0;
    if (!work || e >= d2) return;
    const int y = ediv(e, d), x = e - y * d;
    if (x == y) return;
    const int ux = U[x], ly = L[y];
    if (ux == kAluNegInf || ly == kAluNegInf) return;
    const int s0x = S[x];
    if (s0x < r_le(-ux)) return;
    *se = S[e];
    *th = s0x - r_lt(-ly);
  };
  if (d2 <= kProbeRegs * kGroup) {
    const int nper = (d2 + kGroup - 1) / kGroup;   // group-uniform
    int se[kProbeRegs], th[kProbeRegs], ec[kProbeRegs];
#pragma unroll
    for (int k = 0; k < kProbeRegs; ++k) {
      const int e = t + k * kGroup;
      ec[k] = min(e, d2 - 1);
      terms(e, &se[k], &th[k]);
    }
    for (int m = 0; m < (int)len; ++m) {   // group-uniform loop
      const auto M = StoredZone<NAR>::at(P, P.nodes[P.off[h] + m], d2);
      ++reads;
      bool wit = false;
#pragma unroll
      for (int k = 0; k < kProbeRegs; ++k)
        if (k < nper) {
          const int mp = M[ec[k]];
          wit |= (mp < se[k]) & ((mp & ~1) < th[k]);
        }
      const u64 b = __ballot(wit) & gmask;
      if (b == 0) { cov = true; break; }
    }
  } else {
    for (int m = 0; m < (int)len; ++m) {   // group-uniform loop
      const auto M = StoredZone<NAR>::at(P, P.nodes[P.off[h] + m], d2);
      ++reads;
      bool wit = false;
      for (int e = t; e < d2; e += kGroup) {
        int se, th;
        terms(e, &se, &th);
        const int mp = M[e];
        wit |= (mp < se) & ((mp & ~1) < th);
      }
      const u64 b = __ballot(wit) & gmask;
      if (b == 0) { cov = true; break; }
    }
  }
  if (active && t == 0) covered[g] = cov ? 1 : 0;
  const bool lead = active && t == 0;
  wave_stat(ctr, C_PREADS, lead ? (unsigned long long)reads : 0ull);
  wave_stat(ctr, C_PSUB, (lead && cov) ? 1ull : 0ull);
}

// tickets in the per-key work lists, for the successors the pre-filter kept
__global__ void k_ticket(Passed P, const long* __restrict__ gbk, const int* __restrict__ covered, long G,
                         u32* __restrict__ gticket, u32* __restrict__ touched, Counters* ctr) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long h = g < G ? gbk[g] : -1;
  bool first = false;
  if (h >= 0 && !covered[g]) {
    const u32 t = atomicAdd(&P.ncnt[h], 1u);
    gticket[g] = t;
    first = t == 0;
  }
  const int idx = block_append(&ctr->ntouched, first);
  if (first) touched[idx] = (u32)h;
}

__host__ __device__ inline u64 list_cap(u64 n) {
  if (n <= (u64)kSortLds) return n;
  u64 c = 1;
  while (c < n) c <<= 1;
  return c;
}

// Per touched key: its work-list capacity, and the chain segment it reserves for this level
// (twice the working chain, so that a working chain longer than the LDS copy fits beside its
// cache-slot table).
// (with `ctr`, also clears the light/heavy key counts)
// T: the touched keys, or with Tp an upper bound and the count *Tp in device memory.
__global__ void k_touched_caps(Passed P, const u32* __restrict__ touched, int T, int* __restrict__ caps,
                               int* __restrict__ segs, Counters* ctr, const int* __restrict__ Tp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0 && ctr) { ctr->nlight = 0; ctr->nheavy = 0; ctr->npair = 0; }
  if (Tp) T = min(T, *Tp);
  if (t >= T) return;
  const u32 h = touched[t];
  caps[t] = (int)list_cap(P.ncnt[h]);
  segs[t] = 2 * (int)(P.len[h] + P.ncnt[h]);
}

__global__ void k_touched_offsets(Passed P, const u32* __restrict__ touched, int T,
                                  const long long* __restrict__ off, const long long* __restrict__ soff,
                                  u64 seg_base) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  P.noff[touched[t]] = (u64)off[t];
  P.nseg[touched[t]] = seg_base + (u64)soff[t];
}

__global__ void k_scatter(Passed P, const long* __restrict__ gbk, const int* __restrict__ covered,
                          const u32* __restrict__ gticket, long G, int* __restrict__ wlist) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const long h = gbk[g];
  if (h < 0 || covered[g]) return;
  wlist[P.noff[h] + gticket[g]] = (int)g;
}

// ---------------------------------------------------------------- passed list: merge

template <bool kGlobal>
__device__ __forceinline__ void bitonic_sort(int* a, int n) {   // n a power of two, ascending
  const int l = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (kGlobal) bsync(); else wsync();
      for (int i = l; i < n; i += kWave) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int x = a[i], y = a[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { a[i] = y; a[ixj] = x; }
        }
      }
    }
  if (kGlobal) bsync(); else wsync();
}

constexpr u32 kNewTag = 0x40000000u;     // chain entry refers to a successor stored this level
constexpr u32 kWideTag = 0x20000000u;    // ... whose zone a union widened (set under the pair pre-pass)
constexpr u32 kPosMask = kWideTag - 1;   // work-list position of a new member

// Lane-serial a_LU test (same predicate as z_alu_le, one lane per pair of zones): Z <=_aLU Zp.
template <class ZA, class ZB>
__device__ __forceinline__ bool alu_le_lane(ZA Z, ZB Zp, int d, const int* L, const int* U) {
  for (int x = 0; x < d; ++x) {
    const int ux = U[x];
    if (ux == kAluNegInf) continue;
    const int z0x = Z[x];
    if (z0x < r_le(-ux)) continue;
    for (int y = 0; y < d; ++y) {
      if (y == x) continue;
      const int ly = L[y];
      if (ly == kAluNegInf) continue;
      const int zp = Zp[y * d + x];
      if (zp < Z[y * d + x] && radd(zp, r_lt(-ly)) < z0x) return false;
    }
  }
  return true;
}

// The same predicate for d <= 16 where one zone differs per lane and sits in global memory: a
// column of that zone (16 entries) is loaded before any entry of it is compared, so the lane waits
// on memory once per column instead of once per entry. kLaneZ: the per-lane zone is Z (a kill
// test), else Zp (a cover test).
template <bool kLaneZ, class ZA, class ZB>
__device__ __forceinline__ bool alu_le_lane16(ZA Z, ZB Zp, int d, const int* L, const int* U) {
  for (int x = 0; x < d; ++x) {
    const int ux = U[x];
    if (ux == kAluNegInf) continue;
    if (!kLaneZ && Z[x] < r_le(-ux)) continue;
    int col[16];
#pragma unroll
    for (int y = 0; y < 16; ++y) col[y] = y < d ? (kLaneZ ? Z[y * d + x] : Zp[y * d + x]) : 0;
    const int z0x = kLaneZ ? col[0] : Z[x];
    if (kLaneZ && z0x < r_le(-ux)) continue;
#pragma unroll
    for (int y = 0; y < 16; ++y) {
      if (y >= d || y == x) continue;
      const int ly = L[y];
      if (ly == kAluNegInf) continue;
      const int zp = kLaneZ ? Zp[y * d + x] : col[y];
      const int zv = kLaneZ ? col[y] : Z[y * d + x];
      if (zp < zv && radd(zp, r_lt(-ly)) < z0x) return false;
    }
  }
  return true;
}

// Lane-serial convex-union test (the FA x PA formulation of z_try_union) for d <= 16: the entry
// sets live in four 64-bit masks. A gap between the two zones' projections on one clock is checked
// first: it makes the union non-convex, so it only shortens the answer.
template <class ZA, class ZB>
__device__ __forceinline__ bool union_convex_lane(ZA a, ZB b, int d) {
  for (int i = 1; i < d; ++i) {
    const int bu = b[i * d], al = a[i], au = a[i * d], bl = b[i];
    if (bu < kInf && al < kInf && -r_bnd(al) > r_bnd(bu)) return false;
    if (au < kInf && bl < kInf && -r_bnd(bl) > r_bnd(au)) return false;
  }
  const int d2 = d * d;
  // entry sets as four 64-bit words each; every word index below is a compile-time constant
  // after unrolling, so the words stay in registers
  u64 fa[4], pa[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    u64 mf = 0, mp = 0;
    for (int bit = 0; bit < 64; ++bit) {
      const int e = w * 64 + bit;
      if (e >= d2) break;
      const int av = a[e], bv = b[e];
      if (av < bv) mf |= 1ull << bit;
      else if (av > bv) mp |= 1ull << bit;
    }
    fa[w] = mf;
    pa[w] = mp;
  }
  // pass 1: facets of a (FA) crossed into the hull must land in b, checked on PA entries;
  // pass 2: facets of b (PA) must land in a, checked on FA entries
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    auto z = [&](int e) __attribute__((always_inline)) { return pass == 0 ? a[e] : b[e]; };
    auto o = [&](int e) __attribute__((always_inline)) { return pass == 0 ? b[e] : a[e]; };
#pragma unroll
    for (int wf = 0; wf < 4; ++wf)
      for (u64 mf = pass == 0 ? fa[wf] : pa[wf]; mf; mf &= mf - 1) {
        const int f = wf * 64 + __builtin_ctzll(mf);
        const int i = ediv(f, d), j = f - i * d;
        const int neg = r_neg(z(f));
#pragma unroll
        for (int wp = 0; wp < 4; ++wp)
          for (u64 mp = pass == 0 ? pa[wp] : fa[wp]; mp; mp &= mp - 1) {
            const int e = wp * 64 + __builtin_ctzll(mp);
            const int p = ediv(e, d), q = e - p * d;
            const int hpj = max(a[p * d + j], b[p * d + j]);
            const int hiq = max(a[i * d + q], b[i * d + q]);
            if (radd(radd(hpj, neg), hiq) > o(e)) return false;
          }
      }
  }
  return true;
}

// Heavy keys: the pair pre-pass (k_pairs) runs at d <= 16 for keys with at most kPairMaxN kept
// successors whose sequential fold would test at least kPairMinWork pairs (below that the fold is
// short and the pre-pass's launches cost more than they save).
constexpr int kPairTile = 32;
constexpr int kPairMaxN = 1 << 18;
constexpr int kKillSlots = 4;
constexpr int kWideCap = 1024;  // widened members tracked per key in LDS (more: in device memory)
constexpr u64 kPairMinWork = 8192;
__host__ __device__ inline bool pair_eligible(u32 n, u32 old, int d) {
  return d <= 16 && n <= (u32)kPairMaxN && (u64)n * (u64)(old + n) >= kPairMinWork;
}

// LDS layout of k_merge, in ints: S, M, fa, pa (d2 each); L/U; decoded key; scratch; the sorted
// work list; chain ids and cache slots; widened members (heavy keys); the chain-zone cache (cz
// zones at an odd stride).
struct MergeLayout {
  int d2, zs, cz, sort_cap, chain_cap, wide_cap;
  __host__ __device__ int total() const {
    return 4 * d2 + 2 * HORA_MAX_DIM + HORA_MAX_PROCS + HORA_MAX_INTS + 16 + 32 + sort_cap + 2 * chain_cap + wide_cap +
           cz * zs;
  }
};

__host__ __device__ inline MergeLayout merge_layout(int d2, bool heavy) {
  MergeLayout m;
  m.d2 = d2;
  m.zs = d2 | 1;   // odd stride: lanes reading different cached zones hit different banks
  m.sort_cap = heavy ? kSortLds : 64;
  m.chain_cap = heavy ? 2048 : 64;
  m.wide_cap = (heavy && d2 <= 256) ? kWideCap : 0;   // the pair pre-pass runs at d <= 16
  // LDS per workgroup (ints): the light budget leaves the chain cache at its two-zone minimum and
  // lets registers, not LDS, bound the resident wavefronts; both picked by timing fischer 12
  const int budget = heavy ? 38 * 1024 / 4 : 6 * 1024 / 4;
  const int fixed =
      4 * d2 + 2 * HORA_MAX_DIM + HORA_MAX_PROCS + HORA_MAX_INTS + 16 + 32 + m.sort_cap + 2 * m.chain_cap + m.wide_cap;
  m.cz = max(2, min(heavy ? 1024 : 16, (budget - fixed) / m.zs));
  return m;
}

// One wavefront per touched key: the key's kept successors in (state, recipe) order through
// add()'s folds (passed_set.hpp:327-415). Chain entries are previous-level zone ids, or
// kNewTag|g for a zone stored this level (its matrix is in the successor slot g, rewritten there
// when a union widened it). Chain zones are cached in LDS. Short chains test one stored zone at a
// time with lanes over matrix entries; long chains test up to 64 stored zones at once, one per
// lane, and evaluate the 32-zone union window in one round, taking the first success in
// newest-first order as the sequential fold does. At the end the key's live new zones get zone
// ids in one allocation, move to the arena, and are reported as this level's survivors.
//
// With PRE (heavy keys, when the pair pre-pass ran), the work list arrives sorted, and for keys the pair pre-pass covered,
// its answers replace the two long scans: a successor covered by an earlier successor of the key
// is dropped outright, and a kept successor tests coverage only against members that a union
// widened (the others are previous-level zones, which k_probe cleared, or earlier successors as
// they came out of the expansion, which k_pairs cleared); its kills are the recorded ones among
// unwidened live members, plus the widened members that it covers now.
// Two instances: without the pair pre-pass (light keys, and heavy keys when it did not run) six
// wavefronts per SIMD measured fastest; with it, five leave the batched lane tests (used there,
// where chains are long and mostly outside the LDS cache) their registers.
// Heavy keys' member scans, shared by the wavefronts of a workgroup. Wavefront 0 posts a task in
// LDS: test the successor S (in LDS) against n members, the widened-member list `wl` (W tasks) or
// the chain positions n-1 .. 0 (C tasks); a cover task sets `flag` if some live member covers S,
// a kill task retires every live member S covers and counts them. Wavefront w takes the blocks of
// 64 members w, w + kMergeHelpWaves, ...; the members' answers are independent, so the split gives
// the scans of one wavefront exactly.
constexpr int kMergeHelpWaves = 8;
enum { kTaskCoverW = 0, kTaskKillW = 1, kTaskCoverC = 2, kTaskKillC = 3, kTaskExit = 4 };

struct MergeTask {
  int kind, n, flag, kills;
  const int* wl;   // widened members' chain positions (W tasks)
  u32* ch;         // chain ids
  const int* cs;   // chain cache slots (>= 0 LDS slot, <= -2 successor slot, -1 stored zone)
};

template <bool NAR>
__device__ __forceinline__ void merge_slice(MergeTask& T, int w, const int* S, const int* Lb, const int* Ub, int d,
                                            int d2, const int* cache, int zs, const int* szone, const Passed& P,
                                            int* zdead, u64 lvl_zbase) {
  const int l = lane_id(), kind = T.kind, n = T.n;
  const bool cover = kind == kTaskCoverW || kind == kTaskCoverC;
  const bool wlist = kind == kTaskCoverW || kind == kTaskKillW;
  for (int b0 = w * kWave; b0 < n; b0 += kMergeHelpWaves * kWave) {
    if (cover && *(volatile int*)&T.flag) break;
    int m = -1;
    if (b0 + l < n) m = wlist ? T.wl[b0 + l] : n - 1 - (b0 + l);
    u32 id = 0;
    bool hit = false;
    if (m >= 0) {
      id = T.ch[m];
      if (!(id & kDead)) {
        const int s = T.cs[m];
        auto test = [&](auto Z) __attribute__((always_inline)) -> bool {
          if (cover) return d <= 16 ? alu_le_lane16<false>(S, Z, d, Lb, Ub) : alu_le_lane(S, Z, d, Lb, Ub);
          return d <= 16 ? alu_le_lane16<true>(Z, S, d, Lb, Ub) : alu_le_lane(Z, S, d, Lb, Ub);
        };
        if (NAR && s == -1) {
          hit = test(Z16{(const short*)P.zarena + (size_t)(id & ~kDead) * nstride(d2)});
        } else {
          const int* z = s >= 0 ? cache + s * zs
                         : s <= -2 ? szone + (size_t)(-(s + 2)) * d2
                                   : P.zarena + (size_t)(id & ~kDead) * d2;
          hit = test(z);
        }
      }
    }
    const u64 b = __ballot(hit);
    if (cover) {
      if (b && l == 0) atomicOr(&T.flag, 1);
    } else {
      if (hit) {
        // a stored zone an earlier chunk of this level stored leaves the next frontier
        if (zdead && !(id & (kNewTag | kDead)) && (u64)id >= lvl_zbase) zdead[id - lvl_zbase] = 1;
        T.ch[m] = id | kDead;
      }
      if (b && l == 0) atomicAdd(&T.kills, __popcll(b));
    }
  }
}

template <bool PRE, bool NAR>
__global__ __launch_bounds__(PRE ? kMergeHelpWaves * 64 : 64) __attribute__((amdgpu_waves_per_eu(PRE ? 5 : 6))) void k_merge(const DevNet* __restrict__ Np, Passed P, const u32* __restrict__ list, int nlist,
                                              int* __restrict__ wlist, const int* __restrict__ sj,
                                              int* __restrict__ szone, int* __restrict__ survived,
                                              Counters* ctr, int heavy, int* __restrict__ zscr,
                                              int* __restrict__ used, int* __restrict__ flag,
                                              const int* __restrict__ pdrop, const int* __restrict__ kcnt,
                                              const u32* __restrict__ kslot, const long long* __restrict__ koff,
                                              const u32* __restrict__ klist, int* __restrict__ mpos,
                                              int* __restrict__ zdead, u64 lvl_zbase, int* __restrict__ wlg) {
  extern __shared__ int sm[];
  const DevNet& N = *Np;
  const int d = N.d, d2 = N.d2, l = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const MergeLayout ml = merge_layout(d2, heavy != 0);
  int* S = sm;
  int* M = S + d2;
  int* fa = M + d2;
  int* pa = fa + d2;
  int* Lb = pa + d2;
  int* Ub = Lb + HORA_MAX_DIM;
  int* locs = Ub + HORA_MAX_DIM;
  int* ints = locs + HORA_MAX_PROCS;
  int* misc = ints + HORA_MAX_INTS;          // [0..1] try_union scratch, [2..7] allocations
  int* cand = misc + 16;                      // union window positions
  int* sbuf = cand + 32;
  u32* chl = (u32*)(sbuf + ml.sort_cap);
  int* cslot = (int*)(chl + ml.chain_cap);
  int* wl = cslot + ml.chain_cap;             // widened members' chain positions
  int* cache = wl + ml.wide_cap;
  const int zs = ml.zs, cz = ml.cz;
  if ((int)blockIdx.x >= nlist) return;
  // Heavy keys (PRE): wavefront 0 runs the fold; the other wavefronts only take slices of its
  // member scans (merge_slice), between two workgroup barriers per scan, and leave when wavefront
  // 0 posts kTaskExit after the fold. Wavefront 0 therefore orders its own global memory with a
  // fence (gsync), never with a workgroup barrier, outside those scans.
  __shared__ MergeTask tk;
  auto obsync = [&]() __attribute__((always_inline)) { if (PRE) gsync(); else bsync(); };
  if (PRE && wid > 0) {
    for (;;) {
      __syncthreads();   // the task is posted
      if (tk.kind == kTaskExit) return;
      merge_slice<NAR>(tk, wid, S, Lb, Ub, d, d2, cache, zs, szone, P, zdead, lvl_zbase);
      __syncthreads();   // every slice is done
    }
  }

  const u32 h = list[blockIdx.x];
  const u32 n = P.ncnt[h];
  int* lst = wlist + P.noff[h];
  // order the kept successors by compacted index = (state, recipe) order
  if (PRE) {
    // sorted by k_presort
    if (n <= (u32)ml.sort_cap) {
      for (int k = l; k < (int)n; k += kWave) sbuf[k] = lst[k];
      lst = sbuf;
    }
  } else if (n <= (u32)kWave) {
    const int v = (int)l < (int)n ? lst[l] : 0x7fffffff;
    int rank = 0;
    for (int j = 0; j < (int)n; ++j) rank += (__shfl(v, j) < v) ? 1 : 0;   // lanes from n on hold the maximum
    if ((int)l < (int)n) sbuf[rank] = v;
    lst = sbuf;
  } else if (n <= (u32)ml.sort_cap) {
    int p2 = 1;
    while (p2 < (int)n) p2 <<= 1;
    for (int k = l; k < p2; k += kWave) sbuf[k] = k < (int)n ? lst[k] : 0x7fffffff;
    bitonic_sort<false>(sbuf, p2);
    lst = sbuf;
  } else {
    const u64 cap = list_cap(n);
    for (u64 k = n + l; k < cap; k += kWave) lst[k] = 0x7fffffff;
    bitonic_sort<true>(lst, (int)cap);
  }

  // a_LU bounds of this key (state_bounds, engine.cpp:1487-1497)
  decode_key(N, P.key[h], locs, ints);
  wsync();
  if (l < d) {
    int lb = 0, ub = 0;
    if (l > 0) {
      lb = ub = kAluNegInf;
      for (int p = 0; p < N.np; ++p) {
        const size_t b = (size_t)(N.locbase[p] + locs[p]) * d + l;
        lb = max(lb, N.lloc[b]);
        ub = max(ub, N.uloc[b]);
      }
      lb = max(lb, N.Lq[l]);
      ub = max(ub, N.Uq[l]);
    }
    Lb[l] = lb;
    Ub[l] = ub;
  }

  // working chain: previous level's live chain, then one entry per kept successor; it lives in
  // LDS, or, when longer than the LDS copy, in the chain segment this key reserved for the level
  const u32 old_len = P.len[h];
  const u64 old_off = P.off[h];
  const u64 seg = P.nseg[h];
  u32* ch = chl;
  int* cs = cslot;
  const bool chg = old_len + n > (u32)ml.chain_cap;
  if (chg) {
    ch = P.nodes + seg;
    cs = (int*)(P.nodes + seg + old_len + n);
  }
  // ordering after a chain write: the chain may be in global memory
  auto csync = [&]() __attribute__((always_inline)) { if (chg) obsync(); else wsync(); };
  // cache the newest min(old_len, cz) stored zones; the rest of the slots start free
  const int ncached = min((int)old_len, cz);
  for (u32 k = l; k < old_len; k += kWave) {
    ch[k] = P.nodes[old_off + k];
    const int slot = (int)k - ((int)old_len - ncached);
    cs[k] = slot >= 0 ? slot : -1;
  }
  csync();
  for (int k = 0; k < ncached; ++k) {
    const auto src = StoredZone<NAR>::at(P, ch[old_len - ncached + k], d2);
    for (int e = l; e < d2; e += kWave) cache[k * zs + e] = src[e];
  }
  // cache slots are handed out upwards to new members until the cache is full
  int next_slot = ncached;
  // new members are tagged with their position k in the sorted work list. A member's cache-slot
  // entry is its LDS slot (>= 0), or -2 - j for a new member whose zone is in successor slot j (its
  // own, rewritten there when a union widened it), or -1 for a stored zone read from the arena:
  // no chain of dependent loads to find a zone.
  auto zp = [&](int pos) __attribute__((always_inline)) -> const int* {
    const int s = cs[pos];
    if (s >= 0) return cache + s * zs;
    if (s <= -2) return szone + (size_t)(-(s + 2)) * d2;
    return P.zarena + (size_t)(ch[pos] & ~kDead) * d2;
  };
  // f applied to member pos's zone. With narrow storage an uncached stored member is read through
  // Z16 and every other member as int32: a branch on the member's cache-slot entry (uniform for a
  // wavefront-wide test; per lane for the lane-serial ones, which take no cross-lane step).
  auto zwith = [&](int pos, auto&& f) __attribute__((always_inline)) {
    if constexpr (NAR) {
      if (cs[pos] == -1) return f(Z16{(const short*)P.zarena + (size_t)(ch[pos] & ~kDead) * nstride(d2)});
    }
    return f(zp(pos));
  };
  // a stored zone that an earlier chunk of this level stored, retired here, leaves the next frontier
  auto note_dead = [&](u32 id) __attribute__((always_inline)) {
    if (zdead && !(id & (kNewTag | kDead)) && (u64)id >= lvl_zbase) zdead[id - lvl_zbase] = 1;
  };
  int len = (int)old_len, live = (int)old_len;
  unsigned long long n_sub = 0, n_kill = 0, n_uni = 0, n_reads = 0, n_ins = 0;
  // pair pre-pass answers: `use_pairs` for the drops (exact on their own), `pairs_full` while the
  // widened-member list is complete
  const bool use_pairs = PRE && pair_eligible(n, old_len, d);
  bool pairs_full = use_pairs;
  unsigned long long n_full = 0, n_wovf = 0;
  int nw = 0;
  // the widened members' chain positions: in LDS, or, past its wide_cap entries, in the key's
  // slice of wlg (one int per kept successor, as many as the key can widen)
  int* wlp = wl;
  const u64 pbase = P.noff[h];
  if (use_pairs)
    for (u32 k = l; k < n; k += kWave) mpos[pbase + k] = -1;
  obsync();

  // a member scan of the current successor split over the workgroup's wavefronts (PRE only):
  // returns the cover flag or the number of members retired
  auto helped = [&](int kind, int cnt) __attribute__((always_inline)) -> int {
    wsync();
    if (l == 0) {
      tk.kind = kind;
      tk.n = cnt;
      tk.flag = 0;
      tk.kills = 0;
      tk.wl = wlp;
      tk.ch = ch;
      tk.cs = cs;
    }
    __syncthreads();   // the task is posted
    merge_slice<NAR>(tk, 0, S, Lb, Ub, d, d2, cache, zs, szone, P, zdead, lvl_zbase);
    __syncthreads();   // every slice is done
    return (kind == kTaskCoverW || kind == kTaskCoverC) ? tk.flag : tk.kills;
  };
  for (u32 k = 0; k < n; ++k) {
    if (use_pairs && pdrop[pbase + k]) {
      ++n_sub;
      continue;
    }
    if (PRE && !pairs_full) ++n_full;
    const int g = lst[k];
    int* sz = szone + (size_t)sj[g] * d2;
    wsync();
    for (int e = l; e < d2; e += kWave) S[e] = sz[e];
    wsync();
    bool widened = false;
    if (live > 0) {
      const bool wide = live > 8;
      // (1)+(2) covered by a live stored zone: drop
      bool covered = false;
      if (PRE && (pairs_full || wide)) {
        covered = helped(pairs_full ? kTaskCoverW : kTaskCoverC, pairs_full ? nw : len) != 0;
        n_reads += (unsigned long long)(pairs_full ? nw : live);
      } else if (!wide) {
        for (int m = len - 1; m >= 0 && !covered; --m) {
          if (ch[m] & kDead) continue;
          ++n_reads;
          covered = zwith(m, [&](auto Z) { return z_alu_le(S, Z, d, Lb, Ub); });
        }
      } else {
        for (int base = len - 1; base >= 0 && !covered; base -= kWave) {
          const int m = base - l;
          bool c = false;
          if (m >= 0 && !(ch[m] & kDead))
            c = zwith(m, [&](auto Z) { return alu_le_lane(S, Z, d, Lb, Ub); });
          covered = __any(c);
        }
        n_reads += (unsigned long long)live;
      }
      if (covered) { ++n_sub; continue; }
      // (3) retire stored zones covered by S
      if (pairs_full) {
        // recorded kills among live unwidened members, then the widened members as they are now
        const int kc = kcnt[pbase + k];
        if (kc > 0) obsync();   // member positions that lane 0 wrote to global memory
        const u32* kl = kc <= kKillSlots ? kslot + (pbase + k) * kKillSlots : klist + koff[pbase + k];
        for (int b0 = 0; b0 < kc; b0 += kWave) {
          bool kill = false;
          int m = -1;
          if (b0 + l < kc) {
            const u32 tag = kl[b0 + l];
            m = (tag & kNewTag) ? mpos[pbase + (tag & kPosMask)] : (int)tag;
            if (m >= 0) {
              const u32 id = ch[m];
              kill = !(id & kDead) && !(id & kWideTag);
            }
          }
          const u64 b = __ballot(kill);
          wsync();
          if (kill) {
            note_dead(ch[m]);
            ch[m] |= kDead;
          }
          csync();
          live -= __popcll(b);
          n_kill += __popcll(b);
        }
        n_reads += (unsigned long long)kc;
      } else if (!wide) {
        for (int m = len - 1; m >= 0; --m) {
          const u32 id = ch[m];
          if (id & kDead) continue;
          ++n_reads;
          if (zwith(m, [&](auto Z) { return z_alu_le(Z, S, d, Lb, Ub); })) {
            wsync();
            if (l == 0) {
              note_dead(id);
              ch[m] = id | kDead;
            }
            csync();
            --live;
            ++n_kill;
          }
        }
      } else if (PRE) {
        // the chain's live members, in the helped scan below
      } else {
        for (int base = len - 1; base >= 0; base -= kWave) {
          const int m = base - l;
          bool kill = false;
          if (m >= 0 && !(ch[m] & kDead))
            kill = zwith(m, [&](auto Z) { return alu_le_lane(Z, S, d, Lb, Ub); });
          const u64 b = __ballot(kill);
          wsync();
          if (kill) {
            note_dead(ch[m]);
            ch[m] |= kDead;
          }
          csync();
          live -= __popcll(b);
          n_kill += __popcll(b);
        }
        n_reads += (unsigned long long)live;
      }
      if (PRE && (pairs_full || wide)) {
        // the widened members as they are now (pair pre-pass), or every live chain member
        const int kw = helped(pairs_full ? kTaskKillW : kTaskKillC, pairs_full ? nw : len);
        live -= kw;
        n_kill += kw;
        n_reads += (unsigned long long)(pairs_full ? nw : live);
      }
      // (4) convex union with the kMergeScan newest live zones, restarting after each merge
      for (bool merged = true; merged;) {
        merged = false;
        wsync();
        // window positions, newest first
        if (!PRE || len <= kWave) {
          if (l == 0) {
            int c = 0;
            for (int m = len - 1; m >= 0 && c < kMergeScan; --m)
              if (!(ch[m] & kDead)) cand[c++] = m;
            misc[6] = c;
          }
        } else {
          // a heavy key's long chain: the newest live members, 64 chain positions per step (lane l
          // takes position base - l, so lower lanes are newer and the ranks keep newest-first order)
          int c = 0;
          for (int base = len - 1; base >= 0 && c < kMergeScan; base -= kWave) {
            const int m = base - l;
            const bool lv = m >= 0 && !(ch[m] & kDead);
            const u64 b = __ballot(lv);
            const int r = c + __popcll(b & ((1ull << l) - 1ull));
            if (lv && r < kMergeScan) cand[r] = m;
            c += __popcll(b);
          }
          if (l == 0) misc[6] = min(c, kMergeScan);
        }
        wsync();
        const int nc = misc[6];
        if (nc == 0) break;
        n_reads += (unsigned long long)nc;
        int win = -1;
        if (wide && d <= 16) {
          bool ok = false;
          if (l < nc) ok = zwith(cand[l], [&](auto Z) { return union_convex_lane(S, Z, d); });
          const u64 b = __ballot(ok);
          if (b) win = __builtin_ctzll(b);
          if (win >= 0) {
            wsync();
            zwith(cand[win], [&](auto Z) {
              for (int e = l; e < d2; e += kWave) S[e] = max(S[e], Z[e]);
              return 0;
            });
          }
        } else {
          for (int c = 0; c < nc && win < 0; ++c) {
            wsync();
            zwith(cand[c], [&](auto Z) {
              for (int e = l; e < d2; e += kWave) M[e] = Z[e];
              return 0;
            });
            if (z_try_union(S, M, d, fa, pa, misc)) win = c;
          }
        }
        if (win >= 0) {
          const int m = cand[win];
          wsync();
          if (l == 0) {
            note_dead(ch[m]);
            ch[m] |= kDead;
          }
          csync();
          --live;
          ++n_uni;
          widened = true;
          merged = true;
        }
      }
    }
    // (5) store S as the newest member: in a free cache slot when one is left, and back in its
    // successor slot when a union widened it
    wsync();
    int slot = -1;
    if (next_slot < cz) slot = next_slot++;
    if (slot >= 0)
      for (int e = l; e < d2; e += kWave) cache[slot * zs + e] = S[e];
    if (widened) {
      for (int e = l; e < d2; e += kWave) sz[e] = S[e];
      obsync();   // the slot is read back by other lanes when the zone is not cached
    }
    if (l == 0) {
      ch[len] = kNewTag | k | ((use_pairs && widened) ? kWideTag : 0u);
      cs[len] = slot >= 0 ? slot : -2 - sj[g];
      if (use_pairs) mpos[pbase + k] = len;
    }
    if (pairs_full && widened) {
      if (wlp == wl && nw == ml.wide_cap) {
        // full: keep the members still live
        int w2 = 0;
        for (int b0 = 0; b0 < nw; b0 += kWave) {
          const int m = b0 + l < nw ? wl[b0 + l] : -1;
          const bool keep = m >= 0 && !(ch[m] & kDead);
          const u64 b = __ballot(keep);
          const int r = __popcll(b & ((1ull << l) - 1ull));
          wsync();
          if (keep) wl[w2 + r] = m;
          wsync();
          w2 += __popcll(b);
        }
        nw = w2;
        if (nw == ml.wide_cap && wlg) {
          // still full: the list moves to the key's slice of wlg (an entry per kept successor, and
          // each successor widens at most one member, its own)
          int* gl = wlg + pbase;
          for (int b0 = l; b0 < nw; b0 += kWave) gl[b0] = wl[b0];
          obsync();
          wlp = gl;
        }
      }
      if (wlp != wl || nw < ml.wide_cap) {
        if (l == 0) wlp[nw] = len;
        ++nw;
        if (wlp != wl) obsync();   // every lane reads the list from device memory
      } else {
        pairs_full = false;
        n_wovf = 1;
      }
    }
    ++len;
    ++live;
    ++n_ins;
    csync();
  }
  if (PRE) {
    // the helper wavefronts leave
    wsync();
    if (l == 0) tk.kind = kTaskExit;
    __syncthreads();
  }
  // Live new zones go to the level's scratch slot of their work-list position (noff + k), named
  // kNewTag | slot in the chain and in `survived` until k_zassign gives them compact zone ids; the
  // chain is compacted oldest first into this key's reserved segment.
  wsync();
  u32* out = P.nodes + seg;
  const u32 sk = (u32)P.noff[h];
  int w = 0;
  for (int m = 0; m < len; ++m) {
    const u32 id = ch[m];
    if (id & kDead) continue;
    if (id & kNewTag) {
      const u32 kpos = id & kPosMask;
      const u32 slot = sk + kpos;
      const int* src = zp(m);   // a new member: cached or in its successor slot
      int* dst = zscr + (size_t)slot * d2;
      for (int e = l; e < d2; e += kWave) dst[e] = src[e];
      wsync();
      if (l == 0) {
        out[w] = kNewTag | slot;
        survived[lst[kpos]] = (int)(kNewTag | slot);
        flag[lst[kpos]] = 1;
        used[slot] = 1;
      }
    } else {
      wsync();
      if (l == 0) out[w] = id;
    }
    ++w;
  }
  if (l == 0) {
    P.off[h] = seg;
    P.len[h] = (u32)w;
    P.ncnt[h] = 0;
    stat_add(ctr, C_SUB, n_sub);
    stat_add(ctr, C_KILL, n_kill);
    stat_add(ctr, C_UNI, n_uni);
    stat_add(ctr, C_READS, n_reads);
    stat_add(ctr, C_INS, n_ins);
    stat_add(ctr, C_LIVE, (unsigned long long)((long long)w - (long long)old_len));
    if (n_full) stat_add(ctr, C_FULL, n_full);
    if (n_wovf) stat_add(ctr, C_WOVF, n_wovf);
  }
}

// ---------------------------------------------------------------- heavy keys: pair pre-pass

// Workgroup-wide bitonic sort, ascending, n a power of two (LDS or this workgroup's global segment).
__device__ void block_bitonic(int* a, int n) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      __syncthreads();
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int x = a[i], y = a[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { a[i] = y; a[ixj] = x; }
        }
      }
    }
  __syncthreads();
}

constexpr int kPresortLds = 4096;

// Each heavy key's work list sorted in place into compacted-index = (state, recipe) order, one
// workgroup per key; k_pairs and k_merge then read it as it is.
__global__ __launch_bounds__(1024) void k_presort(Passed P, const u32* __restrict__ list, int nlist,
                                                  int* __restrict__ wlist) {
  __shared__ int s[kPresortLds];
  if ((int)blockIdx.x >= nlist) return;
  const u32 h = list[blockIdx.x];
  const int n = (int)P.ncnt[h];
  int* lst = wlist + P.noff[h];
  const int t = threadIdx.x, nt = blockDim.x;
  if (n <= kPresortLds) {
    int p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (int k = t; k < p2; k += nt) s[k] = k < n ? lst[k] : 0x7fffffff;
    block_bitonic(s, p2);
    for (int k = t; k < n; k += nt) lst[k] = s[k];
  } else {
    // n > kSortLds: the list's capacity is a power of two (list_cap)
    const int cap = (int)list_cap((u64)n);
    for (int k = n + t; k < cap; k += nt) lst[k] = 0x7fffffff;
    block_bitonic(lst, cap);
  }
}

// Pair tiles per heavy key: kPairTile successors (rows) against kPairTile candidates (columns),
// the columns being the key's previous-level zones and then its earlier successors (a triangle).
__global__ void k_pair_tiles(Passed P, const u32* __restrict__ list, int nlist, int d, int* __restrict__ tcount) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlist) return;
  const u32 h = list[t];
  const u32 n = P.ncnt[h];
  if (!pair_eligible(n, P.len[h], d)) {
    tcount[t] = 0;
    return;
  }
  const long long rt = (n + kPairTile - 1) / kPairTile, ot = (P.len[h] + kPairTile - 1) / kPairTile;
  tcount[t] = (int)(rt * ot + rt * (rt + 1) / 2);
}

// The pair pre-pass of the ordered merge, all pairs of one heavy key at once. For successor j (in
// work-list order) and
//   a previous-level zone Z of the key: Z <=_aLU S_j records a kill of Z by S_j;
//   an earlier successor S_i (i < j): S_j <=_aLU S_i drops S_j, else S_i <=_aLU S_j records a kill.
// The drop is what the sequential fold (passed_set.hpp:327-415) does to S_j: S_i, stored or
// dropped, is covered by a live zone of the key from its turn on (zones are retired only by zones
// that cover or contain them), and simulation is transitive. The kills are used for members still
// as stored; a member that a union widened is tested again in k_merge.
// Workgroups walk the tiles of all heavy keys (offsets `toff`, total `*ttotal`). Mode 0 decides the
// drops, counts each successor's kills and keeps the first kKillSlots of them; mode 1 runs again
// for the kept successors with more, writing all of their kills to `klist` at `koff`.
__global__ __launch_bounds__(256) void k_pairs(const DevNet* __restrict__ Np, Passed P, const u32* __restrict__ list,
                                               int nlist, const long long* __restrict__ toff,
                                               const long long* __restrict__ ttotal, const int* __restrict__ wlist,
                                               const int* __restrict__ sj, const int* __restrict__ szone,
                                               int* __restrict__ pdrop, int* __restrict__ kcnt,
                                               u32* __restrict__ kslot, int mode, const long long* __restrict__ koff,
                                               int* __restrict__ kpos, u32* __restrict__ klist) {
  extern __shared__ int sm[];
  const DevNet& N = *Np;
  const int d = N.d, d2 = N.d2, zs = d2 | 1, t = threadIdx.x;
  int* Rz = sm;
  int* Cz = Rz + kPairTile * zs;
  int* Lb = Cz + kPairTile * zs;
  int* Ub = Lb + HORA_MAX_DIM;
  int* locs = Ub + HORA_MAX_DIM;
  int* ints = locs + HORA_MAX_PROCS;
  const long long total = *ttotal;
  for (long long tile = blockIdx.x; tile < total; tile += gridDim.x) {
    // owning key: the last one whose tiles start at or before this tile
    int lo = 0, hi = nlist - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (toff[mid] <= tile) lo = mid; else hi = mid - 1;
    }
    const u32 h = list[lo];
    const int n = (int)P.ncnt[h], old = (int)P.len[h];
    const int rtn = (n + kPairTile - 1) / kPairTile, otn = (old + kPairTile - 1) / kPairTile;
    // sibling tiles first, each row's diagonal tile first (a successor is most often covered by
    // one just before it), then the previous-level tiles: rows dropped by then skip them
    const long long u = tile - toff[lo];
    const long long ntri = (long long)rtn * (rtn + 1) / 2;
    int rt, ct;
    bool sib;
    if (u < ntri) {
      const long long s2 = u;
      rt = (int)((sqrt(8.0 * (double)s2 + 1.0) - 1.0) * 0.5);
      while (rt > 0 && (long long)rt * (rt + 1) / 2 > s2) --rt;
      while ((long long)(rt + 1) * (rt + 2) / 2 <= s2) ++rt;
      ct = rt - (int)(s2 - (long long)rt * (rt + 1) / 2);
      sib = true;
    } else {
      const long long u2 = u - ntri;
      rt = (int)(u2 / otn);
      ct = (int)(u2 - (long long)rt * otn);
      sib = false;
    }
    __syncthreads();   // the previous tile's reads of LDS are done
    {
      // a tile whose rows are all dropped already (mode 1: that need no kill list) has nothing to decide
      int live_row = 0;
      if (t < kPairTile) {
        const int j = rt * kPairTile + t;
        const u64 pj = P.noff[h] + (u64)j;
        live_row = j < n && !pdrop[pj] && (mode == 0 || kcnt[pj] > kKillSlots);
      }
      if (__syncthreads_count(live_row) == 0) continue;
    }
    if (t < kWave) decode_key(N, P.key[h], locs, ints);
    __syncthreads();
    if (t < d) {
      int lb = 0, ub = 0;
      if (t > 0) {
        lb = ub = kAluNegInf;
        for (int p = 0; p < N.np; ++p) {
          const size_t b = (size_t)(N.locbase[p] + locs[p]) * d + t;
          lb = max(lb, N.lloc[b]);
          ub = max(ub, N.uloc[b]);
        }
        lb = max(lb, N.Lq[t]);
        ub = max(ub, N.Uq[t]);
      }
      Lb[t] = lb;
      Ub[t] = ub;
    }
    const int* lst = wlist + P.noff[h];
    for (int e = t; e < kPairTile * d2; e += blockDim.x) {
      const int r = e / d2, x = e - r * d2;
      const int j = rt * kPairTile + r, c = ct * kPairTile + r;
      Rz[r * zs + x] = j < n ? szone[(size_t)sj[lst[j]] * d2 + x] : 0;
      int v = 0;
      if (!sib && c < old) v = stored_entry(P, P.nodes[P.off[h] + c], x, d2);
      if (sib && c < n) v = szone[(size_t)sj[lst[c]] * d2 + x];
      Cz[r * zs + x] = v;
    }
    __syncthreads();
    const int r = t >> 3, j = rt * kPairTile + r;
    if (j >= n) continue;
    const u64 pj = P.noff[h] + (u64)j;
    if (pdrop[pj]) continue;   // already dropped by another tile: its kills no longer matter
    const int* Sj = Rz + r * zs;
    if (mode == 1) {
      const int kc = kcnt[pj];
      if (kc <= kKillSlots) continue;
      auto put = [&](u32 tag) {
        const int s = atomicAdd(&kpos[pj], 1);
        if (s < kc) klist[koff[pj] + s] = tag;
      };
      // a kept successor is covered by none of its earlier successors: only the kill direction
      for (int q = 0; q < kPairTile / 8; ++q) {
        const int cc = (t & 7) + 8 * q, c = ct * kPairTile + cc;
        const int* Zc = Cz + cc * zs;
        if (!sib) {
          if (c < old && alu_le_lane(Zc, Sj, d, Lb, Ub)) put((u32)c);
        } else if (c < j && alu_le_lane(Zc, Sj, d, Lb, Ub)) {
          put(kNewTag | (u32)c);
        }
      }
      continue;
    }
    auto record = [&](u32 tag) {
      const int s = atomicAdd(&kcnt[pj], 1);
      if (s < kKillSlots) kslot[pj * kKillSlots + s] = tag;
    };
    for (int q = 0; q < kPairTile / 8; ++q) {
      const int cc = (t & 7) + 8 * q, c = ct * kPairTile + cc;
      const int* Zc = Cz + cc * zs;
      if (!sib) {
        if (c < old && alu_le_lane(Zc, Sj, d, Lb, Ub)) record((u32)c);
      } else if (c < j) {
        if (alu_le_lane(Sj, Zc, d, Lb, Ub)) {
          pdrop[pj] = 1;
          break;
        }
        if (alu_le_lane(Zc, Sj, d, Lb, Ub)) record(kNewTag | (u32)c);
      }
    }
  }
}

// kill-list lengths: the kept successors whose kills overflow their slots
__global__ void k_kill_need(const int* __restrict__ pdrop, const int* __restrict__ kcnt, long W, int* __restrict__ need) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < W) need[p] = (!pdrop[p] && kcnt[p] > kKillSlots) ? kcnt[p] : 0;
}

// light/heavy classification of touched keys by working-chain size
__global__ void k_classify(Passed P, const u32* __restrict__ touched, int T, u32* __restrict__ light,
                           u32* __restrict__ heavy, Counters* ctr, int light_max, int d, const int* __restrict__ Tp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (Tp) T = min(T, *Tp);
  const u32 h = t < T ? touched[t] : 0u;
  const bool ok = t < T;
  const u32 w = ok ? P.len[h] + P.ncnt[h] : 0u;
  const bool lt = ok && w <= (u32)light_max && P.ncnt[h] <= 64u;
  const int il = block_append(&ctr->nlight, lt);
  const int ih = block_append(&ctr->nheavy, ok && !lt);
  block_append(&ctr->npair, ok && !lt && pair_eligible(P.ncnt[h], P.len[h], d));
  if (lt) light[il] = h;
  if (ok && !lt) heavy[ih] = h;
}

// Zone ids for the level's surviving new zones, compact and in slot order: slot -> zbase + rank.
// One wavefront per touched key walks its chain: each provisional entry kNewTag | slot gets its
// zone id, and the zone moves from the scratch slot into the arena.
// Narrow storage: an entry that does not fit 16 bits sets ERR_NARROW, and the caller searches
// again with int32 storage.
__global__ __launch_bounds__(64) void k_zassign(Passed P, const u32* __restrict__ touched, int T,
                                                const long long* __restrict__ rank, const int* __restrict__ zscr,
                                                u64 zbase, int d2, Counters* ctr) {
  const int t = blockIdx.x;
  if (t >= T) return;
  const u32 h = touched[t];
  u32* c = P.nodes + P.off[h];
  const u32 len = P.len[h];
  const int l = threadIdx.x;
  bool bad = false;
  for (u32 m = 0; m < len; ++m) {
    const u32 id = c[m];
    if (!(id & kNewTag)) continue;
    const u32 slot = id & (kNewTag - 1);
    const u64 zid = zbase + (u64)rank[slot];
    const int* src = zscr + (size_t)slot * d2;
    if (P.nar) {
      const int ns = nstride(d2);
      short* dst = (short*)P.zarena + (size_t)zid * ns;
      for (int e = l; e < ns; e += kWave) {
        const int v = e < d2 ? src[e] : 0;
        bad |= !fits16(v);
        dst[e] = narrow16(v);
      }
    } else {
      int* dst = P.zarena + (size_t)zid * d2;
      for (int e = l; e < d2; e += kWave) dst[e] = src[e];
    }
    if (l == 0) c[m] = (u32)zid;
  }
  if (__any(bad) && l == 0) atomicExch(&ctr->err, ERR_NARROW);
}

// Stored zones [a, b) rewritten from int32 to narrow storage in place, one wavefront per zone.
// Zone z's narrow copy starts at byte 2 * ns * z, below its int32 source at 4 * d2 * z; the
// caller picks rounds in which no zone's copy reaches the source of another zone of the same
// round or of a later one (2 * ns * b <= 4 * d2 * a), and sources of earlier rounds are consumed
// (stream order). A copy may overlap its own source, so the zone is read whole before any write.
__global__ __launch_bounds__(64) void k_narrow_arena(int* arena, u64 a, int d2, Counters* ctr) {
  const u64 z = a + blockIdx.x;
  const int l = threadIdx.x, ns = nstride(d2);
  const int* src = arena + z * (u64)d2;
  int v[HORA_MAX_DIM * HORA_MAX_DIM / kWave];
#pragma unroll
  for (int k = 0; k < HORA_MAX_DIM * HORA_MAX_DIM / kWave; ++k) {
    const int e = l + k * kWave;
    v[k] = e < d2 ? src[e] : 0;
  }
  __syncthreads();
  short* dst = (short*)arena + z * (u64)ns;
  bool bad = false;
#pragma unroll
  for (int k = 0; k < HORA_MAX_DIM * HORA_MAX_DIM / kWave; ++k) {
    const int e = l + k * kWave;
    if (e < ns) {
      bad |= !fits16(v[k]);
      dst[e] = narrow16(v[k]);
    }
  }
  if (__any(bad) && l == 0) atomicExch(&ctr->err, ERR_NARROW);
}

// Chunk boundaries: cuts[k] = first frontier state whose slots start at or after k * budget
// (cuts[0] = 0, cuts[K] = F); `pos` is the exclusive scan of the recipe counts.
__global__ void k_chunk_cuts(const long long* __restrict__ pos, long F, long budget, long K,
                             long long* __restrict__ cuts) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > K) return;
  if (k == 0) { cuts[0] = 0; return; }
  if (k == K) { cuts[K] = F; return; }
  const long long want = (long long)k * budget;
  long lo = 0, hi = F;   // first index in [0, F] with pos >= want (F when none)
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (pos[mid] >= want) hi = mid; else lo = mid + 1;
  }
  cuts[k] = lo;
}

// A chunked level's next frontier: the entries whose zone no later chunk retired.
__global__ void k_frontier_live(const u32* __restrict__ zid, long F2, const int* __restrict__ zdead, u64 zbase,
                                int* __restrict__ flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F2) flag[i] = zdead[zid[i] - zbase] ? 0 : 1;
}

__global__ void k_frontier_compact(const u64* __restrict__ dk, const u32* __restrict__ zid, long F2,
                                   const int* __restrict__ flag, const long long* __restrict__ pos,
                                   u64* __restrict__ dk_out, u32* __restrict__ zid_out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F2 && flag[i]) {
    dk_out[pos[i]] = dk[i];
    zid_out[pos[i]] = zid[i];
  }
}

// next frontier from survivors (flagged by k_merge), in compacted-successor order
__global__ void k_build_frontier(const int* __restrict__ survived, const long long* __restrict__ pos, long G,
                                 const int* __restrict__ sj, const u64* __restrict__ sdk,
                                 u64* __restrict__ fr_dk, u32* __restrict__ fr_zid,
                                 const long long* __restrict__ rank, u64 zbase) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int id = survived[g];
  if (id < 0) return;
  const long p = pos[g];
  fr_dk[p] = sdk[sj[g]];
  fr_zid[p] = (u32)(zbase + (u64)rank[(u32)id & (kNewTag - 1)]);   // survivors are this level's new zones
}

__global__ void k_rehash(Passed Pold, Passed Pnew, u64 Hold, Counters* ctr) {
  const u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= Hold) return;
  const u64 k = Pold.key[s];
  if (k == kEmptyKey) return;
  bool ins = false;
  const long h = find_or_insert(Pnew, k, &ins);
  if (h < 0) { atomicExch(&ctr->err, ERR_HASH_FULL); return; }
  Pnew.off[h] = Pold.off[s];
  Pnew.len[h] = Pold.len[s];
}

// ---------------------------------------------------------------- zone-op microbenchmark kernels

struct ZopArgs {
  int op, d;
  long n;
  const int* a;
  const int* b;
  const int* arg;
  const int* L;
  const int* U;
  int* out;
  int* flag;
};

// Streaming ops: one thread per matrix entry, coalesced over the whole batch.
__global__ void k_zop_stream(ZopArgs A) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = A.d, d2 = d * d;
  if (e >= A.n * d2) return;
  const long z = e / d2;
  const int k = (int)(e - z * d2);
  const int i = ediv(k, d), j = k - i * d;
  const int* src = A.a + z * d2;
  int v = src[k];
  if (A.op == HORA_ZOP_UP) {
    if (j == 0 && i >= 1) v = kInf;
  } else {  // HORA_ZOP_RESET
    const int c = A.arg[z];
    if (i == c && j == c) v = kLe0;
    else if (i == c) v = src[j];
    else if (j == c) v = src[i * d];
  }
  A.out[e] = v;
  if (k == 0 && A.flag) A.flag[z] = 1;
}

// Entry-parallel ops for d in {4, 8, 16, 32}: one thread per matrix entry, a workgroup holds whole
// zones (256 threads, or 1024 for d = 32), per-zone flags reduced in LDS. Every entry is read once
// from HBM; the few other entries of the same zone an entry needs (row 0, column b, ...) come from
// the caches its neighbours just filled. `ld` = log2(d).
template <int OP>
__global__ __launch_bounds__(1024) void k_zop_flat(ZopArgs A, int ld) {
  __shared__ int zflag[256];
  __shared__ int stage[1024];
  const int d = A.d, d2 = d * d;
  const int zpb = (int)blockDim.x >> (2 * ld);
  const int lz = (int)threadIdx.x >> (2 * ld);
  const int e = (int)threadIdx.x & (d2 - 1);
  const long z = (long)blockIdx.x * zpb + lz;
  const bool act = z < A.n;
  const int i = e >> ld, j = e & (d - 1);
  if ((int)threadIdx.x < zpb) zflag[threadIdx.x] = 1;
  // the workgroup's zones, one coalesced load; entries other than a thread's own come from LDS
  const bool staged = OP != HORA_ZOP_INCLUDES;
  if (staged) stage[threadIdx.x] = act ? A.a[z * d2 + e] : 0;
  __syncthreads();
  const int* a = staged ? stage + lz * d2 : A.a + (act ? z : 0) * d2;
  if (act) {
    if (OP == HORA_ZOP_INCLUDES) {
      if (A.b[z * d2 + e] > a[e]) zflag[lz] = 0;
    } else if (OP == HORA_ZOP_ALU_LE) {
      // Z = a, Zp = b: a witness (y, x) = (i, j) refutes Z <=_aLU Zp
      const int* L = A.L + z * d;
      const int* U = A.U + z * d;
      if (i != j) {
        const int ux = U[j], ly = L[i];
        if (ux != kAluNegInf && ly != kAluNegInf) {
          const int z0x = a[j];
          const int zp = A.b[z * d2 + e];
          if (z0x >= r_le(-ux) && zp < a[e] && radd(zp, r_lt(-ly)) < z0x) zflag[lz] = 0;
        }
      }
    } else if (OP == HORA_ZOP_DOWN) {
      int v = a[e];
      if (i == 0 && j >= 1) {
        v = kLe0;
        for (int k = 1; k < d; ++k) v = min(v, a[k * d + j]);
      }
      A.out[z * d2 + e] = v;
    } else if (OP == HORA_ZOP_FREE) {
      const unsigned m = (unsigned)A.arg[z];
      int v;
      if (i == j) v = kLe0;
      else if ((m >> i) & 1) v = kInf;
      else if ((m >> j) & 1) v = a[i * d];
      else v = a[e];
      A.out[z * d2 + e] = v;
    } else if (OP == HORA_ZOP_CONSTRAIN) {
      const int ca = A.arg[3 * z], cb = A.arg[3 * z + 1], r = A.arg[3 * z + 2];
      int v = a[e];
      if (r < a[ca * d + cb]) {
        if (radd(a[cb * d + ca], r) < kLe0) {
          if (e == 0) zflag[lz] = 0;   // empty: the matrix is left as it was
        } else {
          // the incremental closure of one tightened edge: column b first, then every entry via b
          const int col = min(a[i * d + cb], radd(a[i * d + ca], r));
          v = min(v, radd(col, a[cb * d + j]));
        }
      }
      if (A.out) A.out[z * d2 + e] = v;
    }
  }
  __syncthreads();
  if (act && e == 0 && A.flag) A.flag[z] = zflag[lz];
}

// Emptiness for d in {4, 8, 16, 32}: one thread per diagonal entry, 256 threads per workgroup.
__global__ __launch_bounds__(256) void k_zop_empty(ZopArgs A, int ld) {
  __shared__ int zflag[64];
  const int d = A.d, zpb = 256 >> ld;
  const int lz = (int)threadIdx.x >> ld, i = (int)threadIdx.x & (d - 1);
  const long z = (long)blockIdx.x * zpb + lz;
  if ((int)threadIdx.x < zpb) zflag[threadIdx.x] = 0;
  __syncthreads();
  if (z < A.n && A.a[z * d * d + i * d + i] < kLe0) zflag[lz] = 1;
  __syncthreads();
  if (z < A.n && i == 0 && A.flag) A.flag[z] = zflag[lz];
}

// Compute ops: one wavefront per zone (zone staged in LDS).
__global__ __launch_bounds__(64) void k_zop_wave(ZopArgs A) {
  extern __shared__ int sm[];
  const int d = A.d, d2 = d * d, l = threadIdx.x;
  int* Z = sm;
  int* B = Z + d2;
  int* fa = B + d2;
  int* pa = fa + d2;
  int* misc = pa + d2;
  for (long z = blockIdx.x; z < A.n; z += gridDim.x) {
    const int* src = A.a + z * d2;
    wsync();
    for (int e = l; e < d2; e += kWave) Z[e] = src[e];
    if (A.b) {
      const int* sb = A.b + z * d2;
      for (int e = l; e < d2; e += kWave) B[e] = sb[e];
    }
    wsync();
    int flag = 1;
    switch (A.op) {
      case HORA_ZOP_CLOSE:
        z_close(Z, d);
        flag = z_empty(Z, d) ? 0 : 1;
        break;
      case HORA_ZOP_CONSTRAIN:
        flag = z_constrain(Z, d, A.arg[3 * z], A.arg[3 * z + 1], A.arg[3 * z + 2]) ? 1 : 0;
        break;
      case HORA_ZOP_EXTRA_LU:
        z_extra_lu_plus(Z, d, A.L + z * d, A.U + z * d, misc);
        break;
      case HORA_ZOP_INCLUDES:
        flag = z_includes(Z, B, d) ? 1 : 0;
        break;
      case HORA_ZOP_ALU_LE:
        flag = z_alu_le(Z, B, d, A.L + z * d, A.U + z * d) ? 1 : 0;
        break;
      case HORA_ZOP_UNION:
        flag = z_try_union(Z, B, d, fa, pa, misc) ? 1 : 0;
        break;
      case HORA_ZOP_EMPTY:
        flag = z_empty(Z, d) ? 1 : 0;
        break;
      case HORA_ZOP_DOWN:
        z_down(Z, d);
        break;
      case HORA_ZOP_FREE:
        z_free_many(Z, d, (uint64_t)(uint32_t)A.arg[z], fa);
        break;
    }
    wsync();
    const bool flag_only = A.op == HORA_ZOP_INCLUDES || A.op == HORA_ZOP_ALU_LE || A.op == HORA_ZOP_EMPTY;
    if (A.out && !flag_only) {
      int* dst = A.out + z * d2;
      for (int e = l; e < d2; e += kWave) dst[e] = Z[e];
    }
    if (l == 0 && A.flag) A.flag[z] = flag;
  }
}

// Close and Extra_LU+ (then close) for d in {4, 8, 16, 32} with each zone held in registers:
// G = min(64, d*d) lanes per zone (four zones per wavefront at d = 4), EPL = d*d / G entries per
// lane. Lane t of a zone holds entries t + r*G, r < EPL: column j = t % d of the rows
// t / d + r * (G / d), so each register row of a zone is one coalesced load. Pivot k needs z[k][j]
// (entry k*d + j: lane k*d % G + j, register k*d / G, the same register for every lane) and z[i][k]
// for each held row (lane (t / d) * d + k, register r): 1 + EPL cross-lane reads (ds_bpermute) per
// pivot, where the LDS kernel reads three LDS values per entry. Row and column k do not change in
// pivot k on a zone without negative cycles, so all lanes may read before any writes. Each
// wavefront walks zones with a stride, loading the next group of zones while it closes this one.
// ZW: wavefronts per workgroup (picked by zop_tuned).
template <int D, int OP>
__global__ __launch_bounds__(256) void k_zop_reg(ZopArgs A, long ngroups) {
  constexpr int D2 = D * D, G = D2 < kWave ? D2 : kWave, EPL = D2 / G, ZPW = kWave / G, RS = G / D;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int q = lane / G, t = lane % G, j = t % D, i0 = t / D, gb = q * G;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nw = (long)gridDim.x * (blockDim.x >> 6);
  int cur[EPL], nxt[EPL];
  auto load = [&](long g, int* v) __attribute__((always_inline)) {
    const long z = g * ZPW + q;
    const bool ok = g < ngroups && z < A.n;
    const int* src = A.a + (ok ? z : 0) * D2;
#pragma unroll
    for (int r = 0; r < EPL; ++r) v[r] = ok ? src[t + r * G] : kLe0;
  };
  long g = wave;
  if (g < ngroups) load(g, cur);
  for (; g < ngroups; g += nw) {
    if (g + nw < ngroups) load(g + nw, nxt);
    const long z = g * ZPW + q;
    const bool ok = z < A.n;
    if (OP == HORA_ZOP_EXTRA_LU) {
      // STTT 2006 Def. 6 on row 0 as it was before any change (z_extra_lu_plus)
      const int* Lz = A.L + (ok ? z : 0) * D;
      const int* Uz = A.U + (ok ? z : 0) * D;
      const int r0j = __shfl(cur[0], gb + j);
      const int Uj = Uz[j];
      int nv[EPL];
#pragma unroll
      for (int r = 0; r < EPL; ++r) {
        const int i = i0 + r * RS;
        const int r0i = __shfl(cur[0], gb + i);
        const int v = cur[r];
        int w = v;
        if (i != j && v < kInf) {
          const int Li = Lz[i];
          if (r_bnd(v) > Li) w = kInf;
          else if (-r_bnd(r0i) > Li) w = kInf;
          else if (i != 0 && -r_bnd(r0j) > Uj) w = kInf;
          else if (v < r_lt(-Uj)) w = r_lt(-Uj);
        }
        nv[r] = w;
      }
#pragma unroll
      for (int r = 0; r < EPL; ++r) cur[r] = nv[r];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int kj = __shfl(cur[(k * D) / G], gb + (k * D) % G + j);
      int ik[EPL];
#pragma unroll
      for (int r = 0; r < EPL; ++r) ik[r] = __shfl(cur[r], gb + i0 * D + k);
#pragma unroll
      for (int r = 0; r < EPL; ++r) cur[r] = min(cur[r], radd(ik[r], kj));
    }
    bool neg = false;
#pragma unroll
    for (int r = 0; r < EPL; ++r)
      if (i0 + r * RS == j && cur[r] < kLe0) neg = true;
    const u64 gm = (G == kWave) ? ~0ull : (((1ull << G) - 1) << gb);
    const bool empty = (__ballot(neg) & gm) != 0;
    if (ok) {
      if (A.out) {
        int* dst = A.out + z * D2;
#pragma unroll
        for (int r = 0; r < EPL; ++r) dst[t + r * G] = cur[r];
      }
      if (t == 0 && A.flag) A.flag[z] = (OP == HORA_ZOP_CLOSE) ? (empty ? 0 : 1) : 1;
    }
#pragma unroll
    for (int r = 0; r < EPL; ++r) cur[r] = nxt[r];
  }
}

// try_union for d = 4: one lane per zone pair, 64 pairs per wavefront. Both zones of every
// pair are staged in LDS with one coalesced load (odd stride: lanes reading their own zones hit
// different banks); each lane runs the lane-serial convex-union test (the FA x PA formulation);
// the hull (entrywise maximum) or the unchanged first zone goes back with one coalesced store.
template <int D>
__global__ __launch_bounds__(64) void k_zop_union_lane(ZopArgs A) {
  constexpr int D2 = D * D, ZS = D2 + 1;
  __shared__ int sa[kWave * ZS], sb[kWave * ZS];
  const int l = threadIdx.x;
  const long base = (long)blockIdx.x * kWave;
  const long nz = min((long)kWave, A.n - base);
  for (int e = l; e < nz * D2; e += kWave) {
    const int z = e / D2, x = e - z * D2;
    sa[z * ZS + x] = A.a[base * D2 + e];
    sb[z * ZS + x] = A.b[base * D2 + e];
  }
  __syncthreads();
  int* za = sa + l * ZS;
  const int* zb = sb + l * ZS;
  bool ok = false;
  if (l < nz) {
    ok = union_convex_lane(za, zb, D);
    if (ok)
      for (int x = 0; x < D2; ++x) za[x] = max(za[x], zb[x]);
  }
  __syncthreads();
  if (A.out)
    for (int e = l; e < nz * D2; e += kWave) {
      const int z = e / D2, x = e - z * D2;
      A.out[base * D2 + e] = sa[z * ZS + x];
    }
  if (l < nz && A.flag) A.flag[base + l] = ok ? 1 : 0;
}

// ---------------------------------------------------------------- host side

#define HIP_OK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw HipFail(std::string(#call) + ": " + hipGetErrorString(e_));                \
  } while (0)

struct HipFail : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ScopeFail : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct CapFail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The stream of the context the calling thread works for (one thread per context): a buffer
// released while it works is freed after that stream's queued work, not after the whole device's,
// so that other contexts' streams (the serving slots) keep running.
thread_local hipStream_t tl_stream = nullptr;

struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(tl_stream) { tl_stream = s; }
  ~StreamScope() { tl_stream = prev; }
};

// A bump allocator over caller-provided device memory. While one is installed on the calling
// thread (ArenaScope), buffers take their memory from it instead of hipMalloc and give nothing
// back; growing past its end throws ArenaFull.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at > cap || bytes > cap - at) return nullptr;
    used = at + bytes;
    return base + at;
  }
};
thread_local Arena* tl_arena = nullptr;

struct ArenaFull : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct ArenaScope {
  Arena* prev;
  explicit ArenaScope(Arena* a) : prev(tl_arena) { tl_arena = a; }
  ~ArenaScope() { tl_arena = prev; }
};

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  bool borrowed = false;   // memory from an Arena: never freed here
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p && !borrowed) {
      // work queued on the owner's stream may still read the old buffer
      if (tl_stream) (void)hipStreamSynchronize(tl_stream);
      else (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
    cap = 0;
    borrowed = false;
  }
  T* take(size_t c) {
    if (tl_arena) {
      void* q = tl_arena->take(c * sizeof(T));
      if (!q) throw ArenaFull("the query needs more device scratch than the static scratch holds");
      return (T*)q;
    }
    T* q = nullptr;
    HIP_OK(hipMalloc(&q, c * sizeof(T)));
    return q;
  }
  // scratch semantics: contents are not preserved on growth
  void ensure(size_t n) {
    if (n <= cap && p) return;
    release();
    size_t c = std::max<size_t>(n, 1);
    if (!tl_arena) c = c + c / 4 + 1024;   // headroom, except in a fixed scratch
    p = take(c);
    borrowed = tl_arena != nullptr;
    cap = c;
  }
  // persistent semantics: contents preserved on growth
  void grow_keep(size_t n, hipStream_t s) {
    if (n <= cap && p) return;
    size_t c = std::max<size_t>(n, tl_arena ? n : cap * 2);
    T* q = take(c);
    if (p) {
      HIP_OK(hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
      HIP_OK(hipStreamSynchronize(s));
      if (!borrowed) (void)hipFree(p);
    }
    p = q;
    cap = c;
    borrowed = tl_arena != nullptr;
  }
  // synchronous copy on `s` (the owner's stream, so that other streams keep running)
  void upload(const std::vector<T>& v, hipStream_t s) {
    ensure(v.size());
    if (!v.empty()) {
      HIP_OK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
      HIP_OK(hipStreamSynchronize(s));
    }
  }
};

// The stored-zone arena. It only grows, and at its largest it holds most of the GPU's memory, so
// growing it by copying into a larger buffer (old and new alive at once) would halve what fits.
// Where the device supports virtual memory management, the arena is a reserved address range the
// size of the device's memory into which physical memory is mapped in 256 MiB steps as it grows:
// no copy, no double footprint, and the address never moves. Elsewhere (or inside a fixed
// scratch) it falls back to DBuf's copying growth.
struct GrowArena {
  int* p = nullptr;
  size_t cap = 0;                 // ints mapped (or allocated)
  DBuf<int> plain;                // fallback
  void* va = nullptr;
  size_t reserved = 0, mapped = 0, step = 0;
  std::vector<std::pair<hipMemGenericAllocationHandle_t, size_t>> parts;
  int vmm = -1;                   // -1 unknown, 0 no, 1 yes
  hipMemAllocationProp prop{};
  GrowArena() = default;
  GrowArena(const GrowArena&) = delete;
  GrowArena& operator=(const GrowArena&) = delete;
  ~GrowArena() { release(); }
  bool use_vmm() {
    if (tl_arena) return false;
    if (vmm < 0) {
      vmm = 0;
      int dev = 0, ok = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&ok, hipDeviceAttributeVirtualMemoryManagementSupported, dev) == hipSuccess && ok) {
        prop = hipMemAllocationProp{};
        prop.type = hipMemAllocationTypePinned;
        prop.location.type = hipMemLocationTypeDevice;
        prop.location.id = dev;
        size_t gran = 0, total = 0;
        if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended) == hipSuccess &&
            gran > 0 && hipDeviceTotalMem(&total, dev) == hipSuccess && total > 0) {
          step = ((size_t(256) << 20) + gran - 1) / gran * gran;
          const size_t want = (total + step - 1) / step * step;
          if (hipMemAddressReserve(&va, want, gran, nullptr, 0) == hipSuccess) {
            reserved = want;
            vmm = 1;
          } else {
            (void)hipGetLastError();
            va = nullptr;
          }
        } else {
          (void)hipGetLastError();
        }
      } else {
        (void)hipGetLastError();
      }
    }
    return vmm == 1;
  }
  // map physical memory until `bytes` are backed (at least an eighth more than now, so that a
  // growing search maps in few steps)
  void map_to(size_t bytes) {
    if (bytes > reserved) throw HipFail("zone arena: more than the device's memory");
    if (bytes > mapped) bytes = std::min(reserved, std::max(bytes, mapped + mapped / 8));
    while (mapped < bytes) {
      hipMemGenericAllocationHandle_t h;
      HIP_OK(hipMemCreate(&h, step, &prop, 0));
      if (hipMemMap((char*)va + mapped, step, 0, h, 0) != hipSuccess) {
        (void)hipMemRelease(h);
        throw HipFail("zone arena: mapping failed");
      }
      hipMemAccessDesc acc{};
      acc.location = prop.location;
      acc.flags = hipMemAccessFlagsProtReadWrite;
      HIP_OK(hipMemSetAccess((char*)va + mapped, step, &acc, 1));
      parts.emplace_back(h, step);
      mapped += step;
    }
    p = (int*)va;
    cap = mapped / sizeof(int);
  }
  void ensure(size_t n) {
    if (use_vmm()) { map_to(std::max<size_t>(n, 1) * sizeof(int)); return; }
    plain.ensure(n);
    p = plain.p;
    cap = plain.cap;
  }
  void grow_keep(size_t n, hipStream_t s) {
    if (use_vmm()) {
      // mapping more pages leaves the mapped ones, and any kernel reading them, untouched
      map_to(std::max<size_t>(n, 1) * sizeof(int));
      return;
    }
    plain.grow_keep(n, s);
    p = plain.p;
    cap = plain.cap;
  }
  void release() {
    if (vmm == 1 && va) {
      if (tl_stream) (void)hipStreamSynchronize(tl_stream);
      else (void)hipDeviceSynchronize();
      size_t off = 0;
      for (auto& ph : parts) {
        (void)hipMemUnmap((char*)va + off, ph.second);
        (void)hipMemRelease(ph.first);
        off += ph.second;
      }
      parts.clear();
      (void)hipMemAddressFree(va, reserved);
      va = nullptr;
      reserved = mapped = 0;
      vmm = -1;
    }
    plain.release();
    p = nullptr;
    cap = 0;
  }
};

}  // namespace

struct LiveState;             // liveness pass state (liveness.hpp)
void live_free(LiveState* L);

struct hora_hip_ctx {
  std::string err;
  hipStream_t stream = nullptr;
  int d = 0, d2 = 0, np = 0, ni = 0, nclocks = 0, tl = 0, ne = 0;
  // static tables
  DBuf<int> blob, lloc, uloc;
  DevNet dn{};
  DevNet* dnp = nullptr;   // device copy of dn, refreshed when dn changes
  std::vector<int> h_loc_init, h_int_init, h_kshift, h_int_lo;
  // the liveness lane of the reference takes its base construction here (structurally_empty_d,
  // src/liveness/internal.cpp:59-103): 1 time-open network, 2 one possible discrete state
  int base_route = 0;
  std::vector<u64> h_kmask;
  int has_target_clock = 0;
  // search state
  DBuf<u64> ht_key, ht_off, ht_noff, ht_nseg;
  DBuf<u32> ht_len, ht_ncnt;
  DBuf<u32> nodes;
  GrowArena zarena;
  u64 H = 0;
  int hbits = 0;
  DBuf<u64> fr_dk, fr_dk2;
  DBuf<u32> fr_zid, fr_zid2;
  DBuf<int> rec_cnt, rec_off, s_valid, s_zone, s_tgt, sj, caps, segs, wlist, survived, flag_i, off_i, covered;
  DBuf<u64> s_dk;
  DBuf<long> gbk;
  DBuf<u32> gticket, touched, klight, kheavy;
  DBuf<long long> pos, pos2, bsum, scan_total, zrank;
  DBuf<int> zscr, zused;      // a level's surviving new zones before they get zone ids
  // heavy keys' pair pre-pass: drops, recorded kills and member positions per work-list position;
  // tile counts and offsets per heavy key
  DBuf<int> pdrop, kcnt, mpos, tcnt, kneed, kpos;
  DBuf<int> zdead;            // a chunked level's retirement marks for the zones it stored
  DBuf<int> wlg;              // heavy keys' widened-member lists past their LDS capacity
  DBuf<long long> cuts;       // chunk boundaries of a level
  DBuf<u32> kslot, klist;
  DBuf<long long> toff, koff;
  bool pairs = true;          // HORA_HIP_PAIRS=0 turns the pre-pass off (the fold then scans in full)
  DBuf<Counters> ctr;
  // profile of the last search
  hora_reach_profile prof{};
  bool trace = false;
  hipEvent_t ev[8] = {};
  LiveState* live = nullptr;
  bool znar_ok = false;       // zones may be stored narrow (int16): constants small, no overflow seen
  size_t znar_at = 0;         // int32 arena bytes past which a search switches to narrow storage
  int expand_wpb = 0;         // wavefronts per k_expand workgroup, picked by expand_tuned()
  int expand_wpb_d = 0;       // the zone dimension it was picked for
  void* hpin = nullptr;       // pinned host staging for the small counter reads of every level
};

namespace {

constexpr size_t kPinBytes = 64 * 1024;

// A small device-to-host read through the context's pinned staging buffer: a direct DMA into pinned
// memory, not a copy through the runtime's staging of pageable memory.
// Two reads share one synchronisation (n2 = 0: one read).
void read_dev(hora_hip_ctx* c, void* dst, const void* src, size_t n, void* dst2 = nullptr, const void* src2 = nullptr,
              size_t n2 = 0) {
  if (!c->hpin) HIP_OK(hipHostMalloc(&c->hpin, kPinBytes, hipHostMallocDefault));
  const size_t o2 = (n + 255) & ~(size_t)255;
  if (o2 + n2 > kPinBytes) throw HipFail("read_dev: request larger than the staging buffer");
  HIP_OK(hipMemcpyAsync(c->hpin, src, n, hipMemcpyDeviceToHost, c->stream));
  if (n2) HIP_OK(hipMemcpyAsync((char*)c->hpin + o2, src2, n2, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  std::memcpy(dst, c->hpin, n);
  if (n2) std::memcpy(dst2, (char*)c->hpin + o2, n2);
}

}  // namespace

namespace {

// Exclusive scan of in[0, n) into out, the total into scan_total[slot] (no synchronisation). With
// `np`, n is an upper bound and the length is *np, in device memory (written by an earlier kernel
// of the stream, so the host need not read it first).
void scan_launch(hora_hip_ctx* c, const int* in, long n, long long* out, int slot, const int* np = nullptr) {
  const long per = (long)kScanBlock * kScanItems;
  const long nb = std::max<long>(1, (n + per - 1) / per);
  c->scan_total.ensure(2);
  if (n <= per) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kScanBlock), 0, c->stream, in, n, out, c->scan_total.p + slot, np);
    return;
  }
  c->bsum.ensure(nb);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, c->stream, in, n, c->bsum.p, np);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, c->stream, c->bsum.p, nb, c->scan_total.p + slot);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanBlock), 0, c->stream, in, n, c->bsum.p, out, np);
}

long long scan_exclusive(hora_hip_ctx* c, const int* in, long n, long long* out) {
  scan_launch(c, in, n, out, 0);
  long long total = 0;
  read_dev(c, &total, c->scan_total.p, sizeof(total));
  return total;
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)std::max<long>(1, (n + bs - 1) / bs); }

Passed passed_view(hora_hip_ctx* c) {
  Passed P;
  P.key = c->ht_key.p;
  P.off = c->ht_off.p;
  P.len = c->ht_len.p;
  P.ncnt = c->ht_ncnt.p;
  P.noff = c->ht_noff.p;
  P.nseg = c->ht_nseg.p;
  P.nodes = c->nodes.p;
  P.zarena = c->zarena.p;
  P.hmask = c->H - 1;
  P.hshift = 64 - c->hbits;
  P.nar = c->dn.znar;
  P.zcap = P.nar ? c->zarena.cap * 2 / (size_t)nstride(c->d2) : c->zarena.cap / (size_t)c->d2;
  P.ncap = c->nodes.cap;
  return P;
}

// ints of arena storage for `zones` stored zones in the search's current storage
size_t arena_ints(const hora_hip_ctx* c, u64 zones) {
  return c->dn.znar ? (size_t)(zones * (u64)nstride(c->d2) / 2) : (size_t)(zones * (u64)c->d2);
}

void alloc_table(hora_hip_ctx* c, int hbits) {
  c->hbits = hbits;
  c->H = 1ull << hbits;
  c->ht_key.ensure(c->H);
  c->ht_off.ensure(c->H);
  c->ht_noff.ensure(c->H);
  c->ht_nseg.ensure(c->H);
  c->ht_len.ensure(c->H);
  c->ht_ncnt.ensure(c->H);
  HIP_OK(hipMemsetAsync(c->ht_key.p, 0xff, c->H * sizeof(u64), c->stream));
  HIP_OK(hipMemsetAsync(c->ht_len.p, 0, c->H * sizeof(u32), c->stream));
  HIP_OK(hipMemsetAsync(c->ht_ncnt.p, 0, c->H * sizeof(u32), c->stream));
  HIP_OK(hipMemsetAsync(c->ht_off.p, 0, c->H * sizeof(u64), c->stream));
}

void grow_table(hora_hip_ctx* c) {
  DBuf<u64> okey, ooff;
  DBuf<u32> olen;
  std::swap(okey.p, c->ht_key.p); std::swap(okey.cap, c->ht_key.cap);
  std::swap(ooff.p, c->ht_off.p); std::swap(ooff.cap, c->ht_off.cap);
  std::swap(olen.p, c->ht_len.p); std::swap(olen.cap, c->ht_len.cap);
  const u64 Hold = c->H;
  Passed Pold = passed_view(c);
  Pold.key = okey.p; Pold.off = ooff.p; Pold.len = olen.p;
  c->ht_noff.release();
  c->ht_nseg.release();
  c->ht_ncnt.release();
  alloc_table(c, c->hbits + 1);
  Passed Pnew = passed_view(c);
  hipLaunchKernelGGL(k_rehash, dim3(blocks_for((long)Hold, 256)), dim3(256), 0, c->stream, Pold, Pnew, Hold, c->ctr.p);
  HIP_OK(hipStreamSynchronize(c->stream));
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

// Host view of the device counters: stripes summed.
struct Tally {
  unsigned long long v[C_NSTAT] = {};
  int err = 0, first_target = 0x7fffffff, ntouched = 0, nlight = 0, nheavy = 0;
  unsigned long long operator[](int k) const { return v[k]; }
};

Tally tally_of(const Counters& hc) {
  Tally t;
  for (int s = 0; s < kStripes; ++s)
    for (int k = 0; k < C_NSTAT; ++k) t.v[k] += hc.st[s][k];
  t.err = hc.err;
  t.first_target = hc.first_target;
  t.ntouched = hc.ntouched;
  t.nlight = hc.nlight;
  t.nheavy = hc.nheavy;
  return t;
}

Tally read_counters(hora_hip_ctx* c) {
  Counters hc;
  read_dev(c, &hc, c->ctr.p, sizeof(Counters));
  return tally_of(hc);
}

// A scan and the counters in one synchronisation; `slot1`, if given, receives the total of a scan
// launched before into slot 1.
long long scan_and_counters(hora_hip_ctx* c, const int* in, long n, long long* out, Tally* t,
                            long long* slot1 = nullptr) {
  scan_launch(c, in, n, out, 0);
  long long total[2] = {0, 0};
  Counters hc;
  read_dev(c, total, c->scan_total.p, sizeof(total), &hc, c->ctr.p, sizeof(Counters));
  *t = tally_of(hc);
  if (slot1) *slot1 = total[1];
  return total[0];
}

// Per-level device time and algorithmic HBM bytes (int32 entries, 4 B each):
//   probe: the successor's zone plus each previous-level chain zone it is tested against;
//   merge: each kept successor's zone, each chain zone the folds read, and each surviving new
//          zone copied into the arena (read + write).
void account(hora_hip_ctx* c, const Tally& hc, const Tally& prev, long G, int T, long F2) {
  const double z = 4.0 * c->d2;
  c->prof.probe_ms += ev_ms(c->ev[5], c->ev[6]);
  c->prof.probe_bytes += z * ((double)G + (double)(hc[C_PREADS] - prev[C_PREADS]));
  c->prof.probe_launches += 1;
  c->prof.merge_ms += ev_ms(c->ev[3], c->ev[4]);
  const double kept = (double)G - (double)(hc[C_PSUB] - prev[C_PSUB]);
  c->prof.merge_bytes += z * (kept + (double)(hc[C_READS] - prev[C_READS]) + 2.0 * (double)F2);
  c->prof.merge_launches += T > 0 ? 1 : 0;
  c->prof.other_ms += ev_ms(c->ev[2], c->ev[5]) + ev_ms(c->ev[6], c->ev[3]);
}

// Expansion of frontier states [fa, fa + F); slot offsets and counts are the chunk's own.
void launch_expand(hora_hip_ctx* c, int wpb, long F, long fa, ExpandOut eo, bool init) {
  const size_t lds = (size_t)wpb * expand_lds_ints(c->d2, c->dn.rcap, c->dn.icap) * sizeof(int);
  if (c->dn.has_bcast)
    hipLaunchKernelGGL(k_expand<true>, dim3(blocks_for(F, wpb)), dim3(64 * wpb), lds, c->stream, c->dnp,
                       c->fr_dk.p + fa, c->fr_zid.p + fa, c->zarena.p, (int)F, c->pos.p, c->rec_cnt.p, eo, c->ctr.p,
                       init ? 1 : 0);
  else
    hipLaunchKernelGGL(k_expand<false>, dim3(blocks_for(F, wpb)), dim3(64 * wpb), lds, c->stream, c->dnp,
                       c->fr_dk.p + fa, c->fr_zid.p + fa, c->zarena.p, (int)F, c->pos.p, c->rec_cnt.p, eo, c->ctr.p,
                       init ? 1 : 0);
}

// Wavefronts per k_expand workgroup. More wavefronts per workgroup lift the per-CU workgroup limit on
// resident wavefronts; fewer keep LDS allocation granular. The first level with enough states runs
// the expansion once per candidate (the kernel is idempotent: same slots, same values) and keeps
// the fastest for the context and zone dimension. HORA_HIP_EXPAND_WPB=1|2|4 fixes the choice.
int expand_tuned(hora_hip_ctx* c, long F, long fa, ExpandOut eo, bool init) {
  if (c->expand_wpb > 0 && c->expand_wpb_d == c->d) return c->expand_wpb;
  if (const char* e = std::getenv("HORA_HIP_EXPAND_WPB")) {
    const int w = std::atoi(e);
    if (w == 1 || w == 2 || w == 4) { c->expand_wpb = w; c->expand_wpb_d = c->d; return w; }
  }
  if (init || F < 65536) return 4;
  const int cand[3] = {1, 2, 4};
  float best = 1e30f;
  int pick = 4;
  for (int rep = 0; rep < 2; ++rep)
    for (int w : cand) {
      HIP_OK(hipEventRecord(c->ev[6], c->stream));
      launch_expand(c, w, F, fa, eo, init);
      HIP_OK(hipEventRecord(c->ev[7], c->stream));
      HIP_OK(hipEventSynchronize(c->ev[7]));
      const float ms = ev_ms(c->ev[6], c->ev[7]);
      if (ms < best) { best = ms; pick = w; }
    }
  c->expand_wpb = pick;
  c->expand_wpb_d = c->d;
  if (c->trace) std::fprintf(stderr, "k_expand: %d wavefronts per workgroup (%.3f ms at F=%ld)\n", pick, best, F);
  return pick;
}

// Successor slots per chunk of a level: the transient per-successor buffers (the zone, its key and
// flags, its place in the work lists and the merge's scratch copy) are sized for one chunk, about
// 2 x 4 d^2 + 128 bytes a slot, within 16 GiB. HORA_HIP_CHUNK_SLOTS sets another bound.
long chunk_slots(hora_hip_ctx* c) {
  if (const char* e = std::getenv("HORA_HIP_CHUNK_SLOTS")) {
    const long v = std::atol(e);
    if (v > 0) return v;
  }
  return std::max<long>(1L << 16, (long)((16ull << 30) / (8ull * (u64)c->d2 + 128ull)));
}

// Chunk boundaries of a frontier of F states whose slot offsets (exclusive scan of the recipe
// counts) are in c->pos and total R: cut k is the first state whose slots start at or after
// k * budget. Empty chunks are dropped.
std::vector<long> chunk_cuts(hora_hip_ctx* c, long F, long R, long budget) {
  long K = (R + budget - 1) / budget;
  if (K > 4096) {
    budget = (R + 4095) / 4096;
    K = (R + budget - 1) / budget;
  }
  c->cuts.ensure(K + 1);
  hipLaunchKernelGGL(k_chunk_cuts, dim3(blocks_for(K + 1, 256)), dim3(256), 0, c->stream, c->pos.p, F, budget, K,
                     c->cuts.p);
  std::vector<long long> h(K + 1);
  read_dev(c, h.data(), c->cuts.p, (K + 1) * sizeof(long long));
  std::vector<long> cuts{0};
  for (long k = 1; k <= K; ++k)
    if ((long)h[k] > cuts.back()) cuts.push_back((long)h[k]);
  if (cuts.back() != F) cuts.push_back(F);
  return cuts;
}

// One search; kNarrowRetry when a stored zone did not fit narrow storage.
constexpr int kNarrowRetry = 1000;

// Switch the search to narrow storage: stored zones [0, n) rewritten in place, in rounds
// [a, b) with 2 * ns * b <= 4 * d2 * a (k_narrow_arena), then the device tables told.
void narrow_arena(hora_hip_ctx* c, u64 n) {
  const int d2 = c->d2, ns = nstride(d2);
  for (u64 a = 0; a < n;) {
    const u64 b = std::min<u64>(n, std::max<u64>(a + 1, (u64)(4 * d2) * a / (u64)(2 * ns)));
    for (u64 lo = a; lo < b; lo += (1u << 30)) {
      const u64 cnt = std::min<u64>(b - lo, 1u << 30);
      hipLaunchKernelGGL(k_narrow_arena, dim3((unsigned)cnt), dim3(64), 0, c->stream, c->zarena.p, lo, d2, c->ctr.p);
    }
    a = b;
  }
  c->dn.znar = 1;
  HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));
}

int search_once(hora_hip_ctx* c, int32_t* reachable, hora_reach_stats* st, int nar) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  st->emode_switch_level = -1;
  st->zone_entry_bytes = nar ? 2 : 4;
  *reachable = 0;
  c->prof = hora_reach_profile{};
  Tally prev;
  const int d2 = c->d2;
  c->dn.lu_plus = 0;
  c->dn.znar = nar;
  HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));

  // passed list and arenas; a context searched before starts at the table size it grew to (slot
  // positions do not affect any result: work lists, frontiers and counts follow successor order)
  alloc_table(c, std::max(20, c->hbits));
  c->zarena.ensure(arena_ints(c, (u64)1 << 20));
  c->nodes.ensure(1 << 22);
  c->ctr.ensure(1);
  HIP_OK(hipMemsetAsync(c->ctr.p, 0, sizeof(Counters), c->stream));
  {
    const int ft = 0x7fffffff;
    HIP_OK(hipMemcpyAsync(&c->ctr.p->first_target, &ft, sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  u64 zcount = 0, ncount = 0;   // zone ids and chain-node entries handed out so far

  // the initial state travels through the same successor pipeline as a one-slot level 0
  u64 dk0 = 0;
  for (int p = 0; p < c->np; ++p) dk0 |= (u64)(unsigned)c->h_loc_init[p] << c->h_kshift[p];
  for (int v = 0; v < c->ni; ++v) dk0 |= (u64)(unsigned)(c->h_int_init[v] - c->h_int_lo[v]) << c->h_kshift[c->np + v];
  c->fr_dk.ensure(1);
  c->fr_zid.ensure(1);
  HIP_OK(hipMemcpyAsync(c->fr_dk.p, &dk0, sizeof(u64), hipMemcpyHostToDevice, c->stream));

  long F = 1;
  int level = 0;
  bool init = true;
  bool done = false;
  // successor slots per chunk of the frontier (transient buffers are sized for one chunk)
  const long slot_budget = chunk_slots(c);
  for (;;) {
    // ---- recipe counts and offsets of the whole frontier
    long R = 1;
    c->rec_cnt.ensure(F);
    c->rec_off.ensure(F);
    c->pos.ensure(F + 1);
    if (!init) {
      st->explored += (u64)F;
      ++level;
      hipLaunchKernelGGL(k_count, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, c->fr_dk.p, (int)F,
                         c->rec_cnt.p, c->ctr.p);
      R = scan_exclusive(c, c->rec_cnt.p, F, c->pos.p);   // slot offsets, read by k_expand
      st->expanded += (u64)F;
    } else {
      HIP_OK(hipMemsetAsync(c->pos.p, 0, sizeof(long long), c->stream));
    }
    // The level in chunks of consecutive frontier states, each with at most slot_budget successor
    // slots, expanded and merged one after another. Merging chunk by chunk is the sequential
    // fold's order (successors in (state, recipe) order); the cover pre-filter of a later chunk
    // tests against zones an earlier chunk stored, which precede it in that order too. A zone a
    // later chunk retires leaves the next frontier at the end of the level.
    std::vector<long> cuts{0, F};
    if (R > slot_budget) cuts = chunk_cuts(c, F, R, slot_budget);
    const bool chunked = cuts.size() > 2;
    const u64 lvl_zbase = zcount;
    long F2 = 0;
    size_t zdead_n = 0;   // retirement marks cleared so far this level
    bool found = false;
    for (size_t ci = 0; ci + 1 < cuts.size(); ++ci) {
      const long fa = cuts[ci], Fc = cuts[ci + 1] - cuts[ci];
      long Rc = R;
      if (chunked) {
        hipLaunchKernelGGL(k_count, dim3(blocks_for(Fc, 256)), dim3(256), 0, c->stream, c->dnp, c->fr_dk.p + fa,
                           (int)Fc, c->rec_cnt.p, c->ctr.p);
        Rc = scan_exclusive(c, c->rec_cnt.p, Fc, c->pos.p);
      }
      // ---- expansion
      c->s_valid.ensure(Rc);
      c->s_tgt.ensure(Rc);
      c->s_dk.ensure(Rc);
      c->s_zone.ensure((size_t)Rc * d2);
      ExpandOut eo{c->s_valid.p, c->s_dk.p, c->s_zone.p, c->s_tgt.p, nullptr, nullptr};
      const int wpb = Rc > 0 ? expand_tuned(c, Fc, fa, eo, init) : 1;
      HIP_OK(hipEventRecord(c->ev[0], c->stream));
      if (Rc > 0) launch_expand(c, wpb, Fc, fa, eo, init);
      HIP_OK(hipEventRecord(c->ev[1], c->stream));
      // ---- compaction of non-empty successors
      c->pos.ensure(Rc + 1);
      Tally hc;
      const long Gall = scan_and_counters(c, c->s_valid.p, Rc, c->pos.p, &hc);
      c->prof.expand_ms += ev_ms(c->ev[0], c->ev[1]);
      c->prof.expand_bytes += 4.0 * d2 * (double)((init ? 0 : Fc) + Gall);
      c->prof.expand_launches += 1;
      long G = Gall;
      found = hc.first_target != 0x7fffffff;
      if (found) {
        // The sequential engine inserts every successor that precedes the first target, then
        // stops; generated counts those plus the target itself (engine.cpp:2085-2095).
        long long before = 0;
        if (hc.first_target > 0)
          read_dev(c, &before, c->pos.p + hc.first_target, sizeof(before));
        if (!init) st->generated += (u64)before + 1;
        st->found = 1;
        *reachable = 1;
        G = (long)before;
      } else if (!init) {
        st->generated += (u64)G;
      }
      if (G == 0) {
        prev = hc;
        if (found) break;
        continue;
      }
      c->sj.ensure(Gall);
      hipLaunchKernelGGL(k_compact_valid, dim3(blocks_for(Rc, 256)), dim3(256), 0, c->stream, c->s_valid.p, c->pos.p, Rc,
                         c->sj.p, c->ctr.p);
      while ((double)(hc[C_KEYS] + (u64)G) > 0.6 * (double)c->H) grow_table(c);
      Passed P = passed_view(c);
      // ---- keys, cover pre-filter, per-key work lists
      c->gbk.ensure(G);
      c->gticket.ensure(G);
      c->touched.ensure(G);
      c->covered.ensure(G);
      c->survived.ensure(G);
      c->flag_i.ensure(G);
      HIP_OK(hipEventRecord(c->ev[2], c->stream));
      hipLaunchKernelGGL(k_find, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->sj.p, c->s_dk.p, G,
                         c->gbk.p, c->ctr.p, c->survived.p, c->flag_i.p);
      HIP_OK(hipEventRecord(c->ev[5], c->stream));
      hipLaunchKernelGGL(P.nar ? k_probe<true> : k_probe<false>, dim3(blocks_for(G, 4)), dim3(64), 0, c->stream,
                         c->dnp, P, c->gbk.p, c->sj.p, c->s_dk.p, c->s_zone.p, G, c->covered.p, c->ctr.p);
      HIP_OK(hipEventRecord(c->ev[6], c->stream));
      hipLaunchKernelGGL(k_ticket, dim3(blocks_for(G, 1024)), dim3(1024), 0, c->stream, P, c->gbk.p, c->covered.p, G,
                         c->gticket.p, c->touched.p, c->ctr.p);
      // The touched keys' count stays on the device: the per-key kernels and the two scans take G
      // (at least the count) as their size and read the count itself, and the host reads it with
      // the scans' totals, one synchronisation for all.
      const int* Tdev = &c->ctr.p->ntouched;
      c->caps.ensure(G);
      c->segs.ensure(G);
      hipLaunchKernelGGL(k_touched_caps, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->touched.p, (int)G,
                         c->caps.p, c->segs.p, c->ctr.p, Tdev);
      // light and heavy keys (k_merge's two shapes) are known once the chains are counted: the
      // classification rides on the same synchronisation as the two scans
      c->klight.ensure(G);
      c->kheavy.ensure(G);
      hipLaunchKernelGGL(k_classify, dim3(blocks_for(G, 1024)), dim3(1024), 0, c->stream, P, c->touched.p, (int)G,
                         c->klight.p, c->kheavy.p, c->ctr.p, kLightMax, c->d, Tdev);
      c->pos.ensure(G + 1);
      c->pos2.ensure(G + 1);
      long long W = 0, Sseg = 0;
      int T = 0;
      int nlh[3] = {0, 0, 0};   // light keys, heavy keys, heavy keys for the pair pre-pass
      {
        scan_launch(c, c->caps.p, G, c->pos.p, 0, Tdev);
        scan_launch(c, c->segs.p, G, c->pos2.p, 1, Tdev);
        long long tt[2] = {0, 0};
        int tl[4] = {0, 0, 0, 0};   // ntouched, nlight, nheavy, npair: adjacent in Counters
        read_dev(c, tt, c->scan_total.p, sizeof(tt), tl, &c->ctr.p->ntouched, sizeof(tl));
        W = tt[0];
        Sseg = tt[1];
        T = tl[0];
        nlh[0] = tl[1];
        nlh[1] = tl[2];
        nlh[2] = tl[3];
      }
      // capacity: the reserved chain segments, and zone ids for at most W surviving new zones (the
      // ids are handed out compactly after the merge; a chain entry holds an id below kWideTag)
      if (zcount + (u64)W >= (u64)kWideTag) {
        c->err = "more stored zones than zone ids";
        return HORA_HIP_E_MEMORY;
      }
      if (!c->dn.znar && c->znar_ok && (zcount + (u64)W) * (u64)d2 * 4 > c->znar_at) {
        narrow_arena(c, zcount);
        P = passed_view(c);
      }
      if (zcount + (u64)W > P.zcap) c->zarena.grow_keep(arena_ints(c, zcount + (u64)W), c->stream);
      if (ncount + (u64)Sseg > P.ncap) c->nodes.grow_keep((size_t)(ncount + (u64)Sseg) * 2, c->stream);
      P = passed_view(c);
      hipLaunchKernelGGL(k_touched_offsets, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                         c->pos.p, c->pos2.p, ncount);
      c->wlist.ensure(W + 1);
      hipLaunchKernelGGL(k_scatter, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->gbk.p, c->covered.p,
                         c->gticket.p, G, c->wlist.p);
      // ---- ordered merge of the kept successors: light keys, then heavy keys; surviving new zones
      // wait in a scratch slot per work-list position
      c->zscr.ensure((size_t)std::max<long long>(W, 1) * d2);
      c->zused.ensure(W + 1);
      HIP_OK(hipMemsetAsync(c->zused.p, 0, (size_t)(W + 1) * sizeof(int), c->stream));
      const size_t lds_l = merge_layout(d2, false).total() * sizeof(int);
      const size_t lds_h = merge_layout(d2, true).total() * sizeof(int);
      // zones this level stored in earlier chunks, which this chunk's merge may retire
      int* zdead = ci > 0 ? c->zdead.p : nullptr;
      HIP_OK(hipEventRecord(c->ev[3], c->stream));
      // heavy keys: work lists sorted, then the pair pre-pass over each key's successors
      const int pre = (c->pairs && nlh[2] > 0) ? 1 : 0;
      if (pre) {
        c->pdrop.ensure(W + 1);
        c->kcnt.ensure(W + 1);
        c->mpos.ensure(W + 1);
        c->kslot.ensure((size_t)(W + 1) * kKillSlots);
        c->tcnt.ensure(nlh[1]);
        c->toff.ensure(nlh[1] + 1);
        HIP_OK(hipMemsetAsync(c->pdrop.p, 0, (size_t)(W + 1) * sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(c->kcnt.p, 0, (size_t)(W + 1) * sizeof(int), c->stream));
        hipLaunchKernelGGL(k_presort, dim3(nlh[1]), dim3(1024), 0, c->stream, P, c->kheavy.p, nlh[1], c->wlist.p);
        hipLaunchKernelGGL(k_pair_tiles, dim3(blocks_for(nlh[1], 256)), dim3(256), 0, c->stream, P, c->kheavy.p,
                           nlh[1], c->d, c->tcnt.p);
        scan_launch(c, c->tcnt.p, nlh[1], c->toff.p, 0);
        const size_t zs = (size_t)(d2 | 1);
        const size_t lds_p = (2 * kPairTile * zs + 2 * HORA_MAX_DIM + HORA_MAX_PROCS + HORA_MAX_INTS) * sizeof(int);
        if (lds_p > 64 * 1024)
          HIP_OK(hipFuncSetAttribute((const void*)k_pairs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_p));
        hipLaunchKernelGGL(k_pairs, dim3(4096), dim3(256), lds_p, c->stream, c->dnp, P, c->kheavy.p, nlh[1], c->toff.p,
                           c->scan_total.p, c->wlist.p, c->sj.p, c->s_zone.p, c->pdrop.p, c->kcnt.p, c->kslot.p, 0,
                           nullptr, nullptr, nullptr);
        // kept successors with more kills than slots: their full lists
        c->kneed.ensure(W + 1);
        c->koff.ensure(W + 1);
        hipLaunchKernelGGL(k_kill_need, dim3(blocks_for(W, 256)), dim3(256), 0, c->stream, c->pdrop.p, c->kcnt.p,
                           (long)W, c->kneed.p);
        scan_launch(c, c->kneed.p, W, c->koff.p, 1);
        long long K = 0;
        read_dev(c, &K, c->scan_total.p + 1, sizeof(K));
        if (K > 0) {
          c->klist.ensure((size_t)K);
          c->kpos.ensure(W + 1);
          HIP_OK(hipMemsetAsync(c->kpos.p, 0, (size_t)(W + 1) * sizeof(int), c->stream));
          hipLaunchKernelGGL(k_pairs, dim3(4096), dim3(256), lds_p, c->stream, c->dnp, P, c->kheavy.p, nlh[1],
                             c->toff.p, c->scan_total.p, c->wlist.p, c->sj.p, c->s_zone.p, c->pdrop.p, c->kcnt.p,
                             c->kslot.p, 1, c->koff.p, c->kpos.p, c->klist.p);
        }
      }
      HIP_OK(hipEventRecord(c->ev[7], c->stream));
      const auto merge_plain = P.nar ? k_merge<false, true> : k_merge<false, false>;
      const auto merge_pre = P.nar ? k_merge<true, true> : k_merge<true, false>;
      if (nlh[0] > 0)
        hipLaunchKernelGGL(merge_plain, dim3(nlh[0]), dim3(64), lds_l, c->stream, c->dnp, P, c->klight.p, nlh[0],
                           c->wlist.p, c->sj.p, c->s_zone.p, c->survived.p, c->ctr.p, 0, c->zscr.p, c->zused.p,
                           c->flag_i.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, zdead, lvl_zbase,
                           (int*)nullptr);
      if (pre) c->wlg.ensure(W + 1);
      if (nlh[1] > 0)
        hipLaunchKernelGGL(pre ? merge_pre : merge_plain, dim3(nlh[1]), dim3(pre ? 64 * kMergeHelpWaves : 64), lds_h,
                           c->stream, c->dnp, P,
                           c->kheavy.p, nlh[1], c->wlist.p, c->sj.p, c->s_zone.p, c->survived.p, c->ctr.p, 1,
                           c->zscr.p, c->zused.p, c->flag_i.p, c->pdrop.p, c->kcnt.p, c->kslot.p, c->koff.p,
                           c->klist.p, c->mpos.p, zdead, lvl_zbase, pre ? c->wlg.p : (int*)nullptr);
      HIP_OK(hipEventRecord(c->ev[4], c->stream));
      ncount += (u64)Sseg;
      // zone ids: the surviving new zones of the chunk, compact, in slot order (the count is read
      // with the frontier's scan below)
      c->zrank.ensure(W + 1);
      scan_launch(c, c->zused.p, W, c->zrank.p, 1);
      const u64 zbase = zcount;
      if (T > 0)
        hipLaunchKernelGGL(k_zassign, dim3((unsigned)T), dim3(64), 0, c->stream, P, c->touched.p, T, c->zrank.p,
                           c->zscr.p, zbase, d2, c->ctr.p);
      if (found) {
        hc = read_counters(c);
        account(c, hc, prev, G, T, 0);
        if (hc.err == ERR_NARROW) return kNarrowRetry;
        if (hc.err) { c->err = "device capacity exceeded"; return HORA_HIP_E_MEMORY; }
        prev = hc;
        break;
      }
      // ---- the chunk's survivors join the next frontier
      c->pos.ensure(G + 1);
      long long nsurv = 0;
      const long F2c = scan_and_counters(c, c->flag_i.p, G, c->pos.p, &hc, &nsurv);
      zcount += (u64)nsurv;
      c->fr_dk2.grow_keep(F2 + F2c + 1, c->stream);
      c->fr_zid2.grow_keep(F2 + F2c + 1, c->stream);
      hipLaunchKernelGGL(k_build_frontier, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, c->survived.p, c->pos.p,
                         G, c->sj.p, c->s_dk.p, c->fr_dk2.p + F2, c->fr_zid2.p + F2, c->zrank.p, zbase);
      F2 += F2c;
      if (chunked) {
        // retirement marks for the zones stored so far this level, cleared as they are added
        const size_t need = (size_t)(zcount - lvl_zbase) + 1;
        c->zdead.grow_keep(need, c->stream);
        HIP_OK(hipMemsetAsync(c->zdead.p + zdead_n, 0, (need - zdead_n) * sizeof(int), c->stream));
        zdead_n = need;
      }
      account(c, hc, prev, G, T, F2c);
      if (c->trace)
        std::fprintf(stderr, "level %d chunk %zu/%zu F=%ld R=%ld G=%ld probe_cov=%llu T=%d (light %d heavy %d) W=%lld F2=%ld | exp %.3f probe %.3f pairs %.3f merge %.3f ms | reads p=%llu m=%llu unions=%llu full=%llu wovf=%llu\n",
                     level, ci + 1, cuts.size() - 1, Fc, Rc, G, hc[C_PSUB] - prev[C_PSUB], T, nlh[0], nlh[1], W, F2c,
                     ev_ms(c->ev[0], c->ev[1]), ev_ms(c->ev[5], c->ev[6]), ev_ms(c->ev[3], c->ev[7]),
                     ev_ms(c->ev[7], c->ev[4]), hc[C_PREADS] - prev[C_PREADS], hc[C_READS] - prev[C_READS],
                     hc[C_UNI] - prev[C_UNI], hc[C_FULL] - prev[C_FULL], hc[C_WOVF] - prev[C_WOVF]);
      prev = hc;
      if (hc.err == ERR_NARROW) return kNarrowRetry;
      if (hc.err) {
        c->err = hc.err == ERR_HASH_FULL ? "passed-list hash table full"
                 : hc.err == ERR_ZONES_FULL ? "zone arena full"
                 : hc.err == ERR_RECIPES ? "a state has more transitions or invariant constraints than the device holds"
                                         : "chain node arena full";
        return HORA_HIP_E_MEMORY;
      }
    }
    if (found) break;
    if (chunked && F2 > 0) {
      // drop the frontier entries a later chunk retired, keeping storing order
      c->flag_i.ensure(F2);
      c->pos.ensure(F2 + 1);
      hipLaunchKernelGGL(k_frontier_live, dim3(blocks_for(F2, 256)), dim3(256), 0, c->stream, c->fr_zid2.p, F2,
                         c->zdead.p, lvl_zbase, c->flag_i.p);
      const long keep = scan_exclusive(c, c->flag_i.p, F2, c->pos.p);
      c->fr_dk.ensure(keep + 1);
      c->fr_zid.ensure(keep + 1);
      hipLaunchKernelGGL(k_frontier_compact, dim3(blocks_for(F2, 256)), dim3(256), 0, c->stream, c->fr_dk2.p,
                         c->fr_zid2.p, F2, c->flag_i.p, c->pos.p, c->fr_dk.p, c->fr_zid.p);
      F = keep;
    } else {
      std::swap(c->fr_dk.p, c->fr_dk2.p); std::swap(c->fr_dk.cap, c->fr_dk2.cap);
      std::swap(c->fr_zid.p, c->fr_zid2.p); std::swap(c->fr_zid.cap, c->fr_zid2.cap);
      F = F2;
    }
    init = false;
    // adaptive Extra_LU+ switch (engine.cpp:2178-2184)
    if (!c->dn.lu_plus && prev[C_KEYS] && (double)prev[C_LIVE] >= 4.0 * (double)prev[C_KEYS]) {
      c->dn.lu_plus = 1;
      HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));
      HIP_OK(hipStreamSynchronize(c->stream));
      st->emode_switch_level = level;
    }
    if (F == 0) break;
  }
  const Tally hc = read_counters(c);
  st->levels = level;
  st->stored_live = hc[C_LIVE];
  st->keys = hc[C_KEYS];
  st->inserted = hc[C_INS];
  st->subsumed = hc[C_SUB] + hc[C_PSUB];
  st->killed = hc[C_KILL];
  st->unions = hc[C_UNI];
  st->chain_reads = hc[C_READS] + hc[C_PREADS];
  st->zone_entry_bytes = c->dn.znar ? 2 : 4;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Stored zones start int32 and switch to narrow (int16) storage once the arena would pass
// znar_at bytes, when the network's constants allow it (decided at bind): int32 reads cost the
// ordered merge a little less, and narrow storage halves the search's largest structure. If a
// zone does not fit 16 bits after all, the search runs again with int32 storage, and the context
// keeps int32 from then on.
int run_search(hora_hip_ctx* c, int32_t* reachable, hora_reach_stats* st) {
  const auto t0 = std::chrono::steady_clock::now();
  int rc = search_once(c, reachable, st, (c->znar_ok && c->znar_at == 0) ? 1 : 0);
  if (rc == kNarrowRetry) {
    c->znar_ok = false;
    rc = search_once(c, reachable, st, 0);
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  c->dn.znar = 0;
  return rc;
}

template <class T>
std::vector<T> vec(const T* p, long n) {
  return p && n > 0 ? std::vector<T>(p, p + n) : std::vector<T>(std::max<long>(n, 1), T{});
}

// Bind a context to a network and query. Callable again on a bound context (a serving slot moving
// to its next job): the stream, events and every device buffer are kept, only the tables change.
// Upper bounds, from the network's tables, on the transitions of one discrete state (local edges
// plus binary sender/receiver pairs, each process at its busiest location) and on the combined
// invariant of one location vector (each process at its location with the most constraints),
// capped at kRecipeCap and kInvCap: k_count's and k_expand's LDS is sized by them, so a small
// network leaves room for more resident wavefronts. A state beyond the caps is an error as before.
void expand_caps(const hora_net_enc* e, int* rcap, int* icap) {
  const int np = e->nprocs, nch = e->nchans, tl = e->total_locs;
  std::vector<int> lcnt(tl, 0);
  std::vector<int> scnt((size_t)tl * std::max(nch, 1), 0), rcnt((size_t)tl * std::max(nch, 1), 0);
  for (int k = 0; k < e->nedges; ++k) {
    const int g = e->locbase[e->edge_proc[k]] + e->edge_src[k];
    const int ch = e->edge_chan[k];
    if (ch < 0) ++lcnt[g];
    else if (e->edge_dir[k] == 1) ++scnt[(size_t)g * nch + ch];
    else if (e->edge_dir[k] == -1) ++rcnt[(size_t)g * nch + ch];
  }
  long long local = 0, inv = 0, bin = 0;
  std::vector<long long> smax((size_t)np * std::max(nch, 1), 0), rmax((size_t)np * std::max(nch, 1), 0);
  for (int p = 0; p < np; ++p) {
    int lm = 0, im = 0;
    for (int g = e->locbase[p]; g < e->locbase[p + 1]; ++g) {
      lm = std::max(lm, lcnt[g]);
      im = std::max(im, e->inv_off[g + 1] - e->inv_off[g]);
      for (int ch = 0; ch < nch; ++ch) {
        smax[(size_t)p * nch + ch] = std::max<long long>(smax[(size_t)p * nch + ch], scnt[(size_t)g * nch + ch]);
        rmax[(size_t)p * nch + ch] = std::max<long long>(rmax[(size_t)p * nch + ch], rcnt[(size_t)g * nch + ch]);
      }
    }
    local += lm;
    inv += im;
  }
  // a broadcast's combinations: per process, its most options at one location (receiving edges,
  // plus the stay-out conjunctions, at most the product of the non-empty guard sizes: integer
  // guards may disable the unguarded receivers; 1 if it never takes part)
  std::vector<long long> omax((size_t)np * std::max(nch, 1), 1);
  for (int ch = 0; ch < nch; ++ch) {
    if (!(e->chan_bcast && e->chan_bcast[ch])) continue;
    for (int p = 0; p < np; ++p)
      for (int g = e->locbase[p]; g < e->locbase[p + 1]; ++g) {
        long long nr = 0, none = 1;
        for (int k = 0; k < e->nedges; ++k)
          if (e->edge_chan[k] == ch && e->edge_dir[k] == -1 && e->locbase[e->edge_proc[k]] + e->edge_src[k] == g) {
            ++nr;
            none = std::min<long long>(none * std::max(1, e->cg_off[k + 1] - e->cg_off[k]), kRecipeCap + 1);
          }
        if (nr > 0) omax[(size_t)p * nch + ch] = std::max(omax[(size_t)p * nch + ch], nr + none);
      }
  }
  for (int ch = 0; ch < nch; ++ch) {
    if (e->chan_bcast && e->chan_bcast[ch]) {
      for (int p = 0; p < np; ++p) {
        long long prod = smax[(size_t)p * nch + ch];
        for (int q = 0; q < np && prod > 0; ++q)
          if (q != p) prod = std::min<long long>(prod * omax[(size_t)q * nch + ch], kRecipeCap + 1);
        bin = std::min<long long>(bin + prod, kRecipeCap + 1);
      }
      continue;
    }
    long long rtot = 0;
    for (int p = 0; p < np; ++p) rtot += rmax[(size_t)p * nch + ch];
    for (int p = 0; p < np; ++p) bin += smax[(size_t)p * nch + ch] * (rtot - rmax[(size_t)p * nch + ch]);
  }
  *rcap = (int)std::max<long long>(1, std::min<long long>(kRecipeCap, local + bin));
  *icap = (int)std::max<long long>(1, std::min<long long>(kInvCap, inv));
}

void open_ctx(const hora_net_enc* e, const hora_target* t, hora_hip_ctx* c) {
  if (!e) throw ScopeFail("null network");
  c->h_kshift.clear();
  c->h_kmask.clear();
  c->nclocks = e->nclocks;
  c->d = e->nclocks + 1;
  c->d2 = c->d * c->d;
  c->np = e->nprocs;
  c->ni = e->nintvars;
  c->tl = e->total_locs;
  c->ne = e->nedges;
  const int d = c->d;
  if (d > HORA_MAX_DIM) throw ScopeFail("zone dimension above HORA_MAX_DIM");
  if (c->np > HORA_MAX_PROCS || c->np < 1) throw ScopeFail("process count out of range");
  if (c->ni > HORA_MAX_INTS) throw ScopeFail("too many integer variables");
  auto diag = [](const int32_t* data, long n) {
    for (long k = 0; k < n; ++k)
      if (data[3 * k] > 0 && data[3 * k + 1] > 0) return true;
    return false;
  };
  if (diag(e->inv_data, e->inv_off[c->tl]) || diag(e->cg_data, e->cg_off[c->ne]))
    throw ScopeFail("diagonal constraints are not covered by the device path");
  if (t && diag(t->clock_cond, t->nclock_cond)) throw ScopeFail("diagonal target constraint");

  // key codec (DiscKeyCodecT, engine.cpp:981-1001)
  auto fbits = [](long nv) { int b = 0; while ((1LL << b) < nv) ++b; return b; };
  int off = 0;
  for (int p = 0; p < c->np; ++p) {
    const int w = fbits(e->locbase[p + 1] - e->locbase[p]);
    c->h_kshift.push_back(off); c->h_kmask.push_back((1ull << w) - 1); off += w;
  }
  for (int v = 0; v < c->ni; ++v) {
    const int w = fbits((long)e->int_hi[v] - e->int_lo[v] + 1);
    c->h_kshift.push_back(off); c->h_kmask.push_back((1ull << w) - 1); off += w;
  }
  if (off > 63) throw ScopeFail("packed discrete key wider than 63 bits");
  c->h_loc_init = vec(e->loc_init, c->np);
  c->h_int_init = vec(e->int_init, c->ni);
  c->h_int_lo = vec(e->int_lo, c->ni);
  {
    bool time_open = true, one_key = true;
    for (int ch = 0; ch < e->nchans; ++ch)
      if (e->chan_urgent && e->chan_urgent[ch]) time_open = false;
    for (int g = 0; g < c->tl; ++g) {
      if (e->loc_flags[g] & 3) time_open = false;
      for (int k = e->inv_off[g]; k < e->inv_off[g + 1]; ++k)
        if (e->inv_data[3 * k] > 0 && e->inv_data[3 * k + 1] == 0 && e->inv_data[3 * k + 2] < HORA_INF) time_open = false;
    }
    for (int p = 0; p < c->np; ++p)
      if (e->locbase[p + 1] - e->locbase[p] != 1) one_key = false;
    for (int v = 0; v < c->ni; ++v)
      if (e->int_lo[v] != e->int_hi[v]) one_key = false;
    c->base_route = time_open ? 1 : one_key ? 2 : 0;
  }

  // bounds: global L/U widened by the query (extend_bounds, engine.cpp:739-752), query floors
  std::vector<int> Lx = vec(e->lbound, d), Ux = vec(e->ubound, d), Lq(d, HORA_ALU_NEG_INF), Uq(d, HORA_ALU_NEG_INF);
  u64 qbits = 0;
  std::vector<int> lrv, icv, ccv;
  int never = 1, nlr = 0, nic = 0, ncc = 0;
  if (t) {
    never = t->never;
    nlr = t->nloc_req; nic = t->nint_cond; ncc = t->nclock_cond;
    lrv = vec(t->loc_req, 2L * nlr);
    icv = vec(t->int_cond, 3L * nic);
    ccv = vec(t->clock_cond, 3L * ncc);
    for (int k = 0; k < ncc; ++k) {
      const int i = ccv[3 * k], j = ccv[3 * k + 1], v = ccv[3 * k + 2] >> 1;
      if (i > 0) qbits |= 1ull << i;
      if (j > 0) qbits |= 1ull << j;
      if (i > 0 && j == 0) { Ux[i] = std::max(Ux[i], v); Uq[i] = std::max(Uq[i], v); }
      else if (i == 0 && j > 0) { Lx[j] = std::max(Lx[j], -v); Lq[j] = std::max(Lq[j], -v); }
    }
  }
  Lx[0] = Ux[0] = 0;
  {
    // narrow zone storage when the largest clock constant of the model and the query is below
    // 4000 (engine.cpp:1596-1611); the device still checks every entry it narrows
    int maxc = 0;
    auto take = [&](int raw) { if (raw < HORA_INF) maxc = std::max(maxc, std::abs(raw >> 1)); };
    for (int k = 0; k < d; ++k) {
      if (Lx[k] != HORA_ALU_NEG_INF) maxc = std::max(maxc, std::abs(Lx[k]));
      if (Ux[k] != HORA_ALU_NEG_INF) maxc = std::max(maxc, std::abs(Ux[k]));
    }
    for (int k = 0; k < e->inv_off[c->tl]; ++k) take(e->inv_data[3 * k + 2]);
    for (int k = 0; k < e->cg_off[c->ne]; ++k) take(e->cg_data[3 * k + 2]);
    for (int k = 0; k < ncc; ++k) take(ccv[3 * k + 2]);
    // HORA_HIP_NARROW=0 keeps int32; =1 narrows whatever the constants (an entry that does not
    // fit then sends the search back to int32). HORA_HIP_NARROW_AT: the int32 arena size in
    // bytes at which a search switches (0: narrow from the start), 8 GiB by default.
    const char* nv = std::getenv("HORA_HIP_NARROW");
    c->znar_ok = (nv && *nv == '1') || (maxc < 4000 && !(nv && *nv == '0'));
    const char* na = std::getenv("HORA_HIP_NARROW_AT");
    c->znar_at = na ? (size_t)std::strtoull(na, nullptr, 10) : (size_t(8) << 30);
  }

  // edges leaving each global location, ascending global edge id (NProc::edges_of order)
  std::vector<int> le_off(c->tl + 1, 0), le_edges(std::max(c->ne, 1));
  for (int g = 0; g < c->ne; ++g) ++le_off[e->locbase[e->edge_proc[g]] + e->edge_src[g] + 1];
  for (int L = 0; L < c->tl; ++L) le_off[L + 1] += le_off[L];
  {
    std::vector<int> fill(le_off.begin(), le_off.end() - 1);
    for (int g = 0; g < c->ne; ++g) le_edges[fill[e->locbase[e->edge_proc[g]] + e->edge_src[g]]++] = g;
  }
  int has_urgent = 0;
  for (int ch = 0; ch < e->nchans; ++ch)
    if (e->chan_urgent && e->chan_urgent[ch]) has_urgent = 1;

  {
    const char* tr = std::getenv("HORA_HIP_TRACE");
    c->trace = tr && *tr == '1';
    const char* pr = std::getenv("HORA_HIP_PAIRS");
    c->pairs = !(pr && *pr == '0');
  }
  if (!c->stream) {
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& ev : c->ev) HIP_OK(hipEventCreate(&ev));
  }
  if (!c->hpin) HIP_OK(hipHostMalloc(&c->hpin, kPinBytes, hipHostMallocDefault));
  const int ne = std::max(c->ne, 1);
  // all static tables except the per-location bounds go into one blob (u64 tables 8-aligned)
  std::vector<int> blob;
  auto put = [&](const std::vector<int>& v) {
    if (blob.size() & 1) blob.push_back(0);
    const size_t off = blob.size();
    blob.insert(blob.end(), v.begin(), v.end());
    if (v.empty()) blob.push_back(0);
    return off;
  };
  auto put64 = [&](const std::vector<u64>& v) {
    std::vector<int> w(2 * std::max<size_t>(v.size(), 1), 0);
    if (!v.empty()) std::memcpy(w.data(), v.data(), v.size() * sizeof(u64));
    return put(w);
  };
  std::vector<int> bc(std::max(e->nchans, 1), 0), ur(std::max(e->nchans, 1), 0);
  for (int ch = 0; ch < e->nchans; ++ch) {
    bc[ch] = e->chan_bcast ? e->chan_bcast[ch] : 0;
    ur[ch] = e->chan_urgent ? e->chan_urgent[ch] : 0;
  }
  const size_t o_locbase = put(vec(e->locbase, c->np + 1));
  const size_t o_loc_flags = put(vec(e->loc_flags, c->tl));
  const size_t o_inv_off = put(vec(e->inv_off, c->tl + 1));
  const size_t o_inv_data = put(vec(e->inv_data, 3L * e->inv_off[c->tl]));
  const size_t o_le_off = put(le_off);
  const size_t o_le_edges = put(le_edges);
  const size_t o_edge_proc = put(vec(e->edge_proc, ne));
  const size_t o_edge_tgt = put(vec(e->edge_tgt, ne));
  const size_t o_edge_chan = put(vec(e->edge_chan, ne));
  const size_t o_edge_dir = put(vec(e->edge_dir, ne));
  const size_t o_cg_off = put(vec(e->cg_off, c->ne + 1));
  const size_t o_cg_data = put(vec(e->cg_data, 3L * e->cg_off[c->ne]));
  const size_t o_rs_off = put(vec(e->rs_off, c->ne + 1));
  const size_t o_rs_clk = put(vec(e->rs_clk, (long)e->rs_off[c->ne]));
  const size_t o_ig_off = put(vec(e->ig_off, c->ne + 1));
  const size_t o_ig_data = put(vec(e->ig_data, 3L * e->ig_off[c->ne]));
  const size_t o_iu_off = put(vec(e->iu_off, c->ne + 1));
  const size_t o_iu_data = put(vec(e->iu_data, 2L * e->iu_off[c->ne]));
  const size_t o_bc = put(bc);
  const size_t o_ur = put(ur);
  const size_t o_Lx = put(Lx);
  const size_t o_Ux = put(Ux);
  const size_t o_Lq = put(Lq);
  const size_t o_Uq = put(Uq);
  const size_t o_amask = put64(vec(e->active_mask, c->tl));
  const size_t o_kshift = put(c->h_kshift);
  const size_t o_kmask = put64(c->h_kmask);
  const size_t o_int_lo = put(c->h_int_lo);
  const size_t o_lr = put(lrv);
  const size_t o_ic = put(icv);
  const size_t o_cc = put(ccv);
  if (blob.size() & 1) blob.push_back(0);
  c->blob.upload(blob, c->stream);
  c->lloc.upload(vec(e->lbound_loc, (long)c->tl * d), c->stream);
  c->uloc.upload(vec(e->ubound_loc, (long)c->tl * d), c->stream);
  const int* B = c->blob.p;

  DevNet& N = c->dn;
  N.d = d; N.d2 = c->d2; N.np = c->np; N.ni = c->ni; N.nclocks = c->nclocks; N.lu_plus = 0;
  N.locbase = B + o_locbase; N.loc_flags = B + o_loc_flags; N.inv_off = B + o_inv_off; N.inv_data = B + o_inv_data;
  N.le_off = B + o_le_off; N.le_edges = B + o_le_edges; N.edge_proc = B + o_edge_proc; N.edge_tgt = B + o_edge_tgt;
  N.edge_chan = B + o_edge_chan; N.edge_dir = B + o_edge_dir;
  N.cg_off = B + o_cg_off; N.cg_data = B + o_cg_data; N.rs_off = B + o_rs_off; N.rs_clk = B + o_rs_clk;
  N.ig_off = B + o_ig_off; N.ig_data = B + o_ig_data; N.iu_off = B + o_iu_off; N.iu_data = B + o_iu_data;
  N.chan_bcast = B + o_bc; N.chan_urgent = B + o_ur; N.has_urgent_chan = has_urgent;
  N.has_bcast = 0;
  for (int ch = 0; ch < e->nchans; ++ch) N.has_bcast |= bc[ch] != 0;
  N.Lx = B + o_Lx; N.Ux = B + o_Ux; N.lloc = c->lloc.p; N.uloc = c->uloc.p; N.Lq = B + o_Lq; N.Uq = B + o_Uq;
  N.amask = (const u64*)(B + o_amask); N.qbits = qbits; N.kshift = B + o_kshift; N.kmask = (const u64*)(B + o_kmask);
  N.int_lo = B + o_int_lo;
  N.lr = B + o_lr; N.ic = B + o_ic; N.cc = B + o_cc;
  N.blob = B;
  N.blob_ints = (int)blob.size();
  N.stage = 0;   // staging the tables in LDS measured slower than the L1/L2-resident originals
  N.inv_upper = 1;
  for (long k = 0; k < e->inv_off[c->tl]; ++k)
    if (e->inv_data[3 * k + 1] != 0 || e->inv_data[3 * k] <= 0) N.inv_upper = 0;
  N.never = never; N.nlr = nlr; N.nic = nic; N.ncc = ncc;
  expand_caps(e, &N.rcap, &N.icap);
  if (!c->dnp) HIP_OK(hipMalloc(&c->dnp, sizeof(DevNet)));
  HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
}

}  // namespace

#include "liveness.hpp"
// Built as fed_hbm.hip (HORA_FED_HBM), this unit supplies only the liveness kernels' device-memory
// tier; the serving loop, the static scratch and the C ABI belong to the main unit.
#ifndef HORA_FED_HBM
#include "serve.hpp"

// ---------------------------------------------------------------- the allocation-free fixpoint

// Device scratch of hora_hip_gfp_solve_noalloc: every buffer that call needs is carved from it
// (Arena), so it calls no hipMalloc. One array per GPU (each device that loads the library holds
// its own copy until the process ends); the address of this device's copy comes from
// hipGetSymbolAddress, once per device, and is kept in noalloc_base[].
constexpr size_t kNoAllocScratchBytes = 256ull << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocDevices = 64;
std::mutex noalloc_mu;
void* noalloc_base[kNoAllocDevices] = {};

// The device address of this device's scratch, or null.
void* noalloc_scratch(std::string* err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) {
    *err = "no current HIP device for the static scratch";
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_base[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    noalloc_base[dev] = p;
  }
  return noalloc_base[dev];
}

// Largest graph the scratch takes: its per-node arrays and a value pool of a few zones per node.
long long noalloc_max_nodes(int d2) { return (long long)(kNoAllocScratchBytes / (size_t)(80 + 16 * d2)); }

template <class T>
void swap_buf(DBuf<T>& a, DBuf<T>& b) {
  std::swap(a.p, b.p);
  std::swap(a.cap, b.cap);
  std::swap(a.borrowed, b.borrowed);
}

namespace {

// Launch shape of k_zop_reg: wavefronts per workgroup, and the workgroups of the grid (each
// wavefront then walks zone groups with a stride, loading the next while it computes one).
struct ZopShape {
  int wpb;
  long blocks;
};

template <int D, int OP>
void launch_zop_reg(const ZopArgs& A, ZopShape sh, hipStream_t s) {
  constexpr int G = D * D < kWave ? D * D : kWave;
  const long ngroups = (A.n + kWave / G - 1) / (kWave / G);
  const long need = (ngroups + sh.wpb - 1) / sh.wpb;
  const long blocks = std::max<long>(1, std::min(need, sh.blocks));
  hipLaunchKernelGGL((k_zop_reg<D, OP>), dim3((unsigned)blocks), dim3(64 * sh.wpb), 0, s, A, ngroups);
}

template <int OP>
void launch_zop_reg_d(const ZopArgs& A, ZopShape sh, hipStream_t s) {
  switch (A.d) {
    case 4: launch_zop_reg<4, OP>(A, sh, s); break;
    case 8: launch_zop_reg<8, OP>(A, sh, s); break;
    case 16: launch_zop_reg<16, OP>(A, sh, s); break;
    default: launch_zop_reg<32, OP>(A, sh, s); break;
  }
}

void launch_zop_reg_any(const ZopArgs& A, ZopShape sh, hipStream_t s) {
  if (A.op == HORA_ZOP_CLOSE) launch_zop_reg_d<HORA_ZOP_CLOSE>(A, sh, s);
  else launch_zop_reg_d<HORA_ZOP_EXTRA_LU>(A, sh, s);
}

// The autotuner of the register kernels' launch shape, per (operation, d). The first call with at
// least 2^16 zones, an output apart from the input (so that a rerun writes the same values) and a
// stream that is not being captured times every candidate twice on the caller's stream and keeps
// the fastest; every later call (and every call during graph capture) launches the pick.
// HORA_HIP_ZOP_SHAPE=wpb,blocks fixes the shape.
std::mutex g_zop_mu;
std::map<int, ZopShape> g_zop_pick;

ZopShape zop_shape(const ZopArgs& A, hipStream_t s) {
  if (const char* e = std::getenv("HORA_HIP_ZOP_SHAPE")) {
    int w = 0;
    long b = 0;
    if (std::sscanf(e, "%d,%ld", &w, &b) == 2 && (w == 1 || w == 2 || w == 4) && b > 0) return ZopShape{w, b};
  }
  const int key = A.op * 64 + A.d;
  {
    std::lock_guard<std::mutex> g(g_zop_mu);
    auto it = g_zop_pick.find(key);
    if (it != g_zop_pick.end()) return it->second;
  }
  const ZopShape dflt{4, 1L << 30};
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (A.n < (1L << 16) || !A.out || A.out == A.a || hipStreamIsCapturing(s, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusNone)
    return dflt;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return dflt;
  std::vector<ZopShape> cand;
  for (int w : {1, 2, 4})
    for (long per : {2L, 8L, 1L << 30}) cand.push_back(ZopShape{w, per >= (1L << 30) ? per : per * cus * (8 / w)});
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return dflt;
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return dflt; }
  ZopShape best = dflt;
  float best_ms = 1e30f;
  for (int rep = 0; rep < 2; ++rep)
    for (const ZopShape& c : cand) {
      (void)hipEventRecord(e0, s);
      launch_zop_reg_any(A, c, s);
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best_ms) { best_ms = ms; best = c; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> g(g_zop_mu);
  g_zop_pick[key] = best;
  return best;
}

}  // namespace

extern "C" {

int hora_hip_abi_version(void) { return HORA_HIP_ABI_VERSION; }

const char* hora_hip_build_info(void) { return "hora_hip gfx950 (wave64, LDS-staged zones, device passed list)"; }

int hora_hip_open(const hora_net_enc* net, const hora_target* target, hora_hip_ctx** out) {
  if (!out) return HORA_HIP_E_ARG;
  *out = nullptr;
  auto* c = new (std::nothrow) hora_hip_ctx();
  if (!c) return HORA_HIP_E_MEMORY;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw HipFail("no HIP device");
    open_ctx(net, target, c);
    *out = c;
    return HORA_HIP_OK;
  } catch (const ScopeFail& ex) {
    c->err = ex.what();
    *out = c;
    return HORA_HIP_E_SCOPE;
  } catch (const HipFail& ex) {
    c->err = ex.what();
    *out = c;
    return HORA_HIP_E_DEVICE;
  } catch (const std::exception& ex) {
    c->err = ex.what();
    *out = c;
    return HORA_HIP_E_INTERNAL;
  }
}

void hora_hip_close(hora_hip_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->dnp) (void)hipFree(ctx->dnp);
  if (ctx->live) live_free(ctx->live);
  if (ctx->hpin) (void)hipHostFree(ctx->hpin);
  delete ctx;
}

const char* hora_hip_last_error(const hora_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hora_hip_reach(hora_hip_ctx* ctx, int32_t* reachable, hora_reach_stats* stats) {
  if (!ctx || !reachable || !stats || !ctx->stream) return HORA_HIP_E_ARG;
  StreamScope scope(ctx->stream);
  try {
    return run_search(ctx, reachable, stats);
  } catch (const HipFail& ex) {
    ctx->err = ex.what();
    return HORA_HIP_E_DEVICE;
  } catch (const std::exception& ex) {
    ctx->err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

int hora_hip_reach_profile(const hora_hip_ctx* ctx, hora_reach_profile* out) {
  if (!ctx || !out) return HORA_HIP_E_ARG;
  *out = ctx->prof;
  return HORA_HIP_OK;
}

int hora_hip_zone_op(int32_t op, int32_t d, int64_t n, const int32_t* a, const int32_t* b, const int32_t* arg,
                     const int32_t* L, const int32_t* U, int32_t* out, int32_t* out_flag, void* stream) {
  if (d < 1 || d > HORA_MAX_DIM || n < 0 || !a) return HORA_HIP_E_ARG;
  if (op < HORA_ZOP_CLOSE || op > HORA_ZOP_FREE) return HORA_HIP_E_ARG;
  if ((op == HORA_ZOP_RESET || op == HORA_ZOP_CONSTRAIN || op == HORA_ZOP_FREE) && !arg) return HORA_HIP_E_ARG;
  if ((op == HORA_ZOP_EXTRA_LU || op == HORA_ZOP_ALU_LE) && (!L || !U)) return HORA_HIP_E_ARG;
  if ((op == HORA_ZOP_INCLUDES || op == HORA_ZOP_ALU_LE || op == HORA_ZOP_UNION) && !b) return HORA_HIP_E_ARG;
  if (n == 0) return HORA_HIP_OK;
  ZopArgs A{op, d, n, a, b, arg, L, U, out, out_flag};
  hipStream_t s = (hipStream_t)stream;
  const int ld = d == 4 ? 2 : d == 8 ? 3 : d == 16 ? 4 : d == 32 ? 5 : 0;
  const bool flat = ld > 0 && (op == HORA_ZOP_INCLUDES || op == HORA_ZOP_ALU_LE || op == HORA_ZOP_DOWN ||
                               op == HORA_ZOP_FREE || op == HORA_ZOP_CONSTRAIN);
  if ((op == HORA_ZOP_DOWN || op == HORA_ZOP_FREE) && !out) return HORA_HIP_E_ARG;
  if (op == HORA_ZOP_UP || op == HORA_ZOP_RESET) {
    if (!out) return HORA_HIP_E_ARG;
    const long tot = n * (long)d * d;
    hipLaunchKernelGGL(k_zop_stream, dim3(blocks_for(tot, 256)), dim3(256), 0, s, A);
  } else if (flat) {
    const int bs = d == 32 ? 1024 : 256;
    const long zpb = bs / (d * d);
    const dim3 grid(blocks_for(n, (int)zpb)), blk(bs);
    switch (op) {
      case HORA_ZOP_INCLUDES: hipLaunchKernelGGL(k_zop_flat<HORA_ZOP_INCLUDES>, grid, blk, 0, s, A, ld); break;
      case HORA_ZOP_ALU_LE: hipLaunchKernelGGL(k_zop_flat<HORA_ZOP_ALU_LE>, grid, blk, 0, s, A, ld); break;
      case HORA_ZOP_DOWN: hipLaunchKernelGGL(k_zop_flat<HORA_ZOP_DOWN>, grid, blk, 0, s, A, ld); break;
      case HORA_ZOP_FREE: hipLaunchKernelGGL(k_zop_flat<HORA_ZOP_FREE>, grid, blk, 0, s, A, ld); break;
      default: hipLaunchKernelGGL(k_zop_flat<HORA_ZOP_CONSTRAIN>, grid, blk, 0, s, A, ld); break;
    }
  } else if (op == HORA_ZOP_EMPTY && ld > 0) {
    hipLaunchKernelGGL(k_zop_empty, dim3(blocks_for(n, 256 >> ld)), dim3(256), 0, s, A, ld);
  } else if ((op == HORA_ZOP_CLOSE || op == HORA_ZOP_EXTRA_LU) && ld > 0) {
    launch_zop_reg_any(A, zop_shape(A, s), s);
  } else if (op == HORA_ZOP_UNION && d == 4) {
    // (at d = 8 one lane's 64-entry test is slower than a wavefront per pair: 0.8 against 1.3 TB/s)
    hipLaunchKernelGGL(k_zop_union_lane<4>, dim3(blocks_for(n, kWave)), dim3(64), 0, s, A);
  } else {
    const long grid = std::min<long>(n, 1L << 20);
    const size_t lds = (4 * d * d + HORA_MAX_DIM + 8) * sizeof(int);
    hipLaunchKernelGGL(k_zop_wave, dim3(grid), dim3(64), lds, s, A);
  }
  return hipGetLastError() == hipSuccess ? HORA_HIP_OK : HORA_HIP_E_DEVICE;
}

int hora_hip_live(hora_hip_ctx* ctx, const hora_live_query* q, int32_t* holds, hora_live_stats* stats) {
  if (!ctx || !q || !holds || !stats || !ctx->stream) return HORA_HIP_E_ARG;
  StreamScope scope(ctx->stream);
  try {
    return run_live(ctx, q, holds, stats);
  } catch (const HipFail& ex) {
    ctx->err = ex.what();
    return HORA_HIP_E_DEVICE;
  } catch (const std::exception& ex) {
    ctx->err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

int hora_hip_gfp_solve(hora_hip_ctx* ctx, const hora_gfp_graph* g, const hora_live_query* q, int32_t* holds,
                       hora_live_stats* stats, int8_t* w_nonempty) {
  if (!ctx || !q || !holds || !stats || !ctx->stream) return HORA_HIP_E_ARG;
  StreamScope scope(ctx->stream);
  try {
    if (!ctx->live) ctx->live = new LiveState();
    return gfp_solve(ctx, *ctx->live, g, q, holds, stats, w_nonempty);
  } catch (const HipFail& ex) {
    ctx->err = ex.what();
    return HORA_HIP_E_DEVICE;
  } catch (const std::exception& ex) {
    ctx->err = ex.what();
    return HORA_HIP_E_INTERNAL;
  }
}

int hora_hip_gfp_solve_noalloc(hora_hip_ctx* ctx, const hora_gfp_graph* g, const hora_live_query* q,
                               int32_t* holds, hora_live_stats* stats, int8_t* w_nonempty) {
  if (!ctx || !q || !holds || !stats || !ctx->stream || !g) return HORA_HIP_E_ARG;
  std::memset(stats, 0, sizeof(*stats));
  *holds = 0;
  if (g->nodes > noalloc_max_nodes(ctx->d2)) {
    ctx->err = "graph of " + std::to_string(g->nodes) + " nodes beyond the static scratch's maximum of " +
               std::to_string(noalloc_max_nodes(ctx->d2)) + " at this zone dimension";
    return HORA_HIP_E_MEMORY;
  }
  void* base = noalloc_scratch(&ctx->err);
  if (!base) return HORA_HIP_E_DEVICE;
  StreamScope scope(ctx->stream);
  Arena arena;
  arena.base = (char*)base;
  arena.cap = kNoAllocScratchBytes;
  LiveState L;
  // the context's scan buffers give way to ones in the scratch for the call
  DBuf<long long> scan_total, bsum;
  swap_buf(ctx->scan_total, scan_total);
  swap_buf(ctx->bsum, bsum);
  int rc;
  try {
    ArenaScope in_scratch(&arena);
    rc = gfp_solve(ctx, L, g, q, holds, stats, w_nonempty);
  } catch (const ArenaFull& ex) {
    ctx->err = ex.what();
    rc = HORA_HIP_E_MEMORY;
  } catch (const HipFail& ex) {
    ctx->err = ex.what();
    rc = HORA_HIP_E_DEVICE;
  } catch (const std::exception& ex) {
    ctx->err = ex.what();
    rc = HORA_HIP_E_INTERNAL;
  }
  (void)hipStreamSynchronize(ctx->stream);   // nothing of this call still reads the scratch
  swap_buf(ctx->scan_total, scan_total);
  swap_buf(ctx->bsum, bsum);
  return rc;
}

int hora_hip_live_graph(const hora_hip_ctx* ctx, hora_live_graph* g) {
  if (!ctx || !g) return HORA_HIP_E_ARG;
  try {
    return live_graph(ctx, g);
  } catch (const std::exception&) {
    return HORA_HIP_E_DEVICE;
  }
}

int hora_hip_server_open(int32_t slots, int64_t hbm_budget_bytes, hora_hip_server** out) {
  if (!out || slots < 1 || slots > 64) return HORA_HIP_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return HORA_HIP_E_DEVICE;
  auto* S = new (std::nothrow) hora_hip_server();
  if (!S) return HORA_HIP_E_MEMORY;
  S->slots = slots;
  if (hbm_budget_bytes > 0) {
    S->budget = hbm_budget_bytes;
  } else {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete S; return HORA_HIP_E_DEVICE; }
    S->budget = (int64_t)(0.9 * (double)fr);
  }
  for (int k = 0; k < slots; ++k) S->ctxs.push_back(new hora_hip_ctx());
  for (int k = 0; k < slots; ++k) S->workers.emplace_back(worker, S, k);
  *out = S;
  return HORA_HIP_OK;
}

int hora_hip_server_submit_reach(hora_hip_server* s, const hora_net_enc* net, const hora_target* target,
                                 int64_t est_bytes, int64_t* job) {
  if (!s || !net) return HORA_HIP_E_ARG;
  Job* j = new Job();
  j->kind = 0;
  j->data.copy_net(net);
  j->data.copy_target(target);
  j->est_bytes = est_bytes > 0 ? est_bytes : default_footprint(net);
  return submit(s, j, job);
}

int hora_hip_server_submit_live(hora_hip_server* s, const hora_net_enc* net, const hora_live_query* q,
                                int64_t est_bytes, int64_t* job) {
  if (!s || !net || !q) return HORA_HIP_E_ARG;
  Job* j = new Job();
  j->kind = 1;
  j->data.copy_net(net);
  j->data.copy_query(q);
  j->est_bytes = est_bytes > 0 ? est_bytes : default_footprint(net);
  return submit(s, j, job);
}

int hora_hip_server_wait(hora_hip_server* s, hora_serve_job_result* out, int32_t max, int32_t* n,
                         int32_t timeout_ms) {
  if (!s || !out || !n || max < 1) return HORA_HIP_E_ARG;
  std::unique_lock<std::mutex> lk(s->m);
  if (s->done.empty() && timeout_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->done.empty(); });
  const int k = std::min<int>(max, (int)s->done.size());
  for (int i = 0; i < k; ++i) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return HORA_HIP_OK;
}

double hora_hip_server_now(const hora_hip_server* s) { return s ? s->now() : 0.0; }

void hora_hip_server_close(hora_hip_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto* c : s->ctxs) hora_hip_close(c);
  delete s;
}

}  // extern "C"
#endif  // HORA_FED_HBM
