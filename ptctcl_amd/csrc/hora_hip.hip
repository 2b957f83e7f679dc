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
//   k_hash     discrete key -> slot of the device-resident passed list (open addressing over a
//              64-bit CAS), plus a ticket in that slot's per-level work list
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
#include <cstring>
#include <memory>
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

struct DevNet {
  int d, d2, np, ni, nclocks, lu_plus;
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
  const unsigned char* chan_bcast;
  const unsigned char* chan_urgent;
  int has_urgent_chan;
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
  const int* lr;
  const int* ic;
  const int* cc;
};

struct Counters {
  unsigned long long subsumed, killed, unions, reads, inserted, keys, live, zcount, ncount;
  int err;
  int first_target;
  int ntouched;
  int pad;
};

enum { ERR_NONE = 0, ERR_HASH_FULL = 1, ERR_ZONES_FULL = 2, ERR_NODES_FULL = 3 };

// ---------------------------------------------------------------- discrete state helpers

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
    if (!atom_holds(ints[N.ig_data[3 * c]], N.ig_data[3 * c + 1], N.ig_data[3 * c + 2])) return false;
  return true;
}

__device__ __forceinline__ bool committed_at(const DevNet& N, int p, int l) {
  return (N.loc_flags[N.locbase[p] + l] & 1) != 0;
}

// Lanes decode the packed key (DiscKeyCodecT::decode, engine.cpp:1018-1025) into LDS.
__device__ __forceinline__ void decode_key(const DevNet& N, u64 dk, int* locs, int* ints) {
  const int l = threadIdx.x;
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
__device__ bool urgent_sync_enabled(const DevNet& N, const int* locs, const int* ints) {
  if (!N.has_urgent_chan) return false;
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + locs[i];
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || !N.chan_urgent[ch]) continue;
      if (!int_ok(N, se, ints)) continue;
      if (N.chan_bcast[ch]) return true;
      for (int j = 0; j < N.np; ++j) {
        if (j == i) continue;
        const int gj = N.locbase[j] + locs[j];
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
__device__ int frozen_flags(const DevNet& N, const int* locs, const int* ints) {
  int f = 0;
  for (int p = 0; p < N.np; ++p) f |= N.loc_flags[N.locbase[p] + locs[p]] & 3;
  if (!(f & 2) && urgent_sync_enabled(N, locs, ints)) f |= 2;
  return f;
}

// Enumerate transition recipes of a discrete state in gen_succ_reuse's order: local edges
// (process order, then edge order), then binary synchronisations (sender process and edge, then
// receiver process and edge). `fn(se, re)` gets global edge ids (re = -1 for a local step).
// Broadcast channels are refused at hora_hip_open.
template <class F>
__device__ void enum_recipes(const DevNet& N, const int* locs, const int* ints, bool src_c, F&& fn) {
  for (int p = 0; p < N.np; ++p) {
    const int lp = locs[p];
    if (src_c && !committed_at(N, p, lp)) continue;
    const int g = N.locbase[p] + lp;
    for (int t = N.le_off[g]; t < N.le_off[g + 1]; ++t) {
      const int e = N.le_edges[t];
      if (N.edge_chan[e] >= 0) continue;
      if (!int_ok(N, e, ints)) continue;
      fn(e, -1);
    }
  }
  for (int i = 0; i < N.np; ++i) {
    const int gi = N.locbase[i] + locs[i];
    const bool i_comm = committed_at(N, i, locs[i]);
    for (int t = N.le_off[gi]; t < N.le_off[gi + 1]; ++t) {
      const int se = N.le_edges[t];
      const int ch = N.edge_chan[se];
      if (ch < 0 || N.edge_dir[se] != 1 || N.chan_bcast[ch]) continue;
      if (!int_ok(N, se, ints)) continue;
      for (int j = 0; j < N.np; ++j) {
        if (j == i) continue;
        const int gj = N.locbase[j] + locs[j];
        for (int t2 = N.le_off[gj]; t2 < N.le_off[gj + 1]; ++t2) {
          const int re = N.le_edges[t2];
          if (N.edge_chan[re] != ch || N.edge_dir[re] != -1) continue;
          if (src_c && !(i_comm || committed_at(N, j, locs[j]))) continue;
          if (!int_ok(N, re, ints)) continue;
          fn(se, re);
        }
      }
    }
  }
}

// Conjoin the combined invariant of `locs` (all_invariants + apply_all, semantics.hpp:82-86).
__device__ bool apply_inv(const DevNet& N, const int* locs, int* z) {
  for (int p = 0; p < N.np; ++p) {
    const int g = N.locbase[p] + locs[p];
    for (int c = N.inv_off[g]; c < N.inv_off[g + 1]; ++c)
      if (!z_constrain(z, N.d, N.inv_data[3 * c], N.inv_data[3 * c + 1], N.inv_data[3 * c + 2])) return false;
  }
  return true;
}

__device__ bool apply_edge_guard(const DevNet& N, int e, int* z) {
  for (int c = N.cg_off[e]; c < N.cg_off[e + 1]; ++c)
    if (!z_constrain(z, N.d, N.cg_data[3 * c], N.cg_data[3 * c + 1], N.cg_data[3 * c + 2])) return false;
  return true;
}

__device__ void apply_edge_resets(const DevNet& N, int e, int* z) {
  for (int r = N.rs_off[e]; r < N.rs_off[e + 1]; ++r) z_reset(z, N.d, N.rs_clk[r]);
}

// close_succ_zone (semantics.hpp:522-568) followed by the caller's emptiness test.
__device__ bool close_succ(const DevNet& N, const int* locs, const int* ints, int* z, int* row0) {
  if (!apply_inv(N, locs, z)) return false;
  const int fz = frozen_flags(N, locs, ints);
  if (!fz) {
    z_up(z, N.d);
    apply_inv(N, locs, z);
  }
  u64 a = N.qbits;
  for (int p = 0; p < N.np; ++p) a |= N.amask[N.locbase[p] + locs[p]];
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
__device__ bool is_target(const DevNet& N, const int* locs, const int* ints, const int* z, int* tmp) {
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
};

__global__ __launch_bounds__(64) void k_count(DevNet N, const u64* __restrict__ fr_dk, int F,
                                              int* __restrict__ cnt) {
  __shared__ int locs[HORA_MAX_PROCS], ints[HORA_MAX_INTS];
  const int i = blockIdx.x;
  if (i >= F) return;
  decode_key(N, fr_dk[i], locs, ints);
  wsync();
  const bool src_c = (frozen_flags(N, locs, ints) & 1) != 0;
  int r = 0;
  enum_recipes(N, locs, ints, src_c, [&](int, int) { ++r; });
  if (threadIdx.x == 0) cnt[i] = r;
}

// LDS layout of k_expand: 3 zones + row-0 scratch + 2 discrete states.
__host__ __device__ constexpr int expand_lds_ints(int d2) {
  return 3 * d2 + HORA_MAX_DIM + 2 * (HORA_MAX_PROCS + HORA_MAX_INTS);
}

// Successors of frontier states [0, F) into slots rec_off[i] + r. With `init` set, the frontier
// is the single initial state: its zone is the zero zone and its one recipe is "no transition".
__global__ __launch_bounds__(64) void k_expand(DevNet N, const u64* __restrict__ fr_dk,
                                               const u32* __restrict__ fr_zid, const int* __restrict__ zarena,
                                               int F, const int* __restrict__ rec_off, ExpandOut out,
                                               Counters* ctr, int init) {
  extern __shared__ int sm[];
  int* src = sm;
  int* z = src + N.d2;
  int* tmp = z + N.d2;
  int* row0 = tmp + N.d2;
  int* locs = row0 + HORA_MAX_DIM;
  int* ints = locs + HORA_MAX_PROCS;
  int* nl = ints + HORA_MAX_INTS;
  int* niv = nl + HORA_MAX_PROCS;
  const int i = blockIdx.x;
  if (i >= F) return;
  const int l = threadIdx.x, d2 = N.d2;
  decode_key(N, fr_dk[i], locs, ints);
  if (init) {
    for (int e = l; e < d2; e += kWave) src[e] = kLe0;
  } else {
    const int* zs = zarena + (size_t)fr_zid[i] * d2;
    for (int e = l; e < d2; e += kWave) src[e] = zs[e];
  }
  wsync();
  auto emit = [&](int slot, int se, int re) {
    wsync();
    if (l < N.np) nl[l] = locs[l];
    if (l < N.ni) niv[l] = ints[l];
    wsync();
    if (l == 0 && se >= 0) {
      nl[N.edge_proc[se]] = N.edge_tgt[se];
      if (re >= 0) nl[N.edge_proc[re]] = N.edge_tgt[re];
      for (int c = N.iu_off[se]; c < N.iu_off[se + 1]; ++c) niv[N.iu_data[2 * c]] = N.iu_data[2 * c + 1];
      if (re >= 0)
        for (int c = N.iu_off[re]; c < N.iu_off[re + 1]; ++c) niv[N.iu_data[2 * c]] = N.iu_data[2 * c + 1];
    }
    z_copy(z, src, d2);
    bool ok = true;
    if (se >= 0) {
      ok = apply_edge_guard(N, se, z);
      if (ok && re >= 0) ok = apply_edge_guard(N, re, z);
      if (ok) {
        apply_edge_resets(N, se, z);
        if (re >= 0) apply_edge_resets(N, re, z);
      }
    }
    wsync();
    if (ok) ok = close_succ(N, nl, niv, z, row0);
    const bool tg = ok && is_target(N, nl, niv, z, tmp);
    wsync();
    if (l == 0) {
      out.valid[slot] = ok ? 1 : 0;
      out.tgt[slot] = tg ? 1 : 0;
      out.dk[slot] = ok ? encode_key(N, nl, niv) : 0;
      if (tg) atomicMin(&ctr->first_target, slot);
    }
    if (ok) {
      int* dst = out.zone + (size_t)slot * d2;
      for (int e = l; e < d2; e += kWave) dst[e] = z[e];
    }
  };
  if (init) {
    emit(0, -1, -1);
    return;
  }
  const bool src_c = (frozen_flags(N, locs, ints) & 1) != 0;
  int slot = rec_off[i];
  enum_recipes(N, locs, ints, src_c, [&](int se, int re) { emit(slot++, se, re); });
}

// ---------------------------------------------------------------- scans

// Exclusive scan, three phases: per-block totals, a scan of the totals, per-block downsweep.
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int* __restrict__ in, long n,
                                                            long long* __restrict__ bsum) {
  __shared__ long long s[kScanBlock];
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
                                                          long long* __restrict__ out) {
  __shared__ long long s[kScanBlock];
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

__global__ void k_offsets_to_int(const long long* __restrict__ in, long n, int* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int)in[i];
}

// compaction of the non-empty successor slots: sj[pos[r]] = r where valid[r]
__global__ void k_compact_valid(const int* __restrict__ valid, const long long* __restrict__ pos, long R,
                                int* __restrict__ sj) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R && valid[r]) sj[pos[r]] = (int)r;
}

// ---------------------------------------------------------------- passed list: hashing

struct Passed {
  u64* key;        // [H] packed discrete key or kEmptyKey
  u64* off;        // [H] chain segment offset into nodes
  u32* len;        // [H] live chain length
  u32* ncnt;       // [H] successors landing here this level
  u64* noff;       // [H] offset of this level's work list
  u32* nodes;      // chain segments (zone ids, oldest first)
  int* zarena;     // stored zones
  int* zlevel;     // [zones] BFS level a zone was stored at
  int* zsrc;       // [zones] compacted successor index that stored it
  u64 hmask;
  int hshift;
  u64 zcap, ncap;
};

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

__global__ void k_hash(Passed P, const int* __restrict__ sj, const u64* __restrict__ sdk, long G,
                       long* __restrict__ gbk, u32* __restrict__ gticket, u32* __restrict__ touched,
                       Counters* ctr) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const u64 dk = sdk[sj[g]];
  bool ins = false;
  const long h = find_or_insert(P, dk, &ins);
  if (h < 0) { atomicExch(&ctr->err, ERR_HASH_FULL); gbk[g] = -1; return; }
  if (ins) atomicAdd(&ctr->keys, 1ull);
  gbk[g] = h;
  const u32 t = atomicAdd(&P.ncnt[h], 1u);
  gticket[g] = t;
  if (t == 0) touched[atomicAdd(&ctr->ntouched, 1)] = (u32)h;
}

__host__ __device__ inline u64 list_cap(u64 n) {
  if (n <= (u64)kSortLds) return n;
  u64 c = 1;
  while (c < n) c <<= 1;
  return c;
}

__global__ void k_touched_caps(Passed P, const u32* __restrict__ touched, int T, int* __restrict__ caps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T) caps[t] = (int)list_cap(P.ncnt[touched[t]]);
}

__global__ void k_touched_offsets(Passed P, const u32* __restrict__ touched, int T,
                                  const long long* __restrict__ off) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T) P.noff[touched[t]] = (u64)off[t];
}

__global__ void k_scatter(Passed P, const long* __restrict__ gbk, const u32* __restrict__ gticket, long G,
                          int* __restrict__ wlist) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const long h = gbk[g];
  if (h < 0) return;
  wlist[P.noff[h] + gticket[g]] = (int)g;
}

// ---------------------------------------------------------------- passed list: merge

__device__ void bitonic_sort(int* a, int n) {   // n a power of two, ascending
  const int l = threadIdx.x;
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      wsync();
      for (int i = l; i < n; i += kWave) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int x = a[i], y = a[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { a[i] = y; a[ixj] = x; }
        }
      }
    }
  wsync();
}

// LDS layout of k_merge.
__host__ __device__ constexpr int merge_lds_ints(int d2) {
  return 4 * d2 + 2 * HORA_MAX_DIM + HORA_MAX_PROCS + HORA_MAX_INTS + 8 + kSortLds;
}

__global__ __launch_bounds__(64) void k_merge(DevNet N, Passed P, const u32* __restrict__ touched,
                                              int* __restrict__ wlist, const int* __restrict__ sj,
                                              const int* __restrict__ szone, int* __restrict__ survived,
                                              int level, Counters* ctr) {
  extern __shared__ int sm[];
  const int d = N.d, d2 = N.d2, l = threadIdx.x;
  int* S = sm;
  int* M = S + d2;
  int* fa = M + d2;
  int* pa = fa + d2;
  int* Lb = pa + d2;
  int* Ub = Lb + HORA_MAX_DIM;
  int* locs = Ub + HORA_MAX_DIM;
  int* ints = locs + HORA_MAX_PROCS;
  int* misc = ints + HORA_MAX_INTS;          // [0..1] try_union scratch, [2] zone id, [3] flag
  int* sbuf = misc + 8;

  const u32 h = touched[blockIdx.x];
  const u32 n = P.ncnt[h];
  int* lst = wlist + P.noff[h];
  // order this key's new successors by compacted index = (state, recipe) order
  const u64 cap = list_cap(n);
  if (n <= (u32)kSortLds) {
    int p2 = 1;
    while (p2 < (int)n) p2 <<= 1;
    for (int k = l; k < p2; k += kWave) sbuf[k] = k < (int)n ? lst[k] : 0x7fffffff;
    bitonic_sort(sbuf, p2);
    lst = sbuf;
  } else {
    for (u64 k = n + l; k < cap; k += kWave) lst[k] = 0x7fffffff;
    bitonic_sort(lst, (int)cap);
  }

  // a_LU bounds of this key (state_bounds, engine.cpp:1487-1497)
  decode_key(N, P.key[h], locs, ints);
  wsync();
  if (l < d) {
    int lb = kAluNegInf, ub = kAluNegInf;
    if (l > 0) {
      for (int p = 0; p < N.np; ++p) {
        const size_t b = (size_t)(N.locbase[p] + locs[p]) * d + l;
        lb = max(lb, N.lloc[b]);
        ub = max(ub, N.uloc[b]);
      }
      lb = max(lb, N.Lq[l]);
      ub = max(ub, N.Uq[l]);
    } else {
      lb = ub = 0;
    }
    Lb[l] = lb;
    Ub[l] = ub;
  }

  // working chain: a fresh segment holding the old live chain, room for every new zone
  const u32 old_len = P.len[h];
  const u64 old_off = P.off[h];
  if (l == 0) {
    const u64 seg = atomicAdd(&ctr->ncount, (unsigned long long)(old_len + n));
    misc[4] = (int)(seg & 0xffffffffu);
    misc[5] = (int)(seg >> 32);
  }
  wsync();
  const u64 seg = (u64)(u32)misc[4] | ((u64)(u32)misc[5] << 32);
  if (seg + old_len + n > P.ncap) {
    if (l == 0) atomicExch(&ctr->err, ERR_NODES_FULL);
    return;
  }
  u32* ch = P.nodes + seg;
  for (u32 k = l; k < old_len; k += kWave) ch[k] = P.nodes[old_off + k];
  wsync();
  int len = (int)old_len, live = (int)old_len;
  unsigned long long n_sub = 0, n_kill = 0, n_uni = 0, n_reads = 0, n_ins = 0;

  for (u32 k = 0; k < n; ++k) {
    const int g = lst[k];
    const int* sz = szone + (size_t)sj[g] * d2;
    wsync();
    for (int e = l; e < d2; e += kWave) S[e] = sz[e];
    wsync();
    if (live > 0) {
      // (1)+(2) covered by a live stored zone: drop
      bool covered = false;
      for (int m = len - 1; m >= 0 && !covered; --m) {
        const u32 id = ch[m];
        if (id & kDead) continue;
        ++n_reads;
        covered = z_alu_le(S, P.zarena + (size_t)id * d2, d, Lb, Ub);
      }
      if (covered) { ++n_sub; continue; }
      // (3) retire stored zones covered by S
      for (int m = len - 1; m >= 0; --m) {
        const u32 id = ch[m];
        if (id & kDead) continue;
        ++n_reads;
        if (z_alu_le(P.zarena + (size_t)id * d2, S, d, Lb, Ub)) {
          wsync();
          if (l == 0) ch[m] = id | kDead;
          --live;
          ++n_kill;
        }
      }
      // (4) convex union with the kMergeScan newest live zones, restarting after each merge
      for (bool merged = true; merged;) {
        merged = false;
        int seen = 0;
        for (int m = len - 1; m >= 0 && seen < kMergeScan; --m) {
          wsync();
          const u32 id = ch[m];
          if (id & kDead) continue;
          ++seen;
          ++n_reads;
          const int* mz = P.zarena + (size_t)id * d2;
          for (int e = l; e < d2; e += kWave) M[e] = mz[e];
          if (z_try_union(S, M, d, fa, pa, misc)) {
            wsync();
            if (l == 0) ch[m] = id | kDead;
            --live;
            ++n_uni;
            merged = true;
            break;
          }
        }
      }
    }
    // (5) append S
    wsync();
    if (l == 0) {
      const unsigned long long zid = atomicAdd(&ctr->zcount, 1ull);
      misc[2] = (int)zid;
      misc[3] = zid < P.zcap ? 1 : 0;
    }
    wsync();
    if (!misc[3]) {
      if (l == 0) atomicExch(&ctr->err, ERR_ZONES_FULL);
      return;
    }
    const u32 zid = (u32)misc[2];
    int* dst = P.zarena + (size_t)zid * d2;
    for (int e = l; e < d2; e += kWave) dst[e] = S[e];
    if (l == 0) {
      P.zlevel[zid] = level;
      P.zsrc[zid] = g;
      ch[len] = zid;
    }
    ++len;
    ++live;
    ++n_ins;
    wsync();
  }
  // compact the live chain in place, oldest first; report this level's survivors
  wsync();
  if (l == 0) {
    int w = 0;
    for (int m = 0; m < len; ++m) {
      const u32 id = ch[m];
      if (id & kDead) continue;
      ch[w++] = id;
      if (P.zlevel[id] == level) survived[P.zsrc[id]] = (int)id;
    }
    P.off[h] = seg;
    P.len[h] = (u32)w;
    P.ncnt[h] = 0;
    atomicAdd(&ctr->subsumed, n_sub);
    atomicAdd(&ctr->killed, n_kill);
    atomicAdd(&ctr->unions, n_uni);
    atomicAdd(&ctr->reads, n_reads);
    atomicAdd(&ctr->inserted, n_ins);
    atomicAdd(&ctr->live, (unsigned long long)((long long)w - (long long)old_len));
  }
}

// next frontier from survivors, in compacted-successor order
__global__ void k_flag_survivors(const int* __restrict__ survived, long G, int* __restrict__ flag) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) flag[g] = survived[g] >= 0 ? 1 : 0;
}

__global__ void k_build_frontier(const int* __restrict__ survived, const long long* __restrict__ pos, long G,
                                 const int* __restrict__ sj, const u64* __restrict__ sdk,
                                 u64* __restrict__ fr_dk, u32* __restrict__ fr_zid) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int id = survived[g];
  if (id < 0) return;
  const long p = pos[g];
  fr_dk[p] = sdk[sj[g]];
  fr_zid[p] = (u32)id;
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
  const int i = k / d, j = k - i * d;
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
    }
    wsync();
    if (A.out) {
      int* dst = A.out + z * d2;
      for (int e = l; e < d2; e += kWave) dst[e] = Z[e];
    }
    if (l == 0 && A.flag) A.flag[z] = flag;
  }
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

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    // work queued on any stream may still read the old buffer
    if (p) { (void)hipDeviceSynchronize(); (void)hipFree(p); }
    p = nullptr;
    cap = 0;
  }
  // scratch semantics: contents are not preserved on growth
  void ensure(size_t n) {
    if (n <= cap && p) return;
    release();
    size_t c = std::max<size_t>(n, 1);
    c = c + c / 4 + 1024;
    HIP_OK(hipMalloc(&p, c * sizeof(T)));
    cap = c;
  }
  // persistent semantics: contents preserved on growth
  void grow_keep(size_t n, hipStream_t s) {
    if (n <= cap && p) return;
    size_t c = std::max<size_t>(n, cap * 2);
    T* q = nullptr;
    HIP_OK(hipMalloc(&q, c * sizeof(T)));
    if (p) {
      HIP_OK(hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
      HIP_OK(hipStreamSynchronize(s));
      (void)hipFree(p);
    }
    p = q;
    cap = c;
  }
  void upload(const std::vector<T>& v) {
    ensure(v.size());
    if (!v.empty()) HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

}  // namespace

struct hora_hip_ctx {
  std::string err;
  hipStream_t stream = nullptr;
  int d = 0, d2 = 0, np = 0, ni = 0, nclocks = 0, tl = 0, ne = 0;
  // static tables
  DBuf<int> locbase, loc_flags, inv_off, inv_data, le_off, le_edges, edge_proc, edge_tgt, edge_chan, edge_dir,
      cg_off, cg_data, rs_off, rs_clk, ig_off, ig_data, iu_off, iu_data, Lx, Ux, lloc, uloc, Lq, Uq, kshift,
      int_lo, lr, ic, cc;
  DBuf<unsigned char> chan_bcast, chan_urgent;
  DBuf<u64> amask, kmask;
  DevNet dn{};
  std::vector<int> h_loc_init, h_int_init, h_kshift, h_int_lo;
  std::vector<u64> h_kmask;
  int has_target_clock = 0;
  // search state
  DBuf<u64> ht_key, ht_off, ht_noff;
  DBuf<u32> ht_len, ht_ncnt;
  DBuf<u32> nodes;
  DBuf<int> zarena, zlevel, zsrc;
  u64 H = 0;
  int hbits = 0;
  DBuf<u64> fr_dk, fr_dk2;
  DBuf<u32> fr_zid, fr_zid2;
  DBuf<int> rec_cnt, rec_off, s_valid, s_zone, s_tgt, sj, caps, wlist, survived, flag_i, off_i;
  DBuf<u64> s_dk;
  DBuf<long> gbk;
  DBuf<u32> gticket, touched;
  DBuf<long long> pos, bsum, scan_total;
  DBuf<Counters> ctr;
  // profile of the last search
  double t_expand = 0, t_merge = 0, t_other = 0, b_expand = 0, b_merge = 0;
  hipEvent_t ev[8] = {};
};

namespace {

long long scan_exclusive(hora_hip_ctx* c, const int* in, long n, long long* out) {
  const long per = (long)kScanBlock * kScanItems;
  const long nb = std::max<long>(1, (n + per - 1) / per);
  c->bsum.ensure(nb);
  c->scan_total.ensure(1);
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, c->stream, in, n, c->bsum.p);
  } else {
    HIP_OK(hipMemsetAsync(c->bsum.p, 0, sizeof(long long), c->stream));
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, c->stream, c->bsum.p, nb, c->scan_total.p);
  if (n > 0)
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanBlock), 0, c->stream, in, n, c->bsum.p, out);
  long long total = 0;
  HIP_OK(hipMemcpyAsync(&total, c->scan_total.p, sizeof(total), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
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
  P.nodes = c->nodes.p;
  P.zarena = c->zarena.p;
  P.zlevel = c->zlevel.p;
  P.zsrc = c->zsrc.p;
  P.hmask = c->H - 1;
  P.hshift = 64 - c->hbits;
  P.zcap = std::min<u64>(std::min<u64>(c->zarena.cap / (size_t)c->d2, c->zlevel.cap), c->zsrc.cap);
  P.ncap = c->nodes.cap;
  return P;
}

void alloc_table(hora_hip_ctx* c, int hbits) {
  c->hbits = hbits;
  c->H = 1ull << hbits;
  c->ht_key.ensure(c->H);
  c->ht_off.ensure(c->H);
  c->ht_noff.ensure(c->H);
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

int run_search(hora_hip_ctx* c, int32_t* reachable, hora_reach_stats* st) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  st->emode_switch_level = -1;
  *reachable = 0;
  c->t_expand = c->t_merge = c->t_other = c->b_expand = c->b_merge = 0;
  const int d2 = c->d2;
  c->dn.lu_plus = 0;

  // passed list, arenas
  alloc_table(c, 20);
  c->zarena.ensure((size_t)d2 << 20);
  c->zlevel.ensure(1 << 20);
  c->zsrc.ensure(1 << 20);
  c->nodes.ensure(1 << 22);
  c->ctr.ensure(1);
  HIP_OK(hipMemsetAsync(c->ctr.p, 0, sizeof(Counters), c->stream));
  Counters h{};
  h.first_target = 0x7fffffff;
  HIP_OK(hipMemcpyAsync(c->ctr.p, &h, sizeof(Counters), hipMemcpyHostToDevice, c->stream));

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
  std::vector<std::pair<hipEvent_t, hipEvent_t>> dummy;
  for (;;) {
    Counters hc{};
    // ---- recipe counts and offsets
    long R = 1;
    c->rec_cnt.ensure(F);
    c->rec_off.ensure(F);
    c->pos.ensure(F + 1);
    if (!init) {
      st->explored += (u64)F;
      ++level;
      hipLaunchKernelGGL(k_count, dim3(F), dim3(64), 0, c->stream, c->dn, c->fr_dk.p, (int)F, c->rec_cnt.p);
      R = scan_exclusive(c, c->rec_cnt.p, F, c->pos.p);
      hipLaunchKernelGGL(k_offsets_to_int, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->pos.p, F, c->rec_off.p);
    } else {
      HIP_OK(hipMemsetAsync(c->rec_off.p, 0, sizeof(int), c->stream));
    }
    // ---- expansion
    c->s_valid.ensure(R);
    c->s_tgt.ensure(R);
    c->s_dk.ensure(R);
    c->s_zone.ensure((size_t)R * d2);
    ExpandOut eo{c->s_valid.p, c->s_dk.p, c->s_zone.p, c->s_tgt.p};
    const size_t lds_e = expand_lds_ints(d2) * sizeof(int);
    HIP_OK(hipEventRecord(c->ev[0], c->stream));
    if (R > 0)
      hipLaunchKernelGGL(k_expand, dim3(F), dim3(64), lds_e, c->stream, c->dn, c->fr_dk.p, c->fr_zid.p,
                         c->zarena.p, (int)F, c->rec_off.p, eo, c->ctr.p, init ? 1 : 0);
    HIP_OK(hipEventRecord(c->ev[1], c->stream));
    // ---- compaction of non-empty successors
    c->pos.ensure(R + 1);
    const long Gall = R > 0 ? scan_exclusive(c, c->s_valid.p, R, c->pos.p) : 0;
    c->t_expand += ev_ms(c->ev[0], c->ev[1]);
    c->b_expand += 4.0 * d2 * (double)(F + Gall);
    HIP_OK(hipMemcpyAsync(&hc, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (!init) st->expanded += (u64)F;
    long G = Gall;
    const bool found = hc.first_target != 0x7fffffff;
    if (found) {
      // The sequential engine inserts every successor that precedes the first target, then
      // stops; generated counts those plus the target itself (engine.cpp:2085-2095).
      long long before = 0;
      if (hc.first_target > 0)
        HIP_OK(hipMemcpy(&before, c->pos.p + hc.first_target, sizeof(before), hipMemcpyDeviceToHost));
      if (!init) st->generated += (u64)before + 1;
      st->found = 1;
      *reachable = 1;
      G = (long)before;
    } else if (!init) {
      st->generated += (u64)G;
    }
    if (G == 0) break;
    c->sj.ensure(Gall);
    hipLaunchKernelGGL(k_compact_valid, dim3(blocks_for(R, 256)), dim3(256), 0, c->stream, c->s_valid.p, c->pos.p, R, c->sj.p);
    // ---- capacity: hash load, zones, chain nodes
    while ((double)(hc.keys + (u64)G) > 0.6 * (double)c->H) grow_table(c);
    if (hc.zcount + (u64)G > passed_view(c).zcap) {
      const size_t want = (size_t)(hc.zcount + (u64)G) * 2;
      c->zarena.grow_keep(want * d2, c->stream);
      c->zlevel.grow_keep(want, c->stream);
      c->zsrc.grow_keep(want, c->stream);
    }
    if (hc.ncount + hc.live + (u64)G + 1 > c->nodes.cap) {
      // chain segments are rewritten per level; compact by copying only live chains is left to
      // growth here: every live chain is at most `live` entries and a level appends <= live + G
      c->nodes.grow_keep((size_t)(hc.ncount + hc.live + (u64)G + 1) * 2, c->stream);
    }
    Passed P = passed_view(c);
    // ---- hash + per-key work lists
    c->gbk.ensure(G);
    c->gticket.ensure(G);
    c->touched.ensure(G);
    HIP_OK(hipMemsetAsync(&c->ctr.p->ntouched, 0, sizeof(int), c->stream));
    HIP_OK(hipEventRecord(c->ev[2], c->stream));
    hipLaunchKernelGGL(k_hash, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->sj.p, c->s_dk.p, G,
                       c->gbk.p, c->gticket.p, c->touched.p, c->ctr.p);
    int T = 0;
    HIP_OK(hipMemcpyAsync(&T, &c->ctr.p->ntouched, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->caps.ensure(T);
    hipLaunchKernelGGL(k_touched_caps, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T, c->caps.p);
    c->pos.ensure(T + 1);
    const long long W = scan_exclusive(c, c->caps.p, T, c->pos.p);
    hipLaunchKernelGGL(k_touched_offsets, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T, c->pos.p);
    c->wlist.ensure(W + 1);
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->gbk.p, c->gticket.p, G, c->wlist.p);
    // ---- merge
    c->survived.ensure(G);
    HIP_OK(hipMemsetAsync(c->survived.p, 0xff, (size_t)G * sizeof(int), c->stream));
    const size_t lds_m = merge_lds_ints(d2) * sizeof(int);
    HIP_OK(hipEventRecord(c->ev[3], c->stream));
    hipLaunchKernelGGL(k_merge, dim3(T), dim3(64), lds_m, c->stream, c->dn, P, c->touched.p, c->wlist.p, c->sj.p,
                       c->s_zone.p, c->survived.p, level, c->ctr.p);
    HIP_OK(hipEventRecord(c->ev[4], c->stream));
    if (found) {
      HIP_OK(hipMemcpyAsync(&hc, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
      HIP_OK(hipStreamSynchronize(c->stream));
      c->t_merge += ev_ms(c->ev[3], c->ev[4]);
      if (hc.err) { c->err = "device capacity exceeded"; return HORA_HIP_E_MEMORY; }
      break;
    }
    // ---- next frontier
    c->flag_i.ensure(G);
    hipLaunchKernelGGL(k_flag_survivors, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, c->survived.p, G, c->flag_i.p);
    c->pos.ensure(G + 1);
    const long F2 = scan_exclusive(c, c->flag_i.p, G, c->pos.p);
    c->fr_dk2.ensure(F2);
    c->fr_zid2.ensure(F2);
    hipLaunchKernelGGL(k_build_frontier, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, c->survived.p, c->pos.p, G,
                       c->sj.p, c->s_dk.p, c->fr_dk2.p, c->fr_zid2.p);
    HIP_OK(hipMemcpyAsync(&hc, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    const double reads_before = (double)st->chain_reads;
    st->chain_reads = hc.reads;
    c->t_merge += ev_ms(c->ev[3], c->ev[4]);
    c->t_other += ev_ms(c->ev[2], c->ev[3]);
    // merge bytes: new zones read, stored zones read by the folds, zones written
    c->b_merge += 4.0 * d2 * ((double)G + ((double)hc.reads - reads_before)) + 4.0 * d2 * (double)F2;
    if (hc.err) {
      c->err = hc.err == ERR_HASH_FULL ? "passed-list hash table full"
               : hc.err == ERR_ZONES_FULL ? "zone arena full" : "chain node arena full";
      return HORA_HIP_E_MEMORY;
    }
    std::swap(c->fr_dk.p, c->fr_dk2.p); std::swap(c->fr_dk.cap, c->fr_dk2.cap);
    std::swap(c->fr_zid.p, c->fr_zid2.p); std::swap(c->fr_zid.cap, c->fr_zid2.cap);
    F = F2;
    init = false;
    // adaptive Extra_LU+ switch (engine.cpp:2178-2184)
    if (!c->dn.lu_plus && hc.keys && (double)hc.live >= 4.0 * (double)hc.keys) {
      c->dn.lu_plus = 1;
      st->emode_switch_level = level;
    }
    if (F == 0) break;
  }
  Counters hc{};
  HIP_OK(hipMemcpy(&hc, c->ctr.p, sizeof(Counters), hipMemcpyDeviceToHost));
  st->levels = level;
  st->stored_live = hc.live;
  st->keys = hc.keys;
  st->inserted = hc.inserted;
  st->subsumed = hc.subsumed;
  st->killed = hc.killed;
  st->unions = hc.unions;
  st->chain_reads = hc.reads;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

template <class T>
std::vector<T> vec(const T* p, long n) {
  return p && n > 0 ? std::vector<T>(p, p + n) : std::vector<T>(std::max<long>(n, 1), T{});
}

void open_ctx(const hora_net_enc* e, const hora_target* t, hora_hip_ctx* c) {
  if (!e) throw ScopeFail("null network");
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
  for (int ch = 0; ch < e->nchans; ++ch)
    if (e->chan_bcast && e->chan_bcast[ch]) throw ScopeFail("broadcast channels are not covered by the device path");
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

  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& ev : c->ev) HIP_OK(hipEventCreate(&ev));
  const int ne = std::max(c->ne, 1);
  c->locbase.upload(vec(e->locbase, c->np + 1));
  c->loc_flags.upload(vec(e->loc_flags, c->tl));
  c->inv_off.upload(vec(e->inv_off, c->tl + 1));
  c->inv_data.upload(vec(e->inv_data, 3L * e->inv_off[c->tl]));
  c->le_off.upload(le_off);
  c->le_edges.upload(le_edges);
  c->edge_proc.upload(vec(e->edge_proc, ne));
  c->edge_tgt.upload(vec(e->edge_tgt, ne));
  c->edge_chan.upload(vec(e->edge_chan, ne));
  c->edge_dir.upload(vec(e->edge_dir, ne));
  c->cg_off.upload(vec(e->cg_off, c->ne + 1));
  c->cg_data.upload(vec(e->cg_data, 3L * e->cg_off[c->ne]));
  c->rs_off.upload(vec(e->rs_off, c->ne + 1));
  c->rs_clk.upload(vec(e->rs_clk, (long)e->rs_off[c->ne]));
  c->ig_off.upload(vec(e->ig_off, c->ne + 1));
  c->ig_data.upload(vec(e->ig_data, 3L * e->ig_off[c->ne]));
  c->iu_off.upload(vec(e->iu_off, c->ne + 1));
  c->iu_data.upload(vec(e->iu_data, 2L * e->iu_off[c->ne]));
  c->chan_bcast.upload(vec(e->chan_bcast, std::max(e->nchans, 1)));
  c->chan_urgent.upload(vec(e->chan_urgent, std::max(e->nchans, 1)));
  c->Lx.upload(Lx);
  c->Ux.upload(Ux);
  c->lloc.upload(vec(e->lbound_loc, (long)c->tl * d));
  c->uloc.upload(vec(e->ubound_loc, (long)c->tl * d));
  c->Lq.upload(Lq);
  c->Uq.upload(Uq);
  c->amask.upload(vec(e->active_mask, c->tl));
  c->kshift.upload(c->h_kshift);
  c->kmask.upload(c->h_kmask);
  c->int_lo.upload(c->h_int_lo);
  c->lr.upload(lrv);
  c->ic.upload(icv);
  c->cc.upload(ccv);

  DevNet& N = c->dn;
  N.d = d; N.d2 = c->d2; N.np = c->np; N.ni = c->ni; N.nclocks = c->nclocks; N.lu_plus = 0;
  N.locbase = c->locbase.p; N.loc_flags = c->loc_flags.p; N.inv_off = c->inv_off.p; N.inv_data = c->inv_data.p;
  N.le_off = c->le_off.p; N.le_edges = c->le_edges.p; N.edge_proc = c->edge_proc.p; N.edge_tgt = c->edge_tgt.p;
  N.edge_chan = c->edge_chan.p; N.edge_dir = c->edge_dir.p;
  N.cg_off = c->cg_off.p; N.cg_data = c->cg_data.p; N.rs_off = c->rs_off.p; N.rs_clk = c->rs_clk.p;
  N.ig_off = c->ig_off.p; N.ig_data = c->ig_data.p; N.iu_off = c->iu_off.p; N.iu_data = c->iu_data.p;
  N.chan_bcast = c->chan_bcast.p; N.chan_urgent = c->chan_urgent.p; N.has_urgent_chan = has_urgent;
  N.Lx = c->Lx.p; N.Ux = c->Ux.p; N.lloc = c->lloc.p; N.uloc = c->uloc.p; N.Lq = c->Lq.p; N.Uq = c->Uq.p;
  N.amask = c->amask.p; N.qbits = qbits; N.kshift = c->kshift.p; N.kmask = c->kmask.p; N.int_lo = c->int_lo.p;
  N.never = never; N.nlr = nlr; N.nic = nic; N.ncc = ncc; N.lr = c->lr.p; N.ic = c->ic.p; N.cc = c->cc.p;
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
  delete ctx;
}

const char* hora_hip_last_error(const hora_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hora_hip_reach(hora_hip_ctx* ctx, int32_t* reachable, hora_reach_stats* stats) {
  if (!ctx || !reachable || !stats || !ctx->stream) return HORA_HIP_E_ARG;
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

int hora_hip_reach_profile(const hora_hip_ctx* ctx, double* expand_ms, double* merge_ms, double* other_ms,
                           double* expand_bytes, double* merge_bytes) {
  if (!ctx) return HORA_HIP_E_ARG;
  if (expand_ms) *expand_ms = ctx->t_expand;
  if (merge_ms) *merge_ms = ctx->t_merge;
  if (other_ms) *other_ms = ctx->t_other;
  if (expand_bytes) *expand_bytes = ctx->b_expand;
  if (merge_bytes) *merge_bytes = ctx->b_merge;
  return HORA_HIP_OK;
}

int hora_hip_zone_op(int32_t op, int32_t d, int64_t n, const int32_t* a, const int32_t* b, const int32_t* arg,
                     const int32_t* L, const int32_t* U, int32_t* out, int32_t* out_flag, void* stream) {
  if (d < 1 || d > HORA_MAX_DIM || n < 0 || !a) return HORA_HIP_E_ARG;
  if (op < HORA_ZOP_CLOSE || op > HORA_ZOP_EMPTY) return HORA_HIP_E_ARG;
  if ((op == HORA_ZOP_RESET || op == HORA_ZOP_CONSTRAIN) && !arg) return HORA_HIP_E_ARG;
  if ((op == HORA_ZOP_EXTRA_LU || op == HORA_ZOP_ALU_LE) && (!L || !U)) return HORA_HIP_E_ARG;
  if ((op == HORA_ZOP_INCLUDES || op == HORA_ZOP_ALU_LE || op == HORA_ZOP_UNION) && !b) return HORA_HIP_E_ARG;
  if (n == 0) return HORA_HIP_OK;
  ZopArgs A{op, d, n, a, b, arg, L, U, out, out_flag};
  hipStream_t s = (hipStream_t)stream;
  if (op == HORA_ZOP_UP || op == HORA_ZOP_RESET) {
    if (!out) return HORA_HIP_E_ARG;
    const long tot = n * (long)d * d;
    hipLaunchKernelGGL(k_zop_stream, dim3(blocks_for(tot, 256)), dim3(256), 0, s, A);
  } else {
    const long grid = std::min<long>(n, 1L << 20);
    const size_t lds = (4 * d * d + HORA_MAX_DIM + 8) * sizeof(int);
    hipLaunchKernelGGL(k_zop_wave, dim3(grid), dim3(64), lds, s, A);
  }
  return hipGetLastError() == hipSuccess ? HORA_HIP_OK : HORA_HIP_E_DEVICE;
}

}  // extern "C"
