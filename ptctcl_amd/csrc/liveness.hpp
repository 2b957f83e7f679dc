// liveness.hpp — device liveness pass of the HIP back end: `p --> q`, `E[] psi`,
// `A[] not deadlock`. Part of the hora_hip.hip translation unit (it uses that file's DevNet,
// expansion kernels, passed-list hash table and scans).
//
// Forward: the zone graph is built level by level on the device with the expansion pipeline of
// the reachability search, with no extrapolation and with every transition recorded. Two
// constructions run here (the route is chosen as the reference's decide_leads_to does,
// src/liveness/api.cpp:357-447):
//   * the online construction (online.hpp): online_intern's storing rule (a_LU covers by the newest
//     active record, convex unions, retirement of covered records, records retired before their
//     turn not expanded; src/liveness/construction.cpp:480-683), the reference's default;
//   * the a_M graph, when a reachable discrete state can seed a deadlock atom (online_seed_key,
//     src/liveness/refinement.cpp:176-282) or when asked for: each successor goes to the first
//     (oldest) stored node of its discrete state that covers it under a_M (a_LU with per-location
//     bounds raised to M = max(L, U): a bisimulation, so deadlocks survive the quotient;
//     construction.cpp:899-913), else it becomes a new node.
// p --> q on a network where D is provably a_LU takes the reference's base construction
// (construction.cpp:863-1096: a FIFO of nodes, each successor covered by the oldest node of its
// discrete state, else enlarging the oldest node it covers in place, else merged by a convex union,
// else a new node), one node per step (base.hpp).
// Node ids are creation order: frontier index, then transition index, as a sequential pass makes
// them. The a_M graph's kernels:
//
//   k_gfilter  E[]: only psi nodes are expanded (src/liveness/api.cpp:83-96)
//   k_count / k_expand / scan     successors in (state, transition) order, transition recorded
//   k_grows    CSR row offsets of the frontier nodes
//   k_find     discrete key -> hash slot
//   k_gprobe   first covering node among the key's nodes stored before this level
//   k_ticket / scatter            per-key work lists of the successors not covered there
//   k_gmerge   one wavefront per key: the key's successors in order, each to the first node
//              created earlier in this level that covers it, else a new node
//   scan       new-node ranks = node ids
//   k_gassign / k_gcopy / k_glink edges, node zones, the key's node chain
//   k_stop     (early witness) stop set of the level's p /\ !q nodes, or deadlock test
//
// Backward (src/liveness/fixpoint.cpp:48-965): closure S of the query (BFS kernels), stop sets
// NoSucc(v) (k_stop: whole zone when no successor or time diverges, else the valuations that enable
// nothing, src/liveness/construction.cpp:1200-1238, src/reachability/deadlock.cpp:74-268), then
// Jacobi rounds of relax (k_gfp: W[v] := W[v] /\ (NoSucc(v) \/ U_e Pre_e(W[t])), exact predecessor
// of src/liveness/operators.cpp:257-306) on two banks of federations until no value changes. The
// limit is the unique greatest fixpoint, so it equals any sequential schedule's.
#pragma once

namespace {

constexpr int kLiveMaxDim = HORA_MAX_DIM;
// LDS one workgroup may take on gfx950: the full-capacity launches (the nodes that outgrow the small
// and middle lists) take all of it.
constexpr int kLdsWorkgroupMax = 160 * 1024;
constexpr int ERR_CAPPED = 5;
constexpr int kLiveW = 64;      // zones in one value W[v] at most (a u64 mask covers them)
constexpr u32 kProbeSerial = 16;  // node segments up to this length are probed one node at a time

struct DevPred {
  int nterms;
  const int* off;
  const int* atoms;
};

// Truth of a DNF predicate on a packed discrete key (each lane may call it; no memory written).
__device__ __forceinline__ bool pred_eval(const DevNet& N, const DevPred& P, u64 dk) {
  for (int t = 0; t < P.nterms; ++t) {
    bool all = true;
    for (int a = P.off[t]; a < P.off[t + 1] && all; ++a) {
      const int kind = P.atoms[4 * a], idx = P.atoms[4 * a + 1], op = P.atoms[4 * a + 2], val = P.atoms[4 * a + 3];
      int v;
      if (kind == 0) v = (int)((dk >> N.kshift[idx]) & N.kmask[idx]);
      else v = (int)((dk >> N.kshift[N.np + idx]) & N.kmask[N.np + idx]) + N.int_lo[idx];
      all = atom_holds(v, op, val);
    }
    if (all) return true;
  }
  return false;
}

// Everything the liveness kernels read besides the network, in device memory.
struct LiveArgs {
  DevPred p, q;
  int kind, restricted;
  const int* oinv_off;        // unrestricted invariant tables (E[] with a clock part)
  const int* oinv_data;
  const u64* gdk;             // [N] node keys
  const int* gzone;           // [N*d2] node zones
  const long long* grow;      // [N+1] CSR rows
  const int* etgt;
  const int* ese;
  const int* ere;
  const signed char* expanded;
  const int* in_s;            // [N] 1 = in the closure
  const int* sidx;            // [N] index into the S-indexed arrays, or -1
  int kl, kp, kc;             // federation capacities (LDS lists), set from the dimension
  int skp;                    // pieces of a stop set's zone difference (its own LDS budget)
  int kw;                     // zones in one value W[v] (the banks' capacity), set from the dimension
  const int* recnt;           // online construction: [N] successors of a node (its row length), else null
  signed char* whit;          // online construction: [N] early witness found at the node, else null
};

constexpr int kLiveStripes = 64;

struct LiveCtr {
  int hit;                    // early witness / deadlock found
  int err;
  int ns;                     // closure list length
  int changed;                // values changed in the last round
  unsigned long long pre_calls;
  unsigned long long products;
  long long nonempty;
  int fed_max;
  int violated;
  int dcnt;                   // nodes a fixpoint round visits
  int rcnt;                   // of those, nodes sent to the full-capacity launch
  int spool;                  // zones taken from the stop-set pool
  int sfull;                  // a stop set did not fit the pool
  int ptop;                   // zones taken from the value pool
  int pfull;                  // a value did not fit the pool
  int wcnt;                   // early-witness candidates of the level
  // online construction, per interning of a level
  int nev;                    // stored records retired (event list length)
  int nuni;                   // convex unions
  int nret;                   // records retired
  int hzn;                    // widened new zones taken from the pool
  int hzfull;                 // the pool was too small
  int vchanged;               // the skip marks of the frontier changed
  int wfirst;                 // first frontier record that is an early witness
  int nexp;                   // frontier records expanded
  int nkeys;                  // discrete states with an active record
  // the fixpoint's per-node statistics, striped over cache lines (one node per workgroup would
  // otherwise meet every other on one address): kLiveStripes x {pre_calls, products, changed},
  // one 64-byte line each; read_lc sums them into the fields above
  unsigned long long stripe[3][kLiveStripes][8];
};

// ---------------------------------------------------------------- forward construction

// The nodes of [first_id, first_id + n) where an early witness can stand: p /\ !q (leads-to) or
// psi (E[] psi), listed for k_stop's mode 1 so that the others take no workgroup.
__global__ void k_wfilter(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap, long n, long long first_id,
                          int* __restrict__ list, LiveCtr* lc) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const LiveArgs& A = *Ap;
  bool keep = false;
  if (b < n) {
    const u64 dk = A.gdk[first_id + b];
    keep = A.kind == HORA_LIVE_EG ? pred_eval(*Np, A.p, dk) : pred_eval(*Np, A.p, dk) && !pred_eval(*Np, A.q, dk);
  }
  const int idx = block_append(&lc->wcnt, keep);
  if (keep) list[idx] = (int)(first_id + b);
}

__global__ void k_gfilter(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ A, int use_psi,
                          const u64* __restrict__ fr_dk, int F, int* __restrict__ cnt,
                          signed char* __restrict__ expanded, long fbase) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F) return;
  const bool keep = !use_psi || pred_eval(*Np, A->p, fr_dk[i]);
  if (!keep) cnt[i] = 0;
  expanded[fbase + i] = keep ? 1 : 0;
}

// rows of the frontier nodes: node fbase+i owns the non-empty successors of its recipe slots
__global__ void k_grows(const int* __restrict__ rec_off, const long long* __restrict__ vpos, long R, int F,
                        long fbase, long long ebase, long long G, long long* __restrict__ grow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > F) return;
  if (i == F) { grow[fbase + F] = ebase + G; return; }
  const long s = rec_off[i];
  grow[fbase + i] = ebase + (s < R ? vpos[s] : G);
}

__global__ void k_gfill_rows(long long* __restrict__ grow, long from, long to, long long E) {
  const long v = from + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v <= to) grow[v] = E;
}

// Covering bounds of a location vector for clock c (`lg` holds the decoded locations): the a_LU
// state bounds (state_bounds, engine.cpp:1487-1497), or with `alu` = 0 both raised to M = max(L, U).
__device__ __forceinline__ void cover_bounds(const DevNet& N, const int* lg, int c, int alu, int* Lc, int* Uc) {
  if (c == 0) { *Lc = 0; *Uc = 0; return; }
  int lb = kAluNegInf, ub = kAluNegInf;
  for (int p = 0; p < N.np; ++p) {
    const size_t b = (size_t)(N.locbase[p] + lg[p]) * N.d + c;
    lb = max(lb, N.lloc[b]);
    ub = max(ub, N.uloc[b]);
  }
  lb = max(lb, N.Lq[c]);
  ub = max(ub, N.Uq[c]);
  if (!alu) lb = ub = max(lb, ub);
  *Lc = lb;
  *Uc = ub;
}

// The first node stored before this level that covers the successor. Four 16-lane groups per
// wavefront, one successor per group, lanes over the matrix entries (coalesced reads of both
// zones); the group walks the key's node segment oldest first and stops at the first cover.
__global__ __launch_bounds__(64) void k_gprobe(const DevNet* __restrict__ Np, Passed P, const long* __restrict__ gbk,
                                               const int* __restrict__ sj, const u64* __restrict__ sdk,
                                               const int* __restrict__ szone, long G, const int* __restrict__ gzone,
                                               int* __restrict__ covered, int* __restrict__ cov_old, Counters* ctr,
                                               int alu, int newest) {
  __shared__ int Lg[4][HORA_MAX_DIM], Ug[4][HORA_MAX_DIM], lg[4][HORA_MAX_PROCS];
  const DevNet& N = *Np;
  const int lane = threadIdx.x, q = lane / kGroup, t = lane % kGroup;
  const long g = (long)blockIdx.x * 4 + q;
  const bool active = g < G;
  const int d = N.d, d2 = N.d2;
  const long h = active ? gbk[g] : -1;
  const u32 len = (h >= 0) ? P.len[h] : 0u;
  if (len > 0) {
    const u64 dk = sdk[sj[g]];
    for (int p = t; p < N.np; p += kGroup) lg[q][p] = (int)((dk >> N.kshift[p]) & N.kmask[p]);
  }
  __syncthreads();
  if (len > 0)
    for (int c = t; c < d; c += kGroup) cover_bounds(N, lg[q], c, alu, &Lg[q][c], &Ug[q][c]);
  __syncthreads();
  const int* S = szone + (size_t)sj[active ? g : 0] * d2;
  const u64 gmask = 0xffffull << (q * kGroup);
  const int* L = Lg[q];
  const int* U = Ug[q];
  const u32* seg = P.nodes + (len > 0 ? P.off[h] : 0);
  int found = -1, reads = 0;
  for (u32 k = 0; k < len; ++k) {   // group-uniform loop
    const u32 m = newest ? len - 1 - k : k;
    const int* M = gzone + (size_t)seg[m] * d2;
    ++reads;
    bool wit = false;   // S <=_aLU M fails (the predicate of z_alu_le)
    for (int e = t; e < d2; e += kGroup) {
      const int y = ediv(e, d), x = e - y * d;
      if (x == y) continue;
      const int ux = U[x], ly = L[y];
      if (ux == kAluNegInf || ly == kAluNegInf) continue;
      const int s0x = S[x];
      if (s0x < r_le(-ux)) continue;
      const int mp = M[e];
      if (mp < S[e] && radd(mp, r_lt(-ly)) < s0x) wit = true;
    }
    if ((__ballot(wit) & gmask) == 0) { found = (int)seg[m]; break; }
  }
  const bool lead = active && t == 0;
  if (lead) {
    covered[g] = found >= 0 ? 1 : 0;
    cov_old[g] = found;
  }
  wave_stat(ctr, C_PREADS, lead ? (unsigned long long)reads : 0ull);
  wave_stat(ctr, C_PSUB, (lead && found >= 0) ? 1ull : 0ull);
}

// One wavefront per touched key: its remaining successors in (state, transition) order, each to
// the first node created earlier in this level that covers it, else a new node. The creators'
// successor indices are kept in order at the front of the key's list, and their number replaces
// the key's queue length for k_glink.
__global__ __launch_bounds__(64) void k_gmerge(const DevNet* __restrict__ Np, Passed P, const u32* __restrict__ touched,
                                               int T, int* __restrict__ wlist, const int* __restrict__ sj,
                                               const u64* __restrict__ sdk, const int* __restrict__ szone,
                                               int* __restrict__ tslot, int* __restrict__ fnew, Counters* ctr,
                                               int alu) {
  extern __shared__ int sm[];
  const DevNet& N = *Np;
  const int d = N.d, d2 = N.d2, l = threadIdx.x;
  int* S = sm;                       // d2
  int* Lb = S + d2;                  // HORA_MAX_DIM
  int* Ub = Lb + HORA_MAX_DIM;       // HORA_MAX_DIM
  int* lg = Ub + HORA_MAX_DIM;       // HORA_MAX_PROCS
  int* lst = lg + HORA_MAX_PROCS;    // kSortLds
  const int t = blockIdx.x;
  if (t >= T) return;
  const u32 h = touched[t];
  const int n = (int)P.ncnt[h];
  const u64 base = P.noff[h];
  const bool in_lds = n <= kSortLds;
  int* list;
  if (in_lds) {
    int cap = 1;
    while (cap < n) cap <<= 1;
    for (int k = l; k < cap; k += kWave) lst[k] = k < n ? wlist[base + k] : 0x7fffffff;
    bitonic_sort<false>(lst, cap);
    list = lst;
  } else {
    const int cap = (int)list_cap((u64)n);
    for (int k = n + l; k < cap; k += kWave) wlist[base + k] = 0x7fffffff;
    bsync();
    bitonic_sort<true>(wlist + base, cap);
    list = wlist + base;
  }
  {
    const u64 dk = sdk[sj[list[0]]];
    if (l < N.np) lg[l] = (int)((dk >> N.kshift[l]) & N.kmask[l]);
    wsync();
    if (l < d) cover_bounds(N, lg, l, alu, &Lb[l], &Ub[l]);
  }
  int nc = 0;
  unsigned long long cov = 0, reads = 0;
  for (int k = 0; k < n; ++k) {
    if (in_lds) wsync(); else bsync();
    const int g = list[k];
    int first = -1;
    if (nc > 0) {
      wsync();
      const int* src = szone + (size_t)sj[g] * d2;
      for (int e = l; e < d2; e += kWave) S[e] = src[e];
      wsync();
      if (nc <= (int)kProbeSerial) {   // few creators: one at a time, lanes over entries
        for (int c = 0; c < nc && first < 0; ++c) {
          if (l == 0) ++reads;
          if (z_alu_le(S, szone + (size_t)sj[list[c]] * d2, d, Lb, Ub)) first = c;
        }
      } else {
        for (int cb = 0; cb < nc && first < 0; cb += kWave) {
          const int c = cb + l;
          bool ok = false;
          if (c < nc) { ok = alu_le_lane(S, szone + (size_t)sj[list[c]] * d2, d, Lb, Ub); ++reads; }
          const u64 b = __ballot(ok);
          if (b) first = cb + __builtin_ctzll(b);
        }
      }
    }
    if (first >= 0) {
      if (l == 0) tslot[g] = list[first];
      ++cov;
      continue;
    }
    if (in_lds) wsync(); else bsync();
    if (l == 0) {
      list[nc] = g;          // nc <= k: only consumed entries are overwritten
      tslot[g] = g;
      fnew[g] = 1;
    }
    ++nc;
    if (!in_lds) { __threadfence_block(); bsync(); }
  }
  wsync();
  if (in_lds)
    for (int k = l; k < nc; k += kWave) wlist[base + k] = lst[k];
  if (l == 0) {
    P.ncnt[h] = (u32)nc;
    stat_add(ctr, C_SUB, cov);
  }
  wave_stat(ctr, C_READS, reads);
}

// ---- heavy keys: the ordered merge decided in parallel
//
// The cover test is a preorder (Z <=_aM Z' iff Z is inside the a_M closure of Z', a closure
// operator), so a successor creates a node iff no EARLIER successor of its key covers it: an
// earlier successor that covers it is either a node or covered by an earlier node, which then
// covers it too. Each successor is tested against all earlier ones independently; a successor
// that does not create a node then goes to the first created node that covers it. The result is
// the sequential merge's.

constexpr int kLightLive = 64;    // keys with at most this many queued successors merge sequentially

__global__ void k_gsplit(Passed P, const u32* __restrict__ touched, int T, u32* __restrict__ light,
                         u32* __restrict__ heavy, Counters* ctr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = t < T;
  const u32 h = ok ? touched[t] : 0u;
  const bool lt = ok && P.ncnt[h] <= (u32)kLightLive;
  const int il = block_append(&ctr->nlight, lt);
  const int ih = block_append(&ctr->nheavy, ok && !lt);
  if (lt) light[il] = h;
  if (ok && !lt) heavy[ih] = h;
}

// one wavefront per heavy key: sort its list; mark each list position with the key's slot
__global__ __launch_bounds__(64) void k_gsort(Passed P, const u32* __restrict__ heavy, int nh,
                                              int* __restrict__ wlist, long long* __restrict__ wkey) {
  extern __shared__ int sm[];
  const int t = blockIdx.x;
  if (t >= nh) return;
  const int l = threadIdx.x;
  const u32 h = heavy[t];
  const int n = (int)P.ncnt[h];
  const u64 base = P.noff[h];
  if (n <= kSortLds) {
    int cap = 1;
    while (cap < n) cap <<= 1;
    for (int k = l; k < cap; k += kWave) sm[k] = k < n ? wlist[base + k] : 0x7fffffff;
    bitonic_sort<false>(sm, cap);
    for (int k = l; k < n; k += kWave) wlist[base + k] = sm[k];
  } else {
    const int cap = (int)list_cap((u64)n);
    for (int k = n + l; k < cap; k += kWave) wlist[base + k] = 0x7fffffff;
    bsync();
    bitonic_sort<true>(wlist + base, cap);
  }
  for (int k = l; k < n; k += kWave) wkey[base + k] = (long long)h;
}

// one wavefront per list position of a heavy key: does an earlier successor cover this one?
__global__ __launch_bounds__(64) void k_gpair(const DevNet* __restrict__ Np, Passed P, const long long* __restrict__ wkey,
                                              long W, const int* __restrict__ wlist, const int* __restrict__ sj,
                                              const u64* __restrict__ sdk, const int* __restrict__ szone,
                                              int* __restrict__ tslot, int* __restrict__ fnew, Counters* ctr,
                                              int alu) {
  __shared__ int S[HORA_MAX_DIM * HORA_MAX_DIM], Lb[HORA_MAX_DIM], Ub[HORA_MAX_DIM], lg[HORA_MAX_PROCS];
  const long p = blockIdx.x;
  if (p >= W) return;
  const long long hh = wkey[p];
  if (hh < 0) return;
  const DevNet& N = *Np;
  const int l = threadIdx.x, d = N.d, d2 = N.d2;
  const u64 base = P.noff[hh];
  const int k = (int)(p - (long)base);
  const int g = wlist[p];
  const u64 dk = sdk[sj[g]];
  if (l < N.np) lg[l] = (int)((dk >> N.kshift[l]) & N.kmask[l]);
  const int* src = szone + (size_t)sj[g] * d2;
  for (int e = l; e < d2; e += kWave) S[e] = src[e];
  wsync();
  if (l < d) cover_bounds(N, lg, l, alu, &Lb[l], &Ub[l]);
  wsync();
  bool cov = false;
  unsigned long long reads = 0;
  for (int b = 0; b < k && !cov; b += kWave) {
    const int m = b + l;
    bool ok = false;
    if (m < k) { ok = alu_le_lane(S, szone + (size_t)sj[wlist[base + m]] * d2, d, Lb, Ub); ++reads; }
    cov = __ballot(ok) != 0;
  }
  wave_stat(ctr, C_READS, reads);
  if (l == 0) {
    tslot[g] = cov ? -2 : g;
    fnew[g] = cov ? 0 : 1;
  }
}

// one wavefront per heavy key: its creators to the front of its list, in order; their number
// replaces the queue length for k_glink and k_gfirst
__global__ __launch_bounds__(64) void k_gcompact(Passed P, const u32* __restrict__ heavy, int nh,
                                                 int* __restrict__ wlist, const int* __restrict__ fnew) {
  const int t = blockIdx.x;
  if (t >= nh) return;
  const int l = threadIdx.x;
  const u32 h = heavy[t];
  const int n = (int)P.ncnt[h];
  const u64 base = P.noff[h];
  int nc = 0;
  for (int b = 0; b < n; b += kWave) {
    const int k = b + l;
    const int g = k < n ? wlist[base + k] : 0;
    const bool cr = k < n && fnew[g];
    const u64 bal = __ballot(cr);
    wsync();
    if (cr) wlist[base + nc + __popcll(bal & ((1ull << l) - 1))] = g;
    nc += __popcll(bal);
    __threadfence_block();
  }
  if (l == 0) P.ncnt[h] = (u32)nc;
}

// one wavefront per successor still pending: the first created node of its key that covers it
__global__ __launch_bounds__(64) void k_gfirst(const DevNet* __restrict__ Np, Passed P, const long* __restrict__ gbk,
                                               long G, const int* __restrict__ wlist, const int* __restrict__ sj,
                                               const u64* __restrict__ sdk, const int* __restrict__ szone,
                                               int* __restrict__ tslot, Counters* ctr, int alu) {
  __shared__ int S[HORA_MAX_DIM * HORA_MAX_DIM], Lb[HORA_MAX_DIM], Ub[HORA_MAX_DIM], lg[HORA_MAX_PROCS];
  const long g = blockIdx.x;
  if (g >= G || tslot[g] != -2) return;
  const DevNet& N = *Np;
  const int l = threadIdx.x, d = N.d, d2 = N.d2;
  const long h = gbk[g];
  const int nc = (int)P.ncnt[h];
  const u64 base = P.noff[h];
  const u64 dk = sdk[sj[g]];
  if (l < N.np) lg[l] = (int)((dk >> N.kshift[l]) & N.kmask[l]);
  const int* src = szone + (size_t)sj[g] * d2;
  for (int e = l; e < d2; e += kWave) S[e] = src[e];
  wsync();
  if (l < d) cover_bounds(N, lg, l, alu, &Lb[l], &Ub[l]);
  wsync();
  int first = -1;
  unsigned long long reads = 0;
  for (int b = 0; b < nc && first < 0; b += kWave) {
    const int m = b + l;
    bool ok = false;
    if (m < nc) { ok = alu_le_lane(S, szone + (size_t)sj[wlist[base + m]] * d2, d, Lb, Ub); ++reads; }
    const u64 bal = __ballot(ok);
    if (bal) first = b + __builtin_ctzll(bal);
  }
  wave_stat(ctr, C_READS, reads);
  if (l == 0) {
    tslot[g] = first >= 0 ? wlist[base + first] : (int)g;
    if (first < 0) atomicExch(&ctr->err, ERR_CAPPED + 1);   // cannot happen: an earlier cover exists
    else stat_add(ctr, C_SUB, 1);
  }
}

__global__ void k_gassign(long G, long long nbase, long long ebase, int record, const int* __restrict__ cov_old,
                          const int* __restrict__ tslot, const int* __restrict__ fnew,
                          const long long* __restrict__ rank, const int* __restrict__ sj,
                          const int* __restrict__ s_se, const int* __restrict__ s_re, int* __restrict__ etgt,
                          int* __restrict__ ese, int* __restrict__ ere, int* __restrict__ newlist) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int co = cov_old[g];
  const long long tgt = co >= 0 ? (long long)co : nbase + rank[tslot[g]];
  if (record) {
    etgt[ebase + g] = (int)tgt;
    ese[ebase + g] = s_se[sj[g]];
    ere[ebase + g] = s_re[sj[g]];
  }
  if (co < 0 && fnew[g]) newlist[rank[g]] = (int)g;
}

__global__ __launch_bounds__(64) void k_gcopy(long F2, long long nbase, const int* __restrict__ newlist,
                                              const int* __restrict__ sj, const u64* __restrict__ sdk,
                                              const int* __restrict__ szone, int d2, int* __restrict__ gzone,
                                              u64* __restrict__ gdk, u64* __restrict__ fr_dk,
                                              u32* __restrict__ fr_zid) {
  const long b = blockIdx.x;
  if (b >= F2) return;
  const int g = newlist[b];
  const long long id = nbase + b;
  const int* src = szone + (size_t)sj[g] * d2;
  int* dst = gzone + (size_t)id * d2;
  for (int e = threadIdx.x; e < d2; e += kWave) dst[e] = src[e];
  if (threadIdx.x == 0) {
    gdk[id] = sdk[sj[g]];
    fr_dk[b] = sdk[sj[g]];
    fr_zid[b] = (u32)id;
  }
}

// segment sizes of the touched keys for this level: the stored nodes plus the queued successors
__global__ void k_gsegs(Passed P, const u32* __restrict__ touched, int T, int* __restrict__ segs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T) segs[t] = (int)(P.len[touched[t]] + P.ncnt[touched[t]]);
}

// One wavefront per touched key: its node segment moves to the one reserved for this level and the
// key's new nodes are appended in creation order.
__global__ __launch_bounds__(64) void k_glink(Passed P, const u32* __restrict__ touched, int T,
                                              const int* __restrict__ wlist, const long long* __restrict__ rank,
                                              long long nbase) {
  const int t = blockIdx.x;
  if (t >= T) return;
  const int l = threadIdx.x;
  const u32 h = touched[t];
  const u32 nc = P.ncnt[h];
  const u32 len = P.len[h];
  const u64 base = P.noff[h], from = P.off[h], to = P.nseg[h];
  for (u32 k = l; k < len; k += kWave) P.nodes[to + k] = P.nodes[from + k];
  for (u32 c = l; c < nc; c += kWave) P.nodes[to + len + c] = (u32)(nbase + rank[wlist[base + c]]);
  if (l == 0) {
    P.off[h] = to;
    P.len[h] = len + nc;
    P.ncnt[h] = 0;
  }
}

// the first node of every discrete state of the graph
__global__ void k_heads(Passed P, long H, int* __restrict__ heads, LiveCtr* lc) {
  const long h = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = h < H && P.key[h] != kEmptyKey && P.len[h] > 0;
  const int idx = block_append(&lc->ns, ok);
  if (ok) heads[idx] = (int)P.nodes[P.off[h]];
}

// ---------------------------------------------------------------- stop sets and deadlock

// Gather the combined invariant of a location vector from the tables (off, data), one lane per
// process, as (i, j, raw) triples.
__device__ __forceinline__ int wave_gather_inv(const DevNet& N, const int* off, const int* data, const int* locs,
                                               int* inv) {
  const int l = threadIdx.x;
  int ia = 0, ib = 0;
  if (l < N.np) {
    const int g = N.locbase[l] + locs[l];
    ia = off[g];
    ib = off[g + 1];
  }
  int total = 0;
  const int o0 = wave_excl_scan(ib - ia, &total);
  if (total <= kInvCap)
    for (int c = ia; c < ib; ++c) {
      const int o = 3 * (o0 + c - ia);
      inv[o] = data[3 * c];
      inv[o + 1] = data[3 * c + 1];
      inv[o + 2] = data[3 * c + 2];
    }
  wsync();
  return total;
}

__device__ __forceinline__ bool has_upper_list(const int* inv, int n) {
  for (int c = 0; c < n; ++c)
    if (inv[3 * c] > 0 && inv[3 * c + 1] == 0 && inv[3 * c + 2] < kInf) return true;
  return false;
}

__device__ __forceinline__ void zcopy_w(int* dst, const int* src, int d2) {
  wsync();
  for (int e = threadIdx.x; e < d2; e += kWave) dst[e] = src[e];
}

// Drop members of a list contained in another member (the first of equal members stays).
// `drop` is cap ints of LDS scratch. Returns the new length.
__device__ __forceinline__ int fed_compact(int* F, int n, int zs, int d, int* drop) {
  const int l = threadIdx.x, d2 = d * d;
  wsync();
  for (int k = l; k < n; k += kWave) drop[k] = 0;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) {
      if (j == k) continue;
      wsync();
      if (drop[k]) break;
      if (!z_includes(F + j * zs, F + k * zs, d)) continue;
      if (j < k || !z_includes(F + k * zs, F + j * zs, d)) {
        wsync();
        if (l == 0) drop[k] = 1;
      }
    }
  wsync();
  int m = 0;
  for (int k = 0; k < n; ++k) {
    wsync();
    if (drop[k]) continue;
    if (m != k)
      for (int e = l; e < d2; e += kWave) F[m * zs + e] = F[k * zs + e];
    ++m;
  }
  wsync();
  return m;
}

// The reference's federation normalisation (fed_compact, src/liveness/operators.cpp:428-458): each
// member in order is dropped when a kept member includes it, replaces the first kept member it
// includes, or is joined to the first kept member whose union with it is one zone (z_try_union);
// otherwise it is kept. One more pass when anything was reshaped. The set is unchanged; used when
// a list is full after fed_compact (which only drops contained members). `fa`, `pa`: d*d ints of
// LDS scratch each, `cnt` 2 ints. Returns the new length.
__device__ __noinline__ int fed_compact_union(int* F, int n, int zs, int d, int* fa, int* pa, int* cnt) {
  const int d2 = d * d;
  int m = 0;
  bool merged = false;
  for (int k = 0; k < n; ++k) {
    int* z = F + k * zs;
    bool absorbed = false;
    for (int o = 0; o < m && !absorbed; ++o) {
      int* oz = F + o * zs;
      if (z_includes(oz, z, d)) {
        absorbed = true;
      } else if (z_includes(z, oz, d)) {
        zcopy_w(oz, z, d2);
        absorbed = merged = true;
      } else if (z_try_union(oz, z, d, fa, pa, cnt)) {
        absorbed = merged = true;
      }
    }
    if (!absorbed) {
      if (m != k) zcopy_w(F + m * zs, z, d2);
      ++m;
    }
  }
  if (merged && m > 1) {
    int s = 0;
    for (int k = 0; k < m; ++k) {
      int* z = F + k * zs;
      bool absorbed = false;
      for (int o = 0; o < s && !absorbed; ++o) {
        int* oz = F + o * zs;
        absorbed = z_includes(oz, z, d) || z_try_union(oz, z, d, fa, pa, cnt);
      }
      if (!absorbed) {
        if (s != k) zcopy_w(F + s * zs, z, d2);
        ++s;
      }
    }
    m = s;
  }
  wsync();
  return m;
}

// p minus y appended to out (capacity cap). With UNI, when out fills, the pieces from before this
// call are normalised with unions (fed_compact_union) and the subtraction runs again; false when
// there is no room (even then). `cnt`: 2 ints of LDS scratch (r and t serve as the union's).
template <bool UNI>
__device__ __forceinline__ bool subtract_room(const int* p, const int* y, int d, int* out, int zs, int* n, int cap,
                                              int* r, int* t, int* cnt) {
  const int n0 = *n;
  if (z_subtract(p, y, d, out, zs, n, cap, r, t)) return true;
  if (!UNI) return false;
  *n = fed_compact_union(out, n0, zs, d, r, t, cnt);
  return z_subtract(p, y, d, out, zs, n, cap, r, t);
}

// z covered by the union of the list? Subtract every member; `cur`/`nxt` are piece lists of
// capacity kp. Returns 1 covered, 0 not, -1 capacity exceeded.
template <bool UNI>
__device__ __forceinline__ int fed_covers(const int* F, int n, int zs, const int* z, int d, int* cur, int* nxt,
                                          int kp, int* r, int* t, int* cnt) {
  const int d2 = d * d;
  for (int c = 0; c < n; ++c)
    if (z_includes(F + c * zs, z, d)) return 1;
  int nc = 1;
  zcopy_w(cur, z, d2);
  for (int c = 0; c < n && nc > 0; ++c) {
    int nn = 0;
    for (int k = 0; k < nc; ++k)
      if (!subtract_room<UNI>(cur + k * zs, F + c * zs, d, nxt, zs, &nn, kp, r, t, cnt)) return -1;
    int* sw = cur; cur = nxt; nxt = sw;
    nc = nn;
  }
  return nc == 0 ? 1 : 0;
}

// LDS capacities of one k_stop launch: zones of the escape list (kl) and of each subtraction piece
// list (kp), invariant constraints of a location vector (ic), transitions of a state (rc).
// np, ni: the network's processes and integer variables (two discrete states are held).
struct StopCaps {
  int kl, kp, ic, rc, np, ni;
  __host__ __device__ int ints(int d2) const { return (6 + kl + 2 * kp) * d2 + 2 * (np + ni) + 2 * 3 * ic + 2 * rc; }
};

// Stop set of node v (mode 0), early witness test (mode 1: p /\ !q nodes only), deadlock test
// (mode 2), seed test of v's discrete state (mode 3). Writes scnt (mode 0): -1 whole zone, else the
// number of zones in sz; sets ctr->hit for a non-empty answer in modes 1 and 2; counts the seeds
// in ctr->hit in mode 3.
// UNI: the launch with the full lists, which normalises a full list with unions before it gives up.
template <bool UNI>
__global__ __launch_bounds__(64) void k_stop(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap,
                                             const int* __restrict__ list, long n, long long first_id, int mode,
                                             StopCaps K, int* __restrict__ retry, int* __restrict__ scnt,
                                             int* __restrict__ soff, int* __restrict__ sz, int scap, LiveCtr* lc,
                                             int* __restrict__ gl) {
#ifdef HORA_FED_HBM
  int* sm = gl + (size_t)blockIdx.x * (size_t)K.ints(Np->d2);   // this node's lists in device memory
#else
  extern __shared__ int sm[];
#endif
  const DevNet& N = *Np;
  const LiveArgs& A = *Ap;
  const int d = N.d, d2 = N.d2, l = threadIdx.x;
  int* Zd = sm;
  int* g = Zd + d2;
  int* tmp = g + d2;
  int* r = tmp + d2;
  int* t = r + d2;
  int* zu = t + d2;
  int* live = zu + d2;
  int* cur = live + K.kl * d2;
  int* nxt = cur + K.kp * d2;
  int* locs = nxt + K.kp * d2;
  int* ints = locs + K.np;
  int* tl = ints + K.ni;
  int* ti = tl + K.np;
  int* inv = ti + K.ni;
  int* tinv = inv + 3 * K.ic;
  int* rec = tinv + 3 * K.ic;
  const long b = blockIdx.x;
  if (b >= n) return;
  const long v = list ? list[b] : first_id + b;
  const u64 dk = A.gdk[v];
  // mode 1 tests the witness positions: p /\ !q nodes (leads-to), psi nodes (E[] psi)
  if (mode == 1 && !(A.kind == HORA_LIVE_EG ? pred_eval(N, A.p, dk) : pred_eval(N, A.p, dk) && !pred_eval(N, A.q, dk)))
    return;
  // a capacity of K exceeded: rerun the node with the full capacities, or fail without them
  auto overflow = [&](int err) __attribute__((always_inline)) {
    if (retry) { if (l == 0) retry[atomicAdd(&lc->rcnt, 1)] = (int)v; }
    else if (l == 0) atomicExch(&lc->err, err);
  };
  decode_key(N, dk, locs, ints);
  wsync();
  int fz = 0, ninv = 0;
  u64 am = 0;
  wave_loc_info(N, locs, ints, &fz, &am, inv, &ninv, K.ic);   // the network's (restricted) invariant
  if (ninv > K.ic) { overflow(ERR_RECIPES); return; }
  const bool frozen = fz != 0, src_c = (fz & 1) != 0;
  const int* zv = A.gzone + (size_t)v * d2;
  auto out_whole = [&]() __attribute__((always_inline)) {
    if (mode == 0) { if (l == 0) scnt[A.sidx[v]] = -1; }
    else if (l == 0) { atomicExch(&lc->hit, 1); if (A.whit) A.whit[v] = 1; }
  };
  auto out_none = [&]() __attribute__((always_inline)) {
    if (mode == 0 && l == 0) scnt[A.sidx[v]] = 0;
  };
  // time_diverges (construction.cpp:1097-1162) on the network's invariant
  auto time_diverges = [&]() __attribute__((always_inline)) -> bool {
    if (N.nclocks == 0 || has_upper_list(inv, ninv) || frozen) return false;
    zcopy_w(t, zv, d2);
    z_up(t, d);
    if (!z_apply_list(t, d, inv, ninv)) return false;
    wsync();
    bool bounded = false;
    if (l >= 1 && l < d) bounded = t[l * d] != kInf;
    return !__any(bounded);
  };
  // the table the deadlock is judged in: the unrestricted network's when psi restricted it
  const int* doff = A.restricted ? A.oinv_off : N.inv_off;
  const int* ddat = A.restricted ? A.oinv_data : N.inv_data;
  if (mode == 3) {
    // seed test of the discrete state (online_seed_key, refinement.cpp:176-282): the universe
    // zone, judged in the unrestricted network
    const int no = wave_gather_inv(N, doff, ddat, locs, tinv);
    if (no > K.ic) { overflow(ERR_RECIPES); return; }
    if (!frozen && A.kind != HORA_LIVE_DEADLOCK && !has_upper_list(tinv, no)) return;
    wsync();
    for (int e = l; e < d2; e += kWave) {
      const int i = ediv(e, d), j = e - i * d;
      zu[e] = (i == j || i == 0) ? kLe0 : kInf;
    }
    zv = zu;
  }
  // the deadlock scan of a finished graph reads the expanded nodes only
  if (mode == 2 && !list && !A.expanded[v]) return;
  if (mode == 0 || mode == 1) {   // the stop set proper (stop_set); modes 2, 3 test deadlock only
    if (!A.restricted) {
      const bool nosucc = A.recnt ? A.recnt[v] == 0 : A.grow[v + 1] == A.grow[v];
      if (nosucc || time_diverges()) { out_whole(); return; }
    } else {
      if (time_diverges()) { out_whole(); return; }
      // orig_time_open (construction.cpp:1164-1191)
      const int no = wave_gather_inv(N, doff, ddat, locs, tinv);
      if (no > K.ic) { overflow(ERR_RECIPES); return; }
      if (N.nclocks > 0 && !has_upper_list(tinv, no) && !frozen) {
        wsync();
        for (int e = l; e < d2; e += kWave) {
          const int i = ediv(e, d), j = e - i * d;
          t[e] = (i == j || i == 0) ? kLe0 : kInf;
        }
        if (z_apply_list(t, d, tinv, no)) {
          wsync();
          bool bounded = false;
          if (l >= 1 && l < d) bounded = t[l * d] != kInf;
          if (!__any(bounded)) { out_none(); return; }
        }
      }
    }
  }
  // state_has_deadlock (deadlock.cpp:74-268)
  const int nd = wave_gather_inv(N, doff, ddat, locs, tinv);
  if (nd > K.ic) { overflow(ERR_RECIPES); return; }
  zcopy_w(Zd, zv, d2);
  if (!frozen) z_up(Zd, d);
  if (!z_apply_list(Zd, d, tinv, nd)) { out_none(); return; }
  if (mode == 3 && !frozen && A.kind != HORA_LIVE_DEADLOCK) {
    // a time-open state: every valuation delays for ever, so no deadlock completion to refine
    wsync();
    bool bounded = false;
    if (l >= 1 && l < d) bounded = Zd[l * d] != kInf;
    if (!__any(bounded)) return;
  }
  const int nr = build_recipes(N, locs, ints, src_c, rec, K.rc);
  if (nr < 0) { overflow(ERR_RECIPES); return; }
  wsync();
  int nl = 0;
  for (int k = 0; k < nr; ++k) {
    wsync();
    const int se = rec[2 * k], re = rec[2 * k + 1];
    zcopy_w(g, Zd, d2);
    if (!apply_edge_guard(N, se, g)) continue;
    if (re >= 0 && !apply_edge_guard(N, re, g)) continue;
    u64 rm = edge_reset_mask(N, se) | (re >= 0 ? edge_reset_mask(N, re) : 0ull);
    if (re <= -2) {
      // a broadcast (deadlock.cpp:201-235): each part's receiving edge, or its stay-out conjunction
      bool ok = true;
      bcast_walk(N, locs, ints, se, -2 - re, [&](int, int gj, int rj, int q) {
        if (!ok) return;
        if (rj >= 0) {
          ok = apply_edge_guard(N, rj, g);
          rm |= edge_reset_mask(N, rj);
        } else {
          ok = apply_bcast_none(N, gj, N.edge_chan[se], ints, q, g);
        }
      });
      if (!ok) continue;
    }
    wsync();
    if (l < N.np) tl[l] = locs[l];
    wsync();
    if (l == 0) {
      tl[N.edge_proc[se]] = N.edge_tgt[se];
      if (re >= 0) tl[N.edge_proc[re]] = N.edge_tgt[re];
      if (re <= -2)
        bcast_walk(N, locs, ints, se, -2 - re, [&](int j, int, int rj, int) { if (rj >= 0) tl[j] = N.edge_tgt[rj]; });
    }
    wsync();
    const int nt = wave_gather_inv(N, doff, ddat, tl, tinv);
    if (nt > K.ic) { overflow(ERR_RECIPES); return; }
    bool ok = true;
    for (int c = 0; c < nt && ok; ++c) {
      const int ci = tinv[3 * c], cj = tinv[3 * c + 1], cb = tinv[3 * c + 2];
      const int i = ((rm >> ci) & 1) ? 0 : ci, j = ((rm >> cj) & 1) ? 0 : cj;
      if (i == j) { if (cb < kLe0) ok = false; continue; }
      ok = z_constrain(g, d, i, j, cb);
    }
    if (!ok) continue;
    if (z_equal(g, Zd, d)) { out_none(); return; }   // enabled everywhere: no deadlock
    // keep `live` inclusion-minimal
    bool dom = false;
    for (int m = 0; m < nl && !dom; ++m) dom = z_includes(live + m * d2, g, d);
    if (dom) continue;
    int m2 = 0;
    for (int m = 0; m < nl; ++m) {
      if (z_includes(g, live + m * d2, d)) continue;
      if (m2 != m) zcopy_w(live + m2 * d2, live + m * d2, d2);
      ++m2;
    }
    nl = m2;
    if (UNI && nl >= K.kl) nl = fed_compact_union(live, nl, d2, d, r, t, tmp);
    if (nl >= K.kl) { overflow(ERR_CAPPED); return; }
    zcopy_w(live + nl * d2, g, d2);
    ++nl;
  }
  if (!frozen)
    for (int m = 0; m < nl; ++m) z_down(live + m * d2, d);
  for (int m = 0; m < nl; ++m)
    if (z_includes(live + m * d2, Zd, d)) { out_none(); return; }
  // Dead = Zd minus every escape
  int nc = 1;
  zcopy_w(cur, Zd, d2);
  int* c0 = cur;
  int* c1 = nxt;
  for (int m = 0; m < nl && nc > 0; ++m) {
    int nn = 0;
    for (int k = 0; k < nc; ++k)
      if (!subtract_room<UNI>(c0 + k * d2, live + m * d2, d, c1, d2, &nn, K.kp, r, t, tmp)) {
        overflow(ERR_CAPPED);
        return;
      }
    int* sw = c0; c0 = c1; c1 = sw;
    nc = nn;
  }
  if (nc == 0) { out_none(); return; }
  if (mode == 3) {
    // a seed iff some escape meets the invariant region: otherwise Dead = Inv, nothing to refine
    bool meets = false;
    for (int m = 0; m < nl && !meets; ++m) meets = z_intersect(tmp, live + m * d2, Zd, d);
    if (meets && l == 0) atomicAdd(&lc->hit, 1);
    return;
  }
  if (mode != 0) { if (l == 0) { atomicExch(&lc->hit, 1); if (A.whit) A.whit[v] = 1; } return; }
  // the zones go to the stop-set pool; when it is full the caller grows it and runs the phase again
  const int s = A.sidx[v];
  int off = 0;
  if (l == 0) off = atomicAdd(&lc->spool, nc);
  off = __shfl(off, 0);
  if ((long long)off + nc > scap) { if (l == 0) atomicExch(&lc->sfull, 1); return; }
  wsync();
  for (int k = 0; k < nc; ++k)
    for (int e = l; e < d2; e += kWave) sz[((size_t)off + k) * d2 + e] = c0[k * d2 + e];
  if (l == 0) { scnt[s] = nc; soff[s] = off; }
}

// ---------------------------------------------------------------- closure of the backward pass

__global__ void k_sseed(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap, long N, int* __restrict__ in_s,
                        int* __restrict__ slist, LiveCtr* lc) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const LiveArgs& A = *Ap;
  bool seed = false;
  if (v < N && A.expanded[v]) {
    const u64 dk = A.gdk[v];
    if (A.kind == HORA_LIVE_LEADSTO) seed = pred_eval(*Np, A.p, dk) && !pred_eval(*Np, A.q, dk);
    else seed = v == 0 && pred_eval(*Np, A.p, dk);
  }
  if (v < N) in_s[v] = seed ? 1 : 0;
  const int idx = block_append(&lc->ns, seed);
  if (seed) slist[idx] = (int)v;
}

__global__ void k_sstep(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap, const int* __restrict__ slist,
                        int from, int to, int* __restrict__ in_s, int* __restrict__ slist_out, LiveCtr* lc) {
  const LiveArgs& A = *Ap;
  const int k = from + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= to) return;
  const int v = slist[k];
  for (long long e = A.grow[v]; e < A.grow[v + 1]; ++e) {
    const int t = A.etgt[e];
    if (in_s[t] || !A.expanded[t]) continue;
    const u64 dk = A.gdk[t];
    const bool nq = A.kind == HORA_LIVE_LEADSTO ? !pred_eval(*Np, A.q, dk) : pred_eval(*Np, A.p, dk);
    if (!nq) continue;
    if (atomicCAS(&in_s[t], 0, 1) == 0) slist_out[atomicAdd(&lc->ns, 1)] = t;
  }
}

__global__ void k_sindex(const int* __restrict__ slist, int ns, int* __restrict__ sidx) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ns) sidx[slist[k]] = k;
}

// ---------------------------------------------------------------- greatest fixpoint

// W[s] := {Z_v} in bank 0
__global__ __launch_bounds__(64) void k_gseed(const LiveArgs* __restrict__ Ap, const int* __restrict__ slist, int ns,
                                              int d2, int* __restrict__ wcnt0, int* __restrict__ woff0,
                                              int* __restrict__ wz, int* __restrict__ sel, int* __restrict__ chg) {
  const int s = blockIdx.x;
  if (s >= ns) return;
  const int v = slist[s];
  const int* src = Ap->gzone + (size_t)v * d2;
  int* dst = wz + (size_t)s * d2;
  for (int e = threadIdx.x; e < d2; e += kWave) dst[e] = src[e];
  if (threadIdx.x == 0) {
    wcnt0[s] = 1;
    woff0[s] = s;
    sel[s] = 0;
    chg[s] = 1;
  }
}

// Values: node s's value in bank b is wcnt_b[s] zones at zone offset woff_b[s] of the value pool
// `wz` (capacity wcap zones), bump-allocated: a value written in a round takes fresh zones, so the
// values the round reads stay intact. Stop sets: scnt[s] zones at soff[s] of the pool `sz`
// (scnt -1: the whole zone).
struct GfpBanks {
  int* wcnt0;
  int* wcnt1;
  int* woff0;
  int* woff1;
  int* wz;
  int wcap;
  const int* sel_in;
  int* sel_out;
  const int* chg_in;
  int* chg_out;
  const int* scnt;
  const int* soff;
  const int* sz;
};

// LDS capacities of one k_gfp launch: zones of the value W[v] (kw), of the candidate federation
// (kc), of each subtraction piece list (kp), and invariant constraints (ic) of each location vector.
// np, ni: the network's processes and integer variables (two discrete states are held).
struct GfpCaps {
  int kw, kc, kp, ic, np, ni;
  __host__ __device__ int drop() const { return kc > kp ? kc : kp; }
  __host__ __device__ int ints(int d2) const {
    return (4 + kw + kc + 2 * kp) * d2 + 2 * (np + ni) + 2 * 3 * ic + drop();
  }
  // with lists in device memory: a covered-member flag per value member follows
  __host__ __device__ size_t hbm_ints(int d2) const { return (size_t)ints(d2) + (size_t)kw; }
};

// The nodes a Jacobi round must visit: a non-empty value with a successor whose value changed in
// the previous round (every non-empty value in the first round). The others keep their bank here.
__global__ void k_gdirty(const LiveArgs* __restrict__ Ap, const int* __restrict__ slist, int ns, int first,
                         GfpBanks B, int* __restrict__ dlist, LiveCtr* lc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const LiveArgs& A = *Ap;
  bool dirty = false;
  int bi = 0;
  if (s < ns) {
    bi = B.sel_in[s];
    dirty = (bi ? B.wcnt1 : B.wcnt0)[s] > 0;
    if (dirty && !first) {
      dirty = false;
      const int v = slist[s];
      for (long long e = A.grow[v]; e < A.grow[v + 1] && !dirty; ++e) {
        const int tt = A.etgt[e];
        dirty = A.in_s[tt] && B.chg_in[A.sidx[tt]];
      }
    }
  }
  const int idx = block_append(&lc->dcnt, dirty);
  if (s >= ns) return;
  if (dirty) dlist[idx] = s;
  else { B.sel_out[s] = bi; B.chg_out[s] = 0; }
}

// One Jacobi round of relax (fixpoint.cpp:750-864), one wavefront per listed node of the closure.
// Values are read from the bank each node selected at the start of the round and written to the
// other bank, so the order of the nodes does not matter. With `retry` set, a node that outgrows a
// capacity of K is appended to `retry` untouched (for a launch with the full capacities, which
// reads the same banks); without it, that is the federation cap of the reference (no verdict).
// UNI: the launch with the full lists, which normalises a full list with unions before it gives up.
template <bool UNI>
__global__ __launch_bounds__(64) void k_gfp(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap,
                                            const int* __restrict__ slist, const int* __restrict__ list,
                                            GfpCaps K, GfpBanks B, int* __restrict__ retry, LiveCtr* lc,
                                            int* __restrict__ gl) {
#ifdef HORA_FED_HBM
  int* sm = gl + (size_t)blockIdx.x * (size_t)K.hbm_ints(Np->d2);   // this node's lists in device memory
#else
  extern __shared__ int sm[];
#endif
  const DevNet& N = *Np;
  const LiveArgs& A = *Ap;
  const int d = N.d, d2 = N.d2, l = threadIdx.x;
  int* pz = sm;
  int* tmp = pz + d2;
  int* r = tmp + d2;
  int* t = r + d2;
  int* Wv = t + d2;
  int* cand = Wv + K.kw * d2;
  int* cur = cand + K.kc * d2;
  int* nxt = cur + K.kp * d2;
  int* locs = nxt + K.kp * d2;
  int* ints = locs + K.np;
  int* tlocs = ints + K.ni;
  int* tints = tlocs + K.np;
  int* inv = tints + K.ni;
  int* tinv = inv + 3 * K.ic;
  int* drop = tinv + 3 * K.ic;
  const int s = list[blockIdx.x];
  const int v = slist[s];
  const int bi = B.sel_in[s];
  unsigned long long pcalls = 0, prods = 0;
  auto stats = [&]() __attribute__((always_inline)) {
    if (l == 0 && pcalls) atomicAdd(&lc->stripe[0][blockIdx.x % kLiveStripes][0], pcalls);
    if (l == 0 && prods) atomicAdd(&lc->stripe[1][blockIdx.x % kLiveStripes][0], prods);
  };
  auto keep = [&]() __attribute__((always_inline)) {
    stats();
    if (l == 0) { B.sel_out[s] = bi; B.chg_out[s] = 0; }
  };
  auto overflow = [&](int err) __attribute__((always_inline)) {
    if (retry) {
      if (l == 0) retry[atomicAdd(&lc->rcnt, 1)] = s;
    } else {
      if (l == 0) atomicExch(&lc->err, err);
      keep();
    }
  };
  auto wcnt = [&](int b) __attribute__((always_inline)) { return b ? B.wcnt1 : B.wcnt0; };
  auto woff = [&](int b) __attribute__((always_inline)) { return b ? B.woff1 : B.woff0; };
  const long long e0 = A.grow[v], e1 = A.grow[v + 1];
  const int nW = wcnt(bi)[s];
  if (nW == 0) { keep(); return; }
  if (nW > K.kw) { overflow(ERR_CAPPED); return; }
  {
    const int* src = B.wz + (size_t)woff(bi)[s] * d2;
    for (int e = l; e < nW * d2; e += kWave) Wv[e] = src[e];
  }
  decode_key(N, A.gdk[v], locs, ints);
  wsync();
  int fzv = 0, ninv = 0;
  u64 amv = 0;
  wave_loc_info(N, locs, ints, &fzv, &amv, inv, &ninv, K.ic);
  if (ninv > K.ic) { overflow(ERR_RECIPES); return; }
  const u64 cm = ((1ull << N.nclocks) - 1) << 1;
#ifdef HORA_FED_HBM
  // values above 64 members: the covered members as flags after the lists
  int* wcov = drop + K.drop();
  for (int k = l; k < nW; k += kWave) wcov[k] = 0;
  int ncov = 0;
#else
  u64 covm = 0;
  const u64 full = (nW >= 64) ? ~0ull : ((1ull << nW) - 1);
#endif
  int nc = 0;
  for (long long eb = e0; eb < e1; eb += kWave) {
    // One successor edge per lane: its target's value (count, pool offset) in the bank selected
    // for this round, its key, transition and resets, all read in one round of loads; the edges
    // whose target holds a non-empty value are then taken in order.
    const long long e = eb + l;
    int nYl = 0, offl = 0, sel = -1, rel = -1;
    u64 dkl = 0, rml = 0;
    if (e < e1) {
      const int tt = A.etgt[e];
      if (A.in_s[tt]) {
        const int st = A.sidx[tt];
        const int bt = B.sel_in[st];
        nYl = wcnt(bt)[st];
        if (nYl > 0) {
          offl = woff(bt)[st];
          dkl = A.gdk[tt];
          sel = A.ese[e];
          rel = A.ere[e];
          rml = edge_reset_mask(N, sel) | (rel >= 0 ? edge_reset_mask(N, rel) : 0ull);
          if (rel <= -2)   // a broadcast: the participating receivers reset too
            bcast_walk(N, locs, ints, sel, -2 - rel, [&](int, int, int rj, int) {
              if (rj >= 0) rml |= edge_reset_mask(N, rj);
            });
        }
      }
    }
    for (u64 act = __ballot(nYl > 0); act; act &= act - 1) {
      const int q = __builtin_ctzll(act);
      const int nY = bcast(nYl, q), se = bcast(sel, q), re = bcast(rel, q);
      const u64 rm = bcast64(rml, q);
      wsync();
      decode_key(N, bcast64(dkl, q), tlocs, tints);
      wsync();
      int fzt = 0, ntinv = 0;
      u64 amt = 0;
      wave_loc_info(N, tlocs, tints, &fzt, &amt, tinv, &ntinv, K.ic);
      if (ntinv > K.ic) { overflow(ERR_RECIPES); return; }
      const u64 dead_t = ~amt & cm;
      const int* Ys = B.wz + (size_t)bcast(offl, q) * d2;
      for (int y = 0; y < nY; ++y) {
        ++pcalls;
        // exact_predecessor (operators.cpp:257-306)
        zcopy_w(pz, Ys + y * d2, d2);
        z_free_many(pz, d, dead_t, tmp);
        if (!z_apply_list(pz, d, tinv, ntinv)) continue;
        if (!fzt) {
          z_down(pz, d);
          if (!z_apply_list(pz, d, tinv, ntinv)) continue;
        }
        if (rm && !z_zero_free(pz, d, rm, tmp)) continue;   // resets to zero, then freed
        if (!apply_edge_guard(N, se, pz)) continue;
        if (re >= 0 && !apply_edge_guard(N, re, pz)) continue;
        if (re <= -2) {
          // a broadcast's receivers: their guards, and the stay-out conjunctions of those that
          // did not take part (operators.cpp:290-300)
          bool ok = true;
          bcast_walk(N, locs, ints, se, -2 - re, [&](int, int gj, int rj, int q) {
            if (ok) ok = rj >= 0 ? apply_edge_guard(N, rj, pz) : apply_bcast_none(N, gj, N.edge_chan[se], ints, q, pz);
          });
          if (!ok) continue;
        }
        if (!fzv) z_down(pz, d);
        if (!z_apply_list(pz, d, inv, ninv)) continue;
#ifdef HORA_FED_HBM
        for (int k = 0; k < nW; ++k) {
          wsync();
          if (!wcov[k] && z_includes(pz, Wv + k * d2, d)) {
            wsync();
            if (l == 0) wcov[k] = 1;
            ++ncov;
          }
        }
        if (ncov == nW) { keep(); return; }   // every member covered: the value cannot move
#else
        for (int k = 0; k < nW; ++k)
          if (!((covm >> k) & 1) && z_includes(pz, Wv + k * d2, d)) covm |= 1ull << k;
        if (covm == full) { keep(); return; }   // every member covered: the value cannot move
#endif
        if (nc >= K.kc) nc = fed_compact(cand, nc, d2, d, drop);
        if (UNI && nc >= K.kc) nc = fed_compact_union(cand, nc, d2, d, r, t, drop);
        if (nc >= K.kc) { overflow(ERR_CAPPED); return; }
        zcopy_w(cand + nc * d2, pz, d2);
        ++nc;
      }
    }
  }
  nc = fed_compact(cand, nc, d2, d, drop);
  // does cand cover every member of W[v]? -1: capacity exceeded
  auto escapes = [&]() __attribute__((always_inline)) -> int {
    for (int k = 0; k < nW; ++k) {
      const int cv = fed_covers<UNI>(cand, nc, d2, Wv + k * d2, d, cur, nxt, K.kp, r, t, drop);
      if (cv < 0) return -1;
      if (cv == 0) return 1;
    }
    return 0;
  };
  int es = escapes();
  if (es < 0) { overflow(ERR_CAPPED); return; }
  if (es == 0) { keep(); return; }
  const int sc = B.scnt[s];
  if (sc != 0) {
    const int add = sc < 0 ? 1 : sc;
    if (nc + add > K.kc) nc = fed_compact(cand, nc, d2, d, drop);
    if (UNI && nc + add > K.kc) nc = fed_compact_union(cand, nc, d2, d, r, t, drop);
    if (nc + add > K.kc) { overflow(ERR_CAPPED); return; }
    if (sc < 0) zcopy_w(cand + nc * d2, A.gzone + (size_t)v * d2, d2);
    else
      for (int k = 0; k < sc; ++k) zcopy_w(cand + (nc + k) * d2, B.sz + ((size_t)B.soff[s] + k) * d2, d2);
    nc += add;
    nc = fed_compact(cand, nc, d2, d, drop);
    es = escapes();
    if (es < 0) { overflow(ERR_CAPPED); return; }
    if (es == 0) { keep(); return; }
  }
  // W[v] /\ cand, member by member (products into `cur`, capacity kp)
  int nn = 0;
  for (int k = 0; k < nW; ++k) {
    const int* z = Wv + k * d2;
    bool whole = false;
    for (int c = 0; c < nc && !whole; ++c) whole = z_includes(cand + c * d2, z, d);
    if (whole) {
      if (nn >= K.kp) { overflow(ERR_CAPPED); return; }
      zcopy_w(cur + nn * d2, z, d2);
      ++nn;
      continue;
    }
    for (int c = 0; c < nc; ++c) {
      if (!z_intersect(tmp, z, cand + c * d2, d)) continue;
      ++prods;
      if (nn >= K.kp) nn = fed_compact(cur, nn, d2, d, drop);
      if (UNI && nn >= K.kp) nn = fed_compact_union(cur, nn, d2, d, r, t, drop);
      if (nn >= K.kp) { overflow(ERR_CAPPED); return; }
      zcopy_w(cur + nn * d2, tmp, d2);
      ++nn;
    }
  }
  nn = fed_compact(cur, nn, d2, d, drop);
  // the value's capacity: the LDS launches' bank width, or the federation cap in device memory
#ifdef HORA_FED_HBM
  const int vcap = K.kw;
#else
  const int vcap = A.kw;
#endif
  if (UNI && nn > vcap) nn = fed_compact_union(cur, nn, d2, d, r, t, drop);
  if (nn > vcap) { overflow(ERR_CAPPED); return; }
  // fresh zones from the value pool; when it is full the caller compacts it and runs the round again
  int off = 0;
  if (l == 0) off = atomicAdd(&lc->ptop, nn);
  off = __shfl(off, 0);
  if ((long long)off + nn > B.wcap) { if (l == 0) atomicExch(&lc->pfull, 1); return; }
  stats();
  const int bo = 1 - bi;
  int* dst = B.wz + (size_t)off * d2;
  wsync();
  for (int e = l; e < nn * d2; e += kWave) dst[e] = cur[e];
  if (l == 0) {
    wcnt(bo)[s] = nn;
    woff(bo)[s] = off;
    B.sel_out[s] = bo;
    B.chg_out[s] = 1;
    atomicAdd(&lc->stripe[2][blockIdx.x % kLiveStripes][0], 1ull);
  }
}

// Value-pool compaction: each node's current value (the bank `sel` names) moves to `noff[s]` of the
// other pool, in node order.
__global__ void k_wcount(int ns, const int* __restrict__ sel, const int* __restrict__ wcnt0,
                         const int* __restrict__ wcnt1, int* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns) out[s] = (sel[s] ? wcnt1 : wcnt0)[s];
}

__global__ __launch_bounds__(64) void k_wmove(int ns, int d2, const int* __restrict__ sel, const int* __restrict__ wcnt0,
                                              const int* __restrict__ wcnt1, int* __restrict__ woff0,
                                              int* __restrict__ woff1, const long long* __restrict__ noff,
                                              const int* __restrict__ src, int* __restrict__ dst) {
  const int s = blockIdx.x;
  if (s >= ns) return;
  const int b = sel[s];
  const int cnt = (b ? wcnt1 : wcnt0)[s];
  int* wo = b ? woff1 : woff0;
  const int* a = src + (size_t)wo[s] * d2;
  int* o = dst + (size_t)noff[s] * d2;
  for (int e = threadIdx.x; e < cnt * d2; e += kWave) o[e] = a[e];
  if (threadIdx.x == 0) wo[s] = (int)noff[s];
}

__global__ void k_gverdict(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap, const int* __restrict__ slist,
                           int ns, const int* __restrict__ sel, const int* __restrict__ wcnt0,
                           const int* __restrict__ wcnt1, signed char* __restrict__ wne, LiveCtr* lc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const LiveArgs& A = *Ap;
  int cnt = 0;
  bool viol = false;
  if (s < ns) {
    const int v = slist[s];
    cnt = sel[s] ? wcnt1[s] : wcnt0[s];
    wne[v] = cnt > 0 ? 1 : 0;
    if (cnt > 0) {
      if (A.kind == HORA_LIVE_LEADSTO) viol = pred_eval(*Np, A.p, A.gdk[v]);
      else viol = v == 0;
      atomicMax(&lc->fed_max, cnt);
    }
  }
  const u64 b = __ballot(cnt > 0);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd((unsigned long long*)&lc->nonempty, (unsigned long long)__popcll(b));
  if (__any(viol) && (threadIdx.x & 63) == 0) atomicExch(&lc->violated, 1);
}

}  // namespace

// The device-memory tier of k_stop and k_gfp: the launch after the full LDS lists, for the nodes
// whose federations outgrow them, with lists of up to the reference's federation cap (fedcap = 4096
// zones, include/hora/detail/liveness_engine.hpp:420) in a per-workgroup slice of `gl`. Its kernels
// are this source compiled a second time with HORA_FED_HBM (fed_hbm.hip); these launch them. The
// small structs pass as bytes: each unit has its own copy of the same definitions.
namespace hh_fed {
constexpr int kFedCap = 4096;
hipError_t launch_stop_hbm(const void* dnp, const void* argsp, const int* list, long n, int mode, const void* caps,
                           int* scnt, int* soff, int* sz, int scap, void* lc, int* gl, hipStream_t s);
hipError_t launch_gfp_hbm(const void* dnp, const void* argsp, const int* slist, const int* list, int n,
                          const void* caps, const void* banks, void* lc, int* gl, hipStream_t s);
}  // namespace hh_fed

#ifdef HORA_FED_HBM
namespace hh_fed {
hipError_t launch_stop_hbm(const void* dnp, const void* argsp, const int* list, long n, int mode, const void* caps,
                           int* scnt, int* soff, int* sz, int scap, void* lc, int* gl, hipStream_t s) {
  StopCaps K;
  std::memcpy(&K, caps, sizeof(K));
  hipLaunchKernelGGL(k_stop<true>, dim3((unsigned)n), dim3(64), 0, s, (const DevNet*)dnp, (const LiveArgs*)argsp, list,
                     n, (long long)0, mode, K, (int*)nullptr, scnt, soff, sz, scap, (LiveCtr*)lc, gl);
  return hipGetLastError();
}
hipError_t launch_gfp_hbm(const void* dnp, const void* argsp, const int* slist, const int* list, int n,
                          const void* caps, const void* banks, void* lc, int* gl, hipStream_t s) {
  GfpCaps K;
  GfpBanks B;
  std::memcpy(&K, caps, sizeof(K));
  std::memcpy(&B, banks, sizeof(B));
  hipLaunchKernelGGL(k_gfp<true>, dim3((unsigned)n), dim3(64), 0, s, (const DevNet*)dnp, (const LiveArgs*)argsp, slist,
                     list, K, B, (int*)nullptr, (LiveCtr*)lc, gl);
  return hipGetLastError();
}
}  // namespace hh_fed
#endif

#include "online.hpp"
#include "base.hpp"

// ---------------------------------------------------------------- host driver

struct LiveState {
  // query
  DBuf<int> pred;          // term offsets and atoms of p and q
  DBuf<int> oinv_off, oinv_data;
  LiveArgs args{};
  LiveArgs* argsp = nullptr;
  DBuf<LiveCtr> lc;
  // graph
  DBuf<u64> gdk;
  DBuf<long long> grow, segpos, wkey;
  DBuf<int> etgt, ese, ere, s_se, s_re, cov_old, tslot, fnew, newlist;
  DBuf<signed char> expanded, wne;
  long long nodes = 0, edges = 0;
  // online construction: record state, per-level interning scratch
  DBuf<signed char> ractive, whit;
  DBuf<int> rret, recnt, lret, ev, cret, fhz, hz, fskip, fnext, valid2, nflag, wl;
  DBuf<long long> pos2, npos;
  DBuf<u64> soff_save;
  DBuf<u32> slen_save;
  long long hzcap = 0;
  bool graph_valid = false;   // the last query built its graph to the end (readable by live_graph)
  // backward pass
  DBuf<int> in_s, slist, sidx, scnt, soff, sz, wcnt0, wcnt1, woff0, woff1, sel0, sel1, chg0, chg1, heads, dlist, rlist,
      rlist2, rlist3, wlist;
  // base construction: its counters, node links, edge lists, key table and FIFO
  DBuf<BaseCtr> bctr;
  DBuf<int> behead, betail, benext, betgt, bese, bere, bqueue, bqueue2, bpool;
  DBuf<u64> bhkey[2];
  DBuf<long long> bkptr[2];
  DBuf<int> bklen[2], bkcap[2];
  DBuf<signed char> binq;
  int bhbits = 0, bht = 0;   // key table size (bits) and which of the two table buffers is current
  DBuf<int> fedhbm;        // federation lists of the device-memory tier
  long long hbm_nodes = 0; // node launches of the device-memory tier in the last query
  DBuf<int> wpool[2], wtmp;   // value pools (current, compaction target); per-node counts
  DBuf<long long> wnoff;      // compaction offsets
  int wcur = 0;
  long long scap = 0, wcap = 0;
  int ns = 0;
  bool args_borrowed = false;   // argsp taken from an Arena
  ~LiveState() {
    if (argsp && !args_borrowed) { (void)hipDeviceSynchronize(); (void)hipFree(argsp); }
  }
};

#ifndef HORA_FED_HBM
void live_free(LiveState* L) { delete L; }
#endif

namespace {

template <class T>
void grow_keep_to(DBuf<T>& b, size_t n, hipStream_t s) {
  if (n > b.cap || !b.p) b.grow_keep(std::max<size_t>(n, 1024), s);
}

LiveCtr read_lc(hora_hip_ctx* c, LiveState& L) {
  LiveCtr h;
  read_dev(c, &h, L.lc.p, sizeof(LiveCtr));
  unsigned long long sum[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k)
    for (int s = 0; s < kLiveStripes; ++s) sum[k] += h.stripe[k][s][0];
  h.pre_calls = sum[0];
  h.products = sum[1];
  h.changed = (int)sum[2];
  return h;
}

void upload_args(hora_hip_ctx* c, LiveState& L) {
  HIP_OK(hipMemcpyAsync(L.argsp, &L.args, sizeof(LiveArgs), hipMemcpyHostToDevice, c->stream));
}

void sync_graph_args(LiveState& L) {
  L.args.gdk = L.gdk.p;
  L.args.gzone = nullptr;
  L.args.grow = L.grow.p;
  L.args.etgt = L.etgt.p;
  L.args.ese = L.ese.p;
  L.args.ere = L.ere.p;
  L.args.expanded = L.expanded.p;
}

// Full federation capacities from gfx950's whole LDS budget of one workgroup (160 KiB: only the
// nodes that outgrow the small and middle lists run at this size), the same for every list, sized
// so that both StopCaps and GfpCaps at these capacities fit (at most 256 zones: small zone
// dimensions leave room for the many escapes of a broadcast's combinations), and the capacity of
// one value (8 to 32 zones, below). False when fewer than four zones per list fit.
bool live_caps(int d2, LiveArgs& a) {
  const int misc = 2 * (HORA_MAX_PROCS + HORA_MAX_INTS) + 2 * 3 * kInvCap + 2 * kRecipeCap + 128;
  const int zones = (kLdsWorkgroupMax / 4 - misc) / d2;
  // the widest value the lists leave room for: 64 zones while 32 candidate and piece zones still fit
  // (zone dimension up to 15), then 32 while 16 fit (up to 21), then 16 (up to 29), then 8
  a.kw = kLiveW;
  while (a.kw > 8 && (zones - 6 - a.kw) / 3 < a.kw / 2) a.kw /= 2;
  const int k = std::min(256, (zones - 6 - a.kw) / 3);
  const int ks = std::min(256, (zones - 6) / 3);     // k_stop holds no value: its lists take the rest
  a.skp = ks;
  a.kl = ks;
  a.kp = k;
  a.kc = k;
  return k >= 4;
}

// Stop-set launches, like the fixpoint's, run first with small lists and rerun the nodes that outgrow
// them with the full ones. HORA_HIP_STOP_CAPS=kl,kp,ic,rc sets the small lists (escape zones, piece
// zones, invariant constraints, transitions). Invariant and transition lists never exceed the
// network's bounds (DevNet::icap, rcap), and the discrete states take the network's sizes.
StopCaps stop_caps_small(const DevNet& dn) {
  StopCaps k{2, 2, std::min(32, dn.icap), std::min(64, dn.rcap), dn.np, std::max(dn.ni, 1)};
  if (const char* e = std::getenv("HORA_HIP_STOP_CAPS")) {
    StopCaps o = k;
    if (std::sscanf(e, "%d,%d,%d,%d", &o.kl, &o.kp, &o.ic, &o.rc) == 4 && o.kl >= 1 && o.kp >= 1 && o.ic >= 1 &&
        o.ic <= kInvCap && o.rc >= 1 && o.rc <= kRecipeCap && o.ints(dn.d2) <= 64 * 1024 / 4)
      k = o;
  }
  return k;
}

// The full LDS lists; HORA_HIP_STOP_FULL_CAPS=kl,kp shrinks them (for tests: nodes that outgrow them
// go on to the device-memory tier).
StopCaps stop_caps_full(const LiveArgs& a, const DevNet& dn) {
  StopCaps k{a.kl, a.skp, dn.icap, dn.rcap, dn.np, std::max(dn.ni, 1)};
  if (const char* e = std::getenv("HORA_HIP_STOP_FULL_CAPS")) {
    int kl = 0, kp = 0;
    if (std::sscanf(e, "%d,%d", &kl, &kp) == 2 && kl >= 1 && kp >= 1 && kl <= k.kl && kp <= k.kp) {
      k.kl = kl;
      k.kp = kp;
    }
  }
  return k;
}

// k_stop over `n` nodes (list[b], or first_id + b without a list) in `mode`; the caller reads the
// counters afterwards. Mode 0 writes the stop sets of the closure into L's stop-set pool.
// Sets the device-memory tier's node count of a query to zero, and copies it into the query's
// statistics on every return.
struct HbmCount {
  LiveState& L;
  hora_live_stats* st;
  HbmCount(LiveState& l, hora_live_stats* s) : L(l), st(s) { L.hbm_nodes = 0; }
  ~HbmCount() { st->fed_hbm_nodes = L.hbm_nodes; }
};

// HORA_HIP_FED_HBM=0 turns the device-memory tier off: a federation that outgrows the full LDS
// lists then stops the query with HORA_HIP_E_CAPPED. So does the allocation-free entry point (a
// fixed scratch, tl_arena), whose scratch is sized for the graph, not for lists of 4096 zones.
bool fed_hbm_on() {
  if (tl_arena) return false;
  const char* e = std::getenv("HORA_HIP_FED_HBM");
  return !(e && *e == '0');
}

// Nodes per launch of the device-memory tier: as many as keep its lists within 4 GiB.
long fed_hbm_batch(long n, size_t ints_per_node) {
  return std::max<long>(1, std::min<long>(n, (long)((size_t(4) << 30) / (ints_per_node * sizeof(int)))));
}

void launch_stop(hora_hip_ctx* c, LiveState& L, const int* list, long n, long long first_id, int mode) {
  if (n <= 0) return;
  const int d2 = c->d2;
  L.rlist.ensure(n);
  HIP_OK(hipMemsetAsync(&L.lc.p->rcnt, 0, sizeof(int), c->stream));
  const StopCaps KA = stop_caps_small(c->dn), KB = stop_caps_full(L.args, c->dn);
  int* scnt = mode == 0 ? L.scnt.p : nullptr;
  int* soff = mode == 0 ? L.soff.p : nullptr;
  int* sz = mode == 0 ? L.sz.p : nullptr;
  const int scap = mode == 0 ? (int)L.scap : 0;
  hipLaunchKernelGGL(k_stop<false>, dim3(n), dim3(64), KA.ints(d2) * sizeof(int), c->stream, c->dnp, L.argsp, list, n,
                     first_id, mode, KA, L.rlist.p, scnt, soff, sz, scap, L.lc.p, (int*)nullptr);
  LiveCtr h = read_lc(c, L);
  if (h.rcnt <= 0 || h.err) return;
  // a middle tier between the small lists and the full ones: still several wavefronts per CU
  // (not under an override of the small lists, whose runs exercise the full lists, unless
  // HORA_HIP_STOP_MID_CAPS=kl,kp sets the middle tier too)
  StopCaps KM{6, 6, KB.ic, KB.rc, KB.np, KB.ni};
  bool mid = !std::getenv("HORA_HIP_STOP_CAPS");
  if (const char* e = std::getenv("HORA_HIP_STOP_MID_CAPS")) {
    int kl = 0, kp = 0;
    if (std::sscanf(e, "%d,%d", &kl, &kp) == 2 && kl >= KA.kl && kp >= KA.kp) {
      KM.kl = kl;
      KM.kp = kp;
      mid = true;
    }
  }
  const int* full_list = L.rlist.p;
  if (mid && KM.kl < KB.kl && KM.kp < KB.kp && KM.ints(d2) <= 64 * 1024 / 4 / 2) {
    const int nm = h.rcnt;
    L.rlist2.ensure(nm);
    HIP_OK(hipMemsetAsync(&L.lc.p->rcnt, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_stop<false>, dim3(nm), dim3(64), KM.ints(d2) * sizeof(int), c->stream, c->dnp, L.argsp,
                       (const int*)L.rlist.p, (long)nm, (long long)0, mode, KM, L.rlist2.p, scnt, soff, sz, scap,
                       L.lc.p, (int*)nullptr);
    h = read_lc(c, L);
    if (h.rcnt <= 0 || h.err) return;
    full_list = L.rlist2.p;
  }
  // the full LDS lists; the nodes that outgrow them rerun with their lists in device memory
  const int nf = h.rcnt;
  L.rlist3.ensure(nf);
  HIP_OK(hipMemsetAsync(&L.lc.p->rcnt, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_stop<true>, dim3(nf), dim3(64), KB.ints(d2) * sizeof(int), c->stream, c->dnp, L.argsp,
                     full_list, (long)nf, (long long)0, mode, KB, fed_hbm_on() ? L.rlist3.p : (int*)nullptr, scnt,
                     soff, sz, scap, L.lc.p, (int*)nullptr);
  h = read_lc(c, L);
  if (h.rcnt <= 0 || h.err) return;
  const StopCaps KH{hh_fed::kFedCap, hh_fed::kFedCap, KB.ic, KB.rc, KB.np, KB.ni};
  const size_t per = (size_t)KH.ints(d2);
  const long batch = fed_hbm_batch(h.rcnt, per);
  L.fedhbm.ensure(per * (size_t)batch);
  L.hbm_nodes += h.rcnt;
  for (long b0 = 0; b0 < h.rcnt; b0 += batch)
    HIP_OK(hh_fed::launch_stop_hbm(c->dnp, L.argsp, L.rlist3.p + b0, std::min<long>(batch, h.rcnt - b0), mode, &KH,
                                   scnt, soff, sz, scap, L.lc.p, L.fedhbm.p, c->stream));
}

// HORA_HIP_LIVE_POOL=tight: the stop-set and value pools start at the least they can hold and grow
// by exactly what is asked, so that growth, compaction and the rerun of a round all happen on small
// inputs (for tests).
bool pool_tight() {
  const char* e = std::getenv("HORA_HIP_LIVE_POOL");
  return e && std::strcmp(e, "tight") == 0;
}

// Moves every node's current value to the front of the other value pool, sized for `room` more
// zones, and makes that pool current.
void compact_values(hora_hip_ctx* c, LiveState& L, int ns, const int* sel, long long room) {
  L.wtmp.ensure(ns);
  L.wnoff.ensure(ns);
  hipLaunchKernelGGL(k_wcount, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, ns, sel, L.wcnt0.p, L.wcnt1.p,
                     L.wtmp.p);
  const long long live = scan_exclusive(c, L.wtmp.p, ns, L.wnoff.p);
  const long long need = live + room;
  if (need > INT32_MAX) throw HipFail("value pool above 2^31 zones");
  DBuf<int>& dst = L.wpool[1 - L.wcur];
  dst.ensure((size_t)need * c->d2);
  hipLaunchKernelGGL(k_wmove, dim3(ns), dim3(64), 0, c->stream, ns, c->d2, sel, L.wcnt0.p, L.wcnt1.p, L.woff0.p,
                     L.woff1.p, L.wnoff.p, L.wpool[L.wcur].p, dst.p);
  L.wcur = 1 - L.wcur;
  L.wcap = pool_tight() ? need : std::min<long long>((long long)(dst.cap / c->d2), INT32_MAX);
  const int top = (int)live;
  HIP_OK(hipMemcpyAsync(&L.lc.p->ptop, &top, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
}

// Fixpoint launches run first with small LDS lists, so that many wavefronts share a CU; the few
// nodes that outgrow them rerun with the full lists of live_caps. HORA_HIP_GFP_CAPS=kw,kc,kp,ic sets
// the small lists (zones of the value, of the candidates, of each piece list; invariant constraints).
GfpCaps gfp_caps_small(const DevNet& dn) {
  GfpCaps k{2, 8, 4, std::min(32, dn.icap), dn.np, std::max(dn.ni, 1)};
  if (k.ints(dn.d2) > 64 * 1024 / 4) k = GfpCaps{1, 4, 2, k.ic, k.np, k.ni};   // zone dimension above 27
  if (const char* e = std::getenv("HORA_HIP_GFP_CAPS")) {
    GfpCaps o = k;
    if (std::sscanf(e, "%d,%d,%d,%d", &o.kw, &o.kc, &o.kp, &o.ic) == 4 && o.kw >= 1 && o.kw <= kLiveW &&
        o.kc >= 1 && o.kp >= 2 && o.ic >= 1 && o.ic <= kInvCap && o.ints(dn.d2) <= 64 * 1024 / 4)
      k = o;
  }
  return k;
}

// The full LDS lists; HORA_HIP_GFP_FULL_CAPS=kw,kc,kp shrinks them (for tests: nodes that outgrow
// them go on to the device-memory tier).
GfpCaps gfp_caps_full(const LiveArgs& a, const DevNet& dn) {
  GfpCaps k{a.kw, a.kc, a.kp, dn.icap, dn.np, std::max(dn.ni, 1)};
  if (const char* e = std::getenv("HORA_HIP_GFP_FULL_CAPS")) {
    int kw = 0, kc = 0, kp = 0;
    if (std::sscanf(e, "%d,%d,%d", &kw, &kc, &kp) == 3 && kw >= 1 && kc >= 1 && kp >= 2 && kw <= k.kw &&
        kc <= k.kc && kp <= k.kp) {
      k.kw = kw;
      k.kc = kc;
      k.kp = kp;
    }
  }
  return k;
}

// Grows a per-record array to n entries, filling what is new with `fill` bytes.
template <class T>
void grow_rec(DBuf<T>& b, size_t n, int fill, hipStream_t s) {
  if (n <= b.cap && b.p) return;
  const size_t old = b.p ? b.cap : 0;
  b.grow_keep(std::max<size_t>(n, 1024), s);
  HIP_OK(hipMemsetAsync(b.p + old, fill, (b.cap - old) * sizeof(T), s));
}

// The packed discrete key of the initial state.
u64 initial_key(const hora_hip_ctx* c) {
  u64 dk0 = 0;
  for (int p = 0; p < c->np; ++p) dk0 |= (u64)(unsigned)c->h_loc_init[p] << c->h_kshift[p];
  for (int v = 0; v < c->ni; ++v) dk0 |= (u64)(unsigned)(c->h_int_init[v] - c->h_int_lo[v]) << c->h_kshift[c->np + v];
  return dk0;
}

// One construction of the graph by the online rule (online.hpp). Returns HORA_HIP_OK or an error
// code; *early receives the verdict of an early witness, or -1.
int construct_online(hora_hip_ctx* c, LiveState& L, const hora_live_query* q, hora_live_stats* st, int* early) {
  const int d2 = c->d2;
  const bool eg = q->kind == HORA_LIVE_EG;
  *early = -1;
  {
    const int32_t seeds = st->seed_keys;
    std::memset(st, 0, sizeof(*st));
    st->seed_keys = seeds;
  }
  HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
  c->dn.lu_plus = 0;
  c->dn.znar = 0;
  HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));
  alloc_table(c, std::max(16, c->hbits));
  c->ctr.ensure(1);
  HIP_OK(hipMemsetAsync(c->ctr.p, 0, sizeof(Counters), c->stream));
  {
    const int ft = 0x7fffffff;
    HIP_OK(hipMemcpyAsync(&c->ctr.p->first_target, &ft, sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  const size_t ncap0 = 1 << 16, ecap0 = 1 << 18;
  L.gdk.ensure(ncap0);
  L.expanded.ensure(ncap0);
  L.ractive.ensure(ncap0);
  L.rret.ensure(ncap0);
  L.recnt.ensure(ncap0);
  L.lret.ensure(ncap0);
  L.whit.ensure(ncap0);
  HIP_OK(hipMemsetAsync(L.expanded.p, 0, L.expanded.cap, c->stream));
  HIP_OK(hipMemsetAsync(L.whit.p, 0, L.whit.cap, c->stream));
  HIP_OK(hipMemsetAsync(L.recnt.p, 0, L.recnt.cap * sizeof(int), c->stream));
  HIP_OK(hipMemsetAsync(L.lret.p, 0xff, L.lret.cap * sizeof(int), c->stream));
  c->nodes.ensure(1 << 20);
  u64 ncount = 0;
  c->zarena.ensure(ncap0 * d2);
  L.etgt.ensure(ecap0); L.ese.ensure(ecap0); L.ere.ensure(ecap0);
  L.nodes = 0;
  L.edges = 0;
  L.ns = 0;
  if (L.hzcap <= 0) L.hzcap = 1024;
  L.args.recnt = L.recnt.p;
  L.args.whit = L.whit.p;

  const u64 dk0 = initial_key(c);
  c->fr_dk.ensure(1);
  c->fr_zid.ensure(1);
  HIP_OK(hipMemcpyAsync(c->fr_dk.p, &dk0, sizeof(u64), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemsetAsync(c->fr_zid.p, 0, sizeof(u32), c->stream));

  long F = 1;
  bool init = true;
  long long rounds_extra = 0;
  for (;;) {
    L.args.recnt = L.recnt.p;
    L.args.whit = L.whit.p;
    sync_graph_args(L);
    L.args.gzone = c->zarena.p;
    upload_args(c, L);
    // ---- expansion of every frontier record
    long R = 1;
    c->rec_cnt.ensure(F);
    c->rec_off.ensure(F);
    c->pos.ensure(F + 1);
    if (!init) {
      hipLaunchKernelGGL(k_count, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, c->fr_dk.p, (int)F,
                         c->rec_cnt.p, c->ctr.p);
      if (eg)
        hipLaunchKernelGGL(k_ogfilter, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, c->fr_dk.p,
                           (int)F, c->rec_cnt.p);
      R = scan_exclusive(c, c->rec_cnt.p, F, c->pos.p);
      hipLaunchKernelGGL(k_offsets_to_int, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->pos.p, F, c->rec_off.p);
    } else {
      HIP_OK(hipMemsetAsync(c->rec_off.p, 0, sizeof(int), c->stream));
      HIP_OK(hipMemsetAsync(c->pos.p, 0, sizeof(long long), c->stream));   // k_expand's slot offset
    }
    c->s_valid.ensure(R);
    c->s_tgt.ensure(R);
    c->s_dk.ensure(R);
    c->s_zone.ensure((size_t)R * d2);
    L.s_se.ensure(R);
    L.s_re.ensure(R);
    ExpandOut eo{c->s_valid.p, c->s_dk.p, c->s_zone.p, c->s_tgt.p, L.s_se.p, L.s_re.p};
    if (R > 0) launch_expand(c, expand_tuned(c, F, 0, eo, init), F, 0, eo, init);
    c->pos.ensure(R + 1);
    const long long Gall = R > 0 ? scan_exclusive(c, c->s_valid.p, R, c->pos.p) : 0;
    {
      const Tally hc = read_counters(c);
      if (hc.err) { c->err = "a state has more transitions or invariant constraints than the device holds"; return HORA_HIP_E_MEMORY; }
    }
    // ---- early witnesses of the frontier records, each judged at its own turn
    if (!init) {
      hipLaunchKernelGGL(k_ofsucc, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->rec_off.p, (int)F, R, c->pos.p,
                         Gall, c->fr_zid.p, L.recnt.p);
      if (q->early) {
        L.wl.ensure(F);
        HIP_OK(hipMemsetAsync(&L.lc.p->wcnt, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_owlist, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, (int)F,
                           c->fr_dk.p, c->fr_zid.p, L.wl.p, L.lc.p);
        const int nw = read_lc(c, L).wcnt;
        launch_stop(c, L, L.wl.p, (long)nw, 0, q->kind == HORA_LIVE_DEADLOCK ? 2 : 1);
        const LiveCtr h = read_lc(c, L);
        if (h.err) {
          c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a stop set" : "invariant or transition table too large";
          st->capped = h.err == ERR_CAPPED;
          return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
        }
      }
    }
    // ---- interning of the level, repeated until the skip marks of the frontier are stable
    L.fskip.ensure(F);
    L.fnext.ensure(F);
    HIP_OK(hipMemsetAsync(L.fskip.p, 0, (size_t)F * sizeof(int), c->stream));
    HIP_OK(hipMemsetAsync(L.fnext.p, 0, (size_t)F * sizeof(int), c->stream));
    L.valid2.ensure(std::max<long>(R, 1));
    L.pos2.ensure(R + 1);
    long G = 0;
    int T = 0, wfirst = INT32_MAX;
    LiveCtr hl{};
    for (;;) {
      wfirst = INT32_MAX;
      if (!init && q->early) {
        HIP_OK(hipMemcpyAsync(&L.lc.p->wfirst, &wfirst, sizeof(int), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_owfirst, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->fr_zid.p, (int)F, L.fskip.p,
                           L.whit.p, L.lc.p);
        wfirst = read_lc(c, L).wfirst;
      }
      if (R > 0)
        hipLaunchKernelGGL(k_oexcl, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->rec_off.p, (int)F, R,
                           L.fskip.p, wfirst, c->s_valid.p, L.valid2.p);
      G = R > 0 ? (long)scan_exclusive(c, L.valid2.p, R, L.pos2.p) : 0;
      HIP_OK(hipMemsetAsync(&L.lc.p->nev, 0, 5 * sizeof(int), c->stream));   // nev, nuni, nret, hzn, hzfull
      T = 0;
      if (G > 0) {
        c->sj.ensure(G);
        hipLaunchKernelGGL(k_compact_valid, dim3(blocks_for(R, 256)), dim3(256), 0, c->stream, L.valid2.p, L.pos2.p, R,
                           c->sj.p, c->ctr.p);
        {
          const Tally hc = read_counters(c);
          while ((double)(hc[C_KEYS] + (u64)G) > 0.6 * (double)c->H) grow_table(c);
        }
        Passed P = passed_view(c);
        c->gbk.ensure(G);
        c->gticket.ensure(G);
        c->touched.ensure(G);
        c->covered.ensure(G);
        L.cov_old.ensure(G);
        L.tslot.ensure(G);
        L.fnew.ensure(G);
        L.cret.ensure(G);
        L.fhz.ensure(G);
        HIP_OK(hipMemsetAsync(L.fnew.p, 0, (size_t)G * sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(L.fhz.p, 0, (size_t)G * sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(L.tslot.p, 0xff, (size_t)G * sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(L.cret.p, 0xff, (size_t)G * sizeof(int), c->stream));
        hipLaunchKernelGGL(k_find, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->sj.p, c->s_dk.p, G, c->gbk.p,
                           c->ctr.p, (int*)nullptr, (int*)nullptr);
        hipLaunchKernelGGL(k_gprobe, dim3(blocks_for(G, 4)), dim3(64), 0, c->stream, c->dnp, P, c->gbk.p, c->sj.p,
                           c->s_dk.p, c->s_zone.p, G, c->zarena.p, c->covered.p, L.cov_old.p, c->ctr.p, 1, 1);
        hipLaunchKernelGGL(k_ticket, dim3(blocks_for(G, 1024)), dim3(1024), 0, c->stream, P, c->gbk.p, c->covered.p, G,
                           c->gticket.p, c->touched.p, c->ctr.p);
        read_dev(c, &T, &c->ctr.p->ntouched, sizeof(int));
        if (T > 0) {
          c->caps.ensure(T);
          c->segs.ensure(T);
          hipLaunchKernelGGL(k_touched_caps, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                             c->caps.p, c->segs.p, c->ctr.p, (const int*)nullptr);
          c->klight.ensure(T);
          c->kheavy.ensure(T);
          hipLaunchKernelGGL(k_classify, dim3(blocks_for(T, 1024)), dim3(1024), 0, c->stream, P, c->touched.p, T,
                             c->klight.p, c->kheavy.p, c->ctr.p, kLightMax, c->d, (const int*)nullptr);
          c->pos2.ensure(T + 1);
          L.segpos.ensure(T + 1);
          const long long W = scan_exclusive(c, c->caps.p, T, c->pos2.p);
          const long long Sseg = scan_exclusive(c, c->segs.p, T, L.segpos.p);
          int nlh[2] = {0, 0};
          read_dev(c, nlh, &c->ctr.p->nlight, sizeof(nlh));
          if (ncount + (u64)Sseg > c->nodes.cap) c->nodes.grow_keep((size_t)(ncount + (u64)Sseg) * 2, c->stream);
          P = passed_view(c);
          hipLaunchKernelGGL(k_touched_offsets, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                             c->pos2.p, L.segpos.p, ncount);
          ncount += (u64)Sseg;
          L.soff_save.ensure(T);
          L.slen_save.ensure(T);
          hipLaunchKernelGGL(k_osave, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                             L.soff_save.p, L.slen_save.p, 0);
          c->wlist.ensure(W + 1);
          hipLaunchKernelGGL(k_scatter, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->gbk.p, c->covered.p,
                             c->gticket.p, G, c->wlist.p);
          L.hz.ensure((size_t)L.hzcap * d2);
          L.ev.ensure(std::max<long>(G, 1) + (long)std::max<long long>(L.nodes, 1));
          const int hzc = (int)std::min<long long>(L.hzcap, INT32_MAX);
          const size_t lds_l = merge_layout(d2, false).total() * sizeof(int);
          const size_t lds_h = merge_layout(d2, true).total() * sizeof(int);
          if (nlh[0] > 0)
            hipLaunchKernelGGL(k_omerge, dim3(nlh[0]), dim3(64), lds_l, c->stream, c->dnp, P, c->klight.p, nlh[0],
                               c->wlist.p, c->sj.p, c->s_zone.p, L.hz.p, hzc, L.fhz.p, L.tslot.p, L.fnew.p, L.cret.p,
                               L.lret.p, L.ev.p, L.lc.p, c->ctr.p, 0);
          if (nlh[1] > 0)
            hipLaunchKernelGGL(k_omerge, dim3(nlh[1]), dim3(64), lds_h, c->stream, c->dnp, P, c->kheavy.p, nlh[1],
                               c->wlist.p, c->sj.p, c->s_zone.p, L.hz.p, hzc, L.fhz.p, L.tslot.p, L.fnew.p, L.cret.p,
                               L.lret.p, L.ev.p, L.lc.p, c->ctr.p, 1);
          hipLaunchKernelGGL(k_opost, dim3(G), dim3(64), 0, c->stream, c->dnp, P, c->gbk.p, (long)G, c->covered.p,
                             c->wlist.p, c->sj.p, c->s_dk.p, c->s_zone.p, L.hz.p, L.fhz.p, L.tslot.p, c->ctr.p);
        }
      }
      hl = read_lc(c, L);
      bool again = false;
      if (hl.hzfull) {
        L.hzcap = std::max<long long>(2 * L.hzcap, 2LL * hl.hzn);
        again = true;
      } else if (!init && hl.nev > 0) {
        HIP_OK(hipMemsetAsync(&L.lc.p->vchanged, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_ovalid, dim3(blocks_for(hl.nev, 256)), dim3(256), 0, c->stream, L.ev.p, hl.nev, L.lret.p,
                           c->fr_zid.p, (int)F, c->sj.p, c->rec_off.p, L.fnext.p);
        hipLaunchKernelGGL(k_ovcmp, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, L.fskip.p, L.fnext.p, (int)F,
                           L.lc.p);
        again = read_lc(c, L).vchanged != 0;
      }
      if (!again) break;
      // put the level's chains back and intern it again
      ++rounds_extra;
      if (T > 0) {
        Passed P = passed_view(c);
        hipLaunchKernelGGL(k_osave, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                           L.soff_save.p, L.slen_save.p, 1);
      }
      if (hl.nev > 0)
        hipLaunchKernelGGL(k_oclear_ev, dim3(blocks_for(hl.nev, 256)), dim3(256), 0, c->stream, L.ev.p, hl.nev, L.lret.p);
    }
    // ---- the level's records, edges and retirements
    long F2 = 0;
    const long long nbase = L.nodes;
    if (G > 0) {
      c->pos.ensure(G + 1);
      F2 = (long)scan_exclusive(c, L.fnew.p, G, c->pos.p);
      const size_t nneed = (size_t)(nbase + F2 + 1);
      grow_rec(L.gdk, nneed, 0, c->stream);
      grow_rec(L.expanded, nneed, 0, c->stream);
      grow_rec(L.ractive, nneed, 0, c->stream);
      grow_rec(L.rret, nneed, 0xff, c->stream);
      grow_rec(L.recnt, nneed, 0, c->stream);
      grow_rec(L.lret, nneed, 0xff, c->stream);
      grow_rec(L.whit, nneed, 0, c->stream);
      if (nneed * d2 > c->zarena.cap) c->zarena.grow_keep(std::max(nneed * d2, c->zarena.cap * 2), c->stream);
      if (!init && (size_t)(L.edges + G) > L.etgt.cap) {
        const size_t ec2 = std::max((size_t)(L.edges + G), L.etgt.cap * 2);
        grow_keep_to(L.etgt, ec2, c->stream);
        grow_keep_to(L.ese, ec2, c->stream);
        grow_keep_to(L.ere, ec2, c->stream);
      }
      L.newlist.ensure(std::max<long>(F2, 1));
      L.nflag.ensure(std::max<long>(F2, 1));
      hipLaunchKernelGGL(k_oassign, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, (long)G, nbase, L.edges,
                         init ? 0 : 1, L.cov_old.p, L.tslot.p, L.fnew.p, c->pos.p, c->sj.p, L.s_se.p, L.s_re.p,
                         L.etgt.p, L.ese.p, L.ere.p, L.newlist.p);
      if (F2 > 0) {
        hipLaunchKernelGGL(k_ocopy, dim3(F2), dim3(64), 0, c->stream, (long)F2, nbase, L.newlist.p, c->sj.p, c->s_dk.p,
                           c->s_zone.p, L.hz.p, L.fhz.p, L.cret.p, d2, c->zarena.p, L.gdk.p, L.ractive.p, L.rret.p,
                           L.nflag.p);
        hipLaunchKernelGGL(k_oretire_new, dim3(blocks_for(F2, 256)), dim3(256), 0, c->stream, (long)F2, nbase,
                           L.newlist.p, L.cret.p, c->pos.p, L.ractive.p, L.rret.p);
      }
      if (hl.nev > 0) {
        hipLaunchKernelGGL(k_oretire_old, dim3(blocks_for(hl.nev, 256)), dim3(256), 0, c->stream, L.ev.p, hl.nev,
                           L.lret.p, nbase, c->pos.p, L.ractive.p, L.rret.p);
        hipLaunchKernelGGL(k_oclear_ev, dim3(blocks_for(hl.nev, 256)), dim3(256), 0, c->stream, L.ev.p, hl.nev, L.lret.p);
      }
      if (T > 0) {
        Passed P = passed_view(c);
        hipLaunchKernelGGL(k_olink, dim3(T), dim3(64), 0, c->stream, P, c->touched.p, T, c->pos.p, nbase);
      }
    }
    if (!init) {
      HIP_OK(hipMemsetAsync(&L.lc.p->nexp, 0, sizeof(int), c->stream));
      hipLaunchKernelGGL(k_ofront, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, c->rec_off.p,
                         (int)F, R, L.pos2.p, (long long)G, c->fr_zid.p, c->fr_dk.p, L.fskip.p, wfirst, eg ? 1 : 0,
                         L.expanded.p, L.recnt.p, L.lc.p);
    }
    // next frontier: the new records still active, in creation order
    long Fn = 0;
    if (F2 > 0) {
      L.npos.ensure(F2 + 1);
      Fn = (long)scan_exclusive(c, L.nflag.p, F2, L.npos.p);
      c->fr_dk2.ensure(std::max<long>(Fn, 1));
      c->fr_zid2.ensure(std::max<long>(Fn, 1));
      hipLaunchKernelGGL(k_onext, dim3(blocks_for(F2, 256)), dim3(256), 0, c->stream, (long)F2, nbase, L.nflag.p,
                         L.npos.p, L.newlist.p, c->sj.p, c->s_dk.p, c->fr_dk2.p, c->fr_zid2.p);
    }
    const LiveCtr he = read_lc(c, L);
    {
      const Tally hc = read_counters(c);
      if (hc.err) { c->err = hc.err == ERR_HASH_FULL ? "hash table full" : "device capacity exceeded"; return HORA_HIP_E_MEMORY; }
      st->chain_reads = (long long)(hc[C_READS] + hc[C_PREADS]);
    }
    if (!init) {
      st->expanded_nodes += he.nexp;
      if (he.nexp > 0) ++st->levels;
      st->generated += G;
      st->covered += G - F2;
      L.edges += G;
    }
    st->unions += hl.nuni;
    st->retired += hl.nret;
    L.nodes += F2;
    if (c->trace)
      std::fprintf(stderr, "online level F=%ld R=%ld G=%ld T=%d new=%ld next=%ld retired=%d unions=%d extra rounds so far %lld\n",
                   F, R, G, T, F2, Fn, hl.nret, hl.nuni, rounds_extra);
    if (wfirst != INT32_MAX) {
      *early = eg ? 1 : 0;
      st->early = 1;
      break;
    }
    std::swap(c->fr_dk.p, c->fr_dk2.p); std::swap(c->fr_dk.cap, c->fr_dk2.cap);
    std::swap(c->fr_zid.p, c->fr_zid2.p); std::swap(c->fr_zid.cap, c->fr_zid2.cap);
    F = Fn;
    init = false;
    if (F == 0) break;
  }
  // edges to records never expanded go to their active coverer; rows from the edge counts
  if (L.edges > 0)
    hipLaunchKernelGGL(k_oresolve, dim3(blocks_for((long)L.edges, 256)), dim3(256), 0, c->stream, L.edges, L.etgt.p,
                       L.expanded.p, L.rret.p);
  grow_keep_to(L.grow, (size_t)(L.nodes + 1), c->stream);
  if (L.nodes > 0) scan_launch(c, L.recnt.p, (long)L.nodes, L.grow.p, 0);
  HIP_OK(hipMemcpyAsync(L.grow.p + L.nodes, &L.edges, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  {
    HIP_OK(hipMemsetAsync(&L.lc.p->nkeys, 0, sizeof(int), c->stream));
    Passed P = passed_view(c);
    hipLaunchKernelGGL(k_okeys, dim3(blocks_for((long)c->H, 256)), dim3(256), 0, c->stream, P, (long)c->H, L.lc.p);
    st->disc_keys = read_lc(c, L).nkeys;
  }
  st->outer_nodes = L.nodes;
  st->fwd_edges = L.edges;
  return HORA_HIP_OK;
}

BaseState base_view(hora_hip_ctx* c, LiveState& L, int t = -1) {
  if (t < 0) t = L.bht;
  BaseState B;
  B.nkey = L.gdk.p;
  B.zone = c->zarena.p;
  B.expanded = L.expanded.p;
  B.inq = L.binq.p;
  B.ehead = L.behead.p;
  B.etail = L.betail.p;
  B.recnt = L.recnt.p;
  B.enext = L.benext.p;
  B.etgt = L.betgt.p;
  B.ese = L.bese.p;
  B.ere = L.bere.p;
  B.hkey = L.bhkey[t].p;
  B.kptr = L.bkptr[t].p;
  B.klen = L.bklen[t].p;
  B.kcap = L.bkcap[t].p;
  B.pool = L.bpool.p;
  B.queue = L.bqueue.p;
  B.hbits = L.bhbits;
  B.bc = L.bctr.p;
  return B;
}

// Table buffer t sized for 2^bits slots, every slot empty.
void base_table(hora_hip_ctx* c, LiveState& L, int t, int bits) {
  const size_t H = (size_t)1 << bits;
  L.bhkey[t].ensure(H);
  L.bkptr[t].ensure(H);
  L.bklen[t].ensure(H);
  L.bkcap[t].ensure(H);
  HIP_OK(hipMemsetAsync(L.bhkey[t].p, 0xff, H * sizeof(u64), c->stream));   // kEmptyKey
}

// A discrete predicate on a packed key, on the host (the device's pred_eval).
bool host_pred(const hora_hip_ctx* c, const hora_dpred& P, u64 dk) {
  for (int t = 0; t < P.nterms; ++t) {
    bool all = true;
    for (int a = P.term_off[t]; a < P.term_off[t + 1] && all; ++a) {
      const int kind = P.atoms[4 * a], idx = P.atoms[4 * a + 1], op = P.atoms[4 * a + 2], val = P.atoms[4 * a + 3];
      const int f = kind == 0 ? idx : c->np + idx;
      int v = (int)((dk >> c->h_kshift[f]) & c->h_kmask[f]);
      if (kind != 0) v += c->h_int_lo[idx];
      all = op == 0 ? v == val : op == 1 ? v != val : op == 2 ? v < val : op == 3 ? v <= val : op == 4 ? v > val : v >= val;
    }
    if (all) return true;
  }
  return false;
}

// Whether the initial discrete state can seed a deadlock atom (mode 3 of k_stop on the initial key,
// online_seed_key, refinement.cpp:176-282): on a network with one possible discrete state, the
// negation of structurally_empty_d's test (src/liveness/internal.cpp:59-103).
int initial_key_seeds(hora_hip_ctx* c, LiveState& L, hora_live_stats* st, bool* seeds) {
  const u64 dk0 = initial_key(c);
  L.gdk.ensure(1);
  HIP_OK(hipMemcpyAsync(L.gdk.p, &dk0, sizeof(u64), hipMemcpyHostToDevice, c->stream));
  c->zarena.ensure(c->d2);
  sync_graph_args(L);
  L.args.gzone = c->zarena.p;
  L.args.recnt = nullptr;
  L.args.whit = nullptr;
  upload_args(c, L);
  HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
  launch_stop(c, L, nullptr, 1, 0, 3);
  const LiveCtr h = read_lc(c, L);
  if (h.err) {
    c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a seed test" : "invariant or transition table too large";
    st->capped = h.err == ERR_CAPPED;
    return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
  }
  *seeds = h.hit > 0;
  return HORA_HIP_OK;
}

// The base construction (base.hpp): one node expanded per step, in FIFO order. Returns HORA_HIP_OK
// or an error code; *early receives the verdict of an early witness, or -1.
int construct_base(hora_hip_ctx* c, LiveState& L, const hora_live_query* q, hora_live_stats* st, int* early) {
  const int d2 = c->d2;
  *early = -1;
  {
    const int32_t seeds = st->seed_keys;
    std::memset(st, 0, sizeof(*st));
    st->seed_keys = seeds;
  }
  HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
  c->dn.lu_plus = 0;
  c->dn.znar = 0;
  HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));
  c->ctr.ensure(1);
  HIP_OK(hipMemsetAsync(c->ctr.p, 0, sizeof(Counters), c->stream));
  {
    const int ft = 0x7fffffff;
    HIP_OK(hipMemcpyAsync(&c->ctr.p->first_target, &ft, sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  // node, edge and FIFO storage (grown below as the steps need)
  // HORA_HIP_LIVE_POOL=tight: small first sizes, so that every buffer and the key table grow (tests)
  const bool tight = pool_tight();
  const size_t ncap0 = tight ? 64 : 4096, ecap0 = tight ? 64 : 16384;
  const int hbits0 = tight ? 4 : 12;
  L.gdk.ensure(ncap0);
  L.expanded.ensure(ncap0);
  L.recnt.ensure(ncap0);
  L.binq.ensure(ncap0);
  L.bpool.ensure(8 * ncap0);
  L.behead.ensure(ncap0);
  L.betail.ensure(ncap0);
  L.bqueue.ensure(ncap0);
  L.benext.ensure(ecap0);
  L.betgt.ensure(ecap0);
  L.bese.ensure(ecap0);
  L.bere.ensure(ecap0);
  c->zarena.ensure(ncap0 * d2);
  L.bctr.ensure(1);
  HIP_OK(hipMemsetAsync(L.bctr.p, 0, sizeof(BaseCtr), c->stream));
  L.bht = 0;
  L.bhbits = hbits0;
  base_table(c, L, 0, hbits0);
  L.nodes = 0;
  L.edges = 0;
  L.ns = 0;
  auto upload_graph_args = [&]() {
    sync_graph_args(L);
    L.args.gzone = c->zarena.p;
    L.args.recnt = L.recnt.p;   // the stop set of an early test reads the node's successor count
    L.args.whit = nullptr;
    upload_args(c, L);
  };
  upload_graph_args();

  // the initial node: slot 0 of the initial expansion
  const u64 dk0 = initial_key(c);
  c->fr_dk.ensure(1);
  c->fr_zid.ensure(1);
  c->rec_cnt.ensure(1);
  c->rec_off.ensure(1);
  c->pos.ensure(2);
  HIP_OK(hipMemcpyAsync(c->fr_dk.p, &dk0, sizeof(u64), hipMemcpyHostToDevice, c->stream));
  HIP_OK(hipMemsetAsync(c->fr_zid.p, 0, sizeof(u32), c->stream));
  HIP_OK(hipMemsetAsync(c->rec_off.p, 0, sizeof(int), c->stream));
  HIP_OK(hipMemsetAsync(c->pos.p, 0, sizeof(long long), c->stream));
  auto outputs = [&](long R) {
    c->s_valid.ensure(std::max<long>(R, 1));
    c->s_tgt.ensure(std::max<long>(R, 1));
    c->s_dk.ensure(std::max<long>(R, 1));
    c->s_zone.ensure((size_t)std::max<long>(R, 1) * d2);
    L.s_se.ensure(std::max<long>(R, 1));
    L.s_re.ensure(std::max<long>(R, 1));
    return ExpandOut{c->s_valid.p, c->s_dk.p, c->s_zone.p, c->s_tgt.p, L.s_se.p, L.s_re.p};
  };
  {
    const ExpandOut eo = outputs(1);
    launch_expand(c, expand_tuned(c, 1, 0, eo, true), 1, 0, eo, true);
  }
  BaseState B = base_view(c, L);
  hipLaunchKernelGGL(k_base_init, dim3(1), dim3(64), 0, c->stream, B, (const int*)c->s_valid.p, (const u64*)c->s_dk.p,
                     (const int*)c->s_zone.p, d2);
  // each step's first half: the frontier of node u and its successor slot count
  auto stage = [&]() {
    hipLaunchKernelGGL(k_base_frontier, dim3(1), dim3(1), 0, c->stream, B, c->fr_dk.p, c->fr_zid.p);
    hipLaunchKernelGGL(k_count, dim3(1), dim3(256), 0, c->stream, c->dnp, c->fr_dk.p, 1, c->rec_cnt.p, c->ctr.p);
    scan_launch(c, c->rec_cnt.p, 1, c->pos.p, 0);
  };
  stage();
  BaseCtr hb{};
  long long R = 0;
  read_dev(c, &hb, L.bctr.p, sizeof(BaseCtr), &R, c->scan_total.p, sizeof(R));
  long long steps = 0;
  while (hb.u >= 0) {
    // room for the step: R more nodes and edges, the FIFO's pushes (each node at most once), R keys
    bool moved = false;
    const size_t nneed = (size_t)hb.nodes + (size_t)R + 1;
    if (nneed > L.gdk.cap || nneed > L.expanded.cap || nneed > L.recnt.cap || nneed > L.binq.cap ||
        nneed > L.behead.cap || nneed > L.betail.cap || nneed * d2 > c->zarena.cap) {
      const size_t n2 = std::max(nneed, 2 * (size_t)hb.nodes);
      grow_keep_to(L.gdk, n2, c->stream);
      grow_keep_to(L.expanded, n2, c->stream);
      grow_keep_to(L.recnt, n2, c->stream);
      grow_keep_to(L.binq, n2, c->stream);
      grow_keep_to(L.behead, n2, c->stream);
      grow_keep_to(L.betail, n2, c->stream);
      if (n2 * d2 > c->zarena.cap) c->zarena.grow_keep(std::max(n2 * d2, c->zarena.cap * 2), c->stream);
      moved = true;
    }
    const size_t eneed = (size_t)hb.edges + (size_t)R;
    if (eneed > L.benext.cap || eneed > L.betgt.cap || eneed > L.bese.cap || eneed > L.bere.cap) {
      const size_t e2 = std::max(eneed, 2 * (size_t)hb.edges);
      grow_keep_to(L.benext, e2, c->stream);
      grow_keep_to(L.betgt, e2, c->stream);
      grow_keep_to(L.bese, e2, c->stream);
      grow_keep_to(L.bere, e2, c->stream);
      moved = true;
    }
    if ((size_t)hb.qt + nneed > L.bqueue.cap) {
      // the queued part moves to the front of a buffer with room for it and the step's pushes
      const long long len = hb.qt - hb.qh;
      L.bqueue2.ensure(std::max<size_t>(2 * ((size_t)len + nneed), 4096));
      if (len > 0)
        HIP_OK(hipMemcpyAsync(L.bqueue2.p, L.bqueue.p + hb.qh, (size_t)len * sizeof(int), hipMemcpyDeviceToDevice,
                              c->stream));
      std::swap(L.bqueue.p, L.bqueue2.p);
      std::swap(L.bqueue.cap, L.bqueue2.cap);
      const long long qq[2] = {0, len};
      HIP_OK(hipMemcpyAsync(&L.bctr.p->qh, qq, sizeof(qq), hipMemcpyHostToDevice, c->stream));
      HIP_OK(hipStreamSynchronize(c->stream));
      hb.qh = 0;
      hb.qt = len;
      moved = true;
    }
    {
      // the key arrays' pool: in one step a key holding n >= r nodes, r of the step's appends,
      // doubles at most once (2n <= 2 * nodes over all such keys); a smaller or new key takes
      // capacities 8, 16, ... summing to less than twice its last, 4 * max(8, n + r) <= 32 + 8r;
      // so a step takes at most 2 * nodes + 40 * R
      const size_t pneed = (size_t)hb.ptop + 2 * nneed + 40 * (size_t)R + 64;
      if (pneed > L.bpool.cap) {
        grow_keep_to(L.bpool, std::max(pneed, 2 * (size_t)hb.ptop), c->stream);
        moved = true;
      }
    }
    if ((double)(hb.keys + R) > 0.5 * (double)((size_t)1 << L.bhbits)) {
      // a table twice as large (or more), filled from the current one
      const long Ho = 1L << L.bhbits;
      int bits = L.bhbits;
      while ((double)(hb.keys + R) > 0.5 * (double)((size_t)1 << bits)) ++bits;
      const int t2 = 1 - L.bht;
      base_table(c, L, t2, bits);
      const BaseState O = base_view(c, L, L.bht);
      L.bhbits = bits;
      const BaseState Bn = base_view(c, L, t2);
      hipLaunchKernelGGL(k_base_rehash, dim3(blocks_for(Ho, 256)), dim3(256), 0, c->stream, O, Ho, Bn);
      L.bht = t2;
      moved = true;
    }
    if (moved) {
      B = base_view(c, L);
      upload_graph_args();
    }
    // the node's successors
    const ExpandOut eo = outputs(R);
    if (R > 0) launch_expand(c, expand_tuned(c, 1, 0, eo, false), 1, 0, eo, false);
    hipLaunchKernelGGL(k_base_count, dim3(1), dim3(64), 0, c->stream, B, (const int*)c->s_valid.p,
                       (const int*)c->rec_cnt.p);
    if (q->early && host_pred(c, q->p, hb.ukey) && !host_pred(c, q->q, hb.ukey)) {
      launch_stop(c, L, &L.bctr.p->u, 1, 0, 1);
      const LiveCtr h = read_lc(c, L);
      if (h.err) {
        c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a stop set" : "invariant or transition table too large";
        st->capped = h.err == ERR_CAPPED;
        return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
      }
    }
    hipLaunchKernelGGL(k_base_step, dim3(1), dim3(64), 0, c->stream, c->dnp, B, (const int*)c->rec_cnt.p,
                       (const int*)c->s_valid.p, (const u64*)c->s_dk.p, (const int*)c->s_zone.p, (const int*)L.s_se.p,
                       (const int*)L.s_re.p, (const LiveCtr*)L.lc.p, (const Counters*)c->ctr.p, q->early ? 1 : 0);
    stage();
    read_dev(c, &hb, L.bctr.p, sizeof(BaseCtr), &R, c->scan_total.p, sizeof(R));
    ++steps;
    if (c->trace && (steps & 1023) == 0)
      std::fprintf(stderr, "base step %lld nodes=%d edges=%d keys=%d queued=%lld\n", steps, hb.nodes, hb.edges, hb.keys,
                   hb.qt - hb.qh);
  }
  if (hb.err) {
    c->err = hb.err == ERR_HASH_FULL ? "hash table full"
                                     : "a state has more transitions or invariant constraints than the device holds";
    return HORA_HIP_E_MEMORY;
  }
  {
    const Tally hc = read_counters(c);
    if (hc.err) { c->err = "a state has more transitions or invariant constraints than the device holds"; return HORA_HIP_E_MEMORY; }
  }
  if (hb.stop) {
    *early = 0;
    st->early = 1;
  }
  // the graph in CSR form
  const long N = hb.nodes;
  grow_keep_to(L.grow, (size_t)N + 1, c->stream);
  long long E = 0;
  if (N > 0) {
    L.tslot.ensure(N);
    hipLaunchKernelGGL(k_base_rows, dim3(blocks_for(N, 256)), dim3(256), 0, c->stream, B, (int)N, L.tslot.p);
    E = scan_exclusive(c, L.tslot.p, N, L.grow.p);
  }
  HIP_OK(hipMemcpyAsync(L.grow.p + N, &E, sizeof(long long), hipMemcpyHostToDevice, c->stream));
  grow_keep_to(L.etgt, (size_t)std::max<long long>(E, 1), c->stream);
  grow_keep_to(L.ese, (size_t)std::max<long long>(E, 1), c->stream);
  grow_keep_to(L.ere, (size_t)std::max<long long>(E, 1), c->stream);
  if (N > 0)
    hipLaunchKernelGGL(k_base_fill, dim3(blocks_for(N, 256)), dim3(256), 0, c->stream, B, (int)N,
                       (const long long*)L.grow.p, L.etgt.p, L.ese.p, L.ere.p);
  L.nodes = N;
  L.edges = E;
  L.args.recnt = nullptr;   // the backward pass reads the rows
  st->outer_nodes = N;
  st->fwd_edges = E;
  st->disc_keys = hb.keys;
  st->expanded_nodes = (long long)hb.expanded;
  st->generated = (long long)hb.generated;
  st->covered = (long long)hb.covered;
  st->enlarged = (long long)hb.enlarged;
  st->unions = (long long)hb.unions;
  if (c->trace)
    std::fprintf(stderr, "base construction: %lld steps, %ld nodes, %lld edges, %d keys\n", steps, N, E, hb.keys);
  return HORA_HIP_OK;
}

// The query's tables and the kernels' arguments for one liveness query or fixpoint solve.
int live_setup(hora_hip_ctx* c, LiveState& L, const hora_live_query* q) {
  // the full-capacity launches may ask for more than 64 KiB of LDS
  HIP_OK(hipFuncSetAttribute((const void*)k_stop<false>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsWorkgroupMax));
  HIP_OK(hipFuncSetAttribute((const void*)k_stop<true>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsWorkgroupMax));
  HIP_OK(hipFuncSetAttribute((const void*)k_gfp<false>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsWorkgroupMax));
  HIP_OK(hipFuncSetAttribute((const void*)k_gfp<true>, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsWorkgroupMax));
  if (c->dn.ncc > 0) { c->err = "open the context with a target without clock constraints for liveness"; return HORA_HIP_E_ARG; }
  if (q->kind < HORA_LIVE_LEADSTO || q->kind > HORA_LIVE_DEADLOCK) return HORA_HIP_E_ARG;
  const int d2 = c->d2;
  // ---- query tables
  {
    std::vector<int> blob;
    auto put_pred = [&](const hora_dpred& p, DevPred& out, size_t* o_off, size_t* o_at) {
      out.nterms = p.nterms;
      *o_off = blob.size();
      for (int t = 0; t <= p.nterms; ++t) blob.push_back(p.nterms > 0 ? p.term_off[t] : 0);
      *o_at = blob.size();
      const int na = p.nterms > 0 ? p.term_off[p.nterms] : 0;
      for (int a = 0; a < 4 * na; ++a) blob.push_back(p.atoms[a]);
      blob.push_back(0);
    };
    size_t po, pa, qo, qa;
    put_pred(q->p, L.args.p, &po, &pa);
    put_pred(q->q, L.args.q, &qo, &qa);
    L.pred.upload(blob, c->stream);
    L.args.p.off = L.pred.p + po; L.args.p.atoms = L.pred.p + pa;
    L.args.q.off = L.pred.p + qo; L.args.q.atoms = L.pred.p + qa;
  }
  L.args.kind = q->kind;
  L.args.restricted = q->orig_inv_off ? 1 : 0;
  if (L.args.restricted) {
    L.oinv_off.upload(std::vector<int>(q->orig_inv_off, q->orig_inv_off + c->tl + 1), c->stream);
    const int nc3 = 3 * q->orig_inv_off[c->tl];
    L.oinv_data.upload(nc3 > 0 ? std::vector<int>(q->orig_inv_data, q->orig_inv_data + nc3) : std::vector<int>(3, 0),
                       c->stream);
    L.args.oinv_off = L.oinv_off.p;
    L.args.oinv_data = L.oinv_data.p;
  }
  live_caps(d2, L.args);
  if (!L.argsp) {
    if (tl_arena) {
      L.argsp = (LiveArgs*)tl_arena->take(sizeof(LiveArgs));
      if (!L.argsp) throw ArenaFull("the query needs more device scratch than the static scratch holds");
      L.args_borrowed = true;
    } else {
      HIP_OK(hipMalloc(&L.argsp, sizeof(LiveArgs)));
    }
  }
  L.lc.ensure(1);
  HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
  return HORA_HIP_OK;
}

// The backward pass over the graph the kernels' arguments name (N nodes): closure of the query,
// stop sets, Jacobi rounds of relax to the greatest fixpoint, readout (src/liveness/
// fixpoint.cpp:48-965). `early` >= 0 is a verdict the construction already reached.
int solve_graph(hora_hip_ctx* c, LiveState& L, const hora_live_query* q, long long N, int early, int32_t* holds,
                hora_live_stats* st) {
  const int d2 = c->d2;
  L.in_s.ensure(std::max<long long>(N, 1));
  L.wne.ensure(std::max<long long>(N, 1));
  L.sidx.ensure(std::max<long long>(N, 1));
  HIP_OK(hipMemsetAsync(L.in_s.p, 0, sizeof(int) * std::max<long long>(N, 1), c->stream));
  HIP_OK(hipMemsetAsync(L.wne.p, 0, std::max<long long>(N, 1), c->stream));
  HIP_OK(hipMemsetAsync(L.sidx.p, 0xff, sizeof(int) * std::max<long long>(N, 1), c->stream));
  L.args.in_s = L.in_s.p;
  L.args.sidx = L.sidx.p;
  upload_args(c, L);
  if (early >= 0) {
    *holds = early;
  } else if (N == 0) {
    *holds = q->kind == HORA_LIVE_EG ? 0 : 1;
  } else if (q->kind == HORA_LIVE_DEADLOCK) {
    HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
    launch_stop(c, L, nullptr, (long)N, 0, 2);
    const LiveCtr h = read_lc(c, L);
    if (h.err) {
      c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a deadlock test" : "invariant or transition table too large";
      st->capped = h.err == ERR_CAPPED;
      return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
    }
    *holds = h.hit ? 0 : 1;
  } else {
    // ---- closure S
    HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
    L.slist.ensure(std::max<long long>(N, 1));
    hipLaunchKernelGGL(k_sseed, dim3(blocks_for(N, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, (long)N, L.in_s.p,
                       L.slist.p, L.lc.p);
    int from = 0, to = read_lc(c, L).ns;
    while (to > from) {
      hipLaunchKernelGGL(k_sstep, dim3(blocks_for(to - from, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, L.slist.p,
                         from, to, L.in_s.p, L.slist.p, L.lc.p);
      from = to;
      to = read_lc(c, L).ns;
    }
    const int ns = to;
    L.ns = ns;
    st->closure = ns;
    if (ns > 0) {
      hipLaunchKernelGGL(k_sindex, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, L.slist.p, ns, L.sidx.p);
      // ---- stop sets (into a pool: most nodes have none; grown and run again when it overflows)
      L.scnt.ensure(ns);
      L.soff.ensure(ns);
      const bool tight = pool_tight();
      // the pool's room in zones of THIS query's dimension: a context reused for a network of
      // another dimension keeps its buffer, whose capacity in zones then differs
      L.scap = L.sz.p ? std::min<long long>((long long)(L.sz.cap / d2), INT32_MAX) : 0;
      if (tight) {
        L.sz.ensure(1);
        L.scap = 1;
      } else if (L.scap < std::max<long long>(ns / 4, 4096)) {
        L.sz.ensure((size_t)std::max<long long>(ns / 4, 4096) * d2);
        L.scap = std::min<long long>((long long)(L.sz.cap / d2), INT32_MAX);
      }
      const auto ts0 = std::chrono::steady_clock::now();
      for (;;) {
        HIP_OK(hipMemsetAsync(&L.lc.p->spool, 0, 2 * sizeof(int), c->stream));   // spool, sfull
        launch_stop(c, L, L.slist.p, (long)ns, 0, 0);
        const LiveCtr h = read_lc(c, L);
        if (h.err) {
          c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a stop set" : "invariant or transition table too large";
          st->capped = h.err == ERR_CAPPED;
          return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
        }
        if (c->trace)
          std::fprintf(stderr, "stop sets: %d nodes, %d at full capacity, %d pool zones of %lld, %.3f ms\n", ns, h.rcnt,
                       h.spool, L.scap,
                       std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts0).count());
        if (!h.sfull) break;
        L.sz.ensure((size_t)h.spool * d2);
        L.scap = tight ? h.spool : std::min<long long>((long long)(L.sz.cap / d2), INT32_MAX);
      }
      // ---- fixpoint
      L.wcnt0.ensure(ns); L.wcnt1.ensure(ns); L.woff0.ensure(ns); L.woff1.ensure(ns);
      L.sel0.ensure(ns); L.sel1.ensure(ns); L.chg0.ensure(ns); L.chg1.ensure(ns);
      L.wcur = 0;
      // the seeds (one zone per node) and, mostly, room for two zones per node changed in a round
      L.wpool[0].ensure((size_t)(3 * (size_t)ns + ns / 2) * d2);
      L.wcap = tight ? ns : std::min<long long>((long long)(L.wpool[0].cap / d2), INT32_MAX);
      HIP_OK(hipMemsetAsync(L.wcnt1.p, 0, sizeof(int) * ns, c->stream));
      hipLaunchKernelGGL(k_gseed, dim3(ns), dim3(64), 0, c->stream, L.argsp, L.slist.p, ns, d2, L.wcnt0.p, L.woff0.p,
                         L.wpool[0].p, L.sel0.p, L.chg0.p);
      {
        const int top = ns;
        HIP_OK(hipMemcpyAsync(&L.lc.p->ptop, &top, sizeof(int), hipMemcpyHostToDevice, c->stream));
      }
      int* sel_in = L.sel0.p; int* sel_out = L.sel1.p;
      int* chg_in = L.chg0.p; int* chg_out = L.chg1.p;
      L.dlist.ensure(ns); L.rlist.ensure(ns);
      const GfpCaps KA = gfp_caps_small(c->dn), KB = gfp_caps_full(L.args, c->dn);
      // a middle tier between the small lists and the full ones (at most 32 KiB, so that several
      // wavefronts share a CU); not under an override of the small lists, whose runs exercise the
      // full ones, unless HORA_HIP_GFP_MID_CAPS=kw,kc,kp sets the middle tier too
      GfpCaps KM{std::min(4, KB.kw), std::min(16, KB.kc), std::min(8, KB.kp), KB.ic, KB.np, KB.ni};
      bool mid_on = !std::getenv("HORA_HIP_GFP_CAPS");
      if (const char* e = std::getenv("HORA_HIP_GFP_MID_CAPS")) {
        int kw = 0, kc = 0, kp = 0;
        if (std::sscanf(e, "%d,%d,%d", &kw, &kc, &kp) == 3 && kw <= KB.kw && kc <= KB.kc && kp <= KB.kp) {
          KM.kw = kw;
          KM.kc = kc;
          KM.kp = kp;
          mid_on = true;
        }
      }
      const bool gfp_mid = mid_on && KM.kw >= KA.kw && KM.kc >= KA.kc && KM.kp >= KA.kp &&
                           (KM.kc < KB.kc || KM.kp < KB.kp || KM.kw < KB.kw) && KM.ints(d2) <= 64 * 1024 / 4 / 2;
      int round = 0;
      long long retried = 0;
      int compactions = 0;
      for (;;) {
        const auto tr0 = std::chrono::steady_clock::now();
        // changed (its stripes); dcnt and rcnt are adjacent in LiveCtr
        HIP_OK(hipMemsetAsync(&L.lc.p->stripe[2][0][0], 0, sizeof(L.lc.p->stripe[2]), c->stream));
        HIP_OK(hipMemsetAsync(&L.lc.p->dcnt, 0, 2 * sizeof(int), c->stream));
        auto banks = [&]() {
          return GfpBanks{L.wcnt0.p, L.wcnt1.p, L.woff0.p, L.woff1.p, L.wpool[L.wcur].p, (int)L.wcap,
                          sel_in, sel_out, chg_in, chg_out, L.scnt.p, L.soff.p, L.sz.p};
        };
        hipLaunchKernelGGL(k_gdirty, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, L.argsp, L.slist.p, ns,
                           round == 0 ? 1 : 0, banks(), L.dlist.p, L.lc.p);
        const LiveCtr h0 = read_lc(c, L);
        const int nd = h0.dcnt;
        // most values hold one or two zones: room for two per visited node, else compact first
        if (!tight && (long long)h0.ptop + 2LL * nd > L.wcap) {
          compact_values(c, L, ns, sel_in, 2LL * nd + ns / 2);
          ++compactions;
        }
        LiveCtr h{};
        for (;;) {
          if (nd > 0)
            hipLaunchKernelGGL(k_gfp<false>, dim3(nd), dim3(64), KA.ints(d2) * sizeof(int), c->stream, c->dnp, L.argsp,
                               L.slist.p, L.dlist.p, KA, banks(), L.rlist.p, L.lc.p, (int*)nullptr);
          h = read_lc(c, L);
          const int* full_list = L.rlist.p;
          if (h.rcnt > 0 && !h.err && !h.pfull && gfp_mid) {
            const int nm = h.rcnt;
            L.rlist2.ensure(nm);
            HIP_OK(hipMemsetAsync(&L.lc.p->rcnt, 0, sizeof(int), c->stream));
            hipLaunchKernelGGL(k_gfp<false>, dim3(nm), dim3(64), KM.ints(d2) * sizeof(int), c->stream, c->dnp, L.argsp,
                               L.slist.p, L.rlist.p, KM, banks(), L.rlist2.p, L.lc.p, (int*)nullptr);
            h = read_lc(c, L);
            full_list = L.rlist2.p;
          }
          long long hbm_round = 0;
          if (h.rcnt > 0 && !h.err && !h.pfull) {
            // the full LDS lists; the nodes that outgrow them rerun with their lists in device memory
            const int nf = h.rcnt;
            L.rlist3.ensure(nf);
            HIP_OK(hipMemsetAsync(&L.lc.p->rcnt, 0, sizeof(int), c->stream));
            hipLaunchKernelGGL(k_gfp<true>, dim3(nf), dim3(64), KB.ints(d2) * sizeof(int), c->stream, c->dnp, L.argsp,
                               L.slist.p, full_list, KB, banks(), fed_hbm_on() ? L.rlist3.p : (int*)nullptr,
                               L.lc.p, (int*)nullptr);
            h = read_lc(c, L);
            if (h.rcnt > 0 && !h.err && !h.pfull) {
              const GfpCaps KH{hh_fed::kFedCap, hh_fed::kFedCap, hh_fed::kFedCap, KB.ic, KB.np, KB.ni};
              const size_t per = KH.hbm_ints(d2);
              const long batch = fed_hbm_batch(h.rcnt, per);
              L.fedhbm.ensure(per * (size_t)batch);
              hbm_round = h.rcnt;
              L.hbm_nodes += h.rcnt;
              const GfpBanks bk = banks();
              for (long b0 = 0; b0 < h.rcnt; b0 += batch)
                HIP_OK(hh_fed::launch_gfp_hbm(c->dnp, L.argsp, L.slist.p, L.rlist3.p + b0,
                                              (int)std::min<long>(batch, h.rcnt - b0), &KH, &bk, L.lc.p, L.fedhbm.p,
                                              c->stream));
              h = read_lc(c, L);
            }
          }
          if (!h.pfull || h.err) break;
          // the pool ran out: drop the round's writes, compact with room for every value at its
          // largest, and run the round again from the same values
          LiveCtr r = h;
          std::memcpy(r.stripe[0], h0.stripe[0], sizeof(r.stripe[0]));
          std::memcpy(r.stripe[1], h0.stripe[1], sizeof(r.stripe[1]));
          std::memset(r.stripe[2], 0, sizeof(r.stripe[2]));
          r.rcnt = 0;
          r.pfull = 0;
          HIP_OK(hipMemcpyAsync(L.lc.p, &r, sizeof(LiveCtr), hipMemcpyHostToDevice, c->stream));
          compact_values(c, L, ns, sel_in,
                         (long long)L.args.kw * nd + (long long)(hh_fed::kFedCap - L.args.kw) * hbm_round +
                             (tight ? 0 : ns / 2));
          ++compactions;
        }
        retried += h.rcnt;
        if (c->trace)
          std::fprintf(stderr, "gfp round %d: %d visited, %d at full capacity, %d values changed, pre_calls so far %llu,"
                       " pool %d of %lld zones, %.3f ms\n", round, h.dcnt, h.rcnt, h.changed,
                       (unsigned long long)h.pre_calls, h.ptop, L.wcap,
                       std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr0).count());
        ++round;
        if (h.err) {
          st->capped = h.err == ERR_CAPPED;
          c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in the fixpoint" : "invariant table too large";
          return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
        }
        std::swap(sel_in, sel_out);
        std::swap(chg_in, chg_out);
        if (h.changed == 0) break;
      }
      st->iters = round;
      if (c->trace)
        std::fprintf(stderr, "gfp: %d rounds, %lld node visits at full capacity (small lists %d,%d,%d,%d), %d pool"
                     " compactions\n", round, retried, KA.kw, KA.kc, KA.kp, KA.ic, compactions);
      hipLaunchKernelGGL(k_gverdict, dim3(blocks_for(ns, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, L.slist.p, ns,
                         sel_in, L.wcnt0.p, L.wcnt1.p, L.wne.p, L.lc.p);
      const LiveCtr h = read_lc(c, L);
      st->nonempty = h.nonempty;
      st->fed_max = h.fed_max;
      st->pre_calls = (long long)h.pre_calls;
      st->products = (long long)h.products;
      *holds = h.violated ? 0 : 1;
      if (q->kind == HORA_LIVE_EG) *holds = h.violated ? 1 : 0;
    } else {
      *holds = q->kind == HORA_LIVE_EG ? 0 : 1;
    }
  }
  return HORA_HIP_OK;
}

int run_live(hora_hip_ctx* c, const hora_live_query* q, int32_t* holds, hora_live_stats* st) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  *holds = 0;
  {
    LiveArgs probe{};
    if (c->d > kLiveMaxDim || !live_caps(c->d2, probe)) {
      c->err = "liveness on the device covers zone dimension up to 32";
      return HORA_HIP_E_SCOPE;
    }
  }
  if (!c->live) c->live = new LiveState();
  LiveState& L = *c->live;
  HbmCount hbm_count(L, st);
  L.graph_valid = false;
  const int d2 = c->d2;
  {
    const int rc = live_setup(c, L, q);
    if (rc != HORA_HIP_OK) return rc;
  }

  // One construction of the graph under a_M covers (am) or a_LU covers. Returns HORA_HIP_OK or an
  // error code; `early` receives the verdict of an early witness, or -1.
  int early = -1;
  long long rows_done = -1;
  auto construct = [&](bool am) -> int {
    const int alu = am ? 0 : 1;
    early = -1;
    rows_done = -1;
    L.args.recnt = nullptr;
    L.args.whit = nullptr;
    HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));   // a seed test may have counted hits
    {
      const int32_t seeds = st->seed_keys;
      std::memset(st, 0, sizeof(*st));
      st->seed_keys = seeds;
    }
    c->dn.lu_plus = 0;
    c->dn.znar = 0;
    HIP_OK(hipMemcpyAsync(c->dnp, &c->dn, sizeof(DevNet), hipMemcpyHostToDevice, c->stream));
    alloc_table(c, std::max(16, c->hbits));
    c->ctr.ensure(1);
    HIP_OK(hipMemsetAsync(c->ctr.p, 0, sizeof(Counters), c->stream));
    {
      const int ft = 0x7fffffff;
      HIP_OK(hipMemcpyAsync(&c->ctr.p->first_target, &ft, sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    // graph storage
    size_t ncap = 1 << 16, ecap = 1 << 18;
    L.gdk.ensure(ncap); L.grow.ensure(ncap + 1); L.expanded.ensure(ncap);
    c->nodes.ensure(1 << 20);
    u64 ncount = 0;   // node-segment entries handed out so far
    c->zarena.ensure(ncap * d2);
    L.etgt.ensure(ecap); L.ese.ensure(ecap); L.ere.ensure(ecap);
    HIP_OK(hipMemsetAsync(L.expanded.p, 0, L.expanded.cap, c->stream));
    L.nodes = 0;
    L.edges = 0;
    L.ns = 0;
  
    const bool eg = q->kind == HORA_LIVE_EG;
    const u64 dk0 = initial_key(c);
    c->fr_dk.ensure(1);
    c->fr_zid.ensure(1);
    HIP_OK(hipMemcpyAsync(c->fr_dk.p, &dk0, sizeof(u64), hipMemcpyHostToDevice, c->stream));
  
    long F = 1;
    long long fbase = 0;   // rows [0, rows_done] written
    bool init = true;
    Tally prev;
    for (;;) {
      sync_graph_args(L);
      upload_args(c, L);
      long R = 1;
      c->rec_cnt.ensure(F);
      c->rec_off.ensure(F);
      c->pos.ensure(F + 1);
      if (!init) {
        ++st->levels;
        hipLaunchKernelGGL(k_count, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, c->fr_dk.p, (int)F,
                           c->rec_cnt.p, c->ctr.p);
        hipLaunchKernelGGL(k_gfilter, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, eg ? 1 : 0,
                           c->fr_dk.p, (int)F, c->rec_cnt.p, L.expanded.p, (long)fbase);
        R = scan_exclusive(c, c->rec_cnt.p, F, c->pos.p);
        hipLaunchKernelGGL(k_offsets_to_int, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->pos.p, F, c->rec_off.p);
      } else {
        HIP_OK(hipMemsetAsync(c->rec_off.p, 0, sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(c->pos.p, 0, sizeof(long long), c->stream));   // k_expand's slot offset
      }
      c->s_valid.ensure(R);
      c->s_tgt.ensure(R);
      c->s_dk.ensure(R);
      c->s_zone.ensure((size_t)R * d2);
      L.s_se.ensure(R);
      L.s_re.ensure(R);
      ExpandOut eo{c->s_valid.p, c->s_dk.p, c->s_zone.p, c->s_tgt.p, L.s_se.p, L.s_re.p};
      if (R > 0) launch_expand(c, expand_tuned(c, F, 0, eo, init), F, 0, eo, init);
      c->pos.ensure(R + 1);
      const long G = R > 0 ? scan_exclusive(c, c->s_valid.p, R, c->pos.p) : 0;
      if (!init) {
        st->expanded_nodes += F;
        st->generated += G;
        grow_keep_to(L.grow, (size_t)(fbase + F + 1), c->stream);
        hipLaunchKernelGGL(k_grows, dim3(blocks_for(F + 1, 256)), dim3(256), 0, c->stream, c->rec_off.p, c->pos.p, R,
                           (int)F, (long)fbase, L.edges, (long long)G, L.grow.p);
        rows_done = fbase + F;
      }
      {
        const Tally hc = read_counters(c);
        if (hc.err) { c->err = "a state has more transitions or invariant constraints than the device holds"; return HORA_HIP_E_MEMORY; }
      }
      long F2 = 0;
      if (G > 0) {
        c->sj.ensure(G);
        hipLaunchKernelGGL(k_compact_valid, dim3(blocks_for(R, 256)), dim3(256), 0, c->stream, c->s_valid.p, c->pos.p, R, c->sj.p,
                           (Counters*)nullptr);
        {
          const Tally hc = read_counters(c);
          while ((double)(hc[C_KEYS] + (u64)G) > 0.6 * (double)c->H) grow_table(c);
        }
        Passed P = passed_view(c);
        c->gbk.ensure(G);
        c->gticket.ensure(G);
        c->touched.ensure(G);
        c->covered.ensure(G);
        L.cov_old.ensure(G);
        L.tslot.ensure(G);
        L.fnew.ensure(G);
        HIP_OK(hipMemsetAsync(L.fnew.p, 0, (size_t)G * sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(L.tslot.p, 0xff, (size_t)G * sizeof(int), c->stream));
        HIP_OK(hipMemsetAsync(&c->ctr.p->ntouched, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_find, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->sj.p, c->s_dk.p, G, c->gbk.p, c->ctr.p,
                           (int*)nullptr, (int*)nullptr);
        HIP_OK(hipEventRecord(c->ev[1], c->stream));
        hipLaunchKernelGGL(k_gprobe, dim3(blocks_for(G, 4)), dim3(64), 0, c->stream, c->dnp, P, c->gbk.p, c->sj.p, c->s_dk.p,
                           c->s_zone.p, G, c->zarena.p, c->covered.p, L.cov_old.p, c->ctr.p, alu, 0);
        hipLaunchKernelGGL(k_ticket, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->gbk.p, c->covered.p, G,
                           c->gticket.p, c->touched.p, c->ctr.p);
        int T = 0;
        read_dev(c, &T, &c->ctr.p->ntouched, sizeof(int));
        if (T > 0) {
          c->caps.ensure(T);
          c->segs.ensure(T);
          hipLaunchKernelGGL(k_touched_caps, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                             c->caps.p, c->segs.p, (Counters*)nullptr, (const int*)nullptr);
          c->pos2.ensure(T + 1);
          const long long Wl = scan_exclusive(c, c->caps.p, T, c->pos2.p);
          hipLaunchKernelGGL(k_gsegs, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T, c->segs.p);
          L.segpos.ensure(T + 1);
          const long long Sseg = scan_exclusive(c, c->segs.p, T, L.segpos.p);
          if (ncount + (u64)Sseg > c->nodes.cap) c->nodes.grow_keep((size_t)(ncount + (u64)Sseg) * 2, c->stream);
          P = passed_view(c);
          hipLaunchKernelGGL(k_touched_offsets, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                             c->pos2.p, L.segpos.p, ncount);
          ncount += (u64)Sseg;
          c->wlist.ensure(Wl + 1);
          hipLaunchKernelGGL(k_scatter, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, P, c->gbk.p, c->covered.p,
                             c->gticket.p, G, c->wlist.p);
          const size_t lds_m = (d2 + 2 * HORA_MAX_DIM + HORA_MAX_PROCS + kSortLds) * sizeof(int);
          HIP_OK(hipEventRecord(c->ev[2], c->stream));
          c->klight.ensure(T);
          c->kheavy.ensure(T);
          HIP_OK(hipMemsetAsync(&c->ctr.p->nlight, 0, 2 * sizeof(int), c->stream));
          hipLaunchKernelGGL(k_gsplit, dim3(blocks_for(T, 256)), dim3(256), 0, c->stream, P, c->touched.p, T,
                             c->klight.p, c->kheavy.p, c->ctr.p);
          int nlh[2] = {0, 0};
          read_dev(c, nlh, &c->ctr.p->nlight, sizeof(nlh));
          if (nlh[0] > 0)
            hipLaunchKernelGGL(k_gmerge, dim3(nlh[0]), dim3(64), lds_m, c->stream, c->dnp, P, c->klight.p, nlh[0],
                               c->wlist.p, c->sj.p, c->s_dk.p, c->s_zone.p, L.tslot.p, L.fnew.p, c->ctr.p, alu);
          if (nlh[1] > 0) {
            L.wkey.ensure(Wl + 1);
            HIP_OK(hipMemsetAsync(L.wkey.p, 0xff, (size_t)(Wl + 1) * sizeof(long long), c->stream));
            hipLaunchKernelGGL(k_gsort, dim3(nlh[1]), dim3(64), kSortLds * sizeof(int), c->stream, P, c->kheavy.p,
                               nlh[1], c->wlist.p, L.wkey.p);
            hipLaunchKernelGGL(k_gpair, dim3(Wl), dim3(64), 0, c->stream, c->dnp, P, L.wkey.p, (long)Wl, c->wlist.p,
                               c->sj.p, c->s_dk.p, c->s_zone.p, L.tslot.p, L.fnew.p, c->ctr.p, alu);
            hipLaunchKernelGGL(k_gcompact, dim3(nlh[1]), dim3(64), 0, c->stream, P, c->kheavy.p, nlh[1], c->wlist.p,
                               L.fnew.p);
            hipLaunchKernelGGL(k_gfirst, dim3(G), dim3(64), 0, c->stream, c->dnp, P, c->gbk.p, (long)G, c->wlist.p,
                               c->sj.p, c->s_dk.p, c->s_zone.p, L.tslot.p, c->ctr.p, alu);
          }
        }
        c->pos.ensure(G + 1);
        F2 = scan_exclusive(c, L.fnew.p, G, c->pos.p);
        // capacities
        const long long nbase = L.nodes;
        const size_t nneed = (size_t)(nbase + F2 + 1);
        if (nneed > L.gdk.cap) {
          const size_t nc2 = std::max(nneed, (size_t)L.gdk.cap * 2);
          grow_keep_to(L.gdk, nc2, c->stream);
          grow_keep_to(L.grow, nc2 + 1, c->stream);
          const size_t old = L.expanded.cap;
          grow_keep_to(L.expanded, nc2, c->stream);
          HIP_OK(hipMemsetAsync(L.expanded.p + old, 0, L.expanded.cap - old, c->stream));
        }
        if (nneed * d2 > c->zarena.cap) c->zarena.grow_keep(std::max(nneed * d2, c->zarena.cap * 2), c->stream);
        if (!init && (size_t)(L.edges + G) > L.etgt.cap) {
          const size_t ec2 = std::max((size_t)(L.edges + G), L.etgt.cap * 2);
          grow_keep_to(L.etgt, ec2, c->stream);
          grow_keep_to(L.ese, ec2, c->stream);
          grow_keep_to(L.ere, ec2, c->stream);
        }
        L.newlist.ensure(std::max<long>(F2, 1));
        HIP_OK(hipEventRecord(c->ev[3], c->stream));
        hipLaunchKernelGGL(k_gassign, dim3(blocks_for(G, 256)), dim3(256), 0, c->stream, G, nbase, L.edges, init ? 0 : 1,
                           L.cov_old.p, L.tslot.p, L.fnew.p, c->pos.p, c->sj.p, L.s_se.p, L.s_re.p, L.etgt.p, L.ese.p,
                           L.ere.p, L.newlist.p);
        c->fr_dk2.ensure(std::max<long>(F2, 1));
        c->fr_zid2.ensure(std::max<long>(F2, 1));
        if (F2 > 0)
          hipLaunchKernelGGL(k_gcopy, dim3(F2), dim3(64), 0, c->stream, (long)F2, nbase, L.newlist.p, c->sj.p, c->s_dk.p,
                             c->s_zone.p, d2, c->zarena.p, L.gdk.p, c->fr_dk2.p, c->fr_zid2.p);
        if (T > 0)
          hipLaunchKernelGGL(k_glink, dim3(T), dim3(64), 0, c->stream, P, c->touched.p, T, c->wlist.p, c->pos.p, nbase);
        HIP_OK(hipEventRecord(c->ev[4], c->stream));
        if (c->trace) {
          HIP_OK(hipStreamSynchronize(c->stream));
          const Tally ht = read_counters(c);
          std::fprintf(stderr, "live level %lld F=%ld R=%ld G=%ld T=%d F2=%ld | probe %.3f (covered %llu, reads %llu) merge %.3f assign+link %.3f ms\n",
                       (long long)st->levels, F, R, G, T, F2, ev_ms(c->ev[1], c->ev[2]), ht[C_PSUB], ht[C_PREADS],
                       T > 0 ? ev_ms(c->ev[2], c->ev[3]) : 0.0f, ev_ms(c->ev[3], c->ev[4]));
        }
        if (!init) {
          L.edges += G;
          st->fwd_edges = L.edges;
        }
        L.nodes += F2;
      }
      // early witness on the level just expanded: a p /\ !q node with a non-empty stop set refutes
      // p --> q (leadsto_witness, construction.cpp:786-813); a psi node with one ends a maximal
      // psi-run, so E[] psi holds (online_eg_terminal_witness, construction.cpp:748-767); a
      // deadlocked valuation refutes A[] not deadlock
      if (!init && q->early) {
        sync_graph_args(L);
        L.args.gzone = c->zarena.p;
        upload_args(c, L);
        if (q->kind == HORA_LIVE_DEADLOCK) {
          launch_stop(c, L, nullptr, (long)F, (long long)fbase, 2);
        } else {
          L.wlist.ensure(F);
          HIP_OK(hipMemsetAsync(&L.lc.p->wcnt, 0, sizeof(int), c->stream));
          hipLaunchKernelGGL(k_wfilter, dim3(blocks_for(F, 256)), dim3(256), 0, c->stream, c->dnp, L.argsp, (long)F,
                             (long long)fbase, L.wlist.p, L.lc.p);
          const int nw = read_lc(c, L).wcnt;
          launch_stop(c, L, L.wlist.p, (long)nw, 0, 1);
        }
        const LiveCtr h = read_lc(c, L);
        if (h.err) {
          c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a stop set" : "invariant or transition table too large";
          st->capped = h.err == ERR_CAPPED;
          return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
        }
        if (h.hit) { early = q->kind == HORA_LIVE_EG ? 1 : 0; st->early = 1; }
      }
      {
        const Tally hc = read_counters(c);
        if (hc.err) { c->err = hc.err == ERR_HASH_FULL ? "hash table full" : "device capacity exceeded"; return HORA_HIP_E_MEMORY; }
        st->covered = (long long)(hc[C_SUB] + hc[C_PSUB]);
        st->chain_reads = (long long)(hc[C_READS] + hc[C_PREADS]);
        st->disc_keys = (long long)hc[C_KEYS];
        prev = hc;
      }
      if (early >= 0) break;
      std::swap(c->fr_dk.p, c->fr_dk2.p); std::swap(c->fr_dk.cap, c->fr_dk2.cap);
      std::swap(c->fr_zid.p, c->fr_zid2.p); std::swap(c->fr_zid.cap, c->fr_zid2.cap);
      fbase = init ? 0 : fbase + F;
      F = F2;
      init = false;
      if (F == 0) break;
    }
    // rows of nodes never expanded
    if (rows_done < L.nodes) {
      grow_keep_to(L.grow, (size_t)(L.nodes + 1), c->stream);
      const long from = (long)(rows_done + 1);
      hipLaunchKernelGGL(k_gfill_rows, dim3(blocks_for(L.nodes - from + 1, 256)), dim3(256), 0, c->stream, L.grow.p, from,
                         (long)L.nodes, L.edges);
    }
    st->outer_nodes = L.nodes;
    st->fwd_edges = L.edges;
    if (L.nodes > 0) {   // E[] psi: nodes outside psi were stored, not expanded
      HIP_OK(hipMemsetAsync(&L.lc.p->nexp, 0, sizeof(int), c->stream));
      hipLaunchKernelGGL(k_ocount, dim3(blocks_for((long)L.nodes, 256)), dim3(256), 0, c->stream, L.expanded.p,
                         (long)L.nodes, L.lc.p);
      st->expanded_nodes = read_lc(c, L).nexp;
    }
    return HORA_HIP_OK;
  };
  const bool auto_route = q->cover != HORA_COVER_AM;
  // p --> q where D is provably a_LU (structurally_empty_d, src/liveness/internal.cpp:59-103): the
  // base construction (src/liveness/api.cpp:37-117); a time-open network always, a network with one
  // possible discrete state unless that state can seed a deadlock atom
  bool base = false;
  if (q->cover == HORA_COVER_AUTO && q->kind == HORA_LIVE_LEADSTO && c->base_route) {
    bool seeds0 = false;
    if (c->base_route == 2) {
      const int rc = initial_key_seeds(c, L, st, &seeds0);
      if (rc != HORA_HIP_OK) return rc;
    }
    base = !seeds0;
  }
  {
    const int rc = base ? construct_base(c, L, q, st, &early)
                        : auto_route ? construct_online(c, L, q, st, &early) : construct(true);
    if (rc != HORA_HIP_OK) return rc;
  }
  if (auto_route && early < 0 && L.nodes > 0) {
    // the a_LU graph stands unless one of its discrete states can seed a deadlock atom
    sync_graph_args(L);
    L.args.gzone = c->zarena.p;
    upload_args(c, L);
    HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
    if (base) {
      const long H = 1L << L.bhbits;
      L.heads.ensure((size_t)H);
      hipLaunchKernelGGL(k_base_heads, dim3(blocks_for(H, 256)), dim3(256), 0, c->stream, base_view(c, L), H,
                         L.heads.p, L.lc.p);
    } else {
      L.heads.ensure((size_t)std::max<u64>(c->H, 1));
      Passed P = passed_view(c);
      hipLaunchKernelGGL(k_heads, dim3(blocks_for((long)c->H, 256)), dim3(256), 0, c->stream, P, (long)c->H, L.heads.p,
                         L.lc.p);
    }
    const int nh = read_lc(c, L).ns;
    int seeds = 0;
    if (nh > 0) {
      HIP_OK(hipMemsetAsync(L.lc.p, 0, sizeof(LiveCtr), c->stream));
      launch_stop(c, L, L.heads.p, (long)nh, 0, 3);
      const LiveCtr h = read_lc(c, L);
      if (h.err) {
        c->err = h.err == ERR_CAPPED ? "federation capacity exceeded in a seed test" : "invariant or transition table too large";
        st->capped = h.err == ERR_CAPPED;
        return h.err == ERR_CAPPED ? HORA_HIP_E_CAPPED : HORA_HIP_E_MEMORY;
      }
      seeds = h.hit;
    }
    st->seed_keys = seeds;
    if (seeds > 0) {
      const int rc = construct(true);
      if (rc != HORA_HIP_OK) return rc;
      st->route = 1;
    }
  } else {
    st->route = auto_route ? 0 : 1;
  }
  const auto t1 = std::chrono::steady_clock::now();
  sync_graph_args(L);
  L.args.gzone = c->zarena.p;
  {
    const int rc = solve_graph(c, L, q, L.nodes, early, holds, st);
    if (rc != HORA_HIP_OK) return rc;
  }
  L.graph_valid = true;
  const auto t2 = std::chrono::steady_clock::now();
  st->build_seconds = std::chrono::duration<double>(t1 - t0).count();
  st->gfp_seconds = std::chrono::duration<double>(t2 - t1).count();
  st->seconds = std::chrono::duration<double>(t2 - t0).count();
  return HORA_HIP_OK;
}

// hora_hip_gfp_solve: the backward pass over a graph the caller built, in device memory.
int gfp_solve(hora_hip_ctx* c, LiveState& L, const hora_gfp_graph* g, const hora_live_query* q, int32_t* holds,
              hora_live_stats* st, int8_t* w_nonempty) {
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(st, 0, sizeof(*st));
  *holds = 0;
  if (!g || g->nodes < 0 || g->edges < 0 || g->nodes > INT32_MAX ||
      (g->nodes > 0 && (!g->dk || !g->zones || !g->row_off || !g->expanded)) ||
      (g->edges > 0 && (!g->etgt || !g->ese || !g->ere))) {
    c->err = "graph arrays missing or sizes out of range";
    return HORA_HIP_E_ARG;
  }
  {
    LiveArgs probe{};
    if (c->d > kLiveMaxDim || !live_caps(c->d2, probe)) {
      c->err = "liveness on the device covers zone dimension up to 32";
      return HORA_HIP_E_SCOPE;
    }
  }
  HbmCount hbm_count(L, st);
  {
    const int rc = live_setup(c, L, q);
    if (rc != HORA_HIP_OK) return rc;
  }
  L.graph_valid = false;   // live_graph describes the last hora_hip_live's own graph
  L.args.gdk = g->dk;
  L.args.gzone = g->zones;
  L.args.grow = (const long long*)g->row_off;
  L.args.etgt = g->etgt;
  L.args.ese = g->ese;
  L.args.ere = g->ere;
  L.args.expanded = (const signed char*)g->expanded;
  L.args.recnt = nullptr;
  L.args.whit = nullptr;
  {
    const int rc = solve_graph(c, L, q, g->nodes, -1, holds, st);
    if (rc != HORA_HIP_OK) return rc;
  }
  if (w_nonempty && g->nodes > 0)
    HIP_OK(hipMemcpyAsync(w_nonempty, L.wne.p, (size_t)g->nodes, hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  st->outer_nodes = g->nodes;
  st->fwd_edges = g->edges;
  st->gfp_seconds = st->seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return HORA_HIP_OK;
}

int live_graph(const hora_hip_ctx* c, hora_live_graph* g) {
  const LiveState* L = c->live;
  g->d = c->d;
  if (L && !L->graph_valid) {   // the last query stopped before its graph was complete
    g->nodes = 0;
    g->edges = 0;
    return HORA_HIP_E_ARG;
  }
  g->nodes = L ? L->nodes : 0;
  g->edges = L ? L->edges : 0;
  if (!L) return HORA_HIP_OK;
  auto cp = [](void* dst, const void* src, size_t bytes) {
    if (dst && bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  };
  const size_t N = (size_t)L->nodes, E = (size_t)L->edges;
  cp(g->dk, L->gdk.p, N * sizeof(u64));
  cp(g->zones, c->zarena.p, N * c->d2 * sizeof(int));
  cp(g->row_off, L->grow.p, (N + 1) * sizeof(long long));
  cp(g->etgt, L->etgt.p, E * sizeof(int));
  cp(g->ese, L->ese.p, E * sizeof(int));
  cp(g->ere, L->ere.p, E * sizeof(int));
  cp(g->expanded, L->expanded.p, N);
  if (g->in_s) {
    std::vector<int> tmp(N);
    cp(tmp.data(), L->in_s.p, N * sizeof(int));
    for (size_t v = 0; v < N; ++v) g->in_s[v] = (int8_t)tmp[v];
  }
  cp(g->w_nonempty, L->wne.p, N);
  return HORA_HIP_OK;
}

}  // namespace
