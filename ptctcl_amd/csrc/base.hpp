// base.hpp — the reference's base construction of the liveness graph on the device: for `p --> q`
// on a network where D is provably a_LU (structurally_empty_d, src/liveness/internal.cpp:59-103),
// LivenessEngine::base_forward (src/liveness/api.cpp:37-117) with expand and intern
// (src/liveness/construction.cpp:863-1096). Part of the hora_hip.hip translation unit, included by
// liveness.hpp.
//
// A FIFO of nodes to expand; each successor of the node being expanded, in tagged_successors
// order, goes to the OLDEST node of its discrete key that a_LU-covers it; else the oldest node it
// a_LU-covers takes its zone in place; else the oldest node whose union with it is one zone takes
// the hull; else it is appended. A node rewritten by a successor of another node loses its edges
// and is expanded again (a node rewritten by its own successor keeps them). After a node's
// successors, the targets of its edges that are neither expanded nor queued join the FIFO.
//
// Every step reads and rewrites what the steps before it stored, so the construction runs one node
// at a time: the node's expansion is k_count and k_expand on a one-state frontier, and k_base_step
// (one wavefront) interns its successors in order, queues the new work and pops the next node. The
// host launches the steps and reads back one small record per node (whether to go on, and the
// sizes it must grow the buffers to).
//
// A key's nodes are an array of node ids in creation order, so the three searches of an intern
// test 64 nodes at once, one per lane (the first hit in creation order is the lowest lane of the
// first block with a hit), instead of walking a list one dependent load at a time. The arrays
// live in one pool, each doubling its capacity when full (the old array is left behind).
#pragma once

namespace {

// The construction's state in device memory. Nodes: key, zone (the arena, d*d ints each), flags,
// the head and tail of the node's edge list. Edges: target, transition, the next edge of the same
// node. A hash table of the keys (open addressing, linear probing) holds each key's node array
// (offset into the pool, length, capacity). The FIFO is an array read from qh and appended at qt.
struct BaseCtr {
  int u;                      // node to expand next, or -1: the construction is over
  int nodes, edges, keys;
  u64 ukey;                   // node u's discrete key (the host decides from it whether to test u)
  long long qh, qt;
  long long ptop;             // pool ints handed out
  int stop;                   // an early witness ended the construction
  int err;                    // ERR_HASH_FULL, or the expansion's error
  unsigned long long expanded, generated, covered, enlarged, unions;
};

struct BaseState {
  u64* nkey;                  // [N]
  int* zone;                  // [N * d2]
  signed char* expanded;      // [N]
  signed char* inq;           // [N]
  int* ehead;                 // [N]
  int* etail;                 // [N]
  int* recnt;                 // [N] successors of the node's last expansion (the stop-set test reads it)
  int* enext;                 // [E]
  int* etgt;                  // [E]
  int* ese;                   // [E]
  int* ere;                   // [E]
  u64* hkey;                  // [H]
  long long* kptr;            // [H] the key's node array: offset into the pool
  int* klen;                  // [H] ... nodes in it
  int* kcap;                  // [H] ... its capacity
  int* pool;                  // node arrays
  int* queue;                 // [Q]
  int hbits;
  BaseCtr* bc;
};

// the slot of key k (inserted when absent), or -1 when the table is full; lane-uniform
__device__ __forceinline__ int base_slot(const BaseState& B, u64 k, bool* fresh) {
  const u64 H = 1ull << B.hbits;
  u64 h = key_mix(k) >> (64 - B.hbits);
  *fresh = false;
  for (u64 n = 0; n < H; ++n, h = (h + 1) & (H - 1)) {
    const u64 cur = B.hkey[h];
    if (cur == k) return (int)h;
    if (cur == kEmptyKey) {
      *fresh = true;
      return (int)h;
    }
  }
  return -1;
}

// Node v appended to the array of key slot h (lane 0; the caller orders the writes). A full array
// moves to twice its capacity at the pool's top.
__device__ __forceinline__ void base_append(const BaseState& B, int h, int v) {
  BaseCtr& C = *B.bc;
  const int n = B.klen[h];
  if (n == B.kcap[h]) {
    const int cap = max(8, 2 * n);
    const long long off = C.ptop;
    C.ptop += cap;
    for (int k = 0; k < n; ++k) B.pool[off + k] = B.pool[B.kptr[h] + k];
    B.kptr[h] = off;
    B.kcap[h] = cap;
  }
  B.pool[B.kptr[h] + n] = v;
  B.klen[h] = n + 1;
}

// The first node to expand: the initial zone (successor slot 0 of the initial expansion) becomes
// node 0 and is popped.
__global__ __launch_bounds__(64) void k_base_init(BaseState B, const int* __restrict__ svalid,
                                                  const u64* __restrict__ sdk, const int* __restrict__ szone,
                                                  int d2) {
  const int l = threadIdx.x;
  BaseCtr& C = *B.bc;
  if (!svalid[0]) {
    if (l == 0) C.u = -1;
    return;
  }
  for (int e = l; e < d2; e += kWave) B.zone[e] = szone[e];
  if (l == 0) {
    bool fresh = false;
    const int h = base_slot(B, sdk[0], &fresh);
    B.hkey[h] = sdk[0];
    B.klen[h] = 0;
    B.kcap[h] = 0;
    base_append(B, h, 0);
    B.nkey[0] = sdk[0];
    B.expanded[0] = 0;
    B.inq[0] = 0;
    B.ehead[0] = B.etail[0] = -1;
    C.u = 0;
    C.ukey = sdk[0];
    C.nodes = 1;
    C.keys = 1;
    C.edges = 0;
    C.qh = C.qt = 0;
  }
}

// The frontier of the next expansion: node u (none when the construction is over).
__global__ void k_base_frontier(BaseState B, u64* __restrict__ fr_dk, u32* __restrict__ fr_zid) {
  const int u = B.bc->u;
  if (u < 0) return;
  fr_dk[0] = B.nkey[u];
  fr_zid[0] = (u32)u;
}

// recnt[u]: node u's successors (its expansion's valid slots), read by the early-witness test.
__global__ __launch_bounds__(64) void k_base_count(BaseState B, const int* __restrict__ svalid,
                                                   const int* __restrict__ rec_cnt) {
  const int n = rec_cnt[0];
  int c = 0;
  for (int r = threadIdx.x; r < n; r += kWave) c += svalid[r] ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (threadIdx.x == 0 && B.bc->u >= 0) B.recnt[B.bc->u] = c;
}

// The position of the first node of `ids` (in order) for which test(zone) holds, or -1: blocks of
// 64 nodes, one per lane.
template <class Test>
__device__ __forceinline__ int base_first(const BaseState& B, const int* ids, int n, int d2, Test test) {
  const int l = lane_id();
  for (int b0 = 0; b0 < n; b0 += kWave) {
    bool hit = false;
    if (b0 + l < n) hit = test(B.zone + (size_t)ids[b0 + l] * d2);
    const u64 m = __ballot(hit);
    if (m) return b0 + __builtin_ctzll(m);
  }
  return -1;
}

// One node's turn: its successors interned in order (intern, construction.cpp:863-1003), the
// targets of its edges queued, the next node popped. With `early`, a witness at the node
// (leadsto_witness, found by the stop-set launch before this one) ends the construction instead.
__global__ __launch_bounds__(64) void k_base_step(const DevNet* __restrict__ Np, BaseState B,
                                                  const int* __restrict__ rec_cnt, const int* __restrict__ svalid,
                                                  const u64* __restrict__ sdk, const int* __restrict__ szone,
                                                  const int* __restrict__ sse, const int* __restrict__ sre,
                                                  const LiveCtr* __restrict__ lc, const Counters* __restrict__ ctr,
                                                  int early) {
  __shared__ int S[HORA_MAX_DIM * HORA_MAX_DIM], T[HORA_MAX_DIM * HORA_MAX_DIM], fa[HORA_MAX_DIM * HORA_MAX_DIM],
      pa[HORA_MAX_DIM * HORA_MAX_DIM], Lb[HORA_MAX_DIM], Ub[HORA_MAX_DIM], locs[HORA_MAX_PROCS],
      ints[HORA_MAX_INTS], misc[4];
  const DevNet& N = *Np;
  const int l = threadIdx.x, d = N.d, d2 = N.d2;
  BaseCtr& C = *B.bc;
  const int u = C.u;
  if (u < 0) return;
  if (ctr->err) {   // the expansion overflowed the device's per-state tables
    if (l == 0) {
      C.err = ctr->err;
      C.u = -1;
    }
    return;
  }
  if (l == 0) {
    B.expanded[u] = 1;
    ++C.expanded;
  }
  if (early && lc->hit) {
    if (l == 0) {
      C.stop = 1;
      C.u = -1;
    }
    return;
  }
  const int nr = rec_cnt[0];
  for (int r = 0; r < nr; ++r) {
    if (!svalid[r]) continue;
    const u64 k = sdk[r];
    bsync();
    for (int e = l; e < d2; e += kWave) S[e] = szone[(size_t)r * d2 + e];
    decode_key(N, k, locs, ints);
    bsync();
    if (l < d) cover_bounds(N, locs, l, 1, &Lb[l], &Ub[l]);
    bool fresh = false;
    const int h = base_slot(B, k, &fresh);
    if (h < 0) {
      if (l == 0) {
        C.err = ERR_HASH_FULL;
        C.u = -1;
      }
      return;
    }
    bsync();
    int v = -1;
    bool grew = false;
    if (!fresh) {
      const int* ids = B.pool + B.kptr[h];
      const int n = B.klen[h];
      // the oldest node that covers the successor (a lane's zone is read a column at a time when
      // d <= 16: one wait per column)
      int j = base_first(B, ids, n, d2, [&](const int* Z) {
        return d <= 16 ? alu_le_lane16<false>(S, Z, d, Lb, Ub) : alu_le_lane(S, Z, d, Lb, Ub);
      });
      if (j >= 0) {
        v = ids[j];
        if (l == 0) ++C.covered;
      } else {
        // the oldest node the successor covers: enlarged to it in place
        j = base_first(B, ids, n, d2, [&](const int* Z) {
          return d <= 16 ? alu_le_lane16<true>(Z, S, d, Lb, Ub) : alu_le_lane(Z, S, d, Lb, Ub);
        });
        if (j >= 0) {
          v = ids[j];
          int* zj = B.zone + (size_t)v * d2;
          for (int e = l; e < d2; e += kWave) zj[e] = S[e];
          grew = true;
          if (l == 0) ++C.enlarged;
        } else {
          // the oldest node whose union with it is one zone: the hull (entrywise maximum)
          if (d <= 16) {
            j = base_first(B, ids, n, d2, [&](const int* Z) { return union_convex_lane(S, Z, d); });
            if (j >= 0) {
              v = ids[j];
              int* zj = B.zone + (size_t)v * d2;
              for (int e = l; e < d2; e += kWave) zj[e] = max(zj[e], S[e]);
            }
          } else {
            for (int m = 0; m < n && v < 0; ++m) {
              const int* zj = B.zone + (size_t)ids[m] * d2;
              bsync();
              for (int e = l; e < d2; e += kWave) T[e] = zj[e];
              if (z_try_union(T, S, d, fa, pa, misc)) v = ids[m];
            }
            if (v >= 0) {
              int* zj = B.zone + (size_t)v * d2;
              wsync();
              for (int e = l; e < d2; e += kWave) zj[e] = T[e];
            }
          }
          if (v >= 0) {
            grew = true;
            if (l == 0) ++C.unions;
          }
        }
      }
    }
    if (v < 0) {
      // appended: a new node, last of its key
      v = C.nodes;
      int* zv = B.zone + (size_t)v * d2;
      for (int e = l; e < d2; e += kWave) zv[e] = S[e];
      if (l == 0) {
        B.nkey[v] = k;
        B.expanded[v] = 0;
        B.inq[v] = 0;
        B.ehead[v] = B.etail[v] = -1;
        if (fresh) {
          B.hkey[h] = k;
          B.klen[h] = 0;
          B.kcap[h] = 0;
          ++C.keys;
        }
        base_append(B, h, v);
        C.nodes = v + 1;
      }
    }
    if (l == 0) {
      ++C.generated;
      const int e = C.edges++;
      B.etgt[e] = v;
      B.ese[e] = sse[r];
      B.ere[e] = sre[r];
      B.enext[e] = -1;
      if (B.etail[u] < 0) B.ehead[u] = e;
      else B.enext[B.etail[u]] = e;
      B.etail[u] = e;
      if (grew) {
        if (v != u) {
          B.expanded[v] = 0;
          B.ehead[v] = B.etail[v] = -1;
        } else {
          B.expanded[u] = 0;
        }
      }
    }
  }
  // queue the targets of the node's edges, then pop; a popped node already expanded only queues
  // its own targets (base_forward's loop)
  if (l == 0) {
    int w = u;
    for (;;) {
      for (int e = B.ehead[w]; e >= 0; e = B.enext[e]) {
        const int t = B.etgt[e];
        if (!B.expanded[t] && !B.inq[t]) {
          B.inq[t] = 1;
          B.queue[C.qt++] = t;
        }
      }
      if (C.qh == C.qt) {
        C.u = -1;
        break;
      }
      w = B.queue[C.qh++];
      B.inq[w] = 0;
      if (!B.expanded[w]) {
        C.u = w;
        C.ukey = B.nkey[w];
        break;
      }
    }
  }
}

// The key table at a new size: every occupied slot of the old table (O) moved to the new one (B),
// with its node array, one thread per old slot.
__global__ void k_base_rehash(BaseState O, long Ho, BaseState B) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= Ho) return;
  const u64 k = O.hkey[s];
  if (k == kEmptyKey) return;
  const u64 H = 1ull << B.hbits;
  u64 h = key_mix(k) >> (64 - B.hbits);
  for (u64 n = 0; n < H; ++n, h = (h + 1) & (H - 1)) {
    const u64 prev = atomicCAS((unsigned long long*)&B.hkey[h], (unsigned long long)kEmptyKey, (unsigned long long)k);
    if (prev == kEmptyKey) {
      B.kptr[h] = O.kptr[s];
      B.klen[h] = O.klen[s];
      B.kcap[h] = O.kcap[s];
      return;
    }
  }
  atomicExch(&B.bc->err, ERR_HASH_FULL);
}

// The graph in CSR form: a node's edges in list order, nodes in creation order.
__global__ void k_base_rows(BaseState B, int n, int* __restrict__ cnt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int c = 0;
  for (int e = B.ehead[v]; e >= 0; e = B.enext[e]) ++c;
  cnt[v] = c;
}

__global__ void k_base_fill(BaseState B, int n, const long long* __restrict__ row, int* __restrict__ etgt,
                            int* __restrict__ ese, int* __restrict__ ere) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  long long o = row[v];
  for (int e = B.ehead[v]; e >= 0; e = B.enext[e], ++o) {
    etgt[o] = B.etgt[e];
    ese[o] = B.ese[e];
    ere[o] = B.ere[e];
  }
}

// One node of each discrete key (the seed test reads the key only).
__global__ void k_base_heads(BaseState B, long H, int* __restrict__ heads, LiveCtr* lc) {
  const long h = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = h < H && B.hkey[h] != kEmptyKey;
  const int idx = block_append(&lc->ns, ok);
  if (ok) heads[idx] = B.pool[B.kptr[h]];
}

}  // namespace
