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
#pragma once

namespace {

// The construction's state in device memory. Nodes: key, zone (the arena, d*d ints each), flags,
// the next node of the same key (creation order), the head and tail of the node's edge list.
// Edges: target, transition, the next edge of the same node. A hash table of the keys (open
// addressing, linear probing) holds each key's first and last node. The FIFO is an array read
// from qh and appended at qt.
struct BaseCtr {
  int u;                      // node to expand next, or -1: the construction is over
  int nodes, edges, keys;
  long long qh, qt;
  int stop;                   // an early witness ended the construction
  int err;                    // ERR_HASH_FULL, or the expansion's error
  unsigned long long expanded, generated, covered, enlarged, unions;
};

struct BaseState {
  u64* nkey;                  // [N]
  int* zone;                  // [N * d2]
  signed char* expanded;      // [N]
  signed char* inq;           // [N]
  int* knext;                 // [N]
  int* ehead;                 // [N]
  int* etail;                 // [N]
  int* recnt;                 // [N] successors of the node's last expansion (the stop-set test reads it)
  int* enext;                 // [E]
  int* etgt;                  // [E]
  int* ese;                   // [E]
  int* ere;                   // [E]
  u64* hkey;                  // [H]
  int* hhead;                 // [H]
  int* htail;                 // [H]
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
    B.hhead[h] = 0;
    B.htail[h] = 0;
    B.nkey[0] = sdk[0];
    B.expanded[0] = 0;
    B.inq[0] = 0;
    B.knext[0] = -1;
    B.ehead[0] = B.etail[0] = -1;
    C.u = 0;
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
  if (threadIdx.x == 0) B.recnt[B.bc->u] = c;
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
      // oldest node that covers the successor
      for (int j = B.hhead[h]; j >= 0 && v < 0; j = B.knext[j])
        if (z_alu_le(S, B.zone + (size_t)j * d2, d, Lb, Ub)) v = j;
      if (v >= 0) {
        if (l == 0) ++C.covered;
      } else {
        // oldest node the successor covers: enlarged to it in place
        for (int j = B.hhead[h]; j >= 0 && v < 0; j = B.knext[j])
          if (z_alu_le(B.zone + (size_t)j * d2, S, d, Lb, Ub)) v = j;
        if (v >= 0) {
          int* zj = B.zone + (size_t)v * d2;
          for (int e = l; e < d2; e += kWave) zj[e] = S[e];
          grew = true;
          if (l == 0) ++C.enlarged;
        } else {
          // oldest node whose union with it is one zone: the hull
          for (int j = B.hhead[h]; j >= 0 && v < 0; j = B.knext[j]) {
            const int* zj = B.zone + (size_t)j * d2;
            bsync();
            for (int e = l; e < d2; e += kWave) T[e] = zj[e];
            if (z_try_union(T, S, d, fa, pa, misc)) v = j;
          }
          if (v >= 0) {
            int* zj = B.zone + (size_t)v * d2;
            wsync();
            for (int e = l; e < d2; e += kWave) zj[e] = T[e];
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
        B.knext[v] = -1;
        B.ehead[v] = B.etail[v] = -1;
        if (fresh) {
          B.hkey[h] = k;
          B.hhead[h] = v;
          ++C.keys;
        } else {
          B.knext[B.htail[h]] = v;
        }
        B.htail[h] = v;
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
        break;
      }
    }
  }
}

// Rebuild of the key table at a new size (creation order kept: the nodes in order, each appended
// to its key's list).
__global__ __launch_bounds__(64) void k_base_rehash(BaseState B) {
  if (threadIdx.x != 0) return;
  BaseCtr& C = *B.bc;
  for (int v = 0; v < C.nodes; ++v) {
    const u64 k = B.nkey[v];
    bool fresh = false;
    const int h = base_slot(B, k, &fresh);
    if (h < 0) {
      C.err = ERR_HASH_FULL;
      return;
    }
    B.knext[v] = -1;
    if (fresh) {
      B.hkey[h] = k;
      B.hhead[h] = v;
    } else {
      B.knext[B.htail[h]] = v;
    }
    B.htail[h] = v;
  }
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
  if (ok) heads[idx] = B.hhead[h];
}

}  // namespace
