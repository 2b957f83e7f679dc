// online.hpp — the liveness construction's storing rule on the device: the online route of the
// reference (LivenessEngine::online_base_forward, src/liveness/refinement.cpp:1060-1065, with
// online_expand :610-781 and online_intern, src/liveness/construction.cpp:480-683) for a map D
// without atoms. Part of the hora_hip.hip translation unit, included by liveness.hpp.
//
// The reference interns each successor of the record it expands, in FIFO order of records:
//   * the successor goes to the newest ACTIVE record of its discrete key that a_LU-covers it;
//   * otherwise it is joined with every active record but the root whose union with it is one
//     zone, in newest-first passes repeated until a pass joins nothing, and appended as a new
//     active record, which retires every other active record but the root that it covers;
//   * a record retired before its turn is never expanded.
// The device runs that order level by level. A level's frontier is the records the previous
// level created that are still active, in creation order; all of them are expanded at once, and
// the successors are interned per key in (frontier index, transition index) order, which is the
// FIFO order. Two facts make the level parallel without changing any answer:
//   1. the newest record covering a successor is active at the successor's turn (a retired record
//      is covered by the newer record that retired it), so the pre-level part of the cover test
//      is one independent probe per successor (k_gprobe, newest first), and the target of a
//      successor covered there is fixed once the level's new records are known (k_opost);
//   2. whether a frontier record is still active at its turn depends only on the successors of
//      the frontier records before it. The level is interned with every frontier record assumed
//      active; each frontier record retired by a successor of an earlier frontier record is then
//      marked skipped, and the level is interned again, until the marks no longer change. The
//      marks of a prefix of the frontier are final after each round, so this ends; a level is
//      interned once when no record is retired before its turn, the usual case.
// The early witness of the record at its turn (leadsto_witness, online_eg_terminal_witness,
// online_deadlock_at) stops the construction there: the successors of that record and of every
// later frontier record are not interned.
#pragma once

namespace {

// One wavefront per touched key: the key's successors that no record stored before this level
// covers, in (frontier index, transition index) order, through online_intern. The working chain
// is the key's active records (ids, creation order) followed by this level's new records
// (kNewTag | successor index g); retired entries carry kDead. Chain zones are cached in LDS like
// k_merge's. A new record whose zone a union widened keeps it in the pool `hz` (slot fhz[g] - 1).
// Outputs per successor g: tslot[g] = the successor index of the new record it goes to (itself
// when it creates one), fnew[g] = 1 for a new record, cret[g] = the successor index of the new
// record that retired the new record of g; per stored record j retired here: lret[j] = that
// successor index, and j is appended to `ev`. At the end the chain of live entries replaces the
// key's chain (in its reserved segment), and the key's new records (successor indices, in
// order) head its work list, their number in P.ncnt.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_omerge(
    const DevNet* __restrict__ Np, Passed P, const u32* __restrict__ list, int nlist, int* __restrict__ wlist,
    const int* __restrict__ sj, const int* __restrict__ szone, int* __restrict__ hz, int hzcap,
    int* __restrict__ fhz, int* __restrict__ tslot, int* __restrict__ fnew, int* __restrict__ cret,
    int* __restrict__ lret, int* __restrict__ ev, LiveCtr* lc, Counters* ctr, int heavy) {
  extern __shared__ int sm[];
  const DevNet& N = *Np;
  const int d = N.d, d2 = N.d2, l = threadIdx.x;
  const MergeLayout ml = merge_layout(d2, heavy != 0);
  int* S = sm;
  int* M = S + d2;
  int* fa = M + d2;
  int* pa = fa + d2;
  int* Lb = pa + d2;
  int* Ub = Lb + HORA_MAX_DIM;
  int* locs = Ub + HORA_MAX_DIM;
  int* ints = locs + HORA_MAX_PROCS;
  int* misc = ints + HORA_MAX_INTS;
  int* sbuf = misc + 16 + 32;
  u32* chl = (u32*)(sbuf + ml.sort_cap);
  int* cslot = (int*)(chl + ml.chain_cap);
  int* cache = cslot + ml.chain_cap;
  const int zs = ml.zs, cz = ml.cz;
  if ((int)blockIdx.x >= nlist) return;

  const u32 h = list[blockIdx.x];
  const u32 n = P.ncnt[h];
  int* lst = wlist + P.noff[h];
  if (n <= (u32)kWave) {
    const int v = (int)l < (int)n ? lst[l] : 0x7fffffff;
    int rank = 0;
    for (int j = 0; j < (int)n; ++j) rank += (__shfl(v, j) < v) ? 1 : 0;
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
  decode_key(N, P.key[h], locs, ints);
  wsync();
  if (l < d) cover_bounds(N, locs, l, 1, &Lb[l], &Ub[l]);

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
  auto csync = [&]() __attribute__((always_inline)) { if (chg) bsync(); else wsync(); };
  const int ncached = min((int)old_len, cz);
  for (u32 k = l; k < old_len; k += kWave) {
    ch[k] = P.nodes[old_off + k];
    const int slot = (int)k - ((int)old_len - ncached);
    cs[k] = slot >= 0 ? slot : -1;
  }
  csync();
  for (int k = 0; k < ncached; ++k) {
    const int* src = P.zarena + (size_t)ch[old_len - ncached + k] * d2;
    for (int e = l; e < d2; e += kWave) cache[k * zs + e] = src[e];
  }
  int next_slot = ncached;
  auto newz = [&](int g) __attribute__((always_inline)) -> const int* {
    const int s = fhz[g];
    return s > 0 ? hz + (size_t)(s - 1) * d2 : szone + (size_t)sj[g] * d2;
  };
  auto zp = [&](int pos) __attribute__((always_inline)) -> const int* {
    const int s = cs[pos];
    if (s >= 0) return cache + s * zs;
    const u32 id = ch[pos] & ~kDead;
    return (id & kNewTag) ? newz((int)(id & (kNewTag - 1))) : P.zarena + (size_t)id * d2;
  };
  // the root record (id 0) is never joined nor retired (construction.cpp:444, 617)
  auto is_root = [&](u32 e) __attribute__((always_inline)) { return (e & ~kDead) == 0u; };
  auto lane_includes = [&](const int* a, const int* b) __attribute__((always_inline)) {
    for (int e = 0; e < d2; ++e)
      if (b[e] > a[e]) return false;
    return true;
  };
  int len = (int)old_len, nc = 0;
  unsigned long long n_cov = 0, n_reads = 0, n_uni = 0, n_ret = 0;
  wsync();

  for (u32 k = 0; k < n; ++k) {
    const int g = lst[k];
    const int* src = szone + (size_t)sj[g] * d2;
    wsync();
    for (int e = l; e < d2; e += kWave) S[e] = src[e];
    wsync();
    // (a) the newest live record created earlier in this level that covers S
    int found = -1;
    if (len > (int)old_len) {
      const int span = len - (int)old_len;
      if (span <= 8) {
        for (int m = len - 1; m >= (int)old_len && found < 0; --m) {
          if (ch[m] & kDead) continue;
          if (l == 0) ++n_reads;
          if (z_alu_le(S, zp(m), d, Lb, Ub)) found = m;
        }
      } else {
        for (int base = len - 1; base >= (int)old_len && found < 0; base -= kWave) {
          const int m = base - l;
          bool c = false;
          if (m >= (int)old_len && !(ch[m] & kDead)) { c = alu_le_lane(S, zp(m), d, Lb, Ub); ++n_reads; }
          const u64 b = __ballot(c);
          if (b) found = base - __builtin_ctzll(b);
        }
      }
    }
    if (found >= 0) {
      if (l == 0) tslot[g] = (int)(ch[found] & (kNewTag - 1));
      ++n_cov;
      continue;
    }
    // (b) convex unions with the live records but the root, newest first, until a pass joins none
    bool widened = false;
    for (bool again = true; again;) {
      again = false;
      int m = len - 1;
      while (m >= 0) {
        int win = -1;
        if (d <= 16) {
          const int q = m - l;
          bool ok = false;
          if (q >= 0 && !(ch[q] & kDead) && !is_root(ch[q])) {
            const int* z = zp(q);
            ok = !lane_includes(S, z) && union_convex_lane(S, z, d);
            ++n_reads;
          }
          const u64 b = __ballot(ok);
          if (b) win = m - __builtin_ctzll(b);
          else m -= kWave;
        } else {
          if (!(ch[m] & kDead) && !is_root(ch[m])) {
            const int* z = zp(m);
            wsync();
            for (int e = l; e < d2; e += kWave) M[e] = z[e];
            if (l == 0) ++n_reads;
            if (!z_includes(S, M, d) && z_try_union(S, M, d, fa, pa, misc)) { win = m; ++n_uni; again = true; widened = true; }
          }
          --m;
          continue;
        }
        if (win >= 0) {
          const int* z = zp(win);
          wsync();
          for (int e = l; e < d2; e += kWave) S[e] = max(S[e], z[e]);
          wsync();
          ++n_uni;
          widened = true;
          again = true;
          m = win - 1;
        }
      }
    }
    // (c) the new record: in a free cache slot while one is left; a widened zone goes to the pool
    wsync();
    int slot = -1;
    if (next_slot < cz) slot = next_slot++;
    if (slot >= 0)
      for (int e = l; e < d2; e += kWave) cache[slot * zs + e] = S[e];
    if (widened) {
      int hs = 0;
      if (l == 0) hs = atomicAdd(&lc->hzn, 1);
      hs = __shfl(hs, 0);
      if (hs < hzcap) {
        for (int e = l; e < d2; e += kWave) hz[(size_t)hs * d2 + e] = S[e];
        if (l == 0) fhz[g] = hs + 1;
      } else if (l == 0) {
        atomicExch(&lc->hzfull, 1);
      }
      __threadfence_block();
      bsync();
    }
    if (l == 0) {
      ch[len] = kNewTag | (u32)g;
      cs[len] = slot;
      tslot[g] = g;
      fnew[g] = 1;
    }
    ++len;
    ++nc;
    csync();
    // (d) retire the other live records but the root that the new one covers
    const int self = len - 1;
    if (len - 1 <= 8) {
      for (int m = len - 2; m >= 0; --m) {
        const u32 id = ch[m];
        if ((id & kDead) || is_root(id)) continue;
        if (l == 0) ++n_reads;
        if (z_alu_le(zp(m), S, d, Lb, Ub)) {
          wsync();
          if (l == 0) {
            ch[m] = id | kDead;
            if (id & kNewTag) cret[id & (kNewTag - 1)] = g;
            else { lret[id] = g; ev[atomicAdd(&lc->nev, 1)] = (int)id; }
          }
          csync();
          ++n_ret;
        }
      }
    } else {
      for (int base = self - 1; base >= 0; base -= kWave) {
        const int m = base - l;
        bool kill = false;
        u32 id = 0;
        if (m >= 0) {
          id = ch[m];
          if (!(id & kDead) && !is_root(id)) { kill = alu_le_lane(zp(m), S, d, Lb, Ub); ++n_reads; }
        }
        const u64 b = __ballot(kill);
        wsync();
        if (kill) {
          ch[m] = id | kDead;
          if (id & kNewTag) cret[id & (kNewTag - 1)] = g;
          else lret[id] = g;
        }
        const bool pre = kill && !(id & kNewTag);
        const int at = wave_append(&lc->nev, pre);
        if (pre) ev[at] = (int)id;
        csync();
        n_ret += __popcll(b);
      }
    }
  }
  // the live chain replaces the key's chain; the new records head the work list, in order
  wsync();
  u32* out = P.nodes + seg;
  int w = 0, c = 0;
  for (int m = 0; m < len; ++m) {
    const u32 id = ch[m];
    if (l == 0) {
      if (m >= (int)old_len) wlist[P.noff[h] + c] = (int)(id & ~kDead & (kNewTag - 1));
    }
    if (m >= (int)old_len) ++c;
    if (id & kDead) continue;
    if (l == 0) out[w] = id;
    ++w;
  }
  wsync();
  if (l == 0) {
    P.off[h] = seg;
    P.len[h] = (u32)w;
    P.ncnt[h] = (u32)nc;
    atomicAdd(&lc->nuni, (int)n_uni);
    atomicAdd(&lc->nret, (int)n_ret);
    stat_add(ctr, C_SUB, n_cov);
  }
  wave_stat(ctr, C_READS, n_reads);
}

// The target of a successor that a record stored before this level covers: the newest record
// created earlier in this level that covers it, if any (the record found by the probe otherwise).
// One wavefront per successor; the key's new records head its work list in creation order.
__global__ __launch_bounds__(64) void k_opost(const DevNet* __restrict__ Np, Passed P, const long* __restrict__ gbk,
                                              long G, const int* __restrict__ covered, const int* __restrict__ wlist,
                                              const int* __restrict__ sj, const u64* __restrict__ sdk,
                                              const int* __restrict__ szone, const int* __restrict__ hz,
                                              const int* __restrict__ fhz, int* __restrict__ tslot, Counters* ctr) {
  __shared__ int S[HORA_MAX_DIM * HORA_MAX_DIM], Lb[HORA_MAX_DIM], Ub[HORA_MAX_DIM], lg[HORA_MAX_PROCS];
  const long g = blockIdx.x;
  if (g >= G || !covered[g]) return;
  const long h = gbk[g];
  const int nc = (int)P.ncnt[h];
  if (nc == 0) return;
  const int* cl = wlist + P.noff[h];
  // creators before g: cl is ascending
  int lo = 0, hi = nc;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cl[mid] < (int)g) lo = mid + 1; else hi = mid;
  }
  const int before = lo;
  if (before == 0) return;
  const DevNet& N = *Np;
  const int l = threadIdx.x, d = N.d, d2 = N.d2;
  const u64 dk = sdk[sj[g]];
  if (l < N.np) lg[l] = (int)((dk >> N.kshift[l]) & N.kmask[l]);
  const int* src = szone + (size_t)sj[g] * d2;
  for (int e = l; e < d2; e += kWave) S[e] = src[e];
  wsync();
  if (l < d) cover_bounds(N, lg, l, 1, &Lb[l], &Ub[l]);
  wsync();
  int first = -1;
  unsigned long long reads = 0;
  for (int base = before - 1; base >= 0 && first < 0; base -= kWave) {
    const int m = base - l;
    bool ok = false;
    if (m >= 0) {
      const int gc = cl[m];
      const int s = fhz[gc];
      const int* z = s > 0 ? hz + (size_t)(s - 1) * d2 : szone + (size_t)sj[gc] * d2;
      ok = alu_le_lane(S, z, d, Lb, Ub);
      ++reads;
    }
    const u64 b = __ballot(ok);
    if (b) first = cl[base - __builtin_ctzll(b)];
  }
  wave_stat(ctr, C_READS, reads);
  if (l == 0 && first >= 0) tslot[g] = first;
}

// Source frontier index of recipe slot r: the last i with rec_off[i] <= r.
__device__ __forceinline__ int src_of_slot(const int* rec_off, int F, int r) {
  int lo = 0, hi = F - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rec_off[mid] <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// A frontier record retired by a successor of an earlier frontier record was not active at its
// turn: mark it skipped (fnext). fr_zid is ascending.
__global__ void k_ovalid(const int* __restrict__ ev, int nev, const int* __restrict__ lret,
                         const u32* __restrict__ fr_zid, int F, const int* __restrict__ sj,
                         const int* __restrict__ rec_off, int* __restrict__ fnext) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nev) return;
  const u32 j = (u32)ev[t];
  int lo = 0, hi = F;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (fr_zid[mid] < j) lo = mid + 1; else hi = mid;
  }
  if (lo >= F || fr_zid[lo] != j) return;
  const int src = src_of_slot(rec_off, F, sj[lret[j]]);
  if (src < lo) fnext[lo] = 1;
}

// Did the skip marks change? Also moves the new marks into place and clears the next ones.
__global__ void k_ovcmp(int* __restrict__ fcur, int* __restrict__ fnext, int F, LiveCtr* lc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F) return;
  const int a = fcur[i], b = fnext[i];
  if (a != b) { atomicExch(&lc->vchanged, 1); fcur[i] = b; }
  fnext[i] = 0;
}

// The first frontier record, active at its turn, that is an early witness.
__global__ void k_owfirst(const u32* __restrict__ fr_zid, int F, const int* __restrict__ fskip,
                          const signed char* __restrict__ whit, LiveCtr* lc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F && !fskip[i] && whit[fr_zid[i]]) atomicMin(&lc->wfirst, i);
}

// The successor slots of skipped frontier records, and of the witness and everything after it,
// are dropped (valid2 from valid).
__global__ void k_oexcl(const int* __restrict__ rec_off, int F, long R, const int* __restrict__ fskip, int wfirst,
                        const int* __restrict__ valid, int* __restrict__ valid2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F) return;
  const long a = rec_off[i], b = i + 1 < F ? rec_off[i + 1] : R;
  const bool drop = fskip[i] || i >= wfirst;
  for (long r = a; r < b; ++r) valid2[r] = drop ? 0 : valid[r];
}

// Successor counts of the frontier records (all their successors, before any is dropped): the
// "no successor" fact the stop set reads at the record's turn.
__global__ void k_ofsucc(const int* __restrict__ rec_off, int F, long R, const long long* __restrict__ pos,
                         long long total, const u32* __restrict__ fr_zid, int* __restrict__ recnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F) return;
  const long a = rec_off[i], b = i + 1 < F ? rec_off[i + 1] : R;
  recnt[fr_zid[i]] = (int)((b < R ? pos[b] : total) - (a < R ? pos[a] : total));
}

// The frontier records where an early witness can stand: p /\ !q (leads-to), psi (E[] psi), all
// (A[] not deadlock); listed for k_stop so that the others take no workgroup.
__global__ void k_owlist(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap, int F,
                         const u64* __restrict__ fr_dk, const u32* __restrict__ fr_zid, int* __restrict__ list,
                         LiveCtr* lc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const LiveArgs& A = *Ap;
  bool keep = false;
  if (i < F) {
    const u64 dk = fr_dk[i];
    keep = A.kind == HORA_LIVE_DEADLOCK ? true
           : A.kind == HORA_LIVE_EG ? pred_eval(*Np, A.p, dk)
                                    : pred_eval(*Np, A.p, dk) && !pred_eval(*Np, A.q, dk);
  }
  const int idx = block_append(&lc->wcnt, keep);
  if (keep) list[idx] = (int)fr_zid[i];
}

// E[] psi: a frontier record outside psi is stored and never expanded (src/liveness/api.cpp:83-96)
__global__ void k_ogfilter(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ A, const u64* __restrict__ fr_dk,
                           int F, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F && !pred_eval(*Np, A->p, fr_dk[i])) cnt[i] = 0;
}

// The frontier records' final state for the level: expanded iff active at their turn and not
// after the witness (a psi-less record under E[] psi is stored but not expanded); the count of
// successors interned (their edge row).
__global__ void k_ofront(const DevNet* __restrict__ Np, const LiveArgs* __restrict__ Ap, const int* __restrict__ rec_off,
                         int F, long R, const long long* __restrict__ pos2, long long total, const u32* __restrict__ fr_zid,
                         const u64* __restrict__ fr_dk, const int* __restrict__ fskip, int wfirst, int eg,
                         signed char* __restrict__ expanded, int* __restrict__ recnt, LiveCtr* lc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool ex = false;
  if (i < F) {
    ex = !fskip[i] && i <= wfirst && (!eg || pred_eval(*Np, Ap->p, fr_dk[i]));
    const long a = rec_off[i], b = i + 1 < F ? rec_off[i + 1] : R;
    const u32 v = fr_zid[i];
    expanded[v] = ex ? 1 : 0;
    recnt[v] = (ex && i < wfirst) ? (int)((b < R ? pos2[b] : total) - (a < R ? pos2[a] : total)) : 0;
  }
  block_append(&lc->nexp, ex);
}

// Edges of the level's successors: to the new record named by tslot (nbase + its rank), or to the
// stored record the probe found; the new records in creation order.
__global__ void k_oassign(long G, long long nbase, long long ebase, int record, const int* __restrict__ cov_old,
                          const int* __restrict__ tslot, const int* __restrict__ fnew, const long long* __restrict__ rank,
                          const int* __restrict__ sj, const int* __restrict__ s_se, const int* __restrict__ s_re,
                          int* __restrict__ etgt, int* __restrict__ ese, int* __restrict__ ere,
                          int* __restrict__ newlist) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int ts = tslot[g];
  if (record) {
    etgt[ebase + g] = ts >= 0 ? (int)(nbase + rank[ts]) : cov_old[g];
    ese[ebase + g] = s_se[sj[g]];
    ere[ebase + g] = s_re[sj[g]];
  }
  if (fnew[g]) newlist[rank[g]] = (int)g;
}

// The level's new records: zone (widened or as generated), key, state.
__global__ __launch_bounds__(64) void k_ocopy(long F2, long long nbase, const int* __restrict__ newlist,
                                              const int* __restrict__ sj, const u64* __restrict__ sdk,
                                              const int* __restrict__ szone, const int* __restrict__ hz,
                                              const int* __restrict__ fhz, const int* __restrict__ cret, int d2,
                                              int* __restrict__ gzone, u64* __restrict__ gdk,
                                              signed char* __restrict__ ractive, int* __restrict__ rret,
                                              int* __restrict__ nflag) {
  const long b = blockIdx.x;
  if (b >= F2) return;
  const int g = newlist[b];
  const long long id = nbase + b;
  const int s = fhz[g];
  const int* src = s > 0 ? hz + (size_t)(s - 1) * d2 : szone + (size_t)sj[g] * d2;
  int* dst = gzone + (size_t)id * d2;
  for (int e = threadIdx.x; e < d2; e += kWave) dst[e] = src[e];
  if (threadIdx.x == 0) {
    gdk[id] = sdk[sj[g]];
    ractive[id] = 1;
    rret[id] = -1;
    nflag[b] = cret[g] < 0 ? 1 : 0;   // still active at the end of the level: next frontier
  }
}

// Retirements of the level: stored records (from the event list) and new records.
__global__ void k_oretire_old(const int* __restrict__ ev, int nev, const int* __restrict__ lret, long long nbase,
                              const long long* __restrict__ rank, signed char* __restrict__ ractive,
                              int* __restrict__ rret) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nev) return;
  const int j = ev[t];
  ractive[j] = 0;
  rret[j] = (int)(nbase + rank[lret[j]]);
}

__global__ void k_oretire_new(long F2, long long nbase, const int* __restrict__ newlist, const int* __restrict__ cret,
                              const long long* __restrict__ rank, signed char* __restrict__ ractive,
                              int* __restrict__ rret) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= F2) return;
  const int r = cret[newlist[b]];
  if (r < 0) return;
  ractive[nbase + b] = 0;
  rret[nbase + b] = (int)(nbase + rank[r]);
}

// The key's chain: new-record entries get their ids.
__global__ __launch_bounds__(64) void k_olink(Passed P, const u32* __restrict__ touched, int T,
                                              const long long* __restrict__ rank, long long nbase) {
  const int t = blockIdx.x;
  if (t >= T) return;
  const u32 h = touched[t];
  u32* c = P.nodes + P.off[h];
  const u32 len = P.len[h];
  for (u32 m = threadIdx.x; m < len; m += kWave) {
    const u32 id = c[m];
    if (id & kNewTag) c[m] = (u32)(nbase + rank[id & (kNewTag - 1)]);
  }
  if (threadIdx.x == 0) P.ncnt[h] = 0;
}

// Next frontier: the new records still active, in creation order.
__global__ void k_onext(long F2, long long nbase, const int* __restrict__ nflag, const long long* __restrict__ npos,
                        const int* __restrict__ newlist, const int* __restrict__ sj, const u64* __restrict__ sdk,
                        u64* __restrict__ fr_dk, u32* __restrict__ fr_zid) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= F2 || !nflag[b]) return;
  const long p = npos[b];
  fr_dk[p] = sdk[sj[newlist[b]]];
  fr_zid[p] = (u32)(nbase + b);
}

// Edges to records never expanded follow the retirements to the active record that covers them
// (online_finalize_graph, src/liveness/refinement.cpp:1071-1160).
__global__ void k_oresolve(long long E, int* __restrict__ etgt, const signed char* __restrict__ expanded,
                           const int* __restrict__ rret) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int t = etgt[e];
  while (!expanded[t] && rret[t] >= 0) t = rret[t];
  etgt[e] = t;
}

// Nodes flagged expanded.
__global__ void k_ocount(const signed char* __restrict__ flag, long n, LiveCtr* lc) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  block_append(&lc->nexp, v < n && flag[v]);
}

// Discrete states of the graph: hash slots with an active record.
__global__ void k_okeys(Passed P, long H, LiveCtr* lc) {
  const long h = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = h < H && P.key[h] != kEmptyKey && P.len[h] > 0;
  block_append(&lc->nkeys, ok);
}

// Per-key chain state of a level's touched keys, saved before the merge and put back when the
// level is interned again.
__global__ void k_osave(Passed P, const u32* __restrict__ touched, int T, u64* __restrict__ off, u32* __restrict__ len,
                        int restore) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const u32 h = touched[t];
  if (restore) {
    P.off[h] = off[t];
    P.len[h] = len[t];
    P.ncnt[h] = 0;
  } else {
    off[t] = P.off[h];
    len[t] = P.len[h];
  }
}

__global__ void k_oclear_ev(const int* __restrict__ ev, int nev, int* __restrict__ lret) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nev) lret[ev[t]] = -1;
}

}  // namespace
