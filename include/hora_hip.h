/* hora_hip.h — C ABI of the MI355X (gfx950) zone-graph back end for HORA.
 *
 * The back end is selected as `--backend hip` beside the reference's `seq` and `omp` engines.
 * Everything crossing this boundary is plain data: POD structs, fixed-width integers, pointers
 * and sizes. No C++ or torch type appears here, and no exception crosses it: every entry point
 * returns 0 on success and a negative HORA_HIP_E* code on failure, with a message available from
 * hora_hip_last_error().
 *
 * Reference interfaces replaced (paths relative to the HORA source tree):
 *   - include/hora/gpu/network.hpp:38-128   class NetGpuExpander (expand / merge / stage_*)
 *   - include/hora/gpu/network.hpp:121-128  make_net_gpu_{metal,cuda} factories
 *   - include/hora/gpu/encode.hpp:14-47     netgpu::NetEncoding (the flat static network tables)
 *   - include/hora/reach.hpp:10-21          ReachStats, enum class Backend
 *   - include/hora/reachability.hpp:61-62   reachable_net(net, target, backend, stats)
 *   - include/hora/network.hpp:150-163      NetTarget
 *
 * Bound encoding (HDBM raw_t, restated in src/metal/netengine_metal.mm:76-100):
 *   raw = (v << 1) | 1  means  x_i - x_j <= v      raw = v << 1  means  x_i - x_j < v
 *   HORA_INF = 1 << 30 is "no bound"; le(0) = 1.  A zone of dimension d = nclocks + 1 is d*d
 *   int32 raws, row-major, entry [i*d + j] bounding x_i - x_j; clock 0 is the constant 0.
 */
#ifndef HORA_HIP_H
#define HORA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HORA_HIP_ABI_VERSION 1

#define HORA_INF (1 << 30)
#define HORA_ALU_NEG_INF (-(1 << 29)) /* "clock never compared" sentinel in L/U arrays */
#define HORA_MAX_DIM 32                /* largest zone dimension the device kernels accept */
#define HORA_MAX_PROCS 64
#define HORA_MAX_INTS 32

/* Error codes. */
#define HORA_HIP_OK 0
#define HORA_HIP_E_ARG (-1)      /* invalid argument or shape */
#define HORA_HIP_E_SCOPE (-2)    /* network outside the back end's scope: caller falls back */
#define HORA_HIP_E_DEVICE (-3)   /* HIP runtime error / no device */
#define HORA_HIP_E_MEMORY (-4)   /* device memory exhausted or a capacity exceeded */
#define HORA_HIP_E_INTERNAL (-5)
#define HORA_HIP_E_CAPPED (-6)   /* a federation outgrew its capacity: no verdict (GfpStats::capped) */

/* Integer guard operators, same order as enum class IntOp (include/hora/network.hpp:23). */
enum { HORA_OP_EQ = 0, HORA_OP_NE = 1, HORA_OP_LT = 2, HORA_OP_LE = 3, HORA_OP_GT = 4, HORA_OP_GE = 5 };

/* Flat static description of a network of timed automata.
 *
 * The raw tables mirror netgpu::NetEncoding (include/hora/gpu/encode.hpp:14-47) plus the
 * fields of hora::Network it leaves behind (integer ranges, initial values, channel kinds).
 * Edges are numbered globally, process-major, in each process's own edge order; a process's
 * outgoing edges from a location are visited in increasing global edge number, which is the
 * order NProc::edges_of gives (src/model/network.cpp:38-45).
 *
 * The analysis tables are the outputs of Network::compute_bounds / compute_loc_bounds /
 * compute_active (src/model/network.cpp:59-245). The caller supplies them; the back end does not
 * recompute them. */
typedef struct hora_net_enc {
  int32_t nclocks;  /* clocks x1..xn; zone dimension d = nclocks + 1 */
  int32_t nprocs;
  int32_t nintvars;
  int32_t nchans;
  int32_t nedges;
  int32_t total_locs;
  const int32_t* int_lo;    /* [nintvars] declared range and initial value */
  const int32_t* int_hi;
  const int32_t* int_init;
  const uint8_t* chan_bcast;  /* [nchans] 1 = broadcast channel */
  const uint8_t* chan_urgent; /* [nchans] 1 = urgent channel */
  const int32_t* locbase;     /* [nprocs+1] first global location of each process */
  const int32_t* loc_init;    /* [nprocs] initial location (process-local index) */
  const int32_t* loc_flags;   /* [total_locs] bit0 committed, bit1 urgent */
  const int32_t* inv_off;     /* [total_locs+1] offsets into inv_data, in triples */
  const int32_t* inv_data;    /* (i, j, raw) triples */
  const int32_t* edge_proc;   /* [nedges] owning process */
  const int32_t* edge_src;    /* [nedges] source location (process-local) */
  const int32_t* edge_tgt;    /* [nedges] target location (process-local) */
  const int32_t* edge_chan;   /* [nedges] channel id or -1 */
  const int32_t* edge_dir;    /* [nedges] +1 send '!', -1 receive '?', 0 none */
  const int32_t* cg_off;      /* [nedges+1] clock guard triples (i, j, raw) */
  const int32_t* cg_data;
  const int32_t* rs_off;      /* [nedges+1] clock resets (clock index) */
  const int32_t* rs_clk;
  const int32_t* ig_off;      /* [nedges+1] integer guard triples (var, op, val) */
  const int32_t* ig_data;
  const int32_t* iu_off;      /* [nedges+1] integer update pairs (var, val) */
  const int32_t* iu_data;
  /* analyses */
  const int32_t* lbound;      /* [d] global Extra_LU bounds, L[0] = U[0] = 0 */
  const int32_t* ubound;
  const int32_t* lbound_loc;  /* [total_locs * d] per-location a_LU bounds (HORA_ALU_NEG_INF = none) */
  const int32_t* ubound_loc;
  const uint64_t* active_mask; /* [total_locs] bit c set iff clock c is active at the location */
} hora_net_enc;

/* E<> target: a conjunction of location requirements, integer atoms and clock constraints.
 * Mirrors hora::NetTarget (include/hora/network.hpp:150-163). */
typedef struct hora_target {
  int32_t never;        /* 1: the literal `false` (full search, matches nothing) */
  int32_t nloc_req;
  const int32_t* loc_req;    /* (proc, local location) pairs */
  int32_t nint_cond;
  const int32_t* int_cond;   /* (var, op, val) triples */
  int32_t nclock_cond;
  const int32_t* clock_cond; /* (i, j, raw) triples */
} hora_target;

/* Counters of one search. Mirrors hora::ReachStats (include/hora/reach.hpp:10-19) plus the
 * device-side quantities the roofline needs. */
typedef struct hora_reach_stats {
  uint64_t explored;    /* sum of frontier sizes over the levels run (engine.cpp:2055) */
  uint64_t generated;   /* non-empty successors counted as the reference counts them */
  uint64_t expanded;    /* frontier states whose successors were computed */
  uint64_t stored_live; /* live zones in the passed list at the end */
  uint64_t keys;        /* distinct discrete states in the passed list */
  uint64_t inserted;    /* zones ever appended to the passed list */
  uint64_t subsumed;    /* successors dropped as covered */
  uint64_t killed;      /* stored zones retired by a covering successor */
  uint64_t unions;      /* successful convex-union merges */
  uint64_t chain_reads; /* stored zones read by the covering / kill / union folds */
  int32_t levels;
  int32_t found;        /* 1 if a target state was reached */
  int32_t emode_switch_level; /* BFS level at which Extra_LU+ switched on, or -1 */
  int32_t zone_entry_bytes;   /* bytes per stored zone entry: 2 (int16 storage) or 4 */
  double seconds;       /* search wall time, excluding set-up */
} hora_reach_stats;

typedef struct hora_hip_ctx hora_hip_ctx;

/* Library identity; no device needed. */
int hora_hip_abi_version(void);
const char* hora_hip_build_info(void);

/* Open a device context for one network and query clock mask.
 * Replaces make_net_gpu_{metal,cuda}(net, qmask) (include/hora/gpu/network.hpp:121-128,
 * src/cuda/netengine_cuda.cu:427-438): returns HORA_HIP_E_SCOPE for a network the device path
 * does not cover (diagonal constraints, d > HORA_MAX_DIM, > HORA_MAX_PROCS processes, a packed
 * discrete key wider than 63 bits), on which the caller falls back to its CPU engine.
 * `target` supplies the clock constants that floor the a_LU bounds and the query clock mask
 * (engine.cpp:1417-1419,1440-1456); it is copied. */
int hora_hip_open(const hora_net_enc* net, const hora_target* target, hora_hip_ctx** out);
void hora_hip_close(hora_hip_ctx* ctx);
const char* hora_hip_last_error(const hora_hip_ctx* ctx);

/* Full device-resident E<> search: reach_engine's GPU branch (src/reachability/engine.cpp:1411,
 * 1581-1993) and reachable_net(net, target, Backend::Hip, stats) (engine.cpp:2244-2299).
 * Writes 1 to *reachable if a target state is reachable, 0 otherwise. */
int hora_hip_reach(hora_hip_ctx* ctx, int32_t* reachable, hora_reach_stats* stats);

/* Device time and algorithmic HBM traffic of the last hora_hip_reach, per kernel family, summed
 * over its BFS levels (HIP events on the context's stream). Bytes count int32 zone entries at
 * 4 B each: expand = frontier zones read + successor zones written; probe = successor zones +
 * the stored zones each was tested against; merge = kept successor zones + stored zones read by
 * the folds + surviving zones copied to the arena (read and write). */
typedef struct hora_reach_profile {
  double expand_ms, probe_ms, merge_ms, other_ms;
  double expand_bytes, probe_bytes, merge_bytes;
  int64_t expand_launches, probe_launches, merge_launches;
} hora_reach_profile;
int hora_hip_reach_profile(const hora_hip_ctx* ctx, hora_reach_profile* out);

/* ---- batched zone operations (zone-kernel microbenchmark; R1-R7, R11, R16-R18) ----
 * One call applies one operation to n zones of dimension d, stored back to back (n*d*d int32).
 * The flag-only operations (INCLUDES, ALU_LE, EMPTY) leave `out` untouched.
 * Input and output pointers are DEVICE pointers; `stream` is a hipStream_t or NULL for the
 * default stream. The call is asynchronous on that stream. Operation semantics are those of the
 * HDBM calls the reference makes (restated in src/metal/netengine_metal.mm:101-335 and
 * src/metal/engine_metal.mm:80-93). */
enum {
  HORA_ZOP_CLOSE = 0,      /* Floyd-Warshall closure in place (dbm_close)                        */
  HORA_ZOP_UP = 1,         /* delay: z[i][0] = inf for i >= 1 (dbm_up)                            */
  HORA_ZOP_RESET = 2,      /* x_k := 0 for k = arg[z] (dbm_reset); arg: int32[n]                  */
  HORA_ZOP_CONSTRAIN = 3,  /* tighten (i, j, raw) = arg[3z..3z+2] and re-close (dbm_constrain);
                              out_flag[z] = 1 if non-empty                                        */
  HORA_ZOP_EXTRA_LU = 4,   /* Extra_LU+ then close; L/U: int32[n*d] each                         */
  HORA_ZOP_INCLUDES = 5,   /* out_flag[z] = (zone a[z] includes zone b[z]) (dbm_includes)        */
  HORA_ZOP_ALU_LE = 6,     /* out_flag[z] = (a[z] <=_aLU b[z]) under L/U[z] (dbm_is_alu_le)       */
  HORA_ZOP_UNION = 7,      /* out_flag[z] = (a[z] U b[z] convex); out zone = hull when convex,
                              else a copy of a[z] (dbm_try_union)                                 */
  HORA_ZOP_EMPTY = 8,      /* out_flag[z] = (zone is empty) (dbm_empty)                          */
  HORA_ZOP_DOWN = 9,       /* time predecessors: z[0][j] = min(<=0, min_i z[i][j]) (Dbm::down)     */
  HORA_ZOP_FREE = 10       /* forget the clocks of the bit mask arg[z] (dbm_free_many)            */
};
int hora_hip_zone_op(int32_t op, int32_t d, int64_t n, const int32_t* a, const int32_t* b,
                     const int32_t* arg, const int32_t* L, const int32_t* U, int32_t* out,
                     int32_t* out_flag, void* stream);


/* ---- liveness: p --> q, E[] psi, A[] not deadlock (backward greatest fixpoint) ----
 * Replaces leads_to_otf / exists_invariant_run_otf / deadlock_reachable
 * (include/hora/liveness.hpp:37-170) on the reference's a_LU lane: one forward zone graph, then
 * W = W /\ (NoSucc \/ Pre(W)) over the closure the query reads (src/liveness/fixpoint.cpp:48-965).
 *
 * A discrete predicate in disjunctive normal form, the discrete part of hora::DiscPred
 * (include/hora/predicate.hpp:36-80): true iff some term has all its atoms true. An atom is four
 * int32: kind (0 = location of process `index`, 1 = integer variable `index`), index, op
 * (HORA_OP_*; locations take EQ or NE), value. nterms = 0 is `false`; a term with no atom is `true`. */
typedef struct hora_dpred {
  int32_t nterms;
  const int32_t* term_off;   /* [nterms+1] atom offsets */
  const int32_t* atoms;      /* 4 int32 per atom */
} hora_dpred;

enum { HORA_COVER_AUTO = 0, HORA_COVER_AM = 1, HORA_COVER_ONLINE = 2 };

enum {
  HORA_LIVE_LEADSTO = 0,     /* p --> q: holds iff no reachable p-state has a maximal run avoiding q */
  HORA_LIVE_EG = 1,          /* E[] psi: holds iff the initial state has a maximal psi-run          */
  HORA_LIVE_DEADLOCK = 2     /* A[] not deadlock: holds iff no reachable state is deadlocked        */
};

typedef struct hora_live_query {
  int32_t kind;
  int32_t early;             /* settle at the first witness state, checked after each level:
                                LEADSTO a p /\ !q node with a stop valuation (leadsto_witness,
                                src/liveness/construction.cpp:786-813), EG a psi node with one
                                (online_eg_terminal_witness, :748-767), DEADLOCK a deadlock     */
  int32_t cover;             /* HORA_COVER_AUTO: the reference's route (decide_leads_to,
                                src/liveness/api.cpp:357-447): the online construction (a_LU covers,
                                src/liveness/construction.cpp:480-683) unless a reachable discrete
                                state can seed a deadlock atom (online_seed_key, src/liveness/
                                refinement.cpp:176-282), then the a_M graph; p --> q on a network
                                where D is provably a_LU (structurally_empty_d, src/liveness/
                                internal.cpp:59-103) takes the base construction (api.cpp:37-117,
                                construction.cpp:863-1096), one node expanded per step.
                                HORA_COVER_ONLINE: as AUTO, but the online construction there too.
                                HORA_COVER_AM: the a_M graph always. */
  int32_t pad;
  hora_dpred p;              /* LEADSTO: p.  EG: psi's discrete part.  DEADLOCK: unused           */
  hora_dpred q;              /* LEADSTO: q.  Otherwise unused                                     */
  /* EG whose psi has a clock conjunction: the context was opened on the network restricted by it
   * (the conjunction adjoined to every location invariant, analyses recomputed); these are the
   * unrestricted network's invariant tables, in hora_net_enc's layout, for the maximality test
   * (stop_set's restricted branch, src/liveness/construction.cpp:1200-1238). NULL otherwise. */
  const int32_t* orig_inv_off;
  const int32_t* orig_inv_data;
} hora_live_query;

/* Mirrors hora::GfpStats (include/hora/liveness.hpp:48-76) plus construction counters. */
typedef struct hora_live_stats {
  int64_t outer_nodes;       /* zone-graph nodes stored                                         */
  int64_t fwd_edges;
  int64_t disc_keys;
  int64_t expanded_nodes;
  int64_t generated;         /* non-empty successors                                            */
  int64_t covered;           /* successors routed to an existing node                           */
  int64_t levels;            /* forward BFS levels                                              */
  int64_t closure;           /* nodes the fixpoint runs over                                    */
  int64_t nonempty;          /* of those, nodes whose converged value is non-empty              */
  int64_t iters;             /* fixpoint rounds (device) or value-changing relaxations (oracle) */
  int64_t pre_calls;         /* exact predecessor computations                                  */
  int64_t fed_max;           /* widest converged value, in zones                                */
  int64_t products;          /* W /\ cand intersections formed (fixpoint.cpp:842-850)          */
  int64_t chain_reads;       /* stored or queued zones read by the cover tests of the construction */
  int64_t retired;           /* records retired by a newer record of their key that covers them
                                (online_compact_key, src/liveness/construction.cpp:426-477)       */
  int64_t unions;            /* successful convex unions of a successor with an active record    */
  int64_t enlarged;          /* A[] not deadlock: unexpanded records enlarged in place            */
  int64_t fed_hbm_nodes;     /* node launches whose federation lists outgrew LDS (device memory) */
  int32_t early;             /* 1: settled by a witness state during the construction           */
  int32_t capped;            /* 1: a federation outgrew its capacity; no verdict                */
  int32_t route;             /* covering of the graph decided on: 0 a_LU, 1 a_M                 */
  int32_t seed_keys;         /* discrete states that could seed a deadlock atom (0 on a_LU)    */
  double seconds;            /* whole query                                                     */
  double build_seconds;      /* forward construction                                            */
  double gfp_seconds;        /* stop sets, closure and fixpoint                                 */
} hora_live_stats;

/* Decide one liveness query on the context's network. Writes 1 to *holds if the property holds.
 * Returns HORA_HIP_E_CAPPED when a federation outgrew the device capacity. */
int hora_hip_live(hora_hip_ctx* ctx, const hora_live_query* q, int32_t* holds, hora_live_stats* stats);

/* Inspection of the last hora_hip_live's graph and fixpoint (no counterpart in the reference:
 * for parity tests and tools). Writes nodes, edges and d; copies each array that is non-NULL (sized
 * from a first call with NULL arrays). Node ids are creation order. Returns HORA_HIP_E_ARG, with
 * zero counts, when the last hora_hip_live failed before its graph and fixpoint were complete. */
typedef struct hora_live_graph {
  int64_t nodes, edges;
  int32_t d, pad;
  uint64_t* dk;              /* [nodes] packed discrete key                                     */
  int32_t* zones;            /* [nodes*d*d] stored zones                                        */
  int64_t* row_off;          /* [nodes+1] edges of node v are [row_off[v], row_off[v+1])        */
  int32_t* etgt;             /* [edges] target node                                             */
  int32_t* ese;              /* [edges] first participant edge (sender or local)                */
  int32_t* ere;              /* [edges] receiver edge, -1, or -2 - c for broadcast combination c */
  int8_t* expanded;          /* [nodes]                                                         */
  int8_t* in_s;              /* [nodes] in the fixpoint's closure                               */
  int8_t* w_nonempty;        /* [nodes] converged value non-empty                               */
} hora_live_graph;
int hora_hip_live_graph(const hora_hip_ctx* ctx, hora_live_graph* g);

/* ---- backward greatest fixpoint over a caller-supplied zone graph ----
 * Replaces LivenessEngine::solve_greatest_fixpoint (src/liveness/fixpoint.cpp:48-965) with the
 * readouts of certificate_in / accepting_closure (fixpoint.cpp:979-1101) and eg_verdict
 * (src/liveness/api.cpp:290-337), for a caller that built the zone graph itself (the reference's own
 * construction, for instance): the call sites are final_verdict_violated (src/liveness/
 * fixpoint.cpp:1110, called by decide_leads_to and run, src/liveness/api.cpp:459 and :346) and
 * eg_verdict (api.cpp:296-304). Every array is a
 * DEVICE pointer and stays the caller's. Keys use the packing of the context's network
 * (hora_hip_open); node 0 is the initial node; an edge records its transition as hora_live_graph
 * does. */
typedef struct hora_gfp_graph {
  int64_t nodes, edges;
  const uint64_t* dk;        /* [nodes] packed discrete keys                                    */
  const int32_t* zones;      /* [nodes*d*d] node zones                                           */
  const int64_t* row_off;    /* [nodes+1] edges of node v are [row_off[v], row_off[v+1])        */
  const int32_t* etgt;       /* [edges] target node                                             */
  const int32_t* ese;        /* [edges] first participant edge (sender or local)                */
  const int32_t* ere;        /* [edges] receiver edge, -1, or -2 - c for broadcast combination c */
  const int8_t* expanded;    /* [nodes] 1 = the node's row holds all its successors              */
} hora_gfp_graph;

/* Decides q on the graph: W = W /\ (NoSucc \/ Pre(W)) over the closure the query reads (kind
 * LEADSTO, EG or DEADLOCK; q->early and q->cover are not used). Writes the verdict to *holds, the
 * fixpoint's counters to *stats (capped on HORA_HIP_E_CAPPED), and, when w_nonempty is not NULL, per
 * node 1 iff its converged value is non-empty (DEVICE int8[nodes]). */
int hora_hip_gfp_solve(hora_hip_ctx* ctx, const hora_gfp_graph* g, const hora_live_query* q, int32_t* holds,
                       hora_live_stats* stats, int8_t* w_nonempty);

/* hora_hip_gfp_solve without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The same computation and results as hora_hip_gfp_solve, with every working
 * buffer carved from one file-scope device array of 256 MiB per GPU (held until the process ends).
 * A graph of more nodes than fit (256 MiB / (80 + 16 d^2) bytes: 96,421 nodes at zone
 * dimension 13), or a run whose pools outgrow the array, returns HORA_HIP_E_MEMORY with the reason in
 * hora_hip_last_error. Federations are held in LDS only here (no device-memory lists): one that
 * outgrows them returns HORA_HIP_E_CAPPED. Every call on a GPU shares that GPU's array: calls that overlap on one GPU,
 * from any contexts, race and must be serialised by the caller. The context must have been opened
 * (hora_hip_open allocates, once). */
int hora_hip_gfp_solve_noalloc(hora_hip_ctx* ctx, const hora_gfp_graph* g, const hora_live_query* q,
                               int32_t* holds, hora_live_stats* stats, int8_t* w_nonempty);

/* ---- serving: verification jobs of different networks, continuous batching on one GPU ----
 * A server owns `slots` device slots (a HIP stream, device-resident passed list, frontier and
 * successor buffers each, kept across jobs). Submitted jobs wait in arrival order; a job is
 * admitted when a slot is free and the declared HBM footprints of the running jobs plus its own
 * stay within the budget (always when nothing runs). Admitted jobs run concurrently, level by
 * level on their own streams. No counterpart in the reference (its CLI runs one query per
 * process); a job's verdict and counters are those of hora_hip_reach / hora_hip_live alone. */
typedef struct hora_hip_server hora_hip_server;

typedef struct hora_serve_job_result {
  int64_t job;               /* id returned at submission                                       */
  int32_t kind;              /* 0 reachability (hora_hip_reach), 1 liveness (hora_hip_live)       */
  int32_t status;            /* HORA_HIP_OK or the error code of the job                          */
  int32_t verdict;           /* reachability: 1 reachable; liveness: 1 holds                     */
  int32_t slot;
  hora_reach_stats reach;
  hora_live_stats live;
  double t_submit, t_start, t_done;   /* seconds on the server's clock (hora_hip_server_now)  */
} hora_serve_job_result;

/* hbm_budget_bytes <= 0: 90% of the device memory free at open. */
int hora_hip_server_open(int32_t slots, int64_t hbm_budget_bytes, hora_hip_server** out);
/* The job's arrays are copied; est_bytes <= 0 takes a default footprint from the zone dimension. */
int hora_hip_server_submit_reach(hora_hip_server* s, const hora_net_enc* net, const hora_target* target,
                                 int64_t est_bytes, int64_t* job);
int hora_hip_server_submit_live(hora_hip_server* s, const hora_net_enc* net, const hora_live_query* q,
                                int64_t est_bytes, int64_t* job);
/* Completed jobs, up to max; waits up to timeout_ms for at least one (0: do not wait). */
int hora_hip_server_wait(hora_hip_server* s, hora_serve_job_result* out, int32_t max, int32_t* n,
                         int32_t timeout_ms);
double hora_hip_server_now(const hora_hip_server* s);
/* Finishes the queued jobs, then releases the slots. */
void hora_hip_server_close(hora_hip_server* s);

#ifdef __cplusplus
}
#endif

#endif /* HORA_HIP_H */
