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
  int32_t pad;
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
  HORA_ZOP_EMPTY = 8       /* out_flag[z] = (zone is empty) (dbm_empty)                          */
};
int hora_hip_zone_op(int32_t op, int32_t d, int64_t n, const int32_t* a, const int32_t* b,
                     const int32_t* arg, const int32_t* L, const int32_t* U, int32_t* out,
                     int32_t* out_flag, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HORA_HIP_H */
