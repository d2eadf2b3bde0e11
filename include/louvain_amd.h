/*
 * louvain_amd.h - C ABI of the MI355X Louvain engine (liblouvain_amd.so).
 *
 * The reference drives its GPU pipeline through six extern C++ functions that
 * share one mutable StatusCUDA/CUDAGraph pair (LouvainCUDA/LouvainCPP.cpp:129-134)
 * plus AllocateMemory/DeAllocateMemory (LouvainCUDA/LouvainCPP.cpp:136-205).
 * This header replaces that boundary with plain C: an opaque context owns one
 * HBM arena, one HIP stream (plus three it fans independent launches out to,
 * see lv_config.fanout) and the per-level device state; every function
 * takes plain pointers and sizes and returns an lv_status code.  No exception,
 * abort or host allocation crosses the ABI: host output buffers belong to the
 * caller and the library copies into them.
 *
 * Semantics are the reference's (synchronous sweeps, singleton guard,
 * one-pass-late convergence, merge-isolated post-step, first-occurrence
 * renumbering, ordered-pair coarsening), with exact arithmetic: all weight
 * sums are int64 in units of 2^-s (s = the smallest exponent that makes every
 * input weight an integer), gains are FP64 evaluated as W - ((gamma*deg)*f).
 *
 * Threading: one context is used by one host thread at a time; contexts are
 * independent (no globals).
 */
#ifndef LOUVAIN_AMD_H
#define LOUVAIN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lv_ctx lv_ctx;

/* Status codes.  The reference returns none (CUDA errors are printed only in
 * AllocateMemory, LouvainCUDA/LouvainCPP.cpp:140-151). */
enum {
    LV_OK = 0,
    LV_ERR_NOMEM = -1,         /* arena or device memory exhausted */
    LV_ERR_INVALID_INPUT = -2, /* unknown/repeated node id, negative or non-finite weight */
    LV_ERR_RANGE = -3,         /* weight sums exceed the exact int64 range */
    LV_ERR_ARG = -4,           /* bad argument or call out of order */
    LV_ERR_DEVICE = -5,        /* HIP runtime error */
    LV_ERR_NO_TRACE = -6,      /* trace requested without LV_FLAG_TRACE */
    LV_ERR_TOO_LARGE = -7      /* input beyond the static scratch of lv_run_static */
};

/* weight element types for lv_load_edges */
enum { LV_WEIGHT_INT32 = 0, LV_WEIGHT_FLOAT32 = 1 };

/* lv_config.flags */
enum {
    LV_FLAG_TRACE = 1,  /* keep every pass's next[] on the host (parity tests) */
    LV_FLAG_TIMING = 2  /* record HIP events around every phase and sweep launch */
};

typedef struct {
    double gamma;        /* resolution; the reference has none (gamma = 1) */
    double min_gain;     /* convergence threshold, reference __MIN = 1e-7
                            (LouvainCUDA/LouvainCPP.cpp:55) */
    int32_t max_passes;  /* 0 = unbounded, the reference's __PASS_MAX -1 */
    int32_t max_levels;  /* 0 = unbounded */
    int32_t device;      /* HIP device ordinal (the reference uses device 0,
                            LouvainCUDA/LouvainCPP.cpp:774) */
    int32_t flags;       /* LV_FLAG_* */
    uint64_t arena_bytes; /* 0 = sized from the graph at lv_load_edges */
    /* Launch-shape tunables; 0 = the tuning table (tuning.txt beside the library,
     * or the file LV_TUNING_FILE names, written by louvaingpu_amd/tune.py) for
     * the graph's size, else the built-in default.  A row or coarsening segment
     * with more than hub_q candidates (64..4096; default 4096) takes the hub
     * hash partitions of about hub_part_target entries (64..2^30; default
     * 2048), cut into chunks of chunk_mul entries per partition (1..1024;
     * default 32).  Results do not depend on them. */
    int32_t hub_q;
    int32_t hub_part_target;
    int32_t chunk_mul;
    /* Edge-list upload of lv_load_edges (positions as ids): 0 = encoded when
     * the list has at least 2^18 edges and the process may use 4 or more CPUs
     * (host threads turn a source-sorted src into row offsets and pack int
     * weights 2, 4 or 8 bits each, with an escape list, while dst is copied;
     * what does not encode is copied as it is), 1 = plain copies, 2 = encoded
     * at any size.  Results do not depend on it. */
    int32_t upload;
    /* Extra HIP streams for the independent launches of one phase (the sweep
     * tiers of a pass, the coarsening's segment tiers and gathers, the two
     * halves of the level-0 CSR fill): 0 = when the phase holds at least 2^22
     * entries, 1 = never, 2 = always.  LV_FLAG_TIMING and a caller's stream
     * (lv_run_static) keep every launch on one stream.  Results do not depend
     * on it. */
    int32_t fanout;
} lv_config;

/* Per-level summary (the reference keeps only Partition{count,...}). */
typedef struct {
    int32_t n;           /* vertices of the level graph */
    int32_t k;           /* communities after renumbering (0 before lv_renumber) */
    int64_t nnz;         /* directed CSR entries the sweep reads */
    int64_t m_units;     /* total weight m in units */
    int32_t passes;      /* sweeps run */
    int32_t scale;       /* weight unit exponent s */
    double q;            /* value oneLevel returned */
    double lower;        /* clamp passed to oneLevel */
} lv_level_info;

/* Timing counters, filled when LV_FLAG_TIMING is set (milliseconds of device
 * time between HIP events on the context stream). */
#define LV_SWEEP_TIERS 12
typedef struct {
    double csr_ms, init_ms, sweep_ms, apply_ms, renumber_ms, coarsen_ms, flatten_ms;
    double tier_ms[LV_SWEEP_TIERS];
    int64_t tier_launches[LV_SWEEP_TIERS];
    double tier_bytes[LV_SWEEP_TIERS];     /* algorithmic bytes 12*nnz+24*N+8*P, with P the
                                              distinct (vertex, community) candidates
                                              the sweep read a community total for */
    double tier_cand[LV_SWEEP_TIERS];      /* P per tier (summed over launches) */
    double sweep_bytes;                    /* sum over tiers */
    int64_t sweep_entries;                 /* sum of nnz over passes */
    int32_t levels, passes;
    uint64_t arena_capacity;               /* bytes of the HBM arena */
    uint64_t arena_high_water;             /* most bytes allocated at once */
    uint64_t upload_bytes;                 /* host-to-device bytes of the last lv_load_edges
                                              (12 per edge as plain copies) */
} lv_stats;

void lv_config_default(lv_config *cfg);
const char *lv_strerror(int status);

/* Context with its stream and arena.  Replaces AllocateMemory
 * (LouvainCUDA/LouvainCPP.cpp:136-198) and the cached_allocator
 * (LouvainCUDA/cached_allocator.h:12-108). */
int lv_create(lv_ctx **out, const lv_config *cfg);
/* Replaces DeAllocateMemory (LouvainCUDA/LouvainCPP.cpp:200-205). */
int lv_destroy(lv_ctx *ctx);

/* Input: node ids (NULL means 0..n-1) and one entry per edge (src id, dst id,
 * weight).  Replaces Graph::setNodes/setEdges plus the H2D copies of
 * LouvainCUDA/LouvainCPP.cpp:283-287.  Ids are mapped to positions on the
 * device; an edge naming an unknown id, a repeated id, or a negative or
 * non-finite weight gives LV_ERR_INVALID_INPUT.  The edge list stays resident
 * in the arena; each lv_run / level-0 lv_neighbour_compute starts from it. */
int lv_load_edges(lv_ctx *ctx, int32_t n, const int32_t *ids, int64_t e, const int32_t *src,
                  const int32_t *dst, const void *w, int wtype);

/* Phase entry points, in the reference's call order
 * (LouvainCUDA/LouvainCPP.cpp:334-499). */

/* CalculateNeighbour (LouvainCUDA/NeighbourCompute.cu:42-118): at level 0
 * builds the symmetric CSR of the loaded edge list and m; at later levels the
 * CSR was produced by lv_induce and this only checks the call order. */
int lv_neighbour_compute(lv_ctx *ctx);
/* initCUDA (LouvainCUDA/InitKernel.cu:128-222): degrees, self-loop weights,
 * singleton communities, degree buckets for the sweep. */
int lv_init(lv_ctx *ctx);
/* modularityCUDA (LouvainCUDA/OneLevelKernel.cu:965-980): Q of the current
 * community state; valid between lv_init and lv_renumber (renumbering
 * releases the level's community totals). */
int lv_modularity(lv_ctx *ctx, double *q);
/* oneLevelCUDA (LouvainCUDA/OneLevelKernel.cu:518-954): the sweep loop,
 * merge-isolated post-step; *q receives the level's returned modularity.
 * `lower` is the clamp of LouvainCUDA/OneLevelKernel.cu:895. */
int lv_one_level(lv_ctx *ctx, double lower, double *q, int32_t *passes);
/* renumberCUDA (LouvainCUDA/RenumberKernel.cu:63-220): first-occurrence
 * renumbering; appends the level to the device-resident dendrogram.  The host
 * Partition arrays are fetched with lv_level_partition. */
int lv_renumber(lv_ctx *ctx, int32_t *k);
/* inducedGraphCUDA (LouvainCUDA/InducedGraphKernel.cu:50-113): coarsens the
 * level graph into the next level's CSR: each community's entries are
 * gathered in member order and aggregated by neighbour community (LDS hash
 * tables; hash partitions for communities with more than 2048 entries). */
int lv_induce(lv_ctx *ctx);

/* generate_dendogram's level loop (LouvainCUDA/LouvainCPP.cpp:209-553) run
 * entirely on the device, with the same termination rule; the dendrogram
 * stays in HBM.  *levels receives its length. */
int lv_run(lv_ctx *ctx, int32_t *levels);
/* lv_run that also delivers the dendrogram to the host, as generate_dendogram
 * collects one Partition per level (renumberCUDA's copies,
 * LouvainCUDA/RenumberKernel.cu:202-203): each level's Partition (the layout of
 * lv_level_partition) is copied into node[l]/comm[l] (l < max_levels, each of
 * at least n entries of the input) on a second stream while the next level
 * runs, and the final partition into final_comm (n entries).  Page-locked
 * buffers let the copies overlap the run.  Levels beyond max_levels stay on the
 * device for lv_level_partition. */
int lv_run_partitions(lv_ctx *ctx, int32_t max_levels, int32_t *const *node, int32_t *const *comm,
                      int32_t *final_comm, int32_t *levels);

/* Results. */
int lv_num_levels(lv_ctx *ctx, int32_t *levels);
int lv_level_info_get(lv_ctx *ctx, int32_t level, lv_level_info *out);
/* The level's Partition in the reference's layout: (node, community) listed in
 * order of (old community, position) (LouvainCUDA/RenumberKernel.cu:121-133,
 * 191,202-203).  node/comm hold info.n entries. */
int lv_level_partition(lv_ctx *ctx, int32_t level, int32_t *node, int32_t *comm);
/* Renumbered community of every position of the level (n entries). */
int lv_level_assignment(lv_ctx *ctx, int32_t level, int32_t *comm);
/* partitionAtLevel (LouvainCUDA/LouvainCPP.cpp:82-106): level-0 Partition
 * order, communities mapped through levels 1..level. */
int lv_flatten(lv_ctx *ctx, int32_t level, int32_t *node, int32_t *comm);
/* Final community of each level-0 position, flattened at the last level. */
int lv_final_assignment(lv_ctx *ctx, int32_t *comm);
/* Newman modularity of the final partition on the input graph (exact integer
 * aggregates, evaluated in FP64). */
int lv_final_modularity(lv_ctx *ctx, double *q);

/* calculateClusterSize + layoutProcess (LouvainCUDA/LouvainCPP.cpp:555-651) for a
 * partition given per position (comm[i] in [0, ncomm)): cluster_end[c] = the
 * reference's clusterSize[c] (inclusive prefix sum of cluster sizes),
 * index_in_cluster[i] = earlier positions in the same cluster, and the circle
 * layout x = 50 cos(2 pi s / n), y = 50 + 50 sin(2 pi s / n) at slot
 * s = cluster_end[c-1] + index_in_cluster[i] (double, stored as float).
 * Independent of the loaded graph; uses the context's arena. */
int lv_cluster_layout(lv_ctx *ctx, int32_t n, const int32_t *comm, int32_t ncomm,
                      int32_t *cluster_end, int32_t *index_in_cluster, float *x, float *y);

/* Parity instrumentation. */
/* next[] proposed by `pass` (1-based) of `level`, after the guard (and, on the
 * level's last pass, after the merge-isolated step); needs LV_FLAG_TRACE. */
int lv_trace_next(lv_ctx *ctx, int32_t level, int32_t pass, int32_t *next);
/* Q measured at the start of `pass` (the value the break test compares). */
int lv_trace_pass_q(lv_ctx *ctx, int32_t level, int32_t pass, double *q);
/* Current level's CSR after lv_neighbour_compute/lv_induce: rowptr[n+1],
 * col/w[nnz] (self entries are kept per vertex, not in the rows). */
int lv_level_csr(lv_ctx *ctx, int64_t *rowptr, int32_t *col, int64_t *w);
/* Current level's per-vertex degree k, last self-loop weight and sum of
 * self-loop weights (w==0 counted as 1), in units; needs lv_init. */
int lv_level_vertex(lv_ctx *ctx, int64_t *k, int64_t *loop, int64_t *selfsum);
/* Current level's sizes before lv_init/lv_one_level. */
int lv_current_level(lv_ctx *ctx, int32_t *level, int32_t *n, int64_t *nnz);

/* Graph sets (batching).  The reference runs one graph per generate_dendogram
 * call (LouvainCUDA/LouvainCPP.cpp:209-553); these run many independent graphs,
 * each at nres resolutions, as the components of one union graph, so every
 * phase is one launch for the whole set.  Component c is graph c / nres at
 * resolution gamma[c % nres]; each component follows the single-graph
 * semantics exactly (its own m, break test, lower clamp, level loop and
 * termination), and its results equal an lv_run of that graph alone.
 * Edges use local ids 0..n[g]-1 and are listed graph after graph (e[g] each);
 * cfg.gamma is ignored.  Node ids are positions (no id mapping). */
int lv_batch_load(lv_ctx *ctx, int32_t ngraphs, const int32_t *n, const int64_t *e,
                  const int32_t *src, const int32_t *dst, const void *w, int wtype, int32_t nres,
                  const double *gamma);
/* The same with one pointer per graph for each array (no concatenation by the
 * caller); the arrays are read before the call returns. */
int lv_batch_load_graphs(lv_ctx *ctx, int32_t ngraphs, const int32_t *n, const int64_t *e,
                         const int32_t *const *src, const int32_t *const *dst,
                         const void *const *w, int wtype, int32_t nres, const double *gamma);
int lv_batch_run(lv_ctx *ctx);
int lv_batch_num_components(lv_ctx *ctx, int32_t *ncomp);
int lv_batch_num_levels(lv_ctx *ctx, int32_t comp, int32_t *levels);
/* Per-component level summary (nnz is not tracked per component: 0). */
int lv_batch_level_info(lv_ctx *ctx, int32_t comp, int32_t level, lv_level_info *out);
/* Renumbered community of every vertex of the component at that level. */
int lv_batch_level_assignment(lv_ctx *ctx, int32_t comp, int32_t level, int32_t *comm);
/* Final community of every level-0 vertex of the component (comp = -1: all
 * components, concatenated in component order). */
int lv_batch_final_assignment(lv_ctx *ctx, int32_t comp, int32_t *comm);

/* No-allocation run.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops and harnesses that forbid
 * hipMalloc).  lv_load_edges + lv_run + lv_final_assignment (+
 * lv_final_modularity when `modularity` is not NULL) of one graph, with every
 * device buffer - the arena and the context's scalars - carved out of one
 * file-scope device array of LV_STATIC_SCRATCH_BYTES that each GPU holds until
 * the process ends.  No hipMalloc, hipFree, pinned allocation, stream or event
 * is created; kernels and copies go to `stream` (a hipStream_t, NULL = the
 * null stream), which is synchronised before the call returns.  Inputs and
 * outputs are the caller's host arrays (final_comm: n entries).  cfg: NULL =
 * lv_config_default; flags must be 0 and arena_bytes is ignored.  A graph whose
 * footprint bound (200 bytes per edge + 320 per vertex) exceeds the array
 * gives LV_ERR_TOO_LARGE before any device work; LV_ERR_NOMEM if a run still
 * outgrows it.  Every call on one GPU uses that GPU's one array, whose contents
 * persist from the last call (each run initialises what it reads): calls on
 * one device are serialised by a process-wide lock, so calls that overlap on
 * one GPU from this process wait for each other rather than race. */
#define LV_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int lv_run_static(const lv_config *cfg, void *stream, int32_t n, const int32_t *ids, int64_t e,
                  const int32_t *src, const int32_t *dst, const void *w, int wtype,
                  int32_t *final_comm, int32_t *levels, double *modularity);

int lv_get_stats(lv_ctx *ctx, lv_stats *out);
int lv_reset_stats(lv_ctx *ctx);
/* Wall time of the last lv_run in milliseconds (host clock, device synced). */
int lv_last_run_ms(lv_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* LOUVAIN_AMD_H */
