/*
 * tdfs.h -- C ABI of the MI355X T-DFS subgraph-matching library (libtdfs.so).
 *
 * Plain C: fixed-width integers, plain pointers and sizes.  Every function
 * returns 0 on success or a negative TDFS_E* code; nothing asserts or exits
 * across this boundary.  The message of the last failure on the calling
 * thread is available from tdfs_last_error().
 *
 * Reference interfaces replaced (lyuheng/tdfs):
 *   tdfs_pattern_parse / tdfs_pattern_load  STMatch::PatternPreprocessor ctor
 *                                           src/pattern.h:90-101 (+ readfile 110-143)
 *   tdfs_plan                               STMatch::Pattern src/pattern.h:43-62
 *                                           (live fields only) + PatternMultiplicity :69
 *   tdfs_graph_read_meta / tdfs_graph_read  STMatch::GraphPreprocessor::read_bin_file
 *                                           src/graph.h:136-193
 *   tdfs_graph_upload                       GraphPreprocessor::to_gpu src/graph.h:34-50,
 *                                           build_src_vtx src/graph.h:195-212
 *   tdfs_count                              the launch + reduction of cu_test.cu:80-105
 *                                           (kernel _parallel_match, src/gpu_match.cuh:10-13)
 *   tdfs_count_batch, tdfs_serve_*          no counterpart (batched / served forms)
 */
#ifndef TDFS_H_
#define TDFS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDFS_MAX_K 7 /* PAT_SIZE, src/config.h:11 */

/* status codes */
#define TDFS_OK 0
#define TDFS_E_ARG -1       /* bad argument */
#define TDFS_E_IO -2        /* file missing or short */
#define TDFS_E_FORMAT -3    /* malformed .g / CSR (unsorted, duplicate, self-loop, asymmetric) */
#define TDFS_E_HIP -4       /* HIP runtime error */
#define TDFS_E_OOM -5       /* device or host allocation failed */
#define TDFS_E_OVERFLOW -6  /* count does not fit in 64 bits */
#define TDFS_E_BUDGET -7    /* serving admission: request exceeds the memory budget */
#define TDFS_E_STATE -8     /* call not valid in the current state */

/* condition operators, same values as STMatch::CondOperator (src/pattern.h:21) */
#define TDFS_LT 0
#define TDFS_GT 1
#define TDFS_NE 2

/* Matching plan of one pattern: the live fields of STMatch::Pattern
 * (src/pattern.h:43-62).  Vertex u is matched at DFS depth u, in file order.
 * cond[u][j] (j < u) is the symmetry-breaking condition of M(u) against M(j):
 * LT: M(u) < M(j), GT: M(u) > M(j), NE: M(u) != M(j)  (src/pattern.h:416-448).
 * shared_lvl[u] = dep >= 2 means C(u) is computed from C(dep) intersected with
 * N(M(b)) for b in bn_sh[u] (result reuse, src/pattern.h:370-413, 459-480);
 * -1 means no reuse. */
typedef struct tdfs_plan {
  int32_t k;                                /* nnodes, 3..7 */
  int32_t labeled;                          /* 1: lb semantics, 0: ulb */
  int32_t sym01;                            /* partial_ori[0][0] == 1: level-0 keeps r < c */
  int32_t label[TDFS_MAX_K];                /* vertex_labels (all 1 when unlabeled) */
  int32_t deg[TDFS_MAX_K];                  /* pattern degree */
  int32_t num_bn[TDFS_MAX_K];               /* backward neighbours (ascending) */
  int32_t bn[TDFS_MAX_K][TDFS_MAX_K];
  int32_t cond[TDFS_MAX_K][TDFS_MAX_K];     /* TDFS_LT / TDFS_GT / TDFS_NE for j < u */
  int32_t shared_lvl[TDFS_MAX_K];
  int32_t num_bn_sh[TDFS_MAX_K];
  int32_t bn_sh[TDFS_MAX_K][TDFS_MAX_K];
  int32_t adj[TDFS_MAX_K][TDFS_MAX_K];      /* pattern adjacency, 0/1 */
  int32_t npairs;                           /* symmetry-breaking pairs (a,b): M(a) < M(b) */
  int32_t pairs[21][2];
  uint64_t multiplicity;                    /* |Aut(P)|, PatternMultiplicity */
} tdfs_plan;

/* Launch parameters.  Zero fields take defaults (tdfs_tuning_default). */
typedef struct tdfs_tuning {
  int32_t waves_per_block;   /* workgroup size in 64-lane waves (default 4) */
  int32_t blocks_per_cu;     /* resident workgroups per CU for the persistent grid (0 = occupancy max) */
  int32_t chunk_edges;       /* level-0 directed edges per cursor grab (default 64, max 1024) */
  int32_t timeout_us;        /* donate remaining work after this long in one task (<0: never) */
  int32_t split_chunk;       /* iterations per donated task (default 1) */
  int32_t queue_log2;        /* task ring capacity = 2^queue_log2 (default 18, 8..26) */
  int32_t donate_levels;     /* donate only from stack levels < this (default 3) */
  int32_t reuse;             /* 1: result reuse (default); 2: off (testing) */
  int32_t leaf_pair_cap;     /* (parent, candidate) pairs per leaf batch (default 4096, >= 64) */
  int32_t leaf_batch;        /* 1: count the leaf for 64 parents at once (default); 2: one at a time */
  int32_t hub_min_degree;    /* vertices of at least this degree get a neighbour bitmap (default 256) */
  int32_t hub_budget_mb;     /* bitmap memory per uploaded graph (default 4096; <0: none); read at upload */
  int32_t degree_order;      /* 1: also keep a copy renumbered by ascending degree and count on it whenever
                                the result cannot depend on the numbering (default); 2: off.  Read at upload */
  int32_t time_limit_ms;     /* > 0: work still pending this long into a launch is dropped and the plans it
                                belonged to are flagged TDFS_INCOMPLETE (default 0: no limit) */
  int32_t cursor_order;      /* level-0 order of a batch: 1 (default) cursors over (edge chunk, plan): all
                                plans sweep the same region of the graph together, with one cursor per XCD
                                (chunk c goes to workgroups on XCD c mod 8) so that a chunk's plans share
                                that XCD's L2; 3 the same with one cursor for the whole device; 2 one
                                cursor per plan (serving) */
  int32_t waves_per_simd;    /* register budget of the timed kernel: 0 as compiled (default, no spills),
                                4, 5, 6 or 8 waves per SIMD (spills) */
  int32_t hoist_min;         /* the inputs of C(u) that do not involve M(u-1) are intersected once per
                                M(u-2) when the level holds at least this many candidates (default 2;
                                < 0: never) */
  int32_t idle_deadline_ms;  /* a wave that has waited this long for work while no busy wave reported
                                progress ends the launch with a scheduler error (TDFS_E_HIP) instead of
                                waiting on (default 60000; < 0: never) */
  int32_t steal;             /* 1: an idle wave takes the back half of a sibling wave's shallowest open
                                DFS level (same workgroup, LDS atomics; the reference's disabled
                                trans_skt, src/gpu_match.cu:35-125); 2: off (default).  Either way idle
                                waves also take donated tasks from the HBM ring */
  int32_t merge_min;         /* > 0: an extension with two input lists that both hold at least this many
                                candidates (neither more than 8x the other) intersects them by merge path
                                (co-rank split of the merge into 8 steps per lane, tiles staged in LDS)
                                instead of probing the shorter list (default 0: never).  Compiled as
                                kernel variants of their own (default occupancy, and 5 waves per SIMD);
                                it excludes stealing */
  int32_t leaf_memo;         /* 1 (default): when no input of the leaf C(k-1) involves M(k-2), the leaf's
                                intersection over M(0..k-3) is computed once and each sibling M(k-2) only
                                subtracts itself or applies its bound (one search); 2: off (testing) */
  int32_t leaf_hash;         /* 1 (default): a leaf whose inputs are one list shared by all siblings and
                                the sibling's own list holds the shared one in LDS -- as a bitmap when its
                                window spans at most 32,768 ids densely (a sibling with a neighbour bitmap
                                then intersects 32 ids per load: word AND + popcount), else as a hash table
                                when the window holds at most 512 elements -- and tests each sibling's list
                                against it; 3: hash table only; 2: off */
  int32_t ext_hash;          /* 1 (default): an extension holds one input other than the probe that has no
                                neighbour bitmap and at most 512 elements in an LDS hash table, so each
                                candidate's membership test is an LDS lookup instead of a binary search
                                in HBM; 2: off */
  int32_t tails;             /* 1 (default): the timed kernel counts closing vertices without enumerating
                                them -- t >= 2 pairwise non-adjacent last vertices with one backward-
                                neighbour set, totally ordered (unlabeled) or of one label, as C(n, t)
                                or n!/(n-t)! of the first one's n candidates; a closing triangle over one
                                set of earlier neighbours (cliques k >= 5) as the triangles among the
                                first one's candidates when they are at most 64 vertices with neighbour
                                bitmaps; 2: off */
  int32_t free_order;        /* 1 (default): the counting kernel matches a labeled pattern in an order
                                that places a vertex as soon as its pattern neighbours are matched and no
                                unmatched vertex shares its label, when that yields more such free levels
                                (each enumerated for one candidate and multiplied); counts are unchanged,
                                statistics and enumeration keep the plan's order; 2: off */
} tdfs_tuning;

/* Result of one counted pattern. */
typedef struct tdfs_result {
  uint64_t raw;          /* sum of per-wave counts, before multiplicity (cu_test.cu:103) */
  uint64_t count;        /* raw * multiplicity when unlabeled, raw when labeled (cu_test.cu:105) */
  uint64_t alg_bytes;    /* algorithmic bytes (level 0 + every extension), when stats are on */
  uint64_t extensions;   /* DFS extensions performed (stats) */
  uint64_t donated;      /* tasks pushed to the task ring */
  double kernel_ms;      /* device time of the launch (hipEvents on the library stream) */
  uint32_t flags;        /* TDFS_INCOMPLETE: the time limit dropped part of this plan's work */
  uint32_t stolen;       /* work taken by idle waves from sibling waves (tdfs_tuning.steal) */
} tdfs_result;

#define TDFS_INCOMPLETE 1u /* the time limit dropped part of this plan's work: count is a lower bound */
#define TDFS_FAILED 2u     /* serving: the launch failed on the device, or the count overflowed 64 bits */

typedef struct tdfs_ctx tdfs_ctx;
typedef struct tdfs_graph tdfs_graph;

const char *tdfs_last_error(void);
const char *tdfs_version(void);

/* ---------------- host side: patterns ---------------- */
/* Parse the text of a .g file (src/pattern.h:110-143) and build its plan. */
int tdfs_pattern_parse(const char *text, int32_t labeled, tdfs_plan *out);
/* Same, reading the file at path. */
int tdfs_pattern_load(const char *path, int32_t labeled, tdfs_plan *out);
/* The plan the counting kernel runs for this one (tdfs_tuning.free_order):
   the same pattern in another vertex order, or a copy of it.  order (k ints,
   may be null) receives the pattern vertex at each position.  No counterpart
   in the reference: it exposes this library's choice of matching order. */
int tdfs_plan_counting_order(const tdfs_plan *plan, tdfs_plan *out, int32_t *order);
/* Matching-order construction: permutes a connected pattern so that every
 * vertex u >= 1 has a backward neighbour and vertex 2 is adjacent to vertex 1
 * (the precondition of the level-0 prune, src/gpu_match.cu:540-548).
 * order_out[i] = original id placed at position i.  Writes the reordered .g
 * text into text_out (capacity cap). */
int tdfs_pattern_reorder(const char *text, int32_t *order_out, char *text_out, size_t cap);

/* ---------------- host side: graphs (reference binary CSR) ---------------- */
/* meta4 = {n, nnz, vid_size, max_degree} from <prefix>.meta.txt */
int tdfs_graph_read_meta(const char *prefix, int64_t *meta4);
/* Reads <prefix>.vertex.bin / .edge.bin (/.label.bin when labels != NULL)
 * into caller buffers rowptr[n+1], colidx[nnz], labels[n]. */
int tdfs_graph_read(const char *prefix, int64_t n, int64_t nnz, int64_t *rowptr, int32_t *colidx,
                    int32_t *labels);
/* Writes the four files; labels may be NULL (no .label.bin). */
int tdfs_graph_write(const char *prefix, int64_t n, int64_t nnz, const int64_t *rowptr,
                     const int32_t *colidx, const int32_t *labels);
/* Checks strictly ascending lists, no self-loops, symmetry, ids in range. */
int tdfs_graph_validate(int64_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx);

/* ---------------- host side: seeded generators ---------------- */
/* Erdos-Renyi G(n, m): m distinct non-loop undirected edges; nnz = 2m. */
int tdfs_gen_er(int32_t n, int64_t m, uint64_t seed, int64_t *rowptr, int32_t *colidx);
/* Chung-Lu: weights (i+i0)^(-1/(gamma-1)), i0 set so the largest expected
 * degree is dmax; endpoints drawn by weight until m distinct non-loop edges.
 * colidx capacity 2m. */
int tdfs_gen_chung_lu(int32_t n, int64_t m, double gamma, double dmax, uint64_t seed,
                      int64_t *rowptr, int32_t *colidx);
/* labels[v] uniform in [0, nlabels) */
int tdfs_gen_labels(int32_t n, int32_t nlabels, uint64_t seed, int32_t *labels);
/* A connected random pattern with k vertices and m edges
 * (m in [k-1, k(k-1)/2]); labels uniform in [0, nlabels) when nlabels > 0,
 * all 1 otherwise.  Already in matching order.  Writes .g text. */
int tdfs_gen_pattern(int32_t k, int32_t m, int32_t nlabels, uint64_t seed, char *text_out,
                     size_t cap);

/* ---------------- host side: work estimate ---------------- */
/* Expands the counting DFS from nsample directed edges drawn (seeded)
 * uniformly inside 16 equal edge-index strata, and scales each stratum by its
 * size / samples (with rows sorted by degree, the hub rows form the last
 * strata and are always sampled).  out3 = {estimated cost in list
 * elements tested, estimated raw count, 1 if the scaled cost passed cap and
 * the expansion stopped (then the cost is a lower bound)}.  Used for seeded
 * feasibility filters and serving admission. */
int tdfs_estimate_work(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
                       const int32_t *labels, const tdfs_plan *plan, int32_t nsample, uint64_t seed,
                       double cap, double *out3);

/* ---------------- device side ---------------- */
void tdfs_tuning_default(tdfs_tuning *t);
int tdfs_ctx_create(int32_t device, tdfs_ctx **out);
int tdfs_ctx_set_tuning(tdfs_ctx *ctx, const tdfs_tuning *t);
int tdfs_ctx_get_tuning(const tdfs_ctx *ctx, tdfs_tuning *t);
/* stats 1: every launch also accumulates alg_bytes / extensions */
int tdfs_ctx_set_stats(tdfs_ctx *ctx, int32_t on);
void tdfs_ctx_destroy(tdfs_ctx *ctx);
/* Cycle counters of the last launch made with stats on (shader clock,
 * summed over waves): out8 = {level-0 filter, non-leaf extensions, leaf
 * counts, idle waiting, prefix rebuild of received tasks, donation, total,
 * number of waves}. */
int tdfs_ctx_last_profile(const tdfs_ctx *ctx, uint64_t *out8);
/* Operation counters of the last statistics launch (summed over waves):
 * {leaf batches, their parents, their units tested (pairs, bitmap words or
 * list elements), their set-up cycles, extensions, their probe elements,
 * their membership tests by binary search (per 64-lane round and input),
 * memoised-leaf calls}.  No counterpart in the reference (a profiling aid). */
int tdfs_ctx_last_counters(const tdfs_ctx *ctx, uint64_t *out8);
/* device properties: out4 = {compute units, wall-clock kHz, max waves resident, reserved} */
int tdfs_ctx_info(const tdfs_ctx *ctx, int64_t *out4);

/* Copies a validated CSR to HBM (kept resident until tdfs_graph_free).
 * labels: int32[n] or NULL (unlabeled).  Any int32 label values are kept
 * (src/graph.h:153-162 reads 32-bit labels): the distinct values are mapped
 * densely on upload (at most 65,535 of them, else TDFS_E_ARG) and pattern
 * labels go through the same map; a pattern label that no vertex carries
 * matches nothing. */
int tdfs_graph_upload(tdfs_ctx *ctx, int32_t n, int64_t nnz, const int64_t *rowptr,
                      const int32_t *colidx, const int32_t *labels, tdfs_graph **out);
void tdfs_graph_free(tdfs_graph *g);
/* out4 = {n, nnz, has degree-ordered copy, hub bitmaps of the copy counted on} */
int tdfs_graph_info(const tdfs_graph *g, int64_t *out4);
/* perm_out[v] = id of vertex v in the degree-ordered copy (TDFS_E_STATE if none).
 * tdfs_count uses that copy when every plan of the call has vertex 2 adjacent
 * to vertex 1 or dP(2) <= dP(0) (then the count is numbering-independent);
 * edge ranges and statistics then refer to the copy's numbering. */
int tdfs_graph_order(const tdfs_graph *g, int32_t *perm_out);
/* tdfs_estimate_work on the numbering tdfs_count would use for this plan. */
int tdfs_graph_estimate(const tdfs_graph *g, const tdfs_plan *plan, int32_t nsample, uint64_t seed,
                        double cap, double *out3);

/* One pattern, one launch, over directed edges [edge_begin, edge_end)
 * (edge_end < 0: all).  Blocks until done. */
int tdfs_count(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plan, int64_t edge_begin,
               int64_t edge_end, tdfs_result *out);
/* tdfs_count without any allocation at call time.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every buffer of the launch (control block, result word, plan,
 * a task ring of 16,384 entries and the per-wave DFS stacks) lives in a
 * static device array of 256 MiB that each GPU holds from library load until
 * the process ends; its address is looked up once per device by
 * tdfs_ctx_create.  The persistent grid shrinks to the workgroups whose DFS
 * stacks fit (one wave needs 4 x (2,048 + 2 (k-3) x max degree) bytes); a plan
 * of which not even one workgroup fits is refused with TDFS_E_BUDGET.  Same
 * results as tdfs_count (tuning, statistics and edge range included).  All
 * calls on one GPU share the array: calls that overlap on one GPU (two
 * contexts or threads) race; the array keeps the last call's contents, and
 * each launch clears what must start at zero. */
int tdfs_count_noalloc(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plan, int64_t edge_begin,
                       int64_t edge_end, tdfs_result *out);
/* nplans patterns in ONE persistent launch over the resident graph;
 * out[i] per plan, kernel_ms is the launch's total in every entry. */
int tdfs_count_batch(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plans, int32_t nplans,
                     tdfs_result *out);

/* Embedding output (no counterpart in the reference, which only counts:
 * src/gpu_match.cu:331,409,933).  Runs the counting DFS of tdfs_count and
 * writes every embedding it reaches as a row of k vertex ids (M(0), ..,
 * M(k-1), pattern vertices in plan order, ids in the caller's numbering) to
 * rows[0 .. cap*k); the order of rows is unspecified.  Unlabeled plans yield
 * one embedding per automorphism class (the symmetry-broken raw set, whose
 * size times |Aut(P)| is the count); labeled plans every label-preserving
 * one.  *written = min(raw, cap); out->raw / out->count are complete even
 * when rows were cut at cap.  Allocates a device buffer of cap rows per call:
 * meant for small outputs (checks, examples). */
int tdfs_enumerate(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plan, int64_t edge_begin, int64_t edge_end,
                   int32_t *rows, uint64_t cap, uint64_t *written, tdfs_result *out);

/* ---------------- prepared batches (hipGraph) ---------------- */
typedef struct tdfs_batch tdfs_batch;
/* Uploads the plans once and captures the whole launch sequence (reset
 * kernel, persistent match kernel, result copies) into a hipGraph on a stream
 * of the batch's own; every buffer belongs to the batch, so batches replay
 * concurrently with each other and with tdfs_count on the same device.
 * blocks_per_cu > 0 caps the persistent grid (to share the device). */
int tdfs_batch_create(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plans, int32_t nplans,
                      int32_t blocks_per_cu, tdfs_batch **out);
int tdfs_batch_launch(tdfs_batch *b);                 /* start one replay, do not wait */
int tdfs_batch_done(const tdfs_batch *b);             /* 1 finished, 0 running, <0 error */
int tdfs_batch_wait(tdfs_batch *b, tdfs_result *out); /* wait, then results per plan */
int tdfs_batch_run(tdfs_batch *b, tdfs_result *out);  /* launch + wait */
void tdfs_batch_destroy(tdfs_batch *b);

/* ---------------- serving (continuous batching) ---------------- */
typedef struct tdfs_server tdfs_server;
/* `slots` concurrent launches over the resident graph, each on its own
 * stream with its own task ring and a persistent grid of blocks_per_cu
 * workgroups per CU.  Admission is by DFS-stack memory: the server holds an
 * arena of stack_mb MiB (0: enough for every slot to run 7-vertex plans), a
 * launch over plans of at most kmax vertices takes waves x (2,048 + 2 (kmax-3)
 * x dmax) x 4 bytes of it, and waiting queries (a FIFO prefix of at most
 * max_batch) start on an idle slot as soon as such a region is free.  A
 * query whose stacks can never fit is refused with TDFS_E_BUDGET; so is one
 * whose sampled work estimate exceeds budget, when budget > 0.  Device
 * memory is fixed at creation (no allocation while serving).  No counterpart
 * in the reference (one pattern per process, cu_test.cu:36-49 sizes the
 * per-warp stacks). */
int tdfs_serve_create(tdfs_ctx *ctx, const tdfs_graph *g, int32_t slots, int32_t blocks_per_cu,
                      int32_t max_batch, int64_t stack_mb, double budget, tdfs_server **out);
/* Queues one query; returns its ticket. */
int tdfs_serve_submit(tdfs_server *sv, const tdfs_plan *plan, uint64_t *ticket);
/* Same, counting only level-0 directed edges [edge_begin, edge_end) (end < 0:
 * all) in the numbering tdfs_count would use for this plan alone (as
 * tdfs_count's range). */
int tdfs_serve_submit_range(tdfs_server *sv, const tdfs_plan *plan, int64_t edge_begin, int64_t edge_end,
                            uint64_t *ticket);
/* Non-blocking progress: reports finished queries (ticket, result, latency
 * from submit to collection in seconds; at most cap) and starts idle slots on
 * waiting queries.  Each query is reported as soon as its own work has ended,
 * while the rest of its launch may still run (result.kernel_ms is then the
 * host time since its launch started).  Call repeatedly.  A launch that fails
 * on the device reports its queries with TDFS_FAILED and frees its slot; an
 * error return means a launch could not be started and its queries are still
 * queued, at the head, for a later poll (the *nout entries written before are
 * valid either way).  Tests of these paths force them from the host: the
 * environment variables TDFS_SERVE_FAIL_START=n and TDFS_SERVE_FAIL_RUN=n, read
 * at tdfs_serve_create, make the server's n-th launch fail to start, or report
 * its run as failed. */
int tdfs_serve_poll(tdfs_server *sv, uint64_t *tickets, tdfs_result *results, double *latency_s, int32_t cap,
                    int32_t *nout);
/* out4 = {queries waiting, slots busy, slots, queries running and not yet reported} */
int tdfs_serve_status(const tdfs_server *sv, int64_t *out4);
void tdfs_serve_destroy(tdfs_server *sv);

#ifdef __cplusplus
}
#endif
#endif /* TDFS_H_ */
