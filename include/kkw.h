/*
 * kkw.h -- C ABI of the MI355X random-walk engine (libkkw.so).
 *
 * The reference (KnightKing) exposes no C ABI: its boundary is the C++
 * header-only templates WalkEngine / WalkConfig / WalkerConfig /
 * TransitionConfig / SecondOrderTransitionConfig and the app entry functions
 * node2vec(), deepwalk(), ppr(), metapath().  Each entry point below names the
 * reference interface it replaces (paths relative to the KnightKing tree);
 * the headers in include/knightking/ rebuild the C++ templates on top of these calls.
 *
 * Conventions: every function returns KKW_OK (0) or a negative KKW_E* code
 * and, on failure, leaves a message for kkw_last_error() (per thread).  Host
 * arrays passed in are copied; the caller keeps ownership.  Buffers returned in
 * kkw_paths are owned by the engine until kkw_paths_free().  One engine drives
 * one GPU on its own HIP stream; calls on one engine from several threads are
 * serialised internally, except kkw_submit/kkw_poll/kkw_wait, which are made
 * for concurrent callers.
 */
#ifndef KKW_H
#define KKW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KKW_ABI_VERSION 3   /* 2: kkw_timing.prepare_ms; 3: kkw_timing.envelope_violations */

enum kkw_status {
    KKW_OK = 0,
    KKW_E_INVALID = -1,   /* bad argument or descriptor                       */
    KKW_E_NOGRAPH = -2,   /* no graph loaded / alias tables not prepared       */
    KKW_E_NOMEM = -3,     /* device memory (or the admission budget) exhausted */
    KKW_E_DEVICE = -4,    /* HIP runtime error                                 */
    KKW_E_BUSY = -5,      /* serving queue full (back-pressure)                */
    KKW_E_CANCELLED = -6, /* request cancelled                                 */
    KKW_E_TIMEOUT = -7,   /* kkw_wait timed out                                */
    KKW_E_IO = -8         /* file could not be read or written                 */
};

/* Walk kinds: src/apps/deepwalk.hpp:30, ppr.hpp:30, node2vec.hpp:108,
 * metapath.hpp:70. */
enum kkw_kind { KKW_DEEPWALK = 0, KKW_PPR = 1, KKW_NODE2VEC = 2, KKW_METAPATH = 3 };

/* Start-vertex distributions: get_equal_dist_func (include/walk.hpp:403-411,
 * walker w starts at w % |V|), a caller list indexed by w % n (ppr.hpp:38-50),
 * get_uniform_dist_func (walk.hpp:413-421). */
enum kkw_start { KKW_START_EQUAL = 0, KKW_START_LIST = 1, KKW_START_UNIFORM = 2 };

/* kkw_run modes: one plain launch per job, or all jobs in one persistent
 * work-queue launch. */
enum kkw_mode { KKW_MODE_PLAIN = 0, KKW_MODE_PERSISTENT = 1 };

/* where kkw_run leaves the paths */
enum kkw_out { KKW_OUT_HOST = 0, KKW_OUT_DEVICE = 1 };

typedef struct kkw_engine kkw_engine;

typedef struct kkw_engine_opts {
    int32_t device;          /* HIP device ordinal                                 */
    int32_t verbose;         /* 1: per-phase timings on stderr                     */
    uint64_t hbm_budget;     /* admission budget for serving, bytes (0 = free HBM) */
} kkw_engine_opts;

/* One walk job: the parameters of node2vec()/deepwalk()/ppr()/metapath() and of
 * WalkerConfig(walker_num, ..., init_dist) (include/walk.hpp:147-193). */
typedef struct kkw_job {
    int32_t kind;            /* enum kkw_kind                                        */
    int32_t use_weights;     /* static component = edge weight (alias tables);
                                0 = uniform (get_trivial_static_comp, static_comp.hpp:29-42) */
    uint32_t walker_begin;   /* first walker id (epochs of WalkConfig::rate)         */
    uint32_t walker_num;
    uint32_t walk_length;    /* L; for PPR the step cap (paths stop there, flagged)   */
    float terminate_prob;    /* PPR                                                  */
    float p, q;              /* node2vec                                             */
    uint64_t seed;           /* Philox key of this job's stream                      */
    int32_t start_mode;      /* enum kkw_start                                        */
    uint32_t n_starts;
    const uint32_t *starts;  /* host array, KKW_START_LIST                           */
    uint32_t n_schemes;      /* metapath                                             */
    const uint32_t *scheme_len;  /* [n_schemes], 1..8 states each                    */
    const uint8_t *scheme_mask;  /* [n_schemes * 8]: allowed edge types per state    */
} kkw_job;

/* Paths of one job: PathSet (include/path.hpp:46-84) with one segment in
 * walker-id order.  Path i holds (lengths[i] & 0x7fffffff) vertices (steps + 1)
 * of walker walker_begin + i; bit 31 of lengths[i] flags a path stopped by the
 * PPR cap.  Rows start at vertices + i * stride, or, when offsets is not NULL
 * (packed form, used by the serving loop), at vertices + offsets[i]; packed
 * rows are consecutive but offsets[0] need not be 0.  A served result points
 * into the serving loop's pinned staging memory until kkw_paths_free returns
 * it (results held back for long keep that memory from new queries). */
typedef struct kkw_paths {
    uint32_t walker_begin;
    uint32_t walker_num;
    uint32_t stride;         /* entries per row (row-major form)                     */
    uint32_t on_device;      /* 1: vertices/lengths are device pointers              */
    uint32_t *vertices;      /* [walker_num * stride] or [offsets[walker_num]]       */
    uint32_t *lengths;       /* [walker_num]                                         */
    uint64_t steps;          /* accepted transitions (sum of lengths - 1)            */
    uint64_t truncated;      /* walkers stopped by the cap                           */
    uint64_t *offsets;       /* [walker_num + 1] in the packed form, else NULL       */
    void *priv;
} kkw_paths;

typedef struct kkw_timing {
    double kernel_ms;        /* walk kernel(s), HIP events on the engine stream      */
    double total_ms;         /* host wall time of the last kkw_run / batch launch     */
    double alias_ms;         /* last alias-table build (kkw_prepare or first use)    */
    double graph_ms;         /* last graph build                                     */
    double index_ms;         /* last neighbour hash-set build                        */
    /* instrumented runs (KKW_FLAG_STATS / KKW_BATCH_STATS): accepted steps, trials,
     * adjacency searches, elements they read, distinct 64-B lines they read,
     * row_ptr pairs read; and, for node2vec, the adjacency queries KnightKing's loop
     * would send (step >= 1, dart above the lower bound, candidate not the previous
     * vertex) with the sum over them of max(1, ceil(log2(deg(prev)+1)) - 3) */
    uint64_t steps, trials, searches, search_probes, row_loads, search_lines;
    uint64_t ref_queries, ref_query_lines;
    double prepare_ms;       /* last kkw_prepare, host wall time: the alias tables and
                                the neighbour hash sets build concurrently, so this is
                                less than alias_ms + index_ms                        */
    /* last kkw_program_run: trials whose dynamic component left the rejection
     * envelope (the check of include/walk.hpp:1003-1005): an outlier candidate
     * above overflow / static + upper + 1e-6, or, for a program without an
     * outlier pair, a static draw above the upper bound.  0 for a correct
     * program; anything else means biased paths. */
    uint64_t envelope_violations;
} kkw_timing;

enum kkw_flags { KKW_FLAG_STATS = 1 };

const char *kkw_last_error(void);
int kkw_abi_version(void);

/* Engine lifetime.  Replaces WalkEngine() / GraphEngine() (include/walk.hpp:378-389,
 * include/graph.hpp:236-257) and set_concurrency (walk.hpp:391-396, a no-op here).
 * Device memory freed by graph loads, table builds and batches is kept by the
 * library for reuse on that device (freeing large blocks to the runtime can
 * stall the host for up to a second); the kept blocks go back to the runtime
 * when an allocation would otherwise fail and when the last engine on the
 * device is destroyed. */
int kkw_engine_create(const kkw_engine_opts *opts, kkw_engine **out);
int kkw_engine_destroy(kkw_engine *eng);

/* Graph loading.  Replaces GraphEngine::load_graph (include/graph.hpp:408-530):
 * edges are Edge<real_t> records {u32 src, u32 dst, f32 w} (12 B, weighted=1) or
 * Edge<EmptyData> {u32 src, u32 dst} (8 B), include/type.hpp:43-85; mirrored when
 * make_undirected (graph.hpp:430-443); CSR keeps file order per source
 * (graph.hpp:309-330); sort_adj sorts each list by neighbour id, stably, as
 * node2vec() does before walking (src/apps/node2vec.hpp:122-125). */
int kkw_graph_from_edges(kkw_engine *eng, uint32_t v_num, const void *edges,
                         uint64_t n_edges, int weighted, int make_undirected, int sort_adj);
/* The same from structure-of-arrays input, with optional per-edge types (the
 * meta_info of WeightedMetaData / UnweightedMetaData, src/apps/metascheme.hpp:42-67);
 * w and etype may be NULL. */
int kkw_graph_from_arrays(kkw_engine *eng, uint32_t v_num, const uint32_t *src,
                          const uint32_t *dst, const float *w, const uint8_t *etype,
                          uint64_t n_edges, int make_undirected, int sort_adj);
/* Sort every adjacency list by neighbour id in place (stable), as node2vec()
 * does with process_vertices + std::sort (src/apps/node2vec.hpp:122-125). */
int kkw_graph_sort_adj(kkw_engine *eng);
/* Binary edge file (read_graph, include/storage.hpp:35-55) or text edge list
 * (read_edgelist, storage.hpp:124-157: "src dst[ w]" per line). */
int kkw_graph_load_file(kkw_engine *eng, uint32_t v_num, const char *path, int text,
                        int weighted, int make_undirected, int sort_adj);
/* A CSR the caller already holds (row_ptr[v_num+1], col[e], optional w[e]). */
int kkw_graph_from_csr(kkw_engine *eng, uint32_t v_num, const uint64_t *row_ptr,
                       const uint32_t *col, const float *w);
/* Synthetic undirected R-MAT built on the device (no reference counterpart; the
 * benchmark graphs): 2^scale vertices, edge_factor * 2^scale raw edges, quadrant
 * probabilities a, b, c (d = 1-a-b-c), self loops and duplicates removed, each
 * edge mirrored, lists sorted, weights 1 + 4U per undirected edge. */
int kkw_graph_rmat(kkw_engine *eng, int scale, int edge_factor, double a, double b,
                   double c, uint64_t seed, int weighted);
/* Edge types for metapath: WeightedMetaData/UnweightedMetaData::meta_info
 * (src/apps/metascheme.hpp:42-67), one byte per CSR entry (< 8 types). */
int kkw_graph_set_edge_types(kkw_engine *eng, const uint8_t *etype);
/* Synthetic typing: vertex v gets type floor(ntypes * U(seed, v)); edge (u,v)
 * gets the type of v. */
int kkw_graph_type_by_dst(kkw_engine *eng, int ntypes, uint64_t seed);
/* get_vertex_num / get_edge_num (include/graph.hpp:178-185). */
int kkw_graph_info(kkw_engine *eng, uint32_t *v_num, uint64_t *e_num);
/* Copy the CSR back (csr / vertex_out_degree, graph.hpp:152-156); any pointer may
 * be NULL. */
int kkw_graph_download(kkw_engine *eng, uint64_t *row_ptr, uint32_t *col, float *w,
                       uint8_t *etype);

/* Static component.  Replaces init_alias_tables (include/walk.hpp:538-652), kept
 * resident across runs instead of being rebuilt per random_walk call.  Also
 * builds the neighbour hash sets node2vec's adjacency test uses (unless turned
 * off below), at the same time on a second stream of the engine; both are
 * otherwise built on first use.  Blocks until both are built. */
int kkw_prepare(kkw_engine *eng);
/* Neighbour hash sets for node2vec's "is the candidate a neighbour of the
 * previous vertex" test (the respond_query of src/apps/node2vec.hpp:158-166,
 * there a std::binary_search): 1 builds them now, 0 drops them and makes the
 * test use the binary search over sorted lists.  Same answers either way. */
int kkw_graph_adj_index(kkw_engine *eng, int enable);
/* Copy the alias buckets back (creation order per vertex) and the per-vertex
 * sequential weight sums; any pointer may be NULL.  meta[i] is the edge type
 * of the primary | the edge type of the alias << 8 (0 without edge types). */
int kkw_alias_download(kkw_engine *eng, float *p, uint32_t *primary, uint32_t *alias,
                       uint32_t *meta, float *vsum);

/* The start vertex walker `walker` of a job seeded `seed` takes under
 * KKW_START_UNIFORM on a graph of v_num vertices (get_uniform_dist_func,
 * include/walk.hpp:413-421): the device's own draw, evaluated on the host.
 * Returns 0 when v_num is 0. */
uint32_t kkw_uniform_start(uint64_t seed, uint32_t walker, uint32_t v_num);

/* Run jobs to completion.  Replaces WalkEngine::random_walk (include/walk.hpp:423-436)
 * as called by the four apps.  out[i] receives job i's paths. */
int kkw_run(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode, int out_where,
            int flags, kkw_paths *out);
int kkw_paths_free(kkw_paths *paths);

/* No counterpart in the reference.  kkw_run for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the same jobs, modes and paths, written into device
 * buffers the caller owns, with nothing allocated on the device.  Job i writes
 * walker walker_begin + r's row at paths[i] + r * strides[i] (16-byte aligned;
 * strides[i] a multiple of 4, at least the path capacity walk_length + 1, for
 * PPR with walk_length 0 the default cap + 1) and its length word (as
 * kkw_paths.lengths) at lengths[i][r].  The work-queue counter, the job table,
 * start lists and scheme tables live in one fixed device array of
 * KKW_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; jobs that
 * need more return KKW_E_NOMEM before anything runs.  Tables are never built
 * here: weighted jobs need the alias tables resident (kkw_prepare), otherwise
 * KKW_E_NOGRAPH; node2vec without hash sets uses the binary search (same
 * answers).  Blocks until the walk has finished on the engine stream.  Calls
 * on engines of one GPU in one process share the array and are serialised by
 * the library; a call never depends on what an earlier one left there. */
#define KKW_NOALLOC_SCRATCH_BYTES (4u << 20)
int kkw_run_noalloc(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode,
                    uint32_t *const *paths, const uint32_t *strides, uint32_t *const *lengths);
/* Epochs streamed to a consumer on the GPU.  WalkConfig's output consumer
 * receives one PathSet per epoch (include/walk.hpp:133-138, 826-842); here the
 * job's walkers run in epochs of epoch_walkers (the last one shorter), each
 * epoch's paths stay in HBM, and `consumer` is called on the calling thread
 * as soon as an epoch has been walked, with the epoch's device paths
 * (on_device = 1, valid during the call and for the work it enqueues) and a
 * HIP stream (hipStream_t, passed as void *) on which to enqueue its GPU work:
 * that stream is ordered after the epoch's walk, and the next epoch is
 * already walking on the engine stream while the consumer's work runs.  An
 * epoch's buffers are reused two epochs later, once the work the consumer
 * enqueued on the stream has finished.  A nonzero return from the consumer
 * stops the run (KKW_E_CANCELLED).  Blocks until the last consumer's work is
 * done.  kkw_last_timing's kernel_ms is the sum of the epochs' walks. */
typedef int (*kkw_epoch_consumer)(const kkw_paths *paths, void *stream, void *user);
/* A host copy of device paths (an epoch handed to a consumer, a batch's
 * KKW_OUT_DEVICE paths): row-major, the same stride; free with kkw_paths_free.
 * Synchronous. */
int kkw_paths_download(const kkw_paths *device_paths, kkw_paths *host_out);
int kkw_run_epochs(kkw_engine *eng, const kkw_job *job, uint32_t epoch_walkers, int mode,
                   kkw_epoch_consumer consumer, void *user);

/* PathSet::dump text format (include/path.hpp:85-109): per walker, optional
 * "id len", then " v" per vertex, then newline; fopen mode "w" or "a". */
int kkw_paths_dump(const kkw_paths *paths, const char *file, const char *mode,
                   int with_head_info);
int kkw_last_timing(kkw_engine *eng, kkw_timing *t);

/* Resident batches: the job table, path buffers and (persistent mode) a
 * captured hipGraph stay on the device; kkw_batch_launch replays them. */
typedef struct kkw_batch kkw_batch;
enum kkw_batch_flags {
    KKW_BATCH_GRAPH = 1,     /* capture the persistent launch in a hipGraph and replay it */
    KKW_BATCH_STATS = 2      /* instrumented kernels: trial / search / probe counters     */
};
int kkw_batch_create(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode,
                     int batch_flags, kkw_batch **out);
int kkw_batch_set_shape(kkw_batch *b, int block, int blocks_per_cu, int chunk);
int kkw_batch_launch(kkw_batch *b);          /* asynchronous on the engine stream */
int kkw_batch_sync(kkw_batch *b, double *kernel_ms);
int kkw_batch_paths(kkw_batch *b, int job, int out_where, kkw_paths *out);
int kkw_batch_destroy(kkw_batch *b);

/* Autotuner: times candidate (block, blocks_per_cu, chunk) shapes of the
 * persistent kernel on this batch and keeps the fastest (cache file optional). */
int kkw_autotune(kkw_batch *b, const char *cache_file, int *block, int *blocks_per_cu,
                 int *chunk);

/* Walk programs: a first-order TransitionConfig + WalkerConfig of arbitrary
 * user logic (include/walk.hpp:147-271, the generic random_walk of
 * walk.hpp:423-436 with internal_walk_epoch 891-1080), written as device
 * functions of fixed names (include/kkw_program.h) and compiled at run time
 * for the engine's GPU.  The static component is the graph's edge weights
 * (alias tables) or uniform; everything dynamic is the program's. */
typedef struct kkw_program kkw_program;

typedef struct kkw_prog_job {
    uint32_t walker_begin;   /* walker ids [walker_begin, +walker_num)                      */
    uint32_t walker_num;
    uint32_t max_length;     /* path capacity in steps; a walk still going there is cut     */
    int32_t use_weights;     /* static component: 1 edge weights (alias), 0 uniform         */
    uint64_t seed;
    int32_t start_mode;      /* enum kkw_start (get_equal_dist_func / list / uniform)      */
    uint32_t n_starts;
    const uint32_t *starts;  /* host array, KKW_START_LIST                                  */
    int32_t second_order;    /* 1: extension re-tested on every retry (walk.hpp:1161-1173)  */
    int32_t mode;            /* KKW_MODE_PLAIN (0): one lane per walker for the launch;     */
                             /* KKW_MODE_PERSISTENT: lanes take the next walker from a work */
                             /* queue when theirs ends.  Paths are the same.                */
    float param[8];          /* kkw_ctx.param                                               */
} kkw_prog_job;

/* Compile `source` (C++ device code defining the functions kkw_program.h
 * lists) for the engine's GPU; `options` adds compiler options (e.g.
 * "-DTAGS=4"), may be NULL.  A compile error returns KKW_E_INVALID with the
 * compiler's log in kkw_last_error(). */
int kkw_program_create(kkw_engine *eng, const char *source, const char *options, kkw_program **out);
/* Destroy every program (and batch) of an engine before the engine. */
int kkw_program_destroy(kkw_program *prog);
/* Compiler log of the last kkw_program_create of this program. */
const char *kkw_program_log(const kkw_program *prog);
/* Copy `bytes` of caller data to the device as kkw_ctx.array[slot] (0-7),
 * replacing what the slot held; bytes = 0 clears it.  The vertex_tag-like
 * arrays user closures capture by reference. */
int kkw_program_set_array(kkw_program *prog, int slot, const void *data, uint64_t bytes);
/* Run one job to completion; host paths as kkw_run gives them. */
int kkw_program_run(kkw_program *prog, const kkw_prog_job *job, kkw_paths *out);

/* Measurement aid (no counterpart in the reference): a plain streaming copy of
 * `bytes` (read once, written once) repeated `iters` times on the engine stream;
 * gives the achievable HBM bandwidth and the known-traffic kernel that hardware
 * byte counters are scaled by. */
int kkw_diag_stream_copy(kkw_engine *eng, uint64_t bytes, int iters, double *ms_per_iter);

/* Measurement aid (no counterpart in the reference): independent random
 * 16-byte reads, each from its own 128-byte line of a `bytes` buffer,
 * `reads` per thread over enough threads to fill the device, repeated `iters`
 * times; gives the rate (lines per second) at which the memory system serves
 * such reads -- the ceiling of a kernel whose every access is a data-dependent
 * 16-byte read, as a random walk's are. */
int kkw_diag_random_gather(kkw_engine *eng, uint64_t bytes, int reads, int iters, double *lines_per_s);

/* Serving: a resident loop admits queries (usually PPR) by free HBM and slot
 * count into flights -- one persistent-kernel launch sequence each, up to four
 * on the device at once -- and retires queries as their last walker finishes.
 * round_steps = 0: a flight walks to the end in one launch.  round_steps = n:
 * a flight walks in launches (rounds) of at most n steps per walker; walkers
 * still going are suspended at a step boundary and continued by the next
 * launch, so the persistent kernel of a long flight gives the GPU back between
 * rounds and flights admitted meanwhile run in between.  Paths are the same
 * either way.  Every query of a flight returns when its flight ends; a flight
 * holds at most a quarter of max_walkers_in_flight (a larger query goes
 * alone), and queries of one length class only (expected steps per walker at
 * most 16, or more), so short queries do not wait for long ones. */
typedef struct kkw_serve_opts {
    uint32_t max_walkers_in_flight;  /* slot budget (walkers)                        */
    uint32_t round_steps;            /* walker steps per kernel round (0 = to end)    */
    uint32_t queue_capacity;         /* pending queries before KKW_E_BUSY             */
    uint32_t idle_poll_us;           /* with no query left, keep polling for new ones
                                        this many microseconds before sleeping (0 =
                                        sleep at once; polling spends a host core and
                                        saves the wake-up, ~15 us, of an arrival) */
} kkw_serve_opts;
int kkw_serve_start(kkw_engine *eng, const kkw_serve_opts *opts);
int kkw_submit(kkw_engine *eng, const kkw_job *job, uint64_t *ticket);
int kkw_poll(kkw_engine *eng, uint64_t ticket, int *done);
int kkw_wait(kkw_engine *eng, uint64_t ticket, double timeout_s, kkw_paths *out,
             double *latency_ms);
int kkw_cancel(kkw_engine *eng, uint64_t ticket);
int kkw_serve_stop(kkw_engine *eng);

#ifdef __cplusplus
}
#endif
#endif
