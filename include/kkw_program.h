/* kkw_program.h -- what a walk program sees.
 *
 * A walk program is the device form of a first-order TransitionConfig plus its
 * WalkerConfig (include/walk.hpp:147-271 of KnightKing): the parts a user
 * writes as closures there -- walker state init/update, extension_comp,
 * dynamic_comp, its upper/lower bounds and the outlier pair -- are written
 * here as functions of fixed names, compiled for the GPU at run time
 * (kkw_program_create) together with the library's walk loop.  The static
 * component is the graph's edge weights (alias tables) or uniform.
 *
 * The same source compiles on the host (KKW_FN becomes static inline), which
 * is how a CPU replay of a program is built.
 *
 * A program defines:
 *
 *   struct WalkerState { ... };          // trivially copyable, per walker
 *   KKW_FN void init_state(WalkerState &s, const kkw_walker &w, uint32_t start,
 *                          const kkw_ctx &c);                // walker_init_state_func
 *   KKW_FN void update_state(WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                            const kkw_edge &e, const kkw_ctx &c); // walker_update_state_func
 *   KKW_FN float extension(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                          const kkw_ctx &c);                // extension_comp
 *
 * and, after `#define KKW_DYNAMIC 1`:
 *
 *   KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                             const kkw_edge &e, const kkw_ctx &c);
 *   KKW_FN float upper_bound(uint32_t v, const kkw_ctx &c);  // dcomp_upperbound_func
 *
 * optionally `#define KKW_LOWER 1` with
 *
 *   KKW_FN float lower_bound(uint32_t v, const kkw_ctx &c);  // dcomp_lowerbound_func
 *
 * and optionally `#define KKW_OUTLIER 1` with
 *
 *   KKW_FN void outlier_bound(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                             float &overflow, uint32_t &num, const kkw_ctx &c);
 *   KKW_FN int64_t outlier_search(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                                 uint32_t idx, const kkw_ctx &c);  // edge or -1
 *
 * A second-order transition (SecondOrderTransitionConfig, include/walk.hpp:
 * 273-363, the query / response supersteps of walk.hpp:1082-1381) defines,
 * after `#define KKW_SECOND_ORDER 1` (with KKW_DYNAMIC), its query and
 * response data and the three functions that pass them:
 *
 *   struct Query { ... };                // stateQuery<query_data_t>::data
 *   struct Response { ... };             // stateResponse<response_data_t>::data
 *   KKW_FN bool post_query(const WalkerState &s, const kkw_walker &w, uint32_t cur,
 *                          const kkw_edge &e, uint32_t &target, Query &q,
 *                          const kkw_ctx &c);    // post_query_func: true = emit(target, q)
 *   KKW_FN void respond_query(uint32_t v, const Query &q, Response &r,
 *                             const kkw_ctx &c);  // respond_query_func at vertex v
 *   KKW_FN float dynamic_comp(const WalkerState &s, const kkw_walker &w, const Response &r,
 *                             uint32_t cur, const kkw_edge &e, const kkw_ctx &c);
 *
 * in place of the first-order dynamic_comp.  A trial the lower bound does
 * not accept posts its query, the target vertex answers it, and the dart is
 * compared with dynamic_comp given the answer (walk.hpp:1226-1257,
 * 1336-1364).  On one GPU the query and the answer are function calls: the
 * target's data is read directly (kkw_ctx holds the whole graph).  A trial
 * whose post_query sends nothing sees the walker's previous response (value-
 * initialised when the walker starts), as the reference's per-walker
 * response cache holds it.  The extension is re-tested on every retry
 * (walk.hpp:1161-1173), whatever kkw_prog_job.second_order says.
 *
 * `e` is the candidate edge, KnightKing's AdjUnit: e.neighbour the vertex
 * it leads to, e.index its global CSR position -- its weight is
 * c.weight[e.index] (when the graph has weights), its type c.etype[e.index],
 * and per-edge caller data is indexed the same way.  outlier_search returns
 * an index.  In update_state, w.step is already the new step
 * (walk.hpp:1039-1043); in the other functions it is the step being taken.
 *
 * A program whose functions never read e.index may say so with
 * `#define KKW_EDGE_INDEX 0`.  Its walk loop then draws from the engine's own
 * tables, whose buckets hold each candidate's vertex id and CSR row, so a
 * step reads no row_ptr entry of the vertex it moves to; e.index is then
 * KKW_NO_EDGE_INDEX for static draws (an outlier_search result keeps its
 * index).  Paths are the same in both forms.
 */
#ifndef KKW_PROGRAM_H
#define KKW_PROGRAM_H

#if defined(__HIPCC_RTC__)
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#define KKW_FN __device__ __forceinline__
#elif defined(__HIPCC__)
#include <stdint.h>
#define KKW_FN __device__ __forceinline__
#else
#include <stdint.h>
#define KKW_FN static inline
#endif

#define KKW_PROGRAM_ARRAYS 8
#define KKW_PROGRAM_PARAMS 8
#define KKW_NO_EDGE_INDEX 0xFFFFFFFFFFFFFFFFull   /* e.index of a program built with KKW_EDGE_INDEX 0 */

/* The graph and the caller's data, as the program's functions read them.  On
 * the device every pointer is a device address. */
typedef struct kkw_ctx {
    uint32_t v_num;
    uint32_t sorted;           /* 1: every list is sorted by neighbour id */
    uint64_t e_num;
    const uint64_t *row_ptr;   /* [v_num + 1] */
    const uint32_t *col;       /* [e_num] */
    const float *weight;       /* [e_num], or null for an unweighted graph */
    const uint8_t *etype;      /* [e_num], or null */
    const uint32_t *adjset;    /* neighbour hash sets (below), or null */
    const void *array[KKW_PROGRAM_ARRAYS];   /* kkw_program_set_array slots (null if unset) */
    float param[KKW_PROGRAM_PARAMS];         /* kkw_prog_job.param */
    /* Kept by the walk loop for kkw_adjacent, per walker: the CSR row of the
     * vertex the walker has just left (row_v; KKW_NO_VERTEX before its first
     * step), so a second-order program's adjacency test on that vertex reads
     * no row_ptr entry.  Programs do not write them. */
    uint32_t row_v;
    uint32_t row_deg;
    uint64_t row_beg;
} kkw_ctx;

#define KKW_NO_VERTEX 0xFFFFFFFFu

/* 1 if u is in v's list, else 0 -- the adjacency test of a second-order
 * walk (node2vec's respond_query, src/apps/node2vec.hpp:158-166).  Reads the
 * engine's neighbour hash sets when present: v's table of 4 deg(v) + 4 slots
 * starts at 4 (row_ptr[v] + v), slot of a key = (key * 0x9E3779B1 mod 2^32) *
 * size / 2^32, linear probing, 0xFFFFFFFF empty.  Otherwise a binary search
 * (sorted lists) or a scan.  The answer does not depend on the method. */
KKW_FN int kkw_adjacent(uint32_t v, uint32_t u, const struct kkw_ctx &c)
{
    const bool kept = v == c.row_v;   /* the row the walk loop already holds */
    const uint64_t b = kept ? c.row_beg : c.row_ptr[v];
    const uint64_t e = kept ? c.row_beg + c.row_deg : c.row_ptr[v + 1];
    if (e == b) return 0;
    if (c.adjset) {
        const uint32_t size = (uint32_t)(4 * (e - b) + 4);
        const uint64_t tb = 4 * (b + (uint64_t)v);   /* table start in adjset */
        uint32_t s = (uint32_t)(((uint64_t)(u * 0x9E3779B1u) * size) >> 32);
#if defined(__HIPCC_RTC__) || defined(__HIPCC__)
        /* one aligned 16-byte block (four slots) per read, slots in probe order */
        for (uint32_t seen = 0; seen < size;) {
            const uint64_t a = tb + s, blk = a & ~(uint64_t)3;
            const uint4 q = *(const uint4 *)(c.adjset + blk);
            const uint32_t e4[4] = {q.x, q.y, q.z, q.w};
            for (uint32_t k = (uint32_t)(a - blk); k < 4 && blk + k < tb + size; k++, seen++) {
                if (e4[k] == u) return 1;
                if (e4[k] == 0xFFFFFFFFu) return 0;
            }
            s = blk + 4 >= tb + size ? 0u : (uint32_t)(blk + 4 - tb);
        }
#else
        for (uint32_t n = 0; n < size; n++) {
            const uint32_t x = c.adjset[tb + s];
            if (x == u) return 1;
            if (x == 0xFFFFFFFFu) return 0;
            s = s + 1 == size ? 0 : s + 1;
        }
#endif
        return 0;
    }
    if (c.sorted) {
        uint64_t lo = b, len = e - b;
        while (len > 0) {
            const uint64_t half = len >> 1;
            if (c.col[lo + half] < u) {
                lo += half + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        return lo < e && c.col[lo] == u;
    }
    for (uint64_t k = b; k < e; k++)
        if (c.col[k] == u) return 1;
    return 0;
}

/* AdjUnit<T> (include/graph.hpp:49-63): the neighbour and where the edge's
 * data lives. */
typedef struct kkw_edge {
    uint64_t index;       /* global CSR position */
    uint32_t neighbour;   /* c.col[index] */
    uint32_t reserved;
} kkw_edge;

/* Walker<T>::id and ::step (include/walk.hpp:31-50). */
typedef struct kkw_walker {
    uint32_t id;
    uint32_t step;
} kkw_walker;

#endif
