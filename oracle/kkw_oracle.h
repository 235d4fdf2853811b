/*
 * kkw_oracle.h -- CPU restatement of KnightKing's random-walk hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what GPU results are compared
 * against: it is linked or loaded by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg, never by the product library
 * (knightking_amd/libkkw.so), which must run its HIP kernels or fail.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the KnightKing source tree).  The one deliberate difference from the
 * reference is the random number source: KnightKing draws from a per-thread
 * std::mt19937 seeded by std::random_device (include/util.hpp:45-70), which no
 * one can replay.  Here every draw comes from a counter-based Philox4x32-10
 * stream (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11) addressed by (seed, walker, step, trial, tag), so that the
 * GPU kernels and this restatement produce the same paths bit for bit.
 *
 * Pinning (what checks each function, besides bit-exact agreement with the GPU):
 *   orc_philox          Random123 known-answer vectors (tests/golden/philox_kat.txt)
 *   orc_u01/orc_bounded closed forms: exactness of the 24-bit float grid,
 *                       floor(r*n/2^64) against Python big-integer arithmetic
 *   orc_rmat_raw        quadrant frequencies against (a,b,c,d); bit layout
 *   orc_build_undirected  invariants: symmetric, sorted, no self loops or
 *                       duplicates, weights equal in both directions, brute
 *                       force set equality on small inputs
 *   orc_csr_from_edges  file-order CSR and mirroring against a Python restatement
 *                       on karate (tests/golden/karate.txt, the reference's dataset)
 *   orc_alias_build     hand-worked FIFO examples (tests/golden/alias_hand.json),
 *                       mass conservation per edge, p in [0,1)
 *   orc_walk            closed-form transition matrices of the reference tests
 *                       (static walk, node2vec second order, metapath reachability),
 *                       geometric PPR length law, path validity
 *   orc_alias_build_local  the same construction as orc_alias_build (one code
 *                       path, local indices instead of ids): col[b + local]
 *                       equals the id table entry for entry
 *   orc_program_walk (program_oracle.cpp, compiled with a program's source)
 *                       programs restating deepwalk/ppr/node2vec (outlier
 *                       included) reproduce orc_walk bit for bit; the tag walks
 *                       of the reference's bound and outlier tests and the
 *                       return-biased example walk follow their closed forms;
 *                       a planted mistake in a program is rejected
 */
#ifndef KKW_ORACLE_H
#define KKW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DEEPWALK = 0, ORC_PPR = 1, ORC_NODE2VEC = 2, ORC_METAPATH = 3 };
enum { ORC_START_EQUAL = 0, ORC_START_LIST = 1, ORC_START_UNIFORM = 2 };
enum { ORC_TAG_TRIAL = 0, ORC_TAG_EXT = 1, ORC_TAG_INIT = 2, ORC_TAG_OUTLIER = 3 };
enum { ORC_TAG_RMAT = 0x100, ORC_TAG_WEIGHT = 0x101, ORC_TAG_VTYPE = 0x102 };

/* ---- random numbers ---------------------------------------------------- */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_u01(uint32_t u);
uint32_t orc_bounded(uint32_t lo, uint32_t hi, uint32_t n);

/* ---- graphs -------------------------------------------------------------- */
void orc_rmat_raw(int scale, uint64_t n_edges, uint32_t thr_a, uint32_t thr_ab,
                  uint32_t thr_abc, uint64_t seed, uint32_t *src, uint32_t *dst);
float orc_edge_weight(uint64_t seed, uint32_t u, uint32_t v);
uint64_t orc_build_undirected(uint32_t V, const uint32_t *src, const uint32_t *dst,
                              uint64_t n, uint64_t seed, int weighted,
                              uint64_t *row_ptr, uint32_t *col, float *w);
void orc_csr_from_edges(uint32_t V, const uint32_t *src, const uint32_t *dst,
                        const float *w, uint64_t n, int undirected, int sort_adj,
                        uint64_t *row_ptr, uint32_t *col, float *wout);
void orc_vertex_types(uint32_t V, uint32_t ntypes, uint64_t seed, uint8_t *vtype);

/* ---- alias tables -------------------------------------------------------- */
void orc_alias_build(uint32_t V, const uint64_t *row_ptr, const uint32_t *col,
                     const float *w, const uint8_t *etype,
                     float *bp, uint32_t *bprim, uint32_t *balias, uint32_t *bmeta,
                     float *vsum);
/* The same tables with local edge indices in bprim/balias (walk programs). */
void orc_alias_build_local(uint32_t V, const uint64_t *row_ptr, const float *w,
                           float *bp, uint32_t *bprim, uint32_t *balias, float *vsum);

/* ---- walks --------------------------------------------------------------- */
typedef struct orc_graph {
    uint32_t V;
    uint64_t E;
    const uint64_t *row_ptr;
    const uint32_t *col;
    const float *w;          /* may be NULL (unweighted graph)                 */
    const uint8_t *etype;    /* may be NULL (no edge types)                    */
    const float *bp;         /* alias buckets, creation order, per vertex      */
    const uint32_t *bprim;
    const uint32_t *balias;
    const uint32_t *bmeta;   /* primary type | alias type << 8                 */
    const float *vsum;       /* sequential float sum of static weights         */
} orc_graph;

typedef struct orc_job {
    int32_t kind;
    int32_t use_alias;
    uint32_t walker_begin;
    uint32_t walker_num;
    uint32_t walk_length;    /* L, or the step cap for PPR                    */
    uint32_t path_stride;    /* >= walk_length + 1                            */
    uint64_t seed;
    float ppr_ext;           /* 1 - terminate_prob, computed in float         */
    float n2v_upper, n2v_lower, n2v_inv_p, n2v_inv_q, n2v_overflow;
    int32_t start_mode;
    uint32_t n_starts;
    const uint32_t *starts;
    uint32_t n_schemes;
    const uint32_t *scheme_len;   /* [n_schemes]                              */
    const uint8_t *scheme_mask;   /* [n_schemes][8] allowed-type bitmask      */
    const uint8_t *vmask;         /* [V] OR of out-edge type bits             */
    uint32_t walker_stride;       /* row i walks walker id begin + i*stride (0 = 1);
                                     lets the CPU baseline sample a workload    */
} orc_job;

typedef struct orc_stats {
    uint64_t steps;
    uint64_t trials;
    uint64_t searches;       /* adjacency tests that needed the binary search */
    uint64_t search_probes;  /* col[] elements read by those searches         */
    uint64_t truncated;      /* walkers stopped by the step cap               */
} orc_stats;

void orc_node2vec_constants(float p, float q, float *upper, float *lower,
                            float *inv_p, float *inv_q, float *overflow);
void orc_walk(const orc_graph *g, const orc_job *job, uint32_t *paths,
              uint32_t *lens, int threads, orc_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
