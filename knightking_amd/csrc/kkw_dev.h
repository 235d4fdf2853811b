// kkw_dev.h -- device-side definitions shared by the kernels of libkkw.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kkw_rng.h"

namespace kkw {

enum : int { KIND_DEEPWALK = 0, KIND_PPR = 1, KIND_NODE2VEC = 2, KIND_METAPATH = 3, KIND_ANY = -1 };
enum : int { START_EQUAL = 0, START_LIST = 1, START_UNIFORM = 2 };

// One alias bucket (32 B, a quarter of a 128-byte line) in two 16-byte
// halves: {p, primary, primary's packed row} and {alias, meta, alias's packed
// row}.  A draw whose coin picks the primary -- always, for uniform buckets --
// reads only the first half; the vertex id and the row (list begin | degree <<
// 40) of the candidate are both there, so a draw and the move it leads to
// cost one 16-byte random access.  meta = edge type of primary | edge type of
// alias << 8 | (on graphs with edge types) type mask of the primary vertex <<
// 16 | type mask of the alias vertex << 24, a vertex's type mask being the OR
// of its out-edges' types (metapath's termination test).  Rows are 0 when the
// graph cannot pack them (2^40 or more edges, or a degree of 2^24 or more).
struct __align__(32) Bucket {
    float p;
    uint32_t primary;
    uint64_t prow;
    uint32_t alias;
    uint32_t meta;
    uint64_t arow;
};
static_assert(sizeof(Bucket) == 32, "bucket is 32 bytes");

static constexpr int ROW_DEG_SHIFT = 40;
static constexpr uint64_t ROW_BEG_MASK = (1ull << ROW_DEG_SHIFT) - 1;

__host__ __device__ __forceinline__ uint64_t pack_row(uint64_t beg, uint64_t end)
{
    return beg | ((end - beg) << ROW_DEG_SHIFT);
}

struct DevGraph {
    uint32_t V;
    uint64_t E;
    const uint64_t *row_ptr;   // [V+1]
    const uint32_t *col;       // [E], sorted per list when sorted != 0
    const float *w;            // [E] or null
    const uint8_t *etype;      // [E] or null
    const uint8_t *vmask;      // [V] or null: OR of (1 << etype) over out-edges
    const Bucket *bucket;      // [E] or null: alias tables of the edge weights
    const Bucket *ubucket;     // [E] or null: uniform buckets (p = 1, alias = primary)
    const float *vsum;         // [V] or null: sequential float sum of weights
    uint32_t symmetric;        // (u,v) is an edge iff (v,u) is (lists sorted)
    uint32_t packed_rows;      // buckets carry the rows of their candidates
    const uint32_t *adjset;    // null, or per-vertex neighbour hash sets (below)
};

// Neighbour hash sets: vertex v owns slots [4 (row_ptr[v] + v), + 4 deg(v) + 4)
// of adjset, an open-addressing table (linear probing, load below 1/4) of its
// neighbour ids; empty slots hold ADJ_EMPTY.  The table of v is found from
// row_ptr[v] alone, starts on a 16-byte block, and a membership test usually
// reads one block: at load 1/4 a probe leaves its first block about a third
// as often as at 1/2 (measured: 5 % fewer L2 requests and 4 % fewer line
// fetches per node2vec launch, for twice the table memory).
static constexpr uint32_t ADJ_EMPTY = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ uint64_t adj_table_start(uint64_t beg, uint64_t v) { return 4 * (beg + v); }
__host__ __device__ __forceinline__ uint32_t adj_table_size(uint64_t deg) { return (uint32_t)(4 * deg + 4); }

__host__ __device__ __forceinline__ uint32_t adj_slot(uint32_t key, uint32_t size)
{
    return (uint32_t)(((uint64_t)(key * 0x9E3779B1u) * size) >> 32);
}

// Device form of a job (host constants precomputed in float).
struct DevJob {
    int32_t kind;
    int32_t use_alias;
    uint32_t walker_begin;
    uint32_t walker_num;
    uint32_t walk_length;
    uint32_t stride;
    uint64_t seed;
    float ppr_ext;
    float upper, lower, inv_p, inv_q, overflow;
    float acc_lo, acc_hi;      // min/max(1, 1/q): darts at or below acc_lo accept,
                               // above acc_hi reject, without the adjacency search
    int32_t start_mode;
    uint32_t n_starts;
    const uint32_t *starts;
    uint32_t n_schemes;
    const uint32_t *scheme_len;
    const uint8_t *scheme_mask;
    uint32_t *paths;
    uint32_t *lens;
    uint64_t chunk_begin;      // first work-queue chunk of this job (persistent mode)
};

struct DevStats {
    unsigned long long steps, trials, searches, probes, row_loads, search_lines, ref_queries, ref_lines;
};

}  // namespace kkw
