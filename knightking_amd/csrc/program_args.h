// program_args.h -- kernel arguments of a walk program's walk loop; shared by
// the host code of libkkw (hipcc) and the run-time compiled kernel (hipRTC).
#pragma once

#include "kkw_program.h"

// Layout of the library's 32-byte alias bucket.  Walk programs that read edge
// indices use tables whose primary/alias fields hold local edge indices and
// whose prow/arow fields hold the two candidates' vertex ids; programs built
// with KKW_EDGE_INDEX 0 use the engine's tables (vertex ids, and the
// candidates' packed rows when the graph packs them).
struct kkw_prog_bucket {
    float p;
    uint32_t primary;
    uint64_t prow;
    uint32_t alias;
    uint32_t meta;
    uint64_t arow;
};
// the walk loop reads each 16-byte half with one load
static_assert(sizeof(kkw_prog_bucket) == 32 && __builtin_offsetof(kkw_prog_bucket, alias) == 16,
              "bucket halves are {p, primary, prow} and {alias, meta, arow}");

// Persistent form on rows of whole 128-byte lines: path entries wait in LDS in
// 64-byte half lines, KKW_PROG_HALF_LINE + 1 words per lane of a block of
// KKW_PROG_BLOCK lanes.
#define KKW_PROG_BLOCK 256u
#define KKW_PROG_HALF_LINE 16u

// A packed row is beg | deg << KKW_PROG_ROW_DEG_SHIFT (the engine's bucket rows).
#define KKW_PROG_ROW_DEG_SHIFT 40u

struct kkw_prog_args {
    kkw_ctx c;
    const kkw_prog_bucket *lb;   // alias tables of the weights (local indices), or null: uniform;
                                 // KKW_EDGE_INDEX 0: the engine's alias or uniform buckets, never null
    const float *vsum;           // per-vertex sequential weight sums, null when lb is
    uint64_t seed;
    uint32_t walker_begin;
    uint32_t walker_num;
    uint32_t max_len;            // path capacity in steps
    uint32_t stride;             // path row length (entries)
    int32_t start_mode;          // 0: id % V, 1: starts[id % n_starts], 2: uniform draw
    uint32_t n_starts;
    const uint32_t *starts;
    int32_t retest_ext;          // extension re-tested on every retry (second order)
    uint32_t chunk;              // persistent form: walkers per work item
    uint32_t *paths;             // [walker_num * stride]
    uint32_t *lens;              // [walker_num]: vertices in the path | 1 << 31 if cut
    unsigned long long *queue;   // persistent form: next work item (zero at launch)
    uint64_t n_chunks;           // persistent form: work items
    uint32_t line_rows;          // persistent form: rows of whole lines, written from LDS
    uint32_t packed_rows;        // KKW_EDGE_INDEX 0: buckets carry rows, beg | deg << 40
    unsigned long long *violations;   // trials whose dynamic component left its envelope (zero at launch)
};
