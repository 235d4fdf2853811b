// walk.hip -- the walker step loop on the GPU.
//
// One lane owns one walker.  advance() performs one trial of the walker's
// current step: termination test, candidate draw (alias bucket or uniform
// index), and the rejection test against the upper/lower bounds with the
// node2vec adjacency test (binary search in the previous vertex's sorted list)
// or the metapath type test.  It restates, per walker, the first-order loop of
// KnightKing include/walk.hpp:954-1064 and the three-phase second-order loop of
// include/walk.hpp:1135-1379, with the app functions of src/apps/deepwalk.hpp,
// ppr.hpp, node2vec.hpp and metapath.hpp.
//
// Two launch forms:
//  * plain: one launch per job, one walker per lane, the lane exits when its
//    walker ends (lanes of finished walkers idle until the wave ends);
//  * persistent: a grid sized to the machine pulls chunks of walkers from a
//    global work queue that spans all jobs of a batch; whenever lanes of a wave
//    retire a walker (found by ballot), the wave refills them from its chunk, so
//    waves stay full until the queue drains.
#include "kkw_internal.h"

#include <hipcub/hipcub.hpp>

#include <cstdlib>
#include <cstring>

namespace kkw {

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n);

// Path write-out: each lane keeps the current group of four path entries in
// registers and stores it as one aligned 16-byte write when the group fills or
// the walk ends (rows hold a multiple of four entries; entries past a path's
// length are left unspecified).
struct PathGroup {
    uint32_t q0, q1, q2, q3;
    __device__ __forceinline__ void set(uint32_t i, uint32_t v)
    {
        const uint32_t k = i & 3u;
        q0 = k == 0 ? v : q0;
        q1 = k == 1 ? v : q1;
        q2 = k == 2 ? v : q2;
        q3 = k == 3 ? v : q3;
    }
    __device__ __forceinline__ void store(uint32_t *row, uint32_t i) const
    {
        *reinterpret_cast<uint4 *>(row + (i & ~3u)) = make_uint4(q0, q1, q2, q3);
    }
};

// Path write-out in whole 128-byte lines (rows of a multiple of 32 entries):
// the lane's current 32 entries wait in LDS and leave as eight 16-byte stores
// to one line when the line fills, so full lines reach HBM whole -- a line
// written 16 bytes at a time by a lane, microseconds apart, is written back
// piecemeal, and partial-line writes cost HBM far more than their bytes.  The
// line that ends a walk is written up to its last 16-byte group.  Lanes sit 33
// words apart in LDS, so lanes at the same step use different banks.  Entries
// past a path's length are unspecified.
constexpr uint32_t LINE_ENTRIES = 32, LINE_LDS_STRIDE = LINE_ENTRIES + 1;
// The trial-loop kernels buffer 64-byte half lines: 17 words per lane keep
// 256-thread blocks at eight per CU (their best occupancy), where whole lines
// would allow four.
constexpr uint32_t HALF_LINE_ENTRIES = 16;

// N entries per buffered chunk (a power of two, a multiple of 4)
template <uint32_t N>
struct LineBufN {
    uint32_t *s;   // this lane's N entries in LDS
    __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { s[i & (N - 1)] = v; }
    // the groups of i's chunk up to the one holding entry i (the whole chunk
    // when i is its last entry)
    __device__ __forceinline__ void store(uint32_t *row, uint32_t i) const
    {
        uint4 *dst = reinterpret_cast<uint4 *>(row + (i & ~(N - 1)));
        const uint32_t groups = ((i & (N - 1)) >> 2) + 1;
#pragma unroll
        for (uint32_t k = 0; k < N / 4; k++)
            if (k < groups) dst[k] = make_uint4(s[4 * k], s[4 * k + 1], s[4 * k + 2], s[4 * k + 3]);
    }
};
using LineBuf = LineBufN<LINE_ENTRIES>;

// Dynamic LDS of the line-buffered kernels: N + 1 words per thread for chunks
// of N entries.
extern __shared__ uint32_t lines_lds[];

template <uint32_t N = LINE_ENTRIES>
__device__ __forceinline__ LineBufN<N> lane_line() { return LineBufN<N>{lines_lds + threadIdx.x * (N + 1)}; }

// Buffered chunks written by the wave together.  A lane whose chunk of N
// entries is complete, or whose walk has ended, sets `groups` to the number of
// 16-byte groups to write (N / 4 for a full chunk); then, at one point of the
// wave's loop, every pending chunk leaves with N / 4 lanes writing 16 bytes
// each, 64 / (N / 4) chunks per store instruction.  The chunk of lane src is
// still in src's LDS buffer (the wave calls this before any lane buffers
// another entry) and goes to chunk_of() as src evaluates it.  One lane
// writing its own chunk took N / 4 store instructions, each with its own
// address translation request: the state-machine kernel on R-MAT-24 node2vec
// runs 2.5 % faster this way.  (The trial-loop kernel, with half lines, did
// not gain: 64 mixed jobs took 1-4 % longer.)
template <uint32_t N, class ChunkOf>
__device__ __forceinline__ void write_pending_chunks(uint32_t &groups, ChunkOf chunk_of)
{
    constexpr uint32_t PER = N / 4;        // lanes per chunk, one 16-byte group each
    constexpr uint32_t SLOTS = 64 / PER;   // chunks per store instruction
    const uint32_t lane = threadIdx.x & 63;
    uint64_t pm = __ballot(groups != 0);
    if (!pm) return;
    const unsigned long long own = groups ? (unsigned long long)(uintptr_t)chunk_of() : 0ull;
    while (pm) {
        const uint32_t slot = lane / PER, part = lane % PER;
        uint64_t m = pm;   // the slot-th pending lane, if there is one
        for (uint32_t k = 0; k < slot && m; k++) m &= m - 1;
        const int src = m ? (int)__builtin_ctzll(m) : 0;
        const uint32_t grp = __shfl(groups, src, 64);
        const unsigned long long dst = __shfl(own, src, 64);
        if (m && part < grp) {
            const uint32_t *s = lines_lds + ((threadIdx.x & ~63u) + (uint32_t)src) * (N + 1) + part * 4;
            *reinterpret_cast<uint4 *>(reinterpret_cast<uint32_t *>((uintptr_t)dst) + part * 4) =
                make_uint4(s[0], s[1], s[2], s[3]);
        }
        uint64_t rest = pm;   // pending lanes past the first SLOTS
        for (uint32_t k = 0; k < SLOTS && rest; k++) rest &= rest - 1;
        if (((pm & ~rest) >> lane) & 1ull) groups = 0;
        pm = rest;
    }
}

struct Walker {
    uint32_t job;      // job index (persistent batches)
    uint32_t idx;      // row within the job
    uint32_t wid;      // walker id (Philox counter)
    uint32_t cur, prev;
    uint64_t beg, end; // list of cur
    uint64_t pbeg, pend; // list of prev
    uint32_t step, trial;
    uint32_t scheme, state;
    uint32_t cvm;      // metapath: type mask of cur (OR of its out-edges' types)
    uint32_t slen;     // metapath: length of the walker's scheme
    uint64_t smask;    // metapath: the scheme's type masks, state k in bits 8k..8k+7
    bool lines;        // row written through the lane's LDS line buffer
    PathGroup pg;
};

struct LaneStats {
    uint32_t steps, trials, searches, probes, rows, lines;
    uint32_t ref_queries, ref_lines;   // reference byte model: queries KnightKing would send, sum of L_b
};

// Lines the reference's binary search over a list of d entries is charged:
// max(1, ceil(log2(d+1)) - 3) (the last four levels share one 64-byte line).
__device__ __forceinline__ uint32_t ref_search_lines(uint64_t d)
{
    uint32_t bits = d ? 64u - (uint32_t)__clzll(d) : 0u;
    return bits > 4 ? bits - 3 : 1u;
}

static constexpr uint32_t MAX_TRIALS = 1u << 20;  // per step; guards degenerate inputs

// Membership of x in the sorted list col[lo..hi) (std::binary_search of
// node2vec.hpp:158-166).  probes counts elements read, lines the distinct
// 64-byte lines they fall in (instrumented builds only use the counts).
__device__ __forceinline__ bool adj_contains(const uint32_t *__restrict__ col, uint64_t lo,
                                             uint64_t hi, uint32_t x, uint32_t &probes, uint32_t &lines)
{
    const uint32_t *base = col + lo;
    uint32_t b = 0, len = (uint32_t)(hi - lo);
    uintptr_t last = ~(uintptr_t)0;
    while (len > 0) {
        uint32_t half = len >> 1;
        uint32_t v = base[b + half];
        probes++;
        uintptr_t line = (uintptr_t)(base + b + half) >> 6;
        lines += line != last;
        last = line;
        if (v == x) return true;
        if (v < x) {
            b += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return false;
}

// Membership of x among prev's neighbours: the neighbour hash set when the
// graph has one, else the binary search over prev's sorted list.
__device__ __forceinline__ bool adj_member(const DevGraph &g, uint32_t prev, uint64_t pbeg, uint64_t pend,
                                           uint32_t x, uint32_t &probes, uint32_t &lines)
{
    if (g.adjset) {
        const uint32_t *tab = g.adjset + adj_table_start(pbeg, prev);
        const uint32_t size = adj_table_size(pend - pbeg);
        uint32_t s = adj_slot(x, size);
        uintptr_t last = ~(uintptr_t)0;
        for (uint32_t n = 0; n < size; n++) {   // load <= 1/2: an empty slot ends it
            const uint32_t e = tab[s];
            probes++;
            uintptr_t line = (uintptr_t)(tab + s) >> 6;
            lines += line != last;
            last = line;
            if (e == x) return true;
            if (e == ADJ_EMPTY) return false;
            s = s + 1 == size ? 0 : s + 1;
        }
        return false;
    }
    return adj_contains(g.col, pbeg, pend, x, probes, lines);
}

// A metapath walker's scheme, read once when the walker starts or resumes
// (lengths are 1..8, masks 8 per scheme): the masks ride in registers, so a
// trial reads no scheme table.
__device__ __forceinline__ void load_scheme(const DevJob &J, Walker &s)
{
    const uint8_t *m = J.scheme_mask + (uint64_t)s.scheme * 8;
    uint64_t v = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) v |= (uint64_t)m[k] << (8 * k);
    s.smask = v;
    s.slen = J.scheme_len[s.scheme];
}

// the type mask of the walker's current scheme state (metapath.hpp:96-99)
__device__ __forceinline__ uint32_t state_mask(const Walker &s)
{
    return (uint32_t)(s.smask >> (8 * s.state)) & 0xFFu;
}

__device__ __forceinline__ uint32_t start_vertex(const DevGraph &g, const DevJob &J, uint32_t wid)
{
    if (J.start_mode == START_LIST) return J.starts[wid % J.n_starts];
    if (J.start_mode == START_UNIFORM) {
        u32x4 r = draw(J.seed, wid, 0, 0, TAG_INIT);
        return bounded(r.z, r.w, g.V);
    }
    return wid % g.V;
}

// lines_ok: the kernel was launched with the LDS line buffer; rows of a
// multiple of 32 entries then leave as whole lines.
template <int KIND>
__device__ __forceinline__ void init_walker(const DevGraph &g, const DevJob &J, uint32_t job,
                                            uint32_t idx, Walker &s, LaneStats &ls, bool lines_ok = false)
{
    s.job = job;
    s.idx = idx;
    s.wid = J.walker_begin + idx;
    s.cur = start_vertex(g, J, s.wid);
    s.prev = 0;
    s.pbeg = s.pend = 0;
    s.step = 0;
    s.trial = 0;
    s.scheme = 0;
    s.state = 0;
    s.slen = 1;
    s.smask = 0;
    const int kind = KIND == KIND_ANY ? J.kind : KIND;
    s.cvm = 0;
    if (kind == KIND_METAPATH) {
        u32x4 r = draw(J.seed, s.wid, 0, 0, TAG_INIT);
        s.scheme = bounded(r.x, r.y, J.n_schemes);
        load_scheme(J, s);
        s.cvm = g.vmask[s.cur];   // later vertices' masks come with their buckets
    }
    s.beg = g.row_ptr[s.cur];
    s.end = g.row_ptr[s.cur + 1];
    ls.rows++;
    s.lines = lines_ok && J.stride % LINE_ENTRIES == 0;
    if (s.lines) lane_line<HALF_LINE_ENTRIES>().set(0, s.cur);
    else s.pg.q0 = s.pg.q1 = s.pg.q2 = s.pg.q3 = s.cur;
}

__device__ __forceinline__ void finish_walker(const DevJob &J, const Walker &s, bool truncated)
{
    uint32_t *row = J.paths + (uint64_t)s.idx * J.stride;
    if (s.lines) {
        if ((s.step & (HALF_LINE_ENTRIES - 1)) != HALF_LINE_ENTRIES - 1) lane_line<HALF_LINE_ENTRIES>().store(row, s.step);
    } else if ((s.step & 3u) != 3u) {
        s.pg.store(row, s.step);
    }
    J.lens[s.idx] = (s.step + 1) | (truncated ? 0x80000000u : 0u);
}

// Move to cand; its row comes from the bucket when row_ok, else from row_ptr.
__device__ __forceinline__ void move_to(const DevGraph &g, const DevJob &J, Walker &s,
                                        uint32_t cand, LaneStats &ls, bool row_ok = false, uint64_t row = 0)
{
    s.step++;
    if (s.lines) {
        const auto lb = lane_line<HALF_LINE_ENTRIES>();
        lb.set(s.step, cand);
        if ((s.step & (HALF_LINE_ENTRIES - 1)) == HALF_LINE_ENTRIES - 1)
            lb.store(J.paths + (uint64_t)s.idx * J.stride, s.step);
    } else {
        s.pg.set(s.step, cand);
        if ((s.step & 3u) == 3u) s.pg.store(J.paths + (uint64_t)s.idx * J.stride, s.step);
    }
    s.prev = s.cur;
    s.pbeg = s.beg;
    s.pend = s.end;
    s.cur = cand;
    if (row_ok) {
        s.beg = row & ROW_BEG_MASK;
        s.end = s.beg + (row >> ROW_DEG_SHIFT);
    } else {
        s.beg = g.row_ptr[cand];
        s.end = g.row_ptr[cand + 1];
        ls.rows++;
    }
    s.trial = 0;
    ls.steps++;
}

// 16-byte load from a 4-byte aligned address (one global_load_dwordx4)
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ u32x4 load16(const uint32_t *p)
{
    u32x4_a4 v = *reinterpret_cast<const u32x4_a4 *>(p);
    return u32x4{v.x, v.y, v.z, v.w};
}

// Candidate by alias bucket or uniform index (include/walk.hpp:1016-1029).
// Uniform draws read the uniform buckets (p = 1: always the primary, which is
// col[beg + k]) when they exist.  row_ok/row: the candidate's packed row.
// etype (metapath): the edge's type | the candidate's type mask << 8.
template <bool ALIAS_KNOWN, bool ALIAS>
__device__ __forceinline__ uint32_t pick(const DevGraph &g, const DevJob &J, const Walker &s,
                                         uint32_t deg, const u32x4 &r, uint32_t *etype, bool &row_ok,
                                         uint64_t &row)
{
    uint32_t k = bounded(r.x, r.y, deg);
    const bool use_alias = ALIAS_KNOWN ? ALIAS : (J.use_alias != 0);
    const Bucket *tab = use_alias ? g.bucket : g.ubucket;
    if (tab) {
        // first half {p, primary, prow}; the second {alias, meta, arow} only
        // when the coin picks the alias, or -- issued with the first, one
        // memory round trip -- when the edge type is wanted (metapath)
        const uint32_t *bp = reinterpret_cast<const uint32_t *>(tab + s.beg + k);
        const u32x4 a = load16(bp);
        u32x4 h = {0u, 0u, 0u, 0u};
        if (etype) h = load16(bp + 4);
        const bool prim = u01(r.z) < __uint_as_float(a.x);
        row_ok = g.packed_rows != 0;
        if (prim && !etype) {
            row = (uint64_t)a.z | ((uint64_t)a.w << 32);
            return a.y;
        }
        if (!etype) h = load16(bp + 4);
        if (etype) *etype = prim ? ((h.y & 0xffu) | ((h.y >> 8) & 0xff00u)) : (((h.y >> 8) & 0xffu) | ((h.y >> 16) & 0xff00u));
        row = prim ? ((uint64_t)a.z | ((uint64_t)a.w << 32)) : ((uint64_t)h.z | ((uint64_t)h.w << 32));
        return prim ? a.y : h.x;
    }
    row_ok = false;
    row = 0;
    const uint32_t c = g.col[s.beg + k];
    if (etype) *etype = g.etype ? ((uint32_t)g.etype[s.beg + k] | ((uint32_t)g.vmask[c] << 8)) : 0u;
    return c;
}

// One trial.  Returns true when the walker has ended (length written).
template <int KIND, bool ALIAS_KNOWN, bool ALIAS>
__device__ __forceinline__ bool advance(const DevGraph &g, const DevJob &J, Walker &s, LaneStats &ls)
{
    const int kind = KIND == KIND_ANY ? J.kind : KIND;
    const uint32_t deg = (uint32_t)(s.end - s.beg);

    if (kind == KIND_NODE2VEC) {
        // Termination, re-tested on every retry (walk.hpp:1161-1173);
        // extension = step >= L ? 0 : 1 (node2vec.hpp:147-150).
        if (deg == 0 || s.step >= J.walk_length || s.trial >= MAX_TRIALS) {
            finish_walker(J, s, s.trial >= MAX_TRIALS);
            return true;
        }
        ls.trials++;
        uint32_t cand = 0;
        float dart = 0.0f;
        bool have = false, pre = false;
        bool row_ok = false;
        uint64_t crow = 0;
        if (J.overflow > 0.0f && s.step > 0) {
            // Outlier appendix (walk.hpp:1181-1208, node2vec.hpp:47-106).
            const uint32_t *base = g.col + s.beg;
            uint32_t lo = 0, len = deg;
            while (len > 0) {  // lower_bound of prev in the current list
                uint32_t half = len >> 1;
                if (base[lo + half] < s.prev) { lo += half + 1; len -= half + 1; }
                else len = half;
            }
            if (lo < deg && base[lo] == s.prev) {
                const bool use_alias = ALIAS_KNOWN ? ALIAS : (J.use_alias != 0);
                float wr = use_alias ? g.w[s.beg + lo] : 1.0f;
                float ub = use_alias ? wr * J.overflow : J.overflow;
                float regular = use_alias ? g.vsum[s.cur] * J.upper : (float)deg * J.upper;
                if (ub > 0.0f) {
                    u32x4 r = draw(J.seed, s.wid, s.step, s.trial, TAG_OUTLIER);
                    float rv = u01(r.x) * (ub * 1.0f + regular) - regular;
                    if (rv > 0.0f) {
                        uint32_t oidx = (uint32_t)(rv / ub);
                        if (oidx + 1 > 1) oidx = 0;
                        dart = (rv - (float)oidx * ub) / wr + J.upper;
                        cand = s.prev;
                        have = true;
                    }
                }
            }
        }
        if (!have) {
            u32x4 r = draw(J.seed, s.wid, s.step, s.trial, TAG_TRIAL);
            cand = pick<ALIAS_KNOWN, ALIAS>(g, J, s, deg, r, nullptr, row_ok, crow);
            dart = u01(r.w) * J.upper;            // walk.hpp:1225
            pre = dart <= J.lower;                // walk.hpp:1226-1229
        }
        if (!pre && s.step > 0 && cand != s.prev && !have) {
            ls.ref_queries++;
            ls.ref_lines += ref_search_lines(s.pend - s.pbeg);
        }
        bool accept;
        if (pre) {
            accept = true;
        } else if (s.step == 0) {
            accept = dart <= J.upper;             // node2vec.hpp:180-182
        } else if (cand == s.prev) {
            accept = dart <= J.inv_p;             // node2vec.hpp:185-186
        } else if (dart <= J.acc_lo) {
            accept = true;                        // P_d is 1 or 1/q, both >= dart
        } else if (dart > J.acc_hi) {
            accept = false;                       // both < dart
        } else {
            uint32_t pr = 0, ln = 0;
            bool found = adj_member(g, s.prev, s.pbeg, s.pend, cand, pr, ln);
            ls.searches++;
            ls.probes += pr;
            ls.lines += ln;
            accept = dart <= (found ? 1.0f : J.inv_q);   // node2vec.hpp:187-191
        }
        if (accept) move_to(g, J, s, cand, ls, row_ok, crow);
        else s.trial++;
        return false;
    }

    // First-order kinds: termination once per step (walk.hpp:956-973).  The
    // trial's draw does not depend on the termination test's outcome: it is
    // computed first, so that the two Philox blocks of a PPR step overlap
    // (PPR-only batches 4 % faster; a walker that stops discards it).
    const u32x4 r = draw(J.seed, s.wid, s.step, s.trial, TAG_TRIAL);
    if (s.trial == 0) {
        bool stop = deg == 0;
        if (!stop) {
            float ext;
            if (kind == KIND_PPR) {
                ext = J.ppr_ext;                                  // ppr.hpp:51-55
            } else if (kind == KIND_METAPATH) {
                ext = (s.step >= J.walk_length || !(s.cvm & state_mask(s))) ? 0.0f : 1.0f;
            } else {
                ext = s.step >= J.walk_length ? 0.0f : 1.0f;      // deepwalk.hpp:35-37
            }
            if (ext == 0.0f) {
                stop = true;
            } else if (ext < 1.0f) {
                u32x4 r = draw(J.seed, s.wid, s.step, 0, TAG_EXT);
                stop = u01(r.x) > ext;
            }
        }
        if (stop) {
            finish_walker(J, s, false);
            return true;
        }
        if (s.step >= J.walk_length) {   // step cap (only PPR can reach it)
            finish_walker(J, s, true);
            return true;
        }
    } else if (s.trial >= MAX_TRIALS) {
        finish_walker(J, s, true);
        return true;
    }
    ls.trials++;
    bool row_ok = false;
    uint64_t crow = 0;
    if (kind == KIND_METAPATH) {
        uint32_t et = 0;
        uint32_t cand = pick<ALIAS_KNOWN, ALIAS>(g, J, s, deg, r, &et, row_ok, crow);
        float dart = u01(r.w) * 1.0f;
        float pd = ((state_mask(s) >> (et & 0xffu)) & 1u) ? 1.0f : 0.0f;
        if (dart <= pd) {                                          // walk.hpp:1030-1036
            move_to(g, J, s, cand, ls, row_ok, crow);
            s.cvm = et >> 8;                                       // the new vertex's type mask
            s.state = (s.state + 1) % s.slen;                      // metapath.hpp:96-99
        } else {
            s.trial++;
        }
        return false;
    }
    uint32_t cand = pick<ALIAS_KNOWN, ALIAS>(g, J, s, deg, r, nullptr, row_ok, crow);
    move_to(g, J, s, cand, ls, row_ok, crow);
    return false;
}

template <bool STATS>
__device__ __forceinline__ void flush_stats(const LaneStats &ls, DevStats *st)
{
    if (!STATS) return;
    // wave reduction, one atomic per counter per wave
    unsigned long long v[8] = {ls.steps, ls.trials, ls.searches, ls.probes, ls.rows, ls.lines,
                               ls.ref_queries, ls.ref_lines};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&st->steps, v[0]);
        atomicAdd(&st->trials, v[1]);
        atomicAdd(&st->searches, v[2]);
        atomicAdd(&st->probes, v[3]);
        atomicAdd(&st->row_loads, v[4]);
        atomicAdd(&st->search_lines, v[5]);
        atomicAdd(&st->ref_queries, v[6]);
        atomicAdd(&st->ref_lines, v[7]);
    }
}

// ---------------------------------------------------------------- plain ----
template <int KIND, bool ALIAS, bool STATS>
__global__ void __launch_bounds__(256) k_walk_plain(DevGraph g, DevJob J, DevStats *st)
{
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    LaneStats ls = {0, 0, 0, 0, 0, 0, 0, 0};
    if (idx < J.walker_num) {
        Walker s;
        init_walker<KIND>(g, J, 0, idx, s, ls);
        while (!advance<KIND, true, ALIAS>(g, J, s, ls)) {
        }
    }
    flush_stats<STATS>(ls, st);
}

// ----------------------------------------------------------- persistent ----
__device__ __forceinline__ uint32_t find_job(const DevJob *jobs, int n_jobs, uint64_t chunk)
{
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_begin <= chunk) lo = mid;
        else hi = mid - 1;
    }
    return (uint32_t)lo;
}

// The next work item (chunk of walkers) of this wave.  A wave's first item is
// the one its own index names, taken without an atomic; later ones come from
// the queue, numbered past the grid's own items, so every item is taken once.
// When the first items too came from the queue, a wave whose first walkers
// ended at once (isolated starts) took a second item before the last waves
// had their first; with about one item per wave, the waves left without one
// exited and that wave walked the rest of its second item only after its
// first walkers, doubling the launch (R-MAT-14 DeepWalk, 256 items for 256
// waves: 8 waves with two items, 131 us against 59 us for the plain form).
__device__ __forceinline__ unsigned long long next_chunk(unsigned long long *queue, bool &first)
{
    if (first) {
        first = false;
        return (unsigned long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    }
    unsigned long long c = 0;
    if ((threadIdx.x & 63) == 0) c = atomicAdd(queue, 1ull);
    return __shfl(c, 0, 64) + (unsigned long long)gridDim.x * (blockDim.x >> 6);
}

// Job tables of up to LDS_JOBS jobs are copied into LDS by every block of a
// mixed-kind launch.  Every trial reads its walker's job fields (seed, kind,
// length, path row); read from a table in device memory, each such read needs
// an address translation, and the walk's random reads evict the table's page
// from the per-CU translation cache between trials (R-MAT-23, 64 mixed jobs:
// 29.1 -> 25.5 ms per launch with the table in LDS, PPR-only jobs 6.8 -> 5.6 ms).
constexpr int LDS_JOBS = 64;
static_assert(sizeof(DevJob) % 16 == 0, "job tables are copied in 16-byte pieces");

// The block's copy of jobs[0, n_jobs) in dynamic LDS, `offset_words` words in
// (after the line buffer), or the table itself when JOBS_LDS is false.  Every
// thread of the block must call it (it ends with a barrier).
template <bool JOBS_LDS>
__device__ __forceinline__ const DevJob *block_job_table(const DevJob *jobs, int n_jobs, uint32_t offset_words)
{
    if constexpr (!JOBS_LDS) {
        return jobs;
    } else {
        uint4 *dst = reinterpret_cast<uint4 *>(lines_lds + offset_words);
        const uint4 *src = reinterpret_cast<const uint4 *>(jobs);
        const uint32_t n16 = (uint32_t)((uint64_t)n_jobs * sizeof(DevJob) / 16);
        for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
        return reinterpret_cast<const DevJob *>(dst);
    }
}

// SINGLE: one job (passed by value), otherwise a job table in device memory
// (JOBS_LDS: copied into LDS, at most LDS_JOBS jobs).
// LINES: launched with the LDS line buffer (jobs whose rows are whole lines use it).
template <int KIND, bool SINGLE, bool ALIAS_KNOWN, bool ALIAS, bool STATS, bool LINES, bool JOBS_LDS>
__global__ void __launch_bounds__(256) k_walk_persistent(DevGraph g, DevJob J0, const DevJob *jobs_in,
                                                       int n_jobs, uint64_t n_chunks, uint32_t chunk,
                                                       unsigned long long *queue, DevStats *st)
{
    static_assert(!(SINGLE && JOBS_LDS), "a single job is passed by value");
    const DevJob *const jobs =
        block_job_table<JOBS_LDS>(jobs_in, n_jobs, LINES ? blockDim.x * (HALF_LINE_ENTRIES + 1) : 0u);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    LaneStats ls = {0, 0, 0, 0, 0, 0, 0, 0};
    Walker s;
    bool active = false;
    // wave-uniform pool of walker rows [pool_next, pool_end) of job pool_job
    uint32_t pool_job = 0, pool_next = 0, pool_end = 0;
    bool drained = false, first_chunk = true;
    for (;;) {
        uint64_t need = __ballot(!active);
        while (need && !drained) {
            if (pool_next >= pool_end) {
                const unsigned long long c = next_chunk(queue, first_chunk);
                if (c >= n_chunks) {
                    drained = true;
                    break;
                }
                uint32_t j = SINGLE ? 0u : find_job(jobs, n_jobs, c);
                const DevJob &Jp = SINGLE ? J0 : jobs[j];
                pool_job = j;
                pool_next = (uint32_t)(c - Jp.chunk_begin) * chunk;
                uint32_t e = pool_next + chunk;
                pool_end = e < Jp.walker_num ? e : Jp.walker_num;
            }
            uint32_t avail = pool_end - pool_next;
            uint32_t rank = (uint32_t)__popcll(need & lt_mask);
            if (!active && rank < avail) {
                const DevJob &J = SINGLE ? J0 : jobs[pool_job];
                init_walker<KIND>(g, J, pool_job, pool_next + rank, s, ls, LINES);
                active = true;
            }
            uint32_t want = (uint32_t)__popcll(need);
            pool_next += want < avail ? want : avail;
            need = __ballot(!active);
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            const DevJob &J = SINGLE ? J0 : jobs[s.job];
            if (advance<KIND, ALIAS_KNOWN, ALIAS>(g, J, s, ls)) active = false;
        }
    }
    flush_stats<STATS>(ls, st);
}

// Plain streaming copy (16 B per lane per iteration): the known-traffic kernel
// that HBM counters and the achievable bandwidth are calibrated against.
// four 16-byte loads in flight per lane, then their stores
// ------------------------------------------------------ serving rounds ----
// A suspended walker's record: everything its next step needs -- job, row,
// where it is and came from with both lists' bounds, step, scheme and state,
// the type mask of its vertex, and its current 16-byte path group (not yet
// in its row) -- so a resume reads nothing at random.  Plane p of record k is
// rec[p * cap + k].
__device__ __forceinline__ void save_walker(const Walker &s, uint4 *rec, uint64_t cap, uint64_t k)
{
    rec[k] = make_uint4(s.job, s.idx, s.cur, s.prev);
    rec[cap + k] = make_uint4(s.step, (s.scheme << 8) | s.state, s.cvm, (uint32_t)(s.end - s.beg));
    rec[2 * cap + k] = make_uint4((uint32_t)s.beg, (uint32_t)(s.beg >> 32), (uint32_t)s.pbeg, (uint32_t)(s.pbeg >> 32));
    rec[3 * cap + k] = make_uint4((uint32_t)(s.pend - s.pbeg), 0u, 0u, 0u);
    rec[4 * cap + k] = make_uint4(s.pg.q0, s.pg.q1, s.pg.q2, s.pg.q3);
}

__device__ __forceinline__ void resume_walker(const DevJob *jobs, const uint4 &a, const uint4 *rec, uint64_t cap,
                                              uint64_t k, Walker &s)
{
    const uint4 b = rec[cap + k], c = rec[2 * cap + k], d = rec[3 * cap + k], e = rec[4 * cap + k];
    s.job = a.x;
    s.idx = a.y;
    s.cur = a.z;
    s.prev = a.w;
    s.wid = jobs[s.job].walker_begin + s.idx;
    s.step = b.x;
    s.trial = 0;
    s.scheme = b.y >> 8;
    s.state = b.y & 0xffu;
    s.cvm = b.z;
    s.beg = (uint64_t)c.x | ((uint64_t)c.y << 32);
    s.end = s.beg + b.w;
    s.pbeg = (uint64_t)c.z | ((uint64_t)c.w << 32);
    s.pend = s.pbeg + d.x;
    s.lines = false;
    s.slen = 1;
    s.smask = 0;
    if (jobs[s.job].kind == KIND_METAPATH) load_scheme(jobs[s.job], s);
    s.pg.q0 = e.x;
    s.pg.q1 = e.y;
    s.pg.q2 = e.z;
    s.pg.q3 = e.w;
}

// The mixed-kind persistent loop of k_walk_persistent, bounded to
// R.round_steps steps per walker: a walker that reaches the bound at a step
// boundary is saved as a record in the next round's list.  A wave reserves
// list slots ROUND_SLOTS at a time (one atomic on *n_out per reservation, not
// one per suspension: every wave appending to one counter serialises the
// round), and marks the slots it leaves unused as holes (plane 0 .x =
// ROUND_HOLE) when it exits; the next round skips them.  *n_live counts the
// walkers listed and the waves ended (one atomic per wave); the last wave to
// end copies the count to *h_live (host memory).  Draws are keyed by (walker, step,
// trial) and trials restart at every step, so paths are the same with or
// without rounds.
template <bool JOBS_LDS>
__global__ void __launch_bounds__(256) k_walk_rounds(DevGraph g, const DevJob *jobs_in, int n_jobs, uint64_t n_chunks,
                                                   uint32_t chunk, unsigned long long *queue, RoundArgs R)
{
    const DevJob *const jobs = block_job_table<JOBS_LDS>(jobs_in, n_jobs, 0u);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    LaneStats ls = {0, 0, 0, 0, 0, 0, 0, 0};
    Walker s;
    bool active = false;
    uint32_t round_start = 0;
    uint32_t pool_job = 0;
    uint64_t pool_next = 0, pool_end = 0;
    bool drained = false, first_chunk = true;
    uint64_t res_base = 0;         // wave-uniform: next reserved slot, slots left, walkers listed
    uint32_t res_left = 0, listed = 0;
    // a continuation round: the walkers the previous round listed.  The count
    // was made by the previous launch's atomics: read it as an atomic (from
    // L2), never as a plain load the compiler may turn into a scalar-cache read
    const uint64_t n_in =
        R.in ? __hip_atomic_load(R.n_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    if (R.in) n_chunks = (n_in + chunk - 1) / chunk;
    for (;;) {
        uint64_t need = __ballot(!active);
        while (need && !drained) {
            if (pool_next >= pool_end) {
                const unsigned long long c = next_chunk(queue, first_chunk);
                if (c >= n_chunks) {
                    drained = true;
                    break;
                }
                if (R.in) {
                    pool_next = c * chunk;
                    const uint64_t e = pool_next + chunk;
                    pool_end = e < n_in ? e : n_in;
                } else {
                    const uint32_t j = find_job(jobs, n_jobs, c);
                    pool_job = j;
                    pool_next = (c - jobs[j].chunk_begin) * chunk;
                    const uint64_t e = pool_next + chunk;
                    pool_end = e < jobs[j].walker_num ? e : jobs[j].walker_num;
                }
            }
            const uint64_t avail = pool_end - pool_next;
            const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
            if (!active && rank < avail) {
                if (R.in) {
                    const uint64_t k = pool_next + rank;
                    const uint4 a = R.in[k];
                    if (a.x != ROUND_HOLE) {
                        resume_walker(jobs, a, R.in, R.cap, k, s);
                        active = true;
                    }
                } else {
                    init_walker<KIND_ANY>(g, jobs[pool_job], pool_job, (uint32_t)(pool_next + rank), s, ls, false);
                    active = true;
                }
                round_start = s.step;
            }
            const uint32_t want = (uint32_t)__popcll(need);
            pool_next += want < avail ? want : avail;
            need = __ballot(!active);
        }
        if (__ballot(active) == 0ull) break;
        bool suspend = false;
        if (active) {
            const DevJob &J = jobs[s.job];
            // a walker at its step cap ends on its next call: it is never
            // suspended, so ceil(cap / round_steps) rounds end every walker
            if (advance<KIND_ANY, false, false>(g, J, s, ls)) active = false;
            else suspend = s.trial == 0 && s.step - round_start >= R.round_steps && s.step < J.walk_length;
        }
        const uint64_t sm = __ballot(suspend);
        if (sm) {
            const uint32_t m = (uint32_t)__popcll(sm), rank = (uint32_t)__popcll(sm & lt_mask);
            unsigned long long fresh = 0;   // a new reservation when the current one is short
            if (m > res_left) {
                if (lane == 0) fresh = atomicAdd(R.n_out, (unsigned long long)ROUND_SLOTS);
                fresh = __shfl(fresh, 0, 64);
            }
            if (suspend) {
                save_walker(s, R.out, R.cap, rank < res_left ? res_base + rank : fresh + (rank - res_left));
                active = false;
            }
            if (m > res_left) {
                res_base = fresh + (m - res_left);
                res_left = ROUND_SLOTS - (m - res_left);
            } else {
                res_base += m;
                res_left -= m;
            }
            listed += m;
        }
    }
    if (lane < res_left) R.out[res_base + lane] = make_uint4(ROUND_HOLE, 0u, 0u, 0u);
    if (lane == 0) {
        // one atomic per wave adds the wave's end (high bits) and its listed
        // walkers (low bits) to one word: the wave that sees every other end
        // already counted holds the round's total and publishes it to the host
        // (no fence needed: both counts live in the one word)
        const unsigned long long waves = (unsigned long long)gridDim.x * (blockDim.x >> 6);
        const unsigned long long old = atomicAdd(R.n_live, (1ull << ROUND_WAVE_SHIFT) + listed);
        if ((old >> ROUND_WAVE_SHIFT) == waves - 1)
            *R.h_live = (old + listed) & ((1ull << ROUND_WAVE_SHIFT) - 1);
    }
}

void launch_walk_rounds(const DevGraph &g, const DevJob *jobs, int n_jobs, uint64_t n_chunks, uint32_t chunk,
                        unsigned long long *queue, const RoundArgs &R, int grid, int block, hipStream_t stream)
{
    if (n_jobs <= LDS_JOBS)
        hipLaunchKernelGGL(k_walk_rounds<true>, dim3(grid), dim3(block), (size_t)n_jobs * sizeof(DevJob), stream, g,
                           jobs, n_jobs, n_chunks, chunk, queue, R);
    else
        hipLaunchKernelGGL(k_walk_rounds<false>, dim3(grid), dim3(block), 0, stream, g, jobs, n_jobs, n_chunks, chunk,
                           queue, R);
}

__global__ void __launch_bounds__(256) k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

void launch_stream_copy(const void *src, void *dst, uint64_t bytes, int cu_count, hipStream_t stream)
{
    uint64_t n = bytes / 16;
    hipLaunchKernelGGL(k_stream_copy, dim3(cu_count * 64), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
}

// Independent random 16-byte reads, each from its own 128-byte line of a
// buffer of `lines` lines: the rate the memory system serves such reads at is
// the ceiling of a kernel whose accesses are all data-dependent 16-byte reads.
// Each read is one global_load_dwordx4 (load16, all four words consumed): read
// through a uint4 struct the compiler narrows it to the words used, and two
// 4-byte reads of one line are served at a lower line rate.
__global__ void __launch_bounds__(256) k_random_gather(const uint32_t *__restrict__ buf, uint64_t lines,
                                                       uint32_t reads, uint64_t seed, uint32_t *__restrict__ sink)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < reads; i++) {
        const u32x4 r = philox(t, i, 0u, 0x200u, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint64_t h = (uint64_t)r.x | ((uint64_t)r.y << 32);
        const uint64_t line = __umul64hi(h, lines);   // < lines
        const u32x4 v = load16(buf + line * 32 + (r.z & 7u) * 4);
        acc += v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9E3779B9u) sink[t & 63] = acc;   // keeps the reads; practically never taken
}

void launch_random_gather(const void *buf, uint64_t bytes, uint32_t threads, uint32_t reads, uint64_t seed,
                          uint32_t *sink, hipStream_t stream)
{
    hipLaunchKernelGGL(k_random_gather, dim3(threads / 256), dim3(256), 0, stream, (const uint32_t *)buf, bytes / 128,
                       reads, seed, sink);
}

// ----------------------------------------------------- packed output ----
// The job that owns flight row i: the last job whose row0 is <= i.
__device__ __forceinline__ int pack_job_of(const PackJob *__restrict__ jobs, int n_jobs, uint64_t i)
{
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].row0 <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// lengths (without the cap flag) as 64-bit counts, for the offset scan, and
// offs[0] = 0.  A count never exceeds its row (the job's stride), so the
// packed output stays within rows x stride entries whatever a length word
// holds.  With h_lens / h_offs (the caller's pinned copies) the lengths and
// offs[0] also go there.
__global__ void k_lens_to_counts(const PackJob *__restrict__ jobs, int n_jobs, const uint32_t *__restrict__ lens,
                                 uint64_t n, uint64_t *__restrict__ cnt, uint64_t *__restrict__ offs,
                                 uint32_t *__restrict__ h_lens, uint64_t *__restrict__ h_offs)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        offs[0] = 0;
        if (h_offs) h_offs[0] = 0;
    }
    if (i >= n) return;
    const uint32_t raw = lens[i];
    const uint32_t len = raw & 0x7fffffffu, cap = jobs[pack_job_of(jobs, n_jobs, i)].stride;
    cnt[i] = len < cap ? len : cap;
    if (h_lens) h_lens[i] = raw;
}

// Copy every walker's row into the packed array at its offset.  Rows of a
// flight are numbered across its jobs; job j owns rows [row0, row0 + n).  One
// wave packs 64 consecutive rows: their packed rows are contiguous, so the
// wave writes that whole range with consecutive lanes on consecutive
// entries (the output may be pinned host memory, written over PCIe), finding
// each entry's row by a binary search over the 64 offsets staged in LDS.
__device__ __forceinline__ void pack_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// With h_offs, each wave also copies its rows' end offsets there (offs[0]
// comes from k_lens_to_counts), so no copy of the offsets follows.
__global__ void __launch_bounds__(256) k_pack_rows(const PackJob *__restrict__ jobs, int n_jobs,
                                                   const uint64_t *__restrict__ offs, uint64_t n,
                                                   uint32_t *__restrict__ out, uint64_t *__restrict__ h_offs)
{
    __shared__ uint64_t s_off[4][65];
    __shared__ const uint32_t *s_src[4][64];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t i0 = ((uint64_t)blockIdx.x * 4 + wv) * 64;
    if (i0 >= n) return;   // whole wave; nothing below synchronises across waves
    const uint32_t cnt = n - i0 < 64 ? (uint32_t)(n - i0) : 64u;
    if (lane < cnt) {
        const uint64_t i = i0 + lane;
        const int j = pack_job_of(jobs, n_jobs, i);
        s_src[wv][lane] = jobs[j].paths + (i - jobs[j].row0) * jobs[j].stride;
    }
    s_off[wv][lane] = offs[i0 + (lane < cnt ? lane : cnt)];
    if (lane == 0) s_off[wv][64] = offs[i0 + cnt];
    pack_wave_sync();
    if (h_offs && lane < cnt) h_offs[i0 + lane + 1] = s_off[wv][lane + 1];
    const uint64_t base = s_off[wv][0];
    const uint64_t total = s_off[wv][cnt] - base;
    for (uint64_t o = lane; o < total; o += 64) {
        uint32_t lo = 0, hi = cnt - 1;
        while (lo < hi) {   // last row whose packed start is <= o
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_off[wv][mid] - base <= o) lo = mid;
            else hi = mid - 1;
        }
        out[base + o] = s_src[wv][lo][o - (s_off[wv][lo] - base)];
    }
}

void launch_pack(const PackJob *jobs, int n_jobs, const uint32_t *lens, uint64_t *cnt, uint64_t *offs, uint64_t n,
                 uint32_t *out, uint32_t *h_lens, uint64_t *h_offs, void *scan_tmp, size_t scan_bytes, hipStream_t s)
{
    if (n == 0) return;
    unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_lens_to_counts, dim3(g), dim3(256), 0, s, jobs, n_jobs, lens, n, cnt, offs, h_lens, h_offs);
    size_t tb = scan_bytes;
    (void)hipcub::DeviceScan::InclusiveSum(scan_tmp, tb, cnt, offs + 1, n, s);
    hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, jobs, n_jobs, offs, n, out,
                       h_offs);
}

size_t pack_scan_bytes(uint64_t n)
{
    size_t tb = 0;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, tb, (uint64_t *)nullptr, (uint64_t *)nullptr, n, (hipStream_t)0);
    return tb;
}


// ------------------------------------------------- state-machine kernel ----
// Persistent single-job kernel (node2vec and DeepWalk) in which every loop iteration issues exactly
// one 16-byte load per active lane: the row_ptr pair of the current vertex
// (ROW), the first half of a trial's alias bucket (TRIAL) and, when its coin
// picks the alias, the second half (ALIAS), or four consecutive entries of a
// sorted list or hash set (SEARCH, HASH: node2vec adjacency test).  Lanes at different phases share each memory instruction,
// so a wave is not held back by its longest search; the 4-entry probe also
// ends a search up to two levels early.  Paths are identical to advance()'s.
enum : uint32_t { PH_ROW = 0, PH_TRIAL = 1, PH_SEARCH = 2, PH_CROW = 3, PH_HASH = 4, PH_ALIAS = 5 };


// LINE_N: rows written in chunks of LINE_N entries through LineBufN (dynamic
// LDS of LINE_N + 1 words per thread); 0: 16-byte groups (PathGroup).
template <int KIND, bool ALIAS, bool STATS, uint32_t LINE_N>
__global__ void __launch_bounds__(256) k_walk_sm(DevGraph g, DevJob J, uint64_t n_chunks, uint32_t chunk,
                                                 unsigned long long *queue, DevStats *st)
{
    constexpr bool LINES = LINE_N != 0;
    constexpr uint32_t LN = LINES ? LINE_N : 4u;
    const LineBufN<LN> lb = lane_line<LN>();
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    // candidates come from buckets: the alias tables, or the uniform buckets
    // (both 32 B, both carrying the candidates' rows when the graph packs them)
    const Bucket *const tab = ALIAS ? g.bucket : g.ubucket;
    static_assert(KIND == KIND_NODE2VEC || KIND == KIND_DEEPWALK,
                  "PPR and metapath jobs run in the trial-loop form (k_walk_persistent)");
    const bool rows_in_buckets = g.packed_rows != 0;
    LaneStats ls = {0, 0, 0, 0, 0, 0, 0, 0};
    bool active = false;
    uint32_t idx = 0, wid = 0, cur = 0, prev = 0, step = 0, trial = 0, phase = PH_ROW;
    uint64_t beg = 0, end = 0, pbeg = 0, pend = 0, cbeg = 0, cend = 0;
    // A search and a hash probe are never live together, so they share state:
    // sorted-list search: key in [sbase, sbase + slen_all), current range [lo, hi);
    // hash probe: table of prev at toff (= sbase), tsize (= slen_all) slots, at
    // slot hslot (= lo).  Between a trial and its ALIAS load, sbase holds the
    // bucket's index.
    uint64_t sbase = 0;
    uint32_t key = 0, slen_all = 0, lo = 0, hi = 0;
    uint64_t &toff = sbase;
    uint32_t &tsize = slen_all, &hslot = lo;
    uint32_t cand = 0, rz = 0, rw = 0;
    bool row_known = false;    // the candidate's row is already in cbeg/cend
    float dart = 0.0f;
    uintptr_t last_line = 0;   // instrumented builds: distinct 64-B lines per search
    uint32_t pool_next = 0, pool_end = 0;
    bool drained = false, first_chunk = true;
    PathGroup pg = {0u, 0u, 0u, 0u};
    uint32_t *const paths = J.paths;
    const uint32_t stride = J.stride;
    uint32_t flush = 0;   // LINES: 16-byte groups of this lane's buffered line waiting to be written
    for (;;) {
        // lines completed or ended in the last iteration leave before any lane
        // buffers another entry (or a new walker's first) in its line
        if constexpr (LINES)
            write_pending_chunks<LN>(flush, [&] { return paths + (uint64_t)idx * stride + (step & ~(LN - 1)); });
        uint64_t need = __ballot(!active);
        while (need && !drained) {
            if (pool_next >= pool_end) {
                const unsigned long long c = next_chunk(queue, first_chunk);
                if (c >= n_chunks) {
                    drained = true;
                    break;
                }
                pool_next = (uint32_t)c * chunk;
                uint32_t e = pool_next + chunk;
                pool_end = e < J.walker_num ? e : J.walker_num;
            }
            uint32_t avail = pool_end - pool_next;
            uint32_t rank = (uint32_t)__popcll(need & lt_mask);
            if (!active && rank < avail) {
                idx = pool_next + rank;
                wid = J.walker_begin + idx;
                cur = start_vertex(g, J, wid);
                prev = 0;
                pbeg = pend = 0;
                step = 0;
                trial = 0;
                phase = PH_ROW;
                if (LINES) lb.set(0, cur);
                else pg.q0 = pg.q1 = pg.q2 = pg.q3 = cur;
                active = true;
            }
            uint32_t want = (uint32_t)__popcll(need);
            pool_next += want < avail ? want : avail;
            need = __ballot(!active);
        }
        if (__ballot(active) == 0ull) break;
        if (!active) continue;

        // ---- the load of this iteration (a trial without rows in its buckets
        // reads both halves of the bucket at once)
        const uint32_t *addr;
        uint64_t blk = 0;
        if (phase == PH_ROW) {
            addr = reinterpret_cast<const uint32_t *>(g.row_ptr + cur);
        } else if (phase == PH_TRIAL) {
            u32x4 r = draw(J.seed, wid, step, trial, TAG_TRIAL);
            uint32_t k = bounded(r.x, r.y, (uint32_t)(end - beg));
            rz = r.z;
            rw = r.w;
            sbase = beg + k;
            addr = reinterpret_cast<const uint32_t *>(tab + sbase);
        } else if (phase == PH_ALIAS) {
            addr = reinterpret_cast<const uint32_t *>(tab + sbase) + 4;
        } else if (phase == PH_CROW) {
            addr = reinterpret_cast<const uint32_t *>(g.row_ptr + cand);
        } else if (phase == PH_HASH) {
            blk = (toff + hslot) & ~(uint64_t)3;
            addr = g.adjset + blk;
        } else {
            // aligned block of four entries holding the middle of the range
            blk = (sbase + lo + ((hi - lo) >> 1)) & ~(uint64_t)3;
            addr = g.col + blk;
        }
        const u32x4 v = load16(addr);
        u32x4 v2 = {0u, 0u, 0u, 0u};
        if (phase == PH_TRIAL && !rows_in_buckets) v2 = load16(addr + 4);

        // ---- consume it
        bool moved = false, rejected = false;
        if (phase == PH_ROW || phase == PH_CROW) {
            const uint64_t b0 = (uint64_t)v.x | ((uint64_t)v.y << 32);
            const uint64_t e0 = (uint64_t)v.z | ((uint64_t)v.w << 32);
            ls.rows++;
            if (phase == PH_ROW) {
                beg = b0;
                end = e0;
            } else {
                cbeg = b0;
                cend = e0;
                row_known = true;
                // prev's list holds cand iff cand's list holds prev (symmetric
                // graph): search the shorter one
                const uint64_t pl = pend - pbeg, cl = cend - cbeg;
                sbase = cl < pl ? cbeg : pbeg;
                slen_all = (uint32_t)(cl < pl ? cl : pl);
                key = cl < pl ? prev : cand;
                lo = 0;
                hi = slen_all;
                if (hi == 0) {   // an empty list holds nothing
                    moved = dart <= J.inv_q;
                    rejected = !moved;
                } else {
                    phase = PH_SEARCH;
                    ls.searches++;
                    last_line = ~(uintptr_t)0;
                    continue;
                }
            }
        } else if (phase == PH_TRIAL || phase == PH_ALIAS) {
            // bucket halves: v = {p, primary, prow} (TRIAL) or {alias, meta, arow}
            // (ALIAS); without rows in the buckets v2 is the second half
            if (phase == PH_TRIAL) {
                ls.trials++;
                const bool prim = u01(rz) < __uint_as_float(v.x);
                if (rows_in_buckets) {
                    if (!prim) {
                        phase = PH_ALIAS;
                        continue;
                    }
                    cand = v.y;
                } else {
                    cand = prim ? v.y : v2.x;
                }
            } else {
                cand = v.x;
            }
            row_known = rows_in_buckets;
            if (rows_in_buckets) {
                const uint64_t row = (uint64_t)v.z | ((uint64_t)v.w << 32);
                cbeg = row & ROW_BEG_MASK;
                cend = cbeg + (row >> ROW_DEG_SHIFT);
            }
            if (KIND == KIND_NODE2VEC) {
                dart = u01(rw) * J.upper;
                if (STATS && step > 0 && !(dart <= J.lower) && cand != prev) {
                    ls.ref_queries++;
                    ls.ref_lines += ref_search_lines(pend - pbeg);
                }
                if (dart <= J.lower || step == 0) {
                    moved = (dart <= J.lower) || (dart <= J.upper);
                    rejected = !moved;
                } else if (cand == prev) {
                    moved = dart <= J.inv_p;
                    rejected = !moved;
                } else if (dart <= J.acc_lo) {
                    moved = true;
                } else if (dart > J.acc_hi) {
                    rejected = true;
                } else if (g.adjset) {
                    toff = adj_table_start(pbeg, prev);
                    tsize = adj_table_size(pend - pbeg);
                    hslot = adj_slot(cand, tsize);
                    phase = PH_HASH;
                    ls.searches++;
                    last_line = ~(uintptr_t)0;
                    continue;
                } else if (g.symmetric) {
                    if (!row_known) {
                        phase = PH_CROW;
                        continue;
                    }
                    const uint64_t pl = pend - pbeg, cl = cend - cbeg;
                    sbase = cl < pl ? cbeg : pbeg;
                    slen_all = (uint32_t)(cl < pl ? cl : pl);
                    key = cl < pl ? prev : cand;
                    lo = 0;
                    hi = slen_all;
                    if (hi == 0) {
                        moved = dart <= J.inv_q;
                        rejected = !moved;
                    } else {
                        phase = PH_SEARCH;
                        ls.searches++;
                        last_line = ~(uintptr_t)0;
                        continue;
                    }
                } else {
                    sbase = pbeg;
                    slen_all = (uint32_t)(pend - pbeg);
                    key = cand;
                    lo = 0;
                    hi = slen_all;
                    phase = PH_SEARCH;
                    ls.searches++;
                    last_line = ~(uintptr_t)0;
                    continue;
                }
            } else {
                moved = true;   // DeepWalk: a static draw is accepted
            }
        } else if (phase == PH_HASH) {
            // slots from hslot to the end of this aligned block, in probe order
            if (STATS) {
                ls.probes++;
                uintptr_t line = (uintptr_t)(g.adjset + blk) >> 6;
                ls.lines += line != last_line;
                last_line = line;
            }
            const uint32_t e4[4] = {v.x, v.y, v.z, v.w};
            const uint64_t first = toff + hslot, lim = toff + tsize;
            int res = -1;   // -1 undecided, 0 absent, 1 present
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint64_t a = blk + (uint64_t)t;
                if (res < 0 && a >= first && a < lim) {
                    if (e4[t] == cand) res = 1;
                    else if (e4[t] == ADJ_EMPTY) res = 0;
                }
            }
            if (res < 0) {
                const uint64_t nx = blk + 4;
                hslot = nx >= lim ? 0u : (uint32_t)(nx - toff);
                continue;
            }
            moved = dart <= (res ? 1.0f : J.inv_q);
            rejected = !moved;
        } else {
            // PH_SEARCH: entries blk..blk+3 (absolute); those inside the list
            // narrow [lo, hi) from both sides; the middle one is among them
            if (STATS) {
                ls.probes++;
                uintptr_t line = (uintptr_t)(g.col + blk) >> 6;
                ls.lines += line != last_line;
                last_line = line;
            }
            const uint32_t e4[4] = {v.x, v.y, v.z, v.w};
            bool found = false;
            uint32_t nlo = lo, nhi = hi;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint64_t a = blk + (uint64_t)t;
                if (a >= sbase && a < sbase + slen_all) {
                    const uint32_t r = (uint32_t)(a - sbase);
                    if (e4[t] == key) found = true;
                    else if (e4[t] < key) nlo = nlo > r + 1 ? nlo : r + 1;
                    else nhi = nhi < r ? nhi : r;
                }
            }
            if (!found) {
                lo = nlo;
                hi = nhi;
                if (lo < hi) continue;
            }
            moved = dart <= (found ? 1.0f : J.inv_q);
            rejected = !moved;
        }

        if (rejected) {
            trial++;
            phase = PH_TRIAL;
            if (trial >= MAX_TRIALS) {
                if (LINES) {
                    if ((step & (LN - 1)) != LN - 1) flush = ((step & (LN - 1)) >> 2) + 1;
                } else if ((step & 3u) != 3u) {
                    pg.store(paths + (uint64_t)idx * stride, step);
                }
                J.lens[idx] = (step + 1) | 0x80000000u;
                active = false;
            }
            continue;
        }
        if (moved) {
            step++;
            if (LINES) {
                lb.set(step, cand);
                if ((step & (LN - 1)) == LN - 1) flush = LN / 4;
            } else {
                pg.set(step, cand);
                if ((step & 3u) == 3u) pg.store(paths + (uint64_t)idx * stride, step);
            }
            prev = cur;
            pbeg = beg;
            pend = end;
            cur = cand;
            trial = 0;
            ls.steps++;
            if (!row_known) {
                phase = PH_ROW;
                continue;
            }
            beg = cbeg;
            end = cend;
        }
        // row of cur known: termination test of the new step (walk.hpp:956-973;
        // node2vec and DeepWalk end at an isolated vertex or at length L)
        const bool stop = end == beg || step >= J.walk_length;
        if (stop) {
            if (LINES) {
                if ((step & (LN - 1)) != LN - 1) flush = ((step & (LN - 1)) >> 2) + 1;
            } else if ((step & 3u) != 3u) {
                pg.store(paths + (uint64_t)idx * stride, step);
            }
            J.lens[idx] = step + 1;
            active = false;
            continue;
        }
        phase = PH_TRIAL;
    }
    flush_stats<STATS>(ls, st);
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------- launchers ----
template <int KIND, bool ALIAS>
static void launch_plain_t(const DevGraph &g, const DevJob &J, bool stats, DevStats *st,
                           int block, hipStream_t stream)
{
    uint32_t grid = (J.walker_num + block - 1) / block;
    if (grid == 0) return;
    if (stats)
        hipLaunchKernelGGL((k_walk_plain<KIND, ALIAS, true>), dim3(grid), dim3(block), 0, stream, g, J, st);
    else
        hipLaunchKernelGGL((k_walk_plain<KIND, ALIAS, false>), dim3(grid), dim3(block), 0, stream, g, J, st);
}

template <int KIND>
static void launch_plain_k(const DevGraph &g, const DevJob &J, bool stats, DevStats *st, int block,
                           hipStream_t stream)
{
    if (J.use_alias) launch_plain_t<KIND, true>(g, J, stats, st, block, stream);
    else launch_plain_t<KIND, false>(g, J, stats, st, block, stream);
}

void launch_walk_plain(const DevGraph &g, const DevJob &J, bool stats, DevStats *st, int block,
                       hipStream_t stream)
{
    switch (J.kind) {
    case KIND_DEEPWALK: launch_plain_k<KIND_DEEPWALK>(g, J, stats, st, block, stream); break;
    case KIND_PPR: launch_plain_k<KIND_PPR>(g, J, stats, st, block, stream); break;
    case KIND_NODE2VEC: launch_plain_k<KIND_NODE2VEC>(g, J, stats, st, block, stream); break;
    default: launch_plain_k<KIND_METAPATH>(g, J, stats, st, block, stream); break;
    }
}

void launch_fill_u64(unsigned long long *p, unsigned long long v, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 63) / 64), dim3(64), 0, stream, p, v, n);
}

template <int KIND, bool ALIAS>
static void launch_single_t(const DevGraph &g, const DevJob &J, uint64_t n_chunks, uint32_t chunk,
                            unsigned long long *queue, bool stats, DevStats *st, int grid, int block,
                            int cu_count, hipStream_t stream)
{
    // The state-machine kernel for node2vec (unless the walk needs the outlier
    // appendix, handled by advance()) and DeepWalk jobs of at least 16384
    // walkers per CU; PPR, metapath and smaller jobs run the trial-loop form.
    // Measured (one MI355X, 256 CUs): PPR 6.1 vs 6.9 ms and metapath 6.9 vs
    // 8.5 ms at R-MAT-22 (4 M walkers); node2vec and DeepWalk equal at 4 M
    // walkers, the state machine 1-3 % faster at R-MAT-24 (16.8 M) and the
    // trial loop faster below: 0.38 vs 0.56 ms (node2vec) and 0.083 vs 0.185
    // ms (DeepWalk) at R-MAT-14, 0.73 vs 0.86 ms (node2vec) at R-MAT-18, and
    // 2.41 vs 2.43-2.64 ms at R-MAT-20.  With few waves per SIMD nothing hides
    // the state machine's per-iteration work (3x the vector and 9x the scalar
    // instructions of the plain kernel per wave at R-MAT-14).
    // KKW_WALK_KERNEL=trial or =sm (read at every launch) fixes the form for
    // every job size; =trial also for every kind.
    const char *form = getenv("KKW_WALK_KERNEL");
    const bool trial_form = form && strcmp(form, "trial") == 0;
    const bool sm_form = form && strcmp(form, "sm") == 0;
    const bool large = (uint64_t)J.walker_num >= (uint64_t)(cu_count > 0 ? cu_count : 1) * 16384u;
    constexpr bool SM_KIND = KIND == KIND_NODE2VEC || KIND == KIND_DEEPWALK;
    const bool sm = SM_KIND && !trial_form && (large || sm_form) && !(KIND == KIND_NODE2VEC && J.overflow > 0.0f);
    if constexpr (SM_KIND) {
        if (sm) {
            // rows of whole lines: 128-byte aligned rows (the path buffer is 256-byte aligned)
            const bool lines = J.stride % LINE_ENTRIES == 0;
            const size_t lds = lines ? (size_t)block * LINE_LDS_STRIDE * sizeof(uint32_t) : 0;
            if (stats && lines)
                hipLaunchKernelGGL((k_walk_sm<KIND, ALIAS, true, LINE_ENTRIES>), dim3(grid), dim3(block), lds, stream, g,
                                   J, n_chunks, chunk, queue, st);
            else if (stats)
                hipLaunchKernelGGL((k_walk_sm<KIND, ALIAS, true, 0>), dim3(grid), dim3(block), 0, stream, g, J,
                                   n_chunks, chunk, queue, st);
            else if (lines)
                hipLaunchKernelGGL((k_walk_sm<KIND, ALIAS, false, LINE_ENTRIES>), dim3(grid), dim3(block), lds, stream, g,
                                   J, n_chunks, chunk, queue, st);
            else
                hipLaunchKernelGGL((k_walk_sm<KIND, ALIAS, false, 0>), dim3(grid), dim3(block), 0, stream, g, J,
                                   n_chunks, chunk, queue, st);
            return;
        }
    }
    const bool lines = J.stride % LINE_ENTRIES == 0;
    const size_t lds = lines ? (size_t)block * (HALF_LINE_ENTRIES + 1) * sizeof(uint32_t) : 0;
    if (stats && lines)
        hipLaunchKernelGGL((k_walk_persistent<KIND, true, true, ALIAS, true, true, false>), dim3(grid), dim3(block), lds,
                           stream, g, J, (const DevJob *)nullptr, 1, n_chunks, chunk, queue, st);
    else if (stats)
        hipLaunchKernelGGL((k_walk_persistent<KIND, true, true, ALIAS, true, false, false>), dim3(grid), dim3(block), 0,
                           stream, g, J, (const DevJob *)nullptr, 1, n_chunks, chunk, queue, st);
    else if (lines)
        hipLaunchKernelGGL((k_walk_persistent<KIND, true, true, ALIAS, false, true, false>), dim3(grid), dim3(block), lds,
                           stream, g, J, (const DevJob *)nullptr, 1, n_chunks, chunk, queue, st);
    else
        hipLaunchKernelGGL((k_walk_persistent<KIND, true, true, ALIAS, false, false, false>), dim3(grid), dim3(block), 0,
                           stream, g, J, (const DevJob *)nullptr, 1, n_chunks, chunk, queue, st);
}

// One job in a persistent launch (specialised on kind and alias use).
void launch_walk_persistent_single(const DevGraph &g, const DevJob &J, uint64_t n_chunks,
                                   uint32_t chunk, unsigned long long *queue, bool stats,
                                   DevStats *st, int grid, int block, int cu_count, hipStream_t stream)
{
#define KKW_SINGLE(KD)                                                                                       \
    if (J.use_alias)                                                                                         \
        launch_single_t<KD, true>(g, J, n_chunks, chunk, queue, stats, st, grid, block, cu_count, stream);   \
    else                                                                                                     \
        launch_single_t<KD, false>(g, J, n_chunks, chunk, queue, stats, st, grid, block, cu_count, stream);
    switch (J.kind) {
    case KIND_DEEPWALK: KKW_SINGLE(KIND_DEEPWALK) break;
    case KIND_PPR: KKW_SINGLE(KIND_PPR) break;
    case KIND_NODE2VEC: KKW_SINGLE(KIND_NODE2VEC) break;
    default: KKW_SINGLE(KIND_METAPATH) break;
    }
#undef KKW_SINGLE
}

// Mixed jobs from a device job table (copied into LDS by every block when it
// holds at most LDS_JOBS jobs).
template <bool STATS, bool LINES>
static void launch_mixed_t(const DevGraph &g, const DevJob *jobs, int n_jobs, uint64_t n_chunks, uint32_t chunk,
                           unsigned long long *queue, DevStats *st, int grid, int block, hipStream_t stream)
{
    DevJob dummy{};
    const size_t line_lds = LINES ? (size_t)block * (HALF_LINE_ENTRIES + 1) * sizeof(uint32_t) : 0;
    if (n_jobs <= LDS_JOBS)
        hipLaunchKernelGGL((k_walk_persistent<KIND_ANY, false, false, false, STATS, LINES, true>), dim3(grid),
                           dim3(block), line_lds + (size_t)n_jobs * sizeof(DevJob), stream, g, dummy, jobs, n_jobs,
                           n_chunks, chunk, queue, st);
    else
        hipLaunchKernelGGL((k_walk_persistent<KIND_ANY, false, false, false, STATS, LINES, false>), dim3(grid),
                           dim3(block), line_lds, stream, g, dummy, jobs, n_jobs, n_chunks, chunk, queue, st);
}

void launch_walk_persistent_mixed(const DevGraph &g, const DevJob *jobs, int n_jobs, uint64_t n_chunks,
                                  uint32_t chunk, unsigned long long *queue, bool stats, DevStats *st,
                                  int grid, int block, bool lines, hipStream_t stream)
{
    if (stats && lines) launch_mixed_t<true, true>(g, jobs, n_jobs, n_chunks, chunk, queue, st, grid, block, stream);
    else if (stats) launch_mixed_t<true, false>(g, jobs, n_jobs, n_chunks, chunk, queue, st, grid, block, stream);
    else if (lines) launch_mixed_t<false, true>(g, jobs, n_jobs, n_chunks, chunk, queue, st, grid, block, stream);
    else launch_mixed_t<false, false>(g, jobs, n_jobs, n_chunks, chunk, queue, st, grid, block, stream);
}

}  // namespace kkw
