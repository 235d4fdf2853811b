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
#include "kkw_dev.h"

namespace kkw {

struct Walker {
    uint32_t job;      // job index (persistent batches)
    uint32_t idx;      // row within the job
    uint32_t wid;      // walker id (Philox counter)
    uint32_t cur, prev;
    uint64_t beg, end; // list of cur
    uint64_t pbeg, pend; // list of prev
    uint32_t step, trial;
    uint32_t scheme, state;
};

struct LaneStats {
    uint32_t steps, trials, searches, probes, rows, lines;
};

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

__device__ __forceinline__ uint32_t start_vertex(const DevGraph &g, const DevJob &J, uint32_t wid)
{
    if (J.start_mode == START_LIST) return J.starts[wid % J.n_starts];
    if (J.start_mode == START_UNIFORM) {
        u32x4 r = draw(J.seed, wid, 0, 0, TAG_INIT);
        return bounded(r.z, r.w, g.V);
    }
    return wid % g.V;
}

template <int KIND>
__device__ __forceinline__ void init_walker(const DevGraph &g, const DevJob &J, uint32_t job,
                                            uint32_t idx, Walker &s, LaneStats &ls)
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
    const int kind = KIND == KIND_ANY ? J.kind : KIND;
    if (kind == KIND_METAPATH) {
        u32x4 r = draw(J.seed, s.wid, 0, 0, TAG_INIT);
        s.scheme = bounded(r.x, r.y, J.n_schemes);
    }
    s.beg = g.row_ptr[s.cur];
    s.end = g.row_ptr[s.cur + 1];
    ls.rows++;
    J.paths[(uint64_t)idx * J.stride] = s.cur;
}

__device__ __forceinline__ void finish_walker(const DevJob &J, const Walker &s, bool truncated)
{
    J.lens[s.idx] = (s.step + 1) | (truncated ? 0x80000000u : 0u);
}

__device__ __forceinline__ void move_to(const DevGraph &g, const DevJob &J, Walker &s,
                                        uint32_t cand, LaneStats &ls)
{
    s.step++;
    J.paths[(uint64_t)s.idx * J.stride + s.step] = cand;
    s.prev = s.cur;
    s.pbeg = s.beg;
    s.pend = s.end;
    s.cur = cand;
    s.beg = g.row_ptr[cand];
    s.end = g.row_ptr[cand + 1];
    s.trial = 0;
    ls.steps++;
    ls.rows++;
}

// Candidate by alias bucket or uniform index (include/walk.hpp:1016-1029).
template <bool ALIAS_KNOWN, bool ALIAS>
__device__ __forceinline__ uint32_t pick(const DevGraph &g, const DevJob &J, const Walker &s,
                                         uint32_t deg, const u32x4 &r, uint32_t *etype)
{
    uint32_t k = bounded(r.x, r.y, deg);
    const bool use_alias = ALIAS_KNOWN ? ALIAS : (J.use_alias != 0);
    if (use_alias) {
        Bucket b = g.bucket[s.beg + k];
        bool prim = u01(r.z) < b.p;
        if (etype) *etype = prim ? (b.meta & 0xffu) : ((b.meta >> 8) & 0xffu);
        return prim ? b.primary : b.alias;
    }
    if (etype) *etype = g.etype ? g.etype[s.beg + k] : 0u;
    return g.col[s.beg + k];
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
            cand = pick<ALIAS_KNOWN, ALIAS>(g, J, s, deg, r, nullptr);
            dart = u01(r.w) * J.upper;            // walk.hpp:1225
            pre = dart <= J.lower;                // walk.hpp:1226-1229
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
            bool found = adj_contains(g.col, s.pbeg, s.pend, cand, pr, ln);
            ls.searches++;
            ls.probes += pr;
            ls.lines += ln;
            accept = dart <= (found ? 1.0f : J.inv_q);   // node2vec.hpp:187-191
        }
        if (accept) move_to(g, J, s, cand, ls);
        else s.trial++;
        return false;
    }

    // First-order kinds: termination once per step (walk.hpp:956-973).
    if (s.trial == 0) {
        bool stop = deg == 0;
        if (!stop) {
            float ext;
            if (kind == KIND_PPR) {
                ext = J.ppr_ext;                                  // ppr.hpp:51-55
            } else if (kind == KIND_METAPATH) {
                ext = (s.step >= J.walk_length ||
                       !(g.vmask[s.cur] & J.scheme_mask[s.scheme * 8 + s.state])) ? 0.0f : 1.0f;
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
    u32x4 r = draw(J.seed, s.wid, s.step, s.trial, TAG_TRIAL);
    if (kind == KIND_METAPATH) {
        uint32_t et = 0;
        uint32_t cand = pick<ALIAS_KNOWN, ALIAS>(g, J, s, deg, r, &et);
        float dart = u01(r.w) * 1.0f;
        float pd = ((J.scheme_mask[s.scheme * 8 + s.state] >> et) & 1u) ? 1.0f : 0.0f;
        if (dart <= pd) {                                          // walk.hpp:1030-1036
            uint32_t slen = J.scheme_len[s.scheme];
            move_to(g, J, s, cand, ls);
            s.state = (s.state + 1) % slen;                        // metapath.hpp:96-99
        } else {
            s.trial++;
        }
        return false;
    }
    uint32_t cand = pick<ALIAS_KNOWN, ALIAS>(g, J, s, deg, r, nullptr);
    move_to(g, J, s, cand, ls);
    return false;
}

template <bool STATS>
__device__ __forceinline__ void flush_stats(const LaneStats &ls, DevStats *st)
{
    if (!STATS) return;
    // wave reduction, one atomic per counter per wave
    unsigned long long v[6] = {ls.steps, ls.trials, ls.searches, ls.probes, ls.rows, ls.lines};
#pragma unroll
    for (int k = 0; k < 6; k++) {
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
    }
}

// ---------------------------------------------------------------- plain ----
template <int KIND, bool ALIAS, bool STATS>
__global__ void __launch_bounds__(256) k_walk_plain(DevGraph g, DevJob J, DevStats *st)
{
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    LaneStats ls = {0, 0, 0, 0, 0, 0};
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

// SINGLE: one job (passed by value), otherwise a job table in device memory.
template <int KIND, bool SINGLE, bool ALIAS_KNOWN, bool ALIAS, bool STATS>
__global__ void __launch_bounds__(256) k_walk_persistent(DevGraph g, DevJob J0, const DevJob *jobs,
                                                       int n_jobs, uint64_t n_chunks, uint32_t chunk,
                                                       unsigned long long *queue, DevStats *st)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    LaneStats ls = {0, 0, 0, 0, 0, 0};
    Walker s;
    bool active = false;
    // wave-uniform pool of walker rows [pool_next, pool_end) of job pool_job
    uint32_t pool_job = 0, pool_next = 0, pool_end = 0;
    bool drained = false;
    for (;;) {
        uint64_t need = __ballot(!active);
        while (need && !drained) {
            if (pool_next >= pool_end) {
                unsigned long long c = 0;
                if (lane == 0) c = atomicAdd(queue, 1ull);
                c = __shfl(c, 0, 64);
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
                init_walker<KIND>(g, J, pool_job, pool_next + rank, s, ls);
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
                            hipStream_t stream)
{
    if (stats)
        hipLaunchKernelGGL((k_walk_persistent<KIND, true, true, ALIAS, true>), dim3(grid), dim3(block), 0,
                           stream, g, J, (const DevJob *)nullptr, 1, n_chunks, chunk, queue, st);
    else
        hipLaunchKernelGGL((k_walk_persistent<KIND, true, true, ALIAS, false>), dim3(grid), dim3(block), 0,
                           stream, g, J, (const DevJob *)nullptr, 1, n_chunks, chunk, queue, st);
}

// One job in a persistent launch (specialised on kind and alias use).
void launch_walk_persistent_single(const DevGraph &g, const DevJob &J, uint64_t n_chunks,
                                   uint32_t chunk, unsigned long long *queue, bool stats,
                                   DevStats *st, int grid, int block, hipStream_t stream)
{
#define KKW_SINGLE(KD)                                                                            \
    if (J.use_alias) launch_single_t<KD, true>(g, J, n_chunks, chunk, queue, stats, st, grid, block, stream); \
    else launch_single_t<KD, false>(g, J, n_chunks, chunk, queue, stats, st, grid, block, stream);
    switch (J.kind) {
    case KIND_DEEPWALK: KKW_SINGLE(KIND_DEEPWALK) break;
    case KIND_PPR: KKW_SINGLE(KIND_PPR) break;
    case KIND_NODE2VEC: KKW_SINGLE(KIND_NODE2VEC) break;
    default: KKW_SINGLE(KIND_METAPATH) break;
    }
#undef KKW_SINGLE
}

// Mixed jobs from a device job table.
void launch_walk_persistent_mixed(const DevGraph &g, const DevJob *jobs, int n_jobs, uint64_t n_chunks,
                                  uint32_t chunk, unsigned long long *queue, bool stats, DevStats *st,
                                  int grid, int block, hipStream_t stream)
{
    DevJob dummy{};
    if (stats)
        hipLaunchKernelGGL((k_walk_persistent<KIND_ANY, false, false, false, true>), dim3(grid), dim3(block),
                           0, stream, g, dummy, jobs, n_jobs, n_chunks, chunk, queue, st);
    else
        hipLaunchKernelGGL((k_walk_persistent<KIND_ANY, false, false, false, false>), dim3(grid), dim3(block),
                           0, stream, g, dummy, jobs, n_jobs, n_chunks, chunk, queue, st);
}

}  // namespace kkw
