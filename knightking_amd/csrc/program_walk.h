// program_walk.h -- the walk loop of a walk program, compiled at run time
// (hipRTC) after the program's own functions (include/kkw_program.h).
//
// One walker per lane, the first-order epoch of include/walk.hpp:954-1064
// (and, for a program with a query / response pair, the second-order one of
// walk.hpp:1135-1379 with the query answered in place):
// per step the extension test (once, or on every retry for a second-order
// program, walk.hpp:1161-1173), then trials until one is accepted -- the
// outlier appendix first when the program has one (981-1014), else a static
// draw (uniform, or the alias bucket of the edge weights, 1016-1029) and,
// with a dynamic component, the dart against lower bound and dynamic_comp
// (1030-1036); then step++, update_state and the new vertex (1039-1056).
// Draws use the same counters as the library's built-in walks: trial words
// (index, coin, dart) at (walker, step, trial, TRIAL), the outlier word at
// (walker, step, trial, OUTLIER), the extension coin at (walker, step, 0 or
// trial, EXT), a uniform start at (walker, 0, 0, INIT).
#pragma once

#include "kkw_rng.h"
#include "program_args.h"

#ifndef KKW_DYNAMIC
#define KKW_DYNAMIC 0
#endif
#ifndef KKW_LOWER
#define KKW_LOWER 0
#endif
#ifndef KKW_OUTLIER
#define KKW_OUTLIER 0
#endif
#ifndef KKW_EDGE_INDEX
#define KKW_EDGE_INDEX 1
#endif
#ifndef KKW_SECOND_ORDER
#define KKW_SECOND_ORDER 0
#endif
#if KKW_SECOND_ORDER && !KKW_DYNAMIC
#error "a second-order transition needs a dynamic component (walk.hpp:339-351)"
#endif
#if KKW_OUTLIER && !KKW_DYNAMIC
#error "an outlier appendix needs a dynamic component (walk.hpp:251-261)"
#endif
#if KKW_LOWER && !KKW_DYNAMIC
#error "a lower bound needs a dynamic component (walk.hpp:251-261)"
#endif

#define KKW_PROG_MAX_TRIALS (1u << 20)   // per step; guards degenerate programs

// The host reads this after loading the module: 1 when the walk loop draws
// from local-index tables (the program reads e.index), 0 when it draws from
// the engine's tables of vertex ids and rows.
extern "C" {
__device__ int kkw_prog_edge_index = KKW_EDGE_INDEX;
}

typedef uint32_t kkw_prog_u4a __attribute__((ext_vector_type(4), aligned(4)));

// row_ptr[v] and row_ptr[v + 1] in one 16-byte load
__device__ __forceinline__ void kkw_prog_row(const uint64_t *rp, uint32_t v, uint64_t &beg, uint32_t &deg)
{
    const kkw_prog_u4a q = *reinterpret_cast<const kkw_prog_u4a *>(rp + v);
    beg = (uint64_t)q.x | ((uint64_t)q.y << 32);
    deg = (uint32_t)(((uint64_t)q.z | ((uint64_t)q.w << 32)) - beg);
}

// The new vertex's row: from the accepted bucket when it carried one, else
// from row_ptr.
__device__ __forceinline__ void kkw_prog_next_row(const uint64_t *rp, uint32_t v, bool nrow_ok, uint64_t nrow,
                                                  uint64_t &beg, uint32_t &deg)
{
    if (nrow_ok) {
        beg = nrow & ((1ull << KKW_PROG_ROW_DEG_SHIFT) - 1ull);
        deg = (uint32_t)(nrow >> KKW_PROG_ROW_DEG_SHIFT);
    } else {
        kkw_prog_row(rp, v, beg, deg);
    }
}

// The walker leaves v (row beg, deg): the context keeps that row for
// kkw_adjacent.
__device__ __forceinline__ void kkw_prog_keep_row(kkw_ctx &c, uint32_t v, uint64_t beg, uint32_t deg)
{
    c.row_v = v;
    c.row_beg = beg;
    c.row_deg = deg;
}

// The walker's last response (a second-order program) and the dynamic
// component of a candidate: post the query, have the target answer it, and
// evaluate dynamic_comp with the answer (walk.hpp:1226-1257, 1336-1364).
#if KKW_SECOND_ORDER
typedef Response kkw_prog_resp;
#else
struct kkw_prog_resp {
};
#endif
#if KKW_DYNAMIC
__device__ __forceinline__ float kkw_prog_dynamic(const WalkerState &s, const kkw_walker &w, uint32_t cur,
                                                  const kkw_edge &e, kkw_prog_resp &resp, const kkw_ctx &c)
{
#if KKW_SECOND_ORDER
    uint32_t target = 0;
    Query q;
    if (post_query(s, w, cur, e, target, q, c)) respond_query(target, q, resp, c);
    return dynamic_comp(s, w, resp, cur, e, c);
#else
    (void)resp;
    return dynamic_comp(s, w, cur, e, c);
#endif
}
#endif

// four path entries held in registers, stored as one aligned 16-byte group
struct kkw_prog_group {
    uint32_t v[4];
    // v[step & 3] = x by selects, so the group stays in registers (a
    // dynamically indexed array would live in scratch memory)
    __device__ __forceinline__ void set(uint32_t step, uint32_t x)
    {
        const uint32_t k = step & 3u;
        v[0] = k == 0 ? x : v[0];
        v[1] = k == 1 ? x : v[1];
        v[2] = k == 2 ? x : v[2];
        v[3] = k == 3 ? x : v[3];
    }
    __device__ __forceinline__ void store(uint32_t *path, uint32_t step) const
    {
        *reinterpret_cast<uint4 *>(path + (step & ~3u)) = make_uint4(v[0], v[1], v[2], v[3]);
    }
};

__device__ __forceinline__ bool kkw_prog_terminates(const WalkerState &s, const kkw_walker &w, uint32_t cur,
                                                    uint32_t deg, uint32_t trial, const kkw_prog_args &A,
                                                    const kkw_ctx &c)
{
    if (deg == 0) return true;
    const float ext = extension(s, w, cur, c);
    if (ext == 0.0f) return true;
    if (ext < 1.0f) return kkw::u01(kkw::draw(A.seed, w.id, w.step, trial, kkw::TAG_EXT).x) > ext;
    return false;
}

// One trial at cur; returns true with the accepted edge in `acc` and, when
// the bucket carried it, the new vertex's packed row in `nrow` (nrow_ok).
// env counts envelope violations (include/walk.hpp:1003-1005): an outlier
// candidate whose dynamic component exceeds overflow / static + upper (+1e-6,
// the reference's tolerance), or -- for a program without an outlier pair,
// where nothing covers the excess -- a static draw whose dynamic component
// exceeds the upper bound.  Either biases the walk; the count is reported.
__device__ __forceinline__ bool kkw_prog_trial(const WalkerState &s, const kkw_walker &w, uint32_t cur,
                                               uint64_t beg, uint32_t deg, uint32_t trial, const kkw_prog_args &A,
                                               const kkw_ctx &c, kkw_prog_resp &resp, kkw_edge &acc,
                                               uint64_t &nrow, bool &nrow_ok, uint32_t &env)
{
    (void)resp;
    (void)env;
    nrow_ok = false;
    nrow = 0;
#if KKW_DYNAMIC
    const float up = upper_bound(cur, c);
#endif
#if KKW_OUTLIER
    {
        float ov = 0.0f;
        uint32_t num = 0;
        outlier_bound(s, w, cur, ov, num, c);
        if (ov > 0.0f && num > 0) {
            const float regular = (A.vsum ? A.vsum[cur] : (float)deg) * up;
            const float rv = kkw::u01(kkw::draw(A.seed, w.id, w.step, trial, kkw::TAG_OUTLIER).x) *
                                 (ov * (float)num + regular) -
                             regular;
            if (rv > 0.0f) {
                uint32_t idx = (uint32_t)(rv / ov);
                if (idx + 1 > num) idx = num - 1;
                const int64_t ei = outlier_search(s, w, cur, idx, c);
                if (ei < 0) return false;
                kkw_edge e;
                e.index = (uint64_t)ei;
                e.neighbour = c.col[ei];
                e.reserved = 0;
                const float sc = A.vsum ? c.weight[ei] : 1.0f;   // vsum is set exactly for weighted draws
                const float h = (rv - (float)idx * ov) / sc + up;
                const float pd = kkw_prog_dynamic(s, w, cur, e, resp, c);
                env += ov / sc + up + 1e-6f < pd ? 1u : 0u;
                if (!(h <= pd)) return false;
                acc = e;
                return true;
            }
        }
    }
#endif
    const kkw::u32x4 r = kkw::draw(A.seed, w.id, w.step, trial, kkw::TAG_TRIAL);
    const uint32_t k = kkw::bounded(r.x, r.y, deg);
    kkw_edge e;
    e.reserved = 0;
#if !KKW_EDGE_INDEX
    {
        // the engine's buckets: {p, primary id, primary row}, {alias id, meta,
        // alias row}; uniform buckets have p = 1 (always the primary, col[beg + k])
        const uint32_t *bw = reinterpret_cast<const uint32_t *>(A.lb + beg + k);
        const kkw_prog_u4a h0 = *reinterpret_cast<const kkw_prog_u4a *>(bw);
        if (kkw::u01(r.z) < __uint_as_float(h0.x)) {
            e.neighbour = h0.y;
            nrow = (uint64_t)h0.z | ((uint64_t)h0.w << 32);
        } else {
            const kkw_prog_u4a h1 = *reinterpret_cast<const kkw_prog_u4a *>(bw + 4);
            e.neighbour = h1.x;
            nrow = (uint64_t)h1.z | ((uint64_t)h1.w << 32);
        }
        e.index = KKW_NO_EDGE_INDEX;
    }
#else
    if (A.lb) {
        // first half {p, local primary, primary's id} as one 16-byte load; the
        // alias half {local alias, meta, alias's id} only when the coin picks it
        const uint32_t *bw = reinterpret_cast<const uint32_t *>(A.lb + beg + k);
        const kkw_prog_u4a h0 = *reinterpret_cast<const kkw_prog_u4a *>(bw);
        if (kkw::u01(r.z) < __uint_as_float(h0.x)) {
            e.index = beg + h0.y;
            e.neighbour = h0.z;
        } else {
            const kkw_prog_u4a h1 = *reinterpret_cast<const kkw_prog_u4a *>(bw + 4);
            e.index = beg + h1.x;
            e.neighbour = h1.z;
        }
    } else {
        e.index = beg + k;
        e.neighbour = c.col[e.index];
    }
#endif
#if KKW_DYNAMIC
    const float dart = kkw::u01(r.w) * up;
    bool ok = false;
#if KKW_LOWER
    ok = dart <= lower_bound(cur, c);
#endif
    if (!ok) {
        const float pd = kkw_prog_dynamic(s, w, cur, e, resp, c);
        env += !KKW_OUTLIER && pd > up ? 1u : 0u;
        ok = dart <= pd;
    }
    if (!ok) return false;
#endif
    acc = e;
    nrow_ok = !KKW_EDGE_INDEX && A.packed_rows != 0;
    return true;
}

// Start vertex of walker id (get_equal_dist_func / a start list / a uniform draw).
__device__ __forceinline__ uint32_t kkw_prog_start(const kkw_prog_args &A, uint32_t id)
{
    if (A.start_mode == 1) return A.starts[id % A.n_starts];
    if (A.start_mode == 2) {
        const kkw::u32x4 r = kkw::draw(A.seed, id, 0, 0, kkw::TAG_INIT);
        return kkw::bounded(r.z, r.w, A.c.v_num);
    }
    return id % A.c.v_num;
}

// Plain form: one lane per walker for the whole launch.
extern "C" __global__ void __launch_bounds__(KKW_PROG_BLOCK) kkw_program_walk(kkw_prog_args A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.walker_num) return;
    kkw_ctx c = A.c;   // per lane: the kept row changes as the walker moves
    kkw_prog_keep_row(c, KKW_NO_VERTEX, 0, 0);
    kkw_walker w;
    w.id = A.walker_begin + i;
    w.step = 0;
    uint32_t cur = kkw_prog_start(A, w.id);
    WalkerState s;
    init_state(s, w, cur, c);
    kkw_prog_resp resp = kkw_prog_resp();
    const bool retest = A.retest_ext || KKW_SECOND_ORDER;
    uint32_t *path = A.paths + (uint64_t)i * A.stride;   // stride is a multiple of 4
    kkw_prog_group grp;
    grp.v[0] = grp.v[1] = grp.v[2] = grp.v[3] = cur;
    uint64_t beg;
    uint32_t deg;
    kkw_prog_row(c.row_ptr, cur, beg, deg);
    bool capped = false;
    uint32_t trial = 0, env = 0;
    for (;;) {
        if (trial == 0 || retest) {
            if (kkw_prog_terminates(s, w, cur, deg, retest ? trial : 0u, A, c)) break;
            if (w.step >= A.max_len) {
                capped = true;
                break;
            }
        }
        if (trial >= KKW_PROG_MAX_TRIALS) {
            capped = true;
            break;
        }
        kkw_edge e;
        uint64_t nrow;
        bool nrow_ok;
        if (!kkw_prog_trial(s, w, cur, beg, deg, trial, A, c, resp, e, nrow, nrow_ok, env)) {
            trial++;
            continue;
        }
        w.step++;
        update_state(s, w, cur, e, c);
        kkw_prog_keep_row(c, cur, beg, deg);
        cur = e.neighbour;
        grp.set(w.step, cur);
        if ((w.step & 3u) == 3u) grp.store(path, w.step);
        kkw_prog_next_row(c.row_ptr, cur, nrow_ok, nrow, beg, deg);
        trial = 0;
    }
    if ((w.step & 3u) != 3u) grp.store(path, w.step);   // the last, partial group
    A.lens[i] = (w.step + 1) | (capped ? 0x80000000u : 0u);
    if (env) atomicAdd(A.violations, (unsigned long long)env);
}

// Path write-out of the persistent form when rows are whole 128-byte lines
// (A.line_rows): the lane's current 16 entries (a 64-byte half line) wait in
// LDS, 17 words per lane, and leave as four 16-byte stores when the half line
// fills; the half line that ends a walk is written up to its last 16-byte
// group.  Half lines reach HBM whole instead of 16 bytes at a time.
extern __shared__ uint32_t kkw_prog_lds[];

__device__ __forceinline__ void kkw_prog_flush(const uint32_t *buf, uint32_t *path, uint32_t step)
{
    uint4 *dst = reinterpret_cast<uint4 *>(path + (step & ~(KKW_PROG_HALF_LINE - 1)));
    const uint32_t groups = ((step & (KKW_PROG_HALF_LINE - 1)) >> 2) + 1;
#pragma unroll
    for (uint32_t k = 0; k < KKW_PROG_HALF_LINE / 4; k++)
        if (k < groups) dst[k] = make_uint4(buf[4 * k], buf[4 * k + 1], buf[4 * k + 2], buf[4 * k + 3]);
}

// Persistent form: a grid sized to the device; each wave takes work items of
// A.chunk consecutive walkers, and a lane whose walker has ended takes the
// next walker of the wave's item (ballot + prefix count), so lanes freed by
// short walks (isolated starts, early termination) are refilled at once.  A
// wave's first item is the one its own index names; later ones come from
// A.queue, numbered past the grid's own items (with every item taken from
// the queue, a wave whose first walkers ended at once could take a second
// item before the last waves had a first, and walk it after its first
// walkers while those waves left empty).  Each loop iteration runs one trial
// per active lane, with the same draws as the plain form; a wave leaves when
// the queue is drained and its lanes are idle, so every wave reaches the exit.
extern "C" __global__ void __launch_bounds__(KKW_PROG_BLOCK) kkw_program_walk_persistent(kkw_prog_args A)
{
    kkw_ctx c = A.c;   // per lane: the kept row changes as the walker moves
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    bool active = false, drained = false, first_item = true;
    uint32_t pool_next = 0, pool_end = 0;
    uint32_t i = 0, cur = 0, deg = 0, trial = 0, env = 0;
    uint64_t beg = 0;
    uint32_t *path = nullptr;
    kkw_walker w;
    w.id = 0;
    w.step = 0;
    WalkerState s;
    kkw_prog_resp resp = kkw_prog_resp();
    const bool retest = A.retest_ext || KKW_SECOND_ORDER;
    kkw_prog_group grp;
    uint32_t *const lbuf = kkw_prog_lds + threadIdx.x * (KKW_PROG_HALF_LINE + 1);
    const bool lines = A.line_rows != 0;
    for (;;) {
        unsigned long long need = __ballot(!active);
        while (need && !drained) {
            if (pool_next >= pool_end) {
                const unsigned long long waves = (unsigned long long)gridDim.x * (blockDim.x >> 6);
                unsigned long long q = 0;
                if (first_item) {
                    first_item = false;
                    q = (unsigned long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
                } else {
                    if (lane == 0) q = atomicAdd(A.queue, 1ull);
                    q = __shfl(q, 0, 64) + waves;
                }
                if (q >= A.n_chunks) {
                    drained = true;
                    break;
                }
                pool_next = (uint32_t)q * A.chunk;
                const uint32_t e = pool_next + A.chunk;
                pool_end = e < A.walker_num ? e : A.walker_num;
            }
            const uint32_t avail = pool_end - pool_next;
            const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
            if (!active && rank < avail) {
                i = pool_next + rank;
                w.id = A.walker_begin + i;
                w.step = 0;
                cur = kkw_prog_start(A, w.id);
                kkw_prog_keep_row(c, KKW_NO_VERTEX, 0, 0);
                init_state(s, w, cur, c);
                resp = kkw_prog_resp();
                path = A.paths + (uint64_t)i * A.stride;
                if (lines) lbuf[0] = cur;
                else grp.v[0] = grp.v[1] = grp.v[2] = grp.v[3] = cur;
                kkw_prog_row(c.row_ptr, cur, beg, deg);
                trial = 0;
                active = true;
            }
            const uint32_t want = (uint32_t)__popcll(need);
            pool_next += want < avail ? want : avail;
            need = __ballot(!active);
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            bool done = false, capped = false;
            if (trial == 0 || retest) {
                if (kkw_prog_terminates(s, w, cur, deg, retest ? trial : 0u, A, c)) {
                    done = true;
                } else if (w.step >= A.max_len) {
                    done = capped = true;
                }
            }
            if (!done && trial >= KKW_PROG_MAX_TRIALS) done = capped = true;
            if (!done) {
                kkw_edge e;
                uint64_t nrow;
                bool nrow_ok;
                if (kkw_prog_trial(s, w, cur, beg, deg, trial, A, c, resp, e, nrow, nrow_ok, env)) {
                    w.step++;
                    update_state(s, w, cur, e, c);
                    kkw_prog_keep_row(c, cur, beg, deg);
                    cur = e.neighbour;
                    if (lines) {
                        lbuf[w.step & (KKW_PROG_HALF_LINE - 1)] = cur;
                        if ((w.step & (KKW_PROG_HALF_LINE - 1)) == KKW_PROG_HALF_LINE - 1)
                            kkw_prog_flush(lbuf, path, w.step);
                    } else {
                        grp.set(w.step, cur);
                        if ((w.step & 3u) == 3u) grp.store(path, w.step);
                    }
                    kkw_prog_next_row(c.row_ptr, cur, nrow_ok, nrow, beg, deg);
                    trial = 0;
                } else {
                    trial++;
                }
            }
            if (done) {
                if (lines) {
                    if ((w.step & (KKW_PROG_HALF_LINE - 1)) != KKW_PROG_HALF_LINE - 1) kkw_prog_flush(lbuf, path, w.step);
                } else if ((w.step & 3u) != 3u) {
                    grp.store(path, w.step);   // the last, partial group
                }
                A.lens[i] = (w.step + 1) | (capped ? 0x80000000u : 0u);
                active = false;
            }
        }
    }
    if (env) atomicAdd(A.violations, (unsigned long long)env);
}
