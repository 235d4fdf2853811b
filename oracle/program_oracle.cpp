/*
 * program_oracle.cpp -- CPU restatement of a walk program's walk loop.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle for walk programs, compiled by the
 * tests together with one program's source (g++ -DKKW_PROGRAM_SOURCE=...,
 * oracle.program_oracle), never by the product library.
 *
 * Restates the first-order epoch of include/walk.hpp:954-1064 (per walker),
 * and for a program with a query / response pair the second-order one of
 * walk.hpp:1135-1379 (query posted, answered at its target, verdict):
 *   - extension_comp once per step (954-973), or on every retry for a
 *     second-order program (walk.hpp:1161-1173); stop when it is 0, or when it
 *     is below 1 and the extension coin exceeds it;
 *   - trials until one is accepted: the outlier appendix when the program has
 *     one (981-1014): area overflow * num + regular_area[v], regular_area =
 *     (sequential weight sum, or degree) * upper[v] (walk.hpp:758-767); a dart
 *     above the regular area picks outlier idx = dart / overflow (clamped) and
 *     is accepted iff (dart - idx * overflow) / static(edge) + upper[v] <=
 *     dynamic_comp; otherwise a static draw -- uniform, or alias bucket of the
 *     weights: primary iff coin < p (1016-1029) -- and, with a dynamic
 *     component, accept iff dart = U * upper[v] is <= lower[v] (when given) or
 *     <= dynamic_comp (1030-1036);
 *   - step++, update_state(walker, current, edge), move (1039-1056).
 * Random words come from the oracle's Philox stream at the counters the
 * library's built-in walks use (kkw_oracle.h).
 *
 * Pinned by (tests/test_oracle_programs.py): programs restating DeepWalk, PPR
 * and node2vec -- its outlier pair included -- give orc_walk's paths bit for
 * bit; the tag walks of src/tests/test_bound.cpp and test_outlier.cpp and the
 * return-biased walk of src/examples/dynamic_walk.cpp follow their closed-form
 * transition laws; a program with a term dropped is rejected by that check.
 */
#include "kkw_program.h"

#include KKW_PROGRAM_SOURCE

#include <stdlib.h>

extern "C" {
#include "kkw_oracle.h"
}

#ifndef KKW_DYNAMIC
#define KKW_DYNAMIC 0
#endif
#ifndef KKW_LOWER
#define KKW_LOWER 0
#endif
#ifndef KKW_EDGE_INDEX
#define KKW_EDGE_INDEX 1   /* 0: the program never reads e.index of a static draw */
#endif
#ifndef KKW_OUTLIER
#define KKW_OUTLIER 0
#endif
#ifndef KKW_SECOND_ORDER
#define KKW_SECOND_ORDER 0
#endif

namespace {

/* The walker's last response; a second-order program's query is posted to
 * its target and answered there before dynamic_comp sees the answer
 * (walk.hpp:1226-1257 post, 1285-1294 respond, 1336-1364 verdict). */
#if KKW_SECOND_ORDER
typedef Response Resp;
#else
struct Resp {
};
#endif
#if KKW_DYNAMIC
float dynamic_of(const WalkerState &s, const kkw_walker &w, uint32_t cur, const kkw_edge &e, Resp &resp,
                 const kkw_ctx &c)
{
#if KKW_SECOND_ORDER
    uint32_t target = 0;
    Query q;
    const bool sent = post_query(s, w, cur, e, target, q, c);
    if (sent) respond_query(target, q, resp, c);
    return dynamic_comp(s, w, resp, cur, e, c);
#else
    (void)resp;
    return dynamic_comp(s, w, cur, e, c);
#endif
}
#endif

const uint32_t kMaxTrials = 1u << 20;

struct Draw {
    uint32_t w[4];
};

Draw draw(uint64_t seed, uint32_t walker, uint32_t step, uint32_t trial, uint32_t tag)
{
    const uint32_t ctr[4] = {walker, step, trial, tag};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    Draw d;
    orc_philox(ctr, key, d.w);
    return d;
}

struct Job {
    kkw_ctx c;
    const float *bp;          // local-index alias tables, or null: uniform
    const uint32_t *bprim;
    const uint32_t *balias;
    const float *vsum;
    uint64_t seed;
    uint32_t max_len;
    int retest_ext;
};

bool terminates(const WalkerState &s, const kkw_walker &w, uint32_t cur, uint32_t deg, uint32_t trial,
                const Job &J)
{
    if (deg == 0) return true;
    const float ext = extension(s, w, cur, J.c);
    if (ext == 0.0f) return true;
    if (ext < 1.0f) return orc_u01(draw(J.seed, w.id, w.step, trial, ORC_TAG_EXT).w[0]) > ext;
    return false;
}

bool trial_at(const WalkerState &s, const kkw_walker &w, uint32_t cur, uint64_t beg, uint32_t deg,
              uint32_t trial, const Job &J, Resp &resp, kkw_edge &acc)
{
    const kkw_ctx &c = J.c;
    (void)s;
    (void)resp;
#if KKW_DYNAMIC
    const float up = upper_bound(cur, c);
#endif
#if KKW_OUTLIER
    float ov = 0.0f;
    uint32_t num = 0;
    outlier_bound(s, w, cur, ov, num, c);
    if (ov > 0.0f && num > 0) {
        const float regular = (J.vsum ? J.vsum[cur] : (float)deg) * up;
        const float area = ov * (float)num + regular;
        const float rv = orc_u01(draw(J.seed, w.id, w.step, trial, ORC_TAG_OUTLIER).w[0]) * area - regular;
        if (rv > 0.0f) {
            uint32_t idx = (uint32_t)(rv / ov);
            if (idx + 1 > num) idx = num - 1;
            const int64_t ei = outlier_search(s, w, cur, idx, c);
            if (ei < 0) return false;
            kkw_edge e;
            e.index = (uint64_t)ei;
            e.neighbour = c.col[ei];
            e.reserved = 0;
            const float sc = J.bp ? c.weight[ei] : 1.0f;
            const float h = (rv - (float)idx * ov) / sc + up;
            if (!(h <= dynamic_of(s, w, cur, e, resp, c))) return false;
            acc = e;
            return true;
        }
    }
#endif
    const Draw r = draw(J.seed, w.id, w.step, trial, ORC_TAG_TRIAL);
    const uint32_t k = orc_bounded(r.w[0], r.w[1], deg);
    uint32_t local = k;
    if (J.bp) local = orc_u01(r.w[2]) < J.bp[beg + k] ? J.bprim[beg + k] : J.balias[beg + k];
    kkw_edge e;
    e.index = beg + local;
    e.neighbour = c.col[e.index];
    e.reserved = 0;
    if (!KKW_EDGE_INDEX) e.index = KKW_NO_EDGE_INDEX;
#if KKW_DYNAMIC
    const float dart = orc_u01(r.w[3]) * up;
    bool ok = false;
#if KKW_LOWER
    ok = dart <= lower_bound(cur, c);
#endif
    if (!ok) ok = dart <= dynamic_of(s, w, cur, e, resp, c);
    if (!ok) return false;
#endif
    acc = e;
    return true;
}

}  // namespace

/* One call runs walkers [walker_begin, +walker_num); paths row-major with
 * `stride` entries, lens = vertices | 1 << 31 when cut at max_len or by the
 * trial guard.  Returns 0. */
extern "C" int orc_program_walk(uint32_t V, const uint64_t *row_ptr, const uint32_t *col, const float *weight,
                                const uint8_t *etype, int sorted, const void *const *arrays, const float *params,
                                const float *bp, const uint32_t *bprim, const uint32_t *balias,
                                const float *vsum, uint64_t seed, uint32_t walker_begin, uint32_t walker_num,
                                uint32_t max_len, uint32_t stride, int start_mode, uint32_t n_starts,
                                const uint32_t *starts, int retest_ext, uint32_t *paths, uint32_t *lens)
{
    Job J;
    J.c.v_num = V;
    J.c.row_v = KKW_NO_VERTEX;   // no kept row: kkw_adjacent reads row_ptr
    J.c.row_deg = 0;
    J.c.row_beg = 0;
    J.c.sorted = sorted ? 1u : 0u;
    J.c.adjset = 0;   // kkw_adjacent answers by binary search or scan here
    J.c.e_num = row_ptr[V];
    J.c.row_ptr = row_ptr;
    J.c.col = col;
    J.c.weight = weight;
    J.c.etype = etype;
    for (int k = 0; k < KKW_PROGRAM_ARRAYS; k++) J.c.array[k] = arrays ? arrays[k] : 0;
    for (int k = 0; k < KKW_PROGRAM_PARAMS; k++) J.c.param[k] = params ? params[k] : 0.0f;
    J.bp = bp;
    J.bprim = bprim;
    J.balias = balias;
    J.vsum = vsum;
    J.seed = seed;
    J.max_len = max_len;
    J.retest_ext = retest_ext;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)walker_num; ii++) {
        const uint32_t i = (uint32_t)ii;
        kkw_walker w;
        w.id = walker_begin + i;
        w.step = 0;
        uint32_t cur;
        if (start_mode == ORC_START_LIST) {
            cur = starts[w.id % n_starts];
        } else if (start_mode == ORC_START_UNIFORM) {
            const Draw r = draw(seed, w.id, 0, 0, ORC_TAG_INIT);
            cur = orc_bounded(r.w[2], r.w[3], V);
        } else {
            cur = w.id % V;
        }
        WalkerState s;
        init_state(s, w, cur, J.c);
        Resp resp = Resp();
        const int retest = retest_ext || KKW_SECOND_ORDER;
        uint32_t *path = paths + (uint64_t)i * stride;
        path[0] = cur;
        int capped = 0;
        uint32_t trial = 0;
        for (;;) {
            const uint64_t beg = row_ptr[cur];
            const uint32_t deg = (uint32_t)(row_ptr[cur + 1] - beg);
            if (trial == 0 || retest) {
                if (terminates(s, w, cur, deg, retest ? trial : 0u, J)) break;
                if (w.step >= max_len) {
                    capped = 1;
                    break;
                }
            }
            if (trial >= kMaxTrials) {
                capped = 1;
                break;
            }
            kkw_edge e;
            if (!trial_at(s, w, cur, beg, deg, trial, J, resp, e)) {
                trial++;
                continue;
            }
            w.step++;
            update_state(s, w, cur, e, J.c);
            cur = e.neighbour;
            path[w.step] = cur;
            trial = 0;
        }
        lens[i] = (w.step + 1) | (capped ? 0x80000000u : 0u);
    }
    return 0;
}
