/*
 * kkw_oracle.c -- CPU restatement of KnightKing's random-walk hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see kkw_oracle.h): used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, to compare the
 * GPU results against.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * -ffp-contract=off keeps every float operation a separately rounded IEEE
 * single operation, as on the GPU side, so the float comparisons that decide
 * each step come out the same.
 */
#include "kkw_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11, section 4; constants from the paper) */
/* ------------------------------------------------------------------------ */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += PH_W0; k1 += PH_W1; }
        uint64_t p0 = (uint64_t)PH_M0 * c0;
        uint64_t p1 = (uint64_t)PH_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* u32 -> float in [0,1) on the 2^-24 grid: exact in float, never 1. */
float orc_u01(uint32_t u)
{
    return (float)(u >> 8) * 5.9604644775390625e-08f;
}

/* floor(r * n / 2^64) for the 64-bit r = hi:lo -- uniform index in [0,n)
 * (bias below n / 2^64).  Restates gen(n) of include/util.hpp:60-64. */
uint32_t orc_bounded(uint32_t lo, uint32_t hi, uint32_t n)
{
    uint64_t a = (uint64_t)hi * n;
    uint64_t b = (uint64_t)lo * n;
    return (uint32_t)((a + (b >> 32)) >> 32);
}

static inline void draw(uint64_t seed, uint32_t walker, uint32_t step,
                        uint32_t trial, uint32_t tag, uint32_t out[4])
{
    uint32_t ctr[4] = {walker, step, trial, tag};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox(ctr, key, out);
}

/* ------------------------------------------------------------------------ */
/* Synthetic R-MAT graphs                                                    */
/* ------------------------------------------------------------------------ */
/* Edge i, levels 4g..4g+3 use the four words of Philox(key=seed,
 * ctr=(i_lo, i_hi, g, TAG_RMAT)).  Level l sets bit (scale-1-l) of src/dst
 * by the quadrant a (0,0), b (0,1), c (1,0), d (1,1) chosen by integer
 * thresholds on the word.  No noise, no vertex permutation. */
void orc_rmat_raw(int scale, uint64_t n_edges, uint32_t thr_a, uint32_t thr_ab,
                  uint32_t thr_abc, uint64_t seed, uint32_t *src, uint32_t *dst)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n_edges; i++) {
        uint32_t s = 0, d = 0;
        for (int g = 0; g * 4 < scale; g++) {
            uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)g, ORC_TAG_RMAT};
            uint32_t r[4];
            orc_philox(ctr, key, r);
            for (int j = 0; j < 4 && g * 4 + j < scale; j++) {
                int l = g * 4 + j;
                uint32_t bit = 1u << (scale - 1 - l);
                uint32_t u = r[j];
                if (u < thr_a) {
                } else if (u < thr_ab) {
                    d |= bit;
                } else if (u < thr_abc) {
                    s |= bit;
                } else {
                    s |= bit;
                    d |= bit;
                }
            }
        }
        src[i] = s;
        dst[i] = d;
    }
}

/* Weight of the undirected edge {u,v}: 1 + 4*U in [1,5), keyed by the
 * canonical pair so both directions carry the same value. */
float orc_edge_weight(uint64_t seed, uint32_t u, uint32_t v)
{
    uint32_t a = u < v ? u : v, b = u < v ? v : u;
    uint32_t ctr[4] = {a, b, 0, ORC_TAG_WEIGHT};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox(ctr, key, r);
    return 1.0f + 4.0f * orc_u01(r[0]);
}

void orc_vertex_types(uint32_t V, uint32_t ntypes, uint64_t seed, uint8_t *vtype)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t v = 0; v < V; v++) {
        uint32_t ctr[4] = {v, 0, 0, ORC_TAG_VTYPE};
        uint32_t r[4];
        orc_philox(ctr, key, r);
        vtype[v] = (uint8_t)(((uint64_t)r[0] * ntypes) >> 32);
    }
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

/* Drop self loops, canonicalise to (min,max), dedupe, mirror, and order by
 * (src, dst): the undirected simple graph that load_graph(--make-undirected)
 * (include/graph.hpp:430-443) followed by node2vec's per-list sort
 * (src/apps/node2vec.hpp:122-125) would hold for a deduplicated edge file.
 * Returns E; arrays must hold 2*n entries. */
uint64_t orc_build_undirected(uint32_t V, const uint32_t *src, const uint32_t *dst,
                              uint64_t n, uint64_t seed, int weighted,
                              uint64_t *row_ptr, uint32_t *col, float *w)
{
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t a = src[i], b = dst[i];
        if (a == b) continue;
        if (a > b) { uint32_t t = a; a = b; b = t; }
        keys[m++] = ((uint64_t)a << 32) | b;
    }
    qsort(keys, m, sizeof(uint64_t), cmp_u64);
    uint64_t u = 0;
    for (uint64_t i = 0; i < m; i++)
        if (i == 0 || keys[i] != keys[i - 1]) keys[u++] = keys[i];
    uint64_t *dir = (uint64_t *)malloc(sizeof(uint64_t) * (u ? 2 * u : 1));
    for (uint64_t i = 0; i < u; i++) {
        dir[2 * i] = keys[i];
        dir[2 * i + 1] = (keys[i] << 32) | (keys[i] >> 32);
    }
    qsort(dir, 2 * u, sizeof(uint64_t), cmp_u64);
    memset(row_ptr, 0, sizeof(uint64_t) * ((size_t)V + 1));
    for (uint64_t i = 0; i < 2 * u; i++) row_ptr[(dir[i] >> 32) + 1]++;
    for (uint32_t v = 0; v < V; v++) row_ptr[v + 1] += row_ptr[v];
    for (uint64_t i = 0; i < 2 * u; i++) {
        uint32_t s = (uint32_t)(dir[i] >> 32), d = (uint32_t)dir[i];
        col[i] = d;
        if (weighted && w) w[i] = orc_edge_weight(seed, s, d);
    }
    free(keys);
    free(dir);
    return 2 * u;
}

/* CSR in file order (include/graph.hpp:309-330, one partition: shuffle_edges
 * at 332-406 keeps file order).  Mirroring as at graph.hpp:430-443: edge i
 * becomes entries 2i (as read) and 2i+1 (reversed, same data).  sort_adj
 * restates node2vec's per-list std::sort by neighbour (node2vec.hpp:122-125);
 * std::sort is unstable, so among equal neighbours the order is unspecified
 * there and stable (file order) here. */
typedef struct { uint32_t d; uint32_t pos; } nbr_pos;
static int cmp_nbr(const void *a, const void *b)
{
    const nbr_pos *x = (const nbr_pos *)a, *y = (const nbr_pos *)b;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

void orc_csr_from_edges(uint32_t V, const uint32_t *src, const uint32_t *dst,
                        const float *w, uint64_t n, int undirected, int sort_adj,
                        uint64_t *row_ptr, uint32_t *col, float *wout)
{
    uint64_t m = undirected ? 2 * n : n;
    memset(row_ptr, 0, sizeof(uint64_t) * ((size_t)V + 1));
    for (uint64_t i = 0; i < n; i++) {
        row_ptr[src[i] + 1]++;
        if (undirected) row_ptr[dst[i] + 1]++;
    }
    for (uint32_t v = 0; v < V; v++) row_ptr[v + 1] += row_ptr[v];
    uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)V + 1));
    memcpy(fill, row_ptr, sizeof(uint64_t) * ((size_t)V + 1));
    for (uint64_t i = 0; i < n; i++) {
        uint64_t e = fill[src[i]]++;
        col[e] = dst[i];
        if (w && wout) wout[e] = w[i];
        if (undirected) {
            e = fill[dst[i]]++;
            col[e] = src[i];
            if (w && wout) wout[e] = w[i];
        }
    }
    (void)m;
    if (sort_adj) {
        for (uint32_t v = 0; v < V; v++) {
            uint64_t b = row_ptr[v], e = row_ptr[v + 1], d = e - b;
            if (d < 2) continue;
            nbr_pos *t = (nbr_pos *)malloc(sizeof(nbr_pos) * d);
            float *tw = (float *)malloc(sizeof(float) * d);
            for (uint64_t k = 0; k < d; k++) { t[k].d = col[b + k]; t[k].pos = (uint32_t)k; }
            qsort(t, d, sizeof(nbr_pos), cmp_nbr);
            if (w && wout) for (uint64_t k = 0; k < d; k++) tw[k] = wout[b + t[k].pos];
            for (uint64_t k = 0; k < d; k++) {
                col[b + k] = t[k].d;
                if (w && wout) wout[b + k] = tw[k];
            }
            free(t);
            free(tw);
        }
    }
    free(fill);
}

/* ------------------------------------------------------------------------ */
/* Alias tables: include/walk.hpp:538-652                                    */
/* ------------------------------------------------------------------------ */
/* Per vertex with d > 0: prob[e] = P_s(e) (the edge weight,
 * src/apps/static_comp.hpp:36-41); sum is the sequential float sum
 * (walk.hpp:591-595); ave = sum / d (walk.hpp:600); edges split in order
 * into small (< ave) and large FIFO queues (601-615); each small is paired
 * with the next large: bucket = {p = prob[s]/ave, primary s, alias l},
 * prob[l] -= ave - prob[s], l re-queued by its new value (617-634); leftovers
 * get p = 1 and no alias (635-647).  Buckets are stored in creation order.
 * "No alias" is stored as alias = primary with p = 1, which selects the
 * primary exactly as `q_ptr == nullptr` does (p < 1 for every paired bucket,
 * and U < 1 always).  bmeta packs the edge types of primary | alias << 8. */
/* local != 0: bprim/balias receive local edge indices (position in the list)
 * instead of neighbour ids -- the reference's p_ptr/q_ptr point at edges. */
static void alias_build(uint32_t V, const uint64_t *row_ptr, const uint32_t *col,
                        const float *w, const uint8_t *etype,
                        float *bp, uint32_t *bprim, uint32_t *balias, uint32_t *bmeta,
                        float *vsum, int local)
{
    uint64_t maxd = 0;
    for (uint32_t v = 0; v < V; v++) {
        uint64_t d = row_ptr[v + 1] - row_ptr[v];
        if (d > maxd) maxd = d;
    }
    #pragma omp parallel
    {
        float *prob = (float *)malloc(sizeof(float) * (maxd ? maxd : 1));
        uint32_t *ps = (uint32_t *)malloc(sizeof(uint32_t) * (maxd ? maxd : 1));
        uint32_t *qs = (uint32_t *)malloc(sizeof(uint32_t) * (maxd ? maxd : 1));
        #pragma omp for schedule(dynamic, 128)
        for (int64_t vi = 0; vi < (int64_t)V; vi++) {
            uint32_t v = (uint32_t)vi;
            uint64_t b = row_ptr[v];
            uint32_t d = (uint32_t)(row_ptr[v + 1] - b);
            if (d == 0) { if (vsum) vsum[v] = 0.0f; continue; }
            float sum = 0.0f;
            for (uint32_t e = 0; e < d; e++) {
                prob[e] = w[b + e];
                sum += prob[e];
            }
            if (vsum) vsum[v] = sum;
            float ave = sum / (float)d;
            uint32_t ph = 0, pt = 0, qh = 0, qt = 0;
            for (uint32_t e = 0; e < d; e++) {
                if (prob[e] < ave) ps[pt++] = e;
                else qs[qt++] = e;
            }
            uint64_t o = b;
            while (pt != ph) {
                uint32_t pi = ps[ph++];
                if (qt != qh) {
                    uint32_t qi = qs[qh++];
                    bp[o] = prob[pi] / ave;
                    bprim[o] = local ? pi : col[b + pi];
                    balias[o] = local ? qi : col[b + qi];
                    bmeta[o] = etype ? ((uint32_t)etype[b + pi] | ((uint32_t)etype[b + qi] << 8)) : 0u;
                    float diff = ave - prob[pi];
                    prob[qi] -= diff;
                    if (prob[qi] < ave) ps[pt++] = qi;
                    else qs[qt++] = qi;
                } else {
                    bp[o] = 1.0f;
                    bprim[o] = local ? pi : col[b + pi];
                    balias[o] = bprim[o];
                    bmeta[o] = etype ? ((uint32_t)etype[b + pi] * 0x101u) : 0u;
                }
                o++;
            }
            while (qt != qh) {
                uint32_t qi = qs[qh++];
                bp[o] = 1.0f;
                bprim[o] = local ? qi : col[b + qi];
                balias[o] = bprim[o];
                bmeta[o] = etype ? ((uint32_t)etype[b + qi] * 0x101u) : 0u;
                o++;
            }
        }
        free(prob);
        free(ps);
        free(qs);
    }
}

void orc_alias_build(uint32_t V, const uint64_t *row_ptr, const uint32_t *col,
                     const float *w, const uint8_t *etype,
                     float *bp, uint32_t *bprim, uint32_t *balias, uint32_t *bmeta,
                     float *vsum)
{
    alias_build(V, row_ptr, col, w, etype, bp, bprim, balias, bmeta, vsum, 0);
}

void orc_alias_build_local(uint32_t V, const uint64_t *row_ptr, const float *w,
                           float *bp, uint32_t *bprim, uint32_t *balias, float *vsum)
{
    uint64_t E = row_ptr[V];
    uint32_t *meta = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    alias_build(V, row_ptr, NULL, w, NULL, bp, bprim, balias, meta, vsum, 1);
    free(meta);
}

/* ------------------------------------------------------------------------ */
/* Walks                                                                     */
/* ------------------------------------------------------------------------ */
/* node2vec bound constants exactly as src/apps/node2vec.hpp:126-127 and 136
 * compute them: std::max/min over doubles, stored as real_t (float); the
 * dynamic components `1 / p` and `1 / q` (node2vec.hpp:182-191) are float
 * divisions. */
void orc_node2vec_constants(float p, float q, float *upper, float *lower,
                            float *inv_p, float *inv_q, float *overflow)
{
    double iq = 1.0 / (double)q, ip = 1.0 / (double)p;
    double up = 1.0 > iq ? 1.0 : iq;
    double m1 = 1.0 < iq ? 1.0 : iq;
    double lo = ip < m1 ? ip : m1;
    *upper = (float)up;
    *lower = (float)lo;
    *inv_p = 1.0f / p;
    *inv_q = 1.0f / q;
    *overflow = (float)(1.0 / (double)p - (double)(float)up);
}

/* lower_bound over col[b..e); returns 1 when x is present. */
static int adj_search(const uint32_t *col, uint64_t b, uint64_t e, uint32_t x,
                      uint64_t *probes, uint64_t *pos)
{
    uint64_t lo = b, len = e - b;
    while (len > 0) {
        uint64_t half = len >> 1;
        uint64_t mid = lo + half;
        (*probes)++;
        if (col[mid] < x) {
            lo = mid + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    if (pos) *pos = lo;
    return lo < e && col[lo] == x;
}

static uint32_t start_vertex(const orc_graph *g, const orc_job *job, uint32_t wid)
{
    /* walk.hpp:403-421 (equal / uniform), ppr.hpp:38-50 (list) */
    if (job->start_mode == ORC_START_LIST) return job->starts[wid % job->n_starts];
    if (job->start_mode == ORC_START_UNIFORM) {
        uint32_t r[4];
        draw(job->seed, wid, 0, 0, ORC_TAG_INIT, r);
        return orc_bounded(r[2], r[3], g->V);
    }
    return wid % g->V;
}

/* Candidate by alias bucket or uniform index: walk.hpp:1016-1029. */
static inline uint32_t pick_candidate(const orc_graph *g, const orc_job *job,
                                      uint64_t beg, uint32_t deg, const uint32_t r[4],
                                      uint32_t *etype_out)
{
    uint32_t idx = orc_bounded(r[0], r[1], deg);
    if (!job->use_alias) {
        if (etype_out) *etype_out = g->etype ? g->etype[beg + idx] : 0;
        return g->col[beg + idx];
    }
    uint64_t k = beg + idx;
    int take_primary = orc_u01(r[2]) < g->bp[k];
    if (etype_out) *etype_out = take_primary ? (g->bmeta[k] & 0xff) : ((g->bmeta[k] >> 8) & 0xff);
    return take_primary ? g->bprim[k] : g->balias[k];
}

/* First-order walk: include/walk.hpp:954-1064 for one walker, with the
 * extension functions of deepwalk.hpp:35-37, ppr.hpp:51-55 and
 * metapath.hpp:102-121, the metapath dynamic component and bound
 * (metapath.hpp:107-120) and state update (metapath.hpp:96-99). */
/* Trials per step before a walker is stopped and flagged; only degenerate
 * inputs (every allowed edge of zero weight) come near it. */
#define ORC_MAX_TRIALS (1u << 20)
#define ORC_TRUNCATED 0x80000000u

static void walk_first_order(const orc_graph *g, const orc_job *job, uint32_t wid,
                             uint32_t *path, uint32_t *len_out, orc_stats *st)
{
    uint32_t flag = 0;
    uint32_t cur = start_vertex(g, job, wid);
    uint32_t step = 0;
    uint32_t scheme = 0, state = 0;
    if (job->kind == ORC_METAPATH) {
        uint32_t r[4];
        draw(job->seed, wid, 0, 0, ORC_TAG_INIT, r);
        scheme = orc_bounded(r[0], r[1], job->n_schemes);   /* metapath.hpp:93 */
    }
    path[0] = cur;
    for (;;) {
        uint64_t beg = g->row_ptr[cur];
        uint32_t deg = (uint32_t)(g->row_ptr[cur + 1] - beg);
        if (deg == 0) break;
        float ext;
        if (job->kind == ORC_PPR) ext = job->ppr_ext;
        else if (job->kind == ORC_METAPATH)
            ext = (step >= job->walk_length ||
                   !(job->vmask[cur] & job->scheme_mask[scheme * 8 + state])) ? 0.0f : 1.0f;
        else ext = step >= job->walk_length ? 0.0f : 1.0f;
        if (ext == 0.0f) break;
        if (ext < 1.0f) {
            uint32_t r[4];
            draw(job->seed, wid, step, 0, ORC_TAG_EXT, r);
            if (orc_u01(r[0]) > ext) break;
        }
        if (step >= job->walk_length) {  /* step cap (PPR only reaches this) */
            st->truncated++;
            flag = ORC_TRUNCATED;
            break;
        }
        uint32_t cand = 0;
        for (uint32_t trial = 0;; trial++) {
            if (trial >= ORC_MAX_TRIALS) {
                st->truncated++;
                *len_out = (step + 1) | ORC_TRUNCATED;
                return;
            }
            uint32_t r[4];
            draw(job->seed, wid, step, trial, ORC_TAG_TRIAL, r);
            uint32_t et = 0;
            st->trials++;
            cand = pick_candidate(g, job, beg, deg, r, &et);
            if (job->kind == ORC_METAPATH) {
                float dart = orc_u01(r[3]) * 1.0f;
                float pd = ((job->scheme_mask[scheme * 8 + state] >> et) & 1u) ? 1.0f : 0.0f;
                if (dart <= pd) break;
            } else {
                break;
            }
        }
        step++;
        if (job->kind == ORC_METAPATH) state = (state + 1) % job->scheme_len[scheme];
        path[step] = cand;
        cur = cand;
        st->steps++;
    }
    *len_out = (step + 1) | flag;
}

/* Second-order walk for node2vec: the per-walker sequence that the three
 * phases of include/walk.hpp:1135-1379 produce for one walker, with the
 * node2vec functions of src/apps/node2vec.hpp:128-196 (query = binary search
 * of the candidate in the previous vertex's sorted list, 158-166; P_d at
 * 170-193) and the outlier appendix of walk.hpp:1181-1208 with
 * Node2vecOutlierFuncWrapper (node2vec.hpp:47-106) when 1/p exceeds the upper
 * bound.  After a rejection the walker is re-tried with a fresh extension
 * test, as the reference re-emits it (walk.hpp:1361-1364, 1161-1173); the
 * extension coin of each trial therefore has its own counter. */
static void walk_node2vec(const orc_graph *g, const orc_job *job, uint32_t wid,
                          uint32_t *path, uint32_t *len_out, orc_stats *st)
{
    uint32_t cur = start_vertex(g, job, wid);
    uint32_t prev = 0;
    uint64_t pbeg = 0, pend = 0;
    uint32_t step = 0, flag = 0;
    const float upper = job->n2v_upper, lower = job->n2v_lower;
    const int outlier = job->n2v_overflow > 0.0f;
    path[0] = cur;
    for (;;) {
        uint64_t beg = g->row_ptr[cur];
        uint64_t end = g->row_ptr[cur + 1];
        uint32_t deg = (uint32_t)(end - beg);
        for (uint32_t trial = 0;; trial++) {
            /* termination test, re-run on every retry (walk.hpp:1161-1173);
             * extension_comp = step >= L ? 0 : 1 (node2vec.hpp:147-150) */
            if (deg == 0 || step >= job->walk_length) goto done;
            if (trial >= ORC_MAX_TRIALS) {
                st->truncated++;
                flag = ORC_TRUNCATED;
                goto done;
            }
            st->trials++;
            uint32_t cand = 0;
            float dart = 0.0f;
            int have = 0, pre = 0;
            if (outlier && step > 0) {
                /* Appendix (walk.hpp:1181-1208): outlier_upperbound_func gives
                 * (w(return edge) or 1) * overflow and one outlier
                 * (node2vec.hpp:49-84); the regular area is sum(P_s) * upper
                 * (walk.hpp:756-767).  The return edge exists in an undirected
                 * graph; where it does not, the appendix is empty. */
                uint64_t rpos = 0, probes = 0;
                if (adj_search(g->col, beg, end, prev, &probes, &rpos)) {
                    float ub = job->use_alias ? g->w[rpos] * job->n2v_overflow : job->n2v_overflow;
                    float regular = job->use_alias ? g->vsum[cur] * upper : (float)deg * upper;
                    if (ub > 0.0f) {
                        uint32_t r[4];
                        draw(job->seed, wid, step, trial, ORC_TAG_OUTLIER, r);
                        float rv = orc_u01(r[0]) * (ub * 1.0f + regular) - regular;
                        if (rv > 0.0f) {
                            uint32_t oidx = (uint32_t)(rv / ub);
                            if (oidx + 1 > 1) oidx = 0;
                            float sc = job->use_alias ? g->w[rpos] : 1.0f;
                            dart = (rv - (float)oidx * ub) / sc + upper;
                            cand = g->col[rpos];
                            have = 1;
                        }
                    }
                }
            }
            if (!have) {
                uint32_t r[4];
                draw(job->seed, wid, step, trial, ORC_TAG_TRIAL, r);
                cand = pick_candidate(g, job, beg, deg, r, NULL);
                dart = orc_u01(r[3]) * upper;                  /* walk.hpp:1225 */
                pre = dart <= lower;                           /* walk.hpp:1226-1229 */
            }
            int accept;
            if (pre) {
                accept = 1;
            } else if (step == 0) {
                accept = dart <= upper;                        /* node2vec.hpp:180-182 */
            } else if (cand == prev) {
                accept = dart <= job->n2v_inv_p;               /* node2vec.hpp:185-186 */
            } else {
                uint64_t probes = 0;
                int found = adj_search(g->col, pbeg, pend, cand, &probes, NULL);
                st->searches++;
                st->search_probes += probes;
                accept = dart <= (found ? 1.0f : job->n2v_inv_q);  /* node2vec.hpp:187-191 */
            }
            if (accept) {
                step++;
                path[step] = cand;
                prev = cur;                                    /* node2vec.hpp:132-135 */
                pbeg = beg;
                pend = end;
                cur = cand;
                st->steps++;
                break;
            }
        }
    }
done:
    *len_out = (step + 1) | flag;
}

void orc_walk(const orc_graph *g, const orc_job *job, uint32_t *paths,
              uint32_t *lens, int threads, orc_stats *stats)
{
    orc_stats total = {0, 0, 0, 0, 0};
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    threads = 1;
#endif
    #pragma omp parallel num_threads(threads)
    {
        orc_stats st = {0, 0, 0, 0, 0};
        #pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)job->walker_num; i++) {
            uint32_t wid = job->walker_begin + (uint32_t)i * (job->walker_stride ? job->walker_stride : 1u);
            uint32_t *path = paths + (uint64_t)i * job->path_stride;
            if (job->kind == ORC_NODE2VEC)
                walk_node2vec(g, job, wid, path, &lens[i], &st);
            else
                walk_first_order(g, job, wid, path, &lens[i], &st);
        }
        #pragma omp critical
        {
            total.steps += st.steps;
            total.trials += st.trials;
            total.searches += st.searches;
            total.search_probes += st.search_probes;
            total.truncated += st.truncated;
        }
    }
    if (stats) *stats = total;
}
