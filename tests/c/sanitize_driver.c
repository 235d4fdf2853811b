/* Runs the oracle's functions on small inputs so that a build under
 * -fsanitize=address,undefined checks them for memory and undefined-behaviour
 * errors: Philox, R-MAT generation and the undirected build, the file-order
 * CSR, both alias constructions, every walk kind (node2vec with and without
 * the outlier appendix, uniform and list starts, sampled walker ids) on one
 * and two threads, and -- when built with WITH_PROGRAM -- the walk-program
 * loop with a second-order program.  Exit status 0 when every walk follows
 * CSR edges. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kkw_oracle.h"

#ifdef WITH_PROGRAM
int orc_program_walk(uint32_t V, const uint64_t *row_ptr, const uint32_t *col, const float *weight,
                     const uint8_t *etype, int sorted, const void *const *arrays, const float *params,
                     const float *bp, const uint32_t *bprim, const uint32_t *balias, const float *vsum,
                     uint64_t seed, uint32_t walker_begin, uint32_t walker_num, uint32_t max_len, uint32_t stride,
                     int start_mode, uint32_t n_starts, const uint32_t *starts, int retest_ext, uint32_t *paths,
                     uint32_t *lens);
#endif

static int is_edge(const uint64_t *rp, const uint32_t *col, uint32_t u, uint32_t v)
{
    for (uint64_t e = rp[u]; e < rp[u + 1]; e++)
        if (col[e] == v) return 1;
    return 0;
}

static int check_paths(const uint64_t *rp, const uint32_t *col, const uint32_t *paths, const uint32_t *lens,
                       uint32_t n, uint32_t stride)
{
    for (uint32_t i = 0; i < n; i++) {
        uint32_t len = lens[i] & 0x7fffffffu;
        if (len == 0 || len > stride) return 1;
        for (uint32_t k = 1; k < len; k++)
            if (!is_edge(rp, col, paths[(uint64_t)i * stride + k - 1], paths[(uint64_t)i * stride + k])) return 1;
    }
    return 0;
}

int main(void)
{
    const uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
    uint32_t out[4];
    orc_philox(ctr, key, out);
    if (out[0] != 0x6627e8d5u) return 10;   /* Random123 philox4x32-10 known answer */

    const int scale = 9;
    const uint32_t V = 1u << scale;
    const uint64_t n = 16ull << scale;
    uint32_t *src = malloc(n * 4), *dst = malloc(n * 4);
    orc_rmat_raw(scale, n, 2448131358u, 3264175144u, 4080218931u, 1, src, dst);
    uint64_t *rp = calloc(V + 1, 8);
    uint32_t *col = malloc(2 * n * 4);
    float *w = malloc(2 * n * 4);
    const uint64_t E = orc_build_undirected(V, src, dst, n, 1, 1, rp, col, w);
    uint8_t *vt = malloc(V), *et = malloc(E ? E : 1), *vm = calloc(V, 1);
    orc_vertex_types(V, 5, 3, vt);
    for (uint32_t v = 0; v < V; v++)
        for (uint64_t e = rp[v]; e < rp[v + 1]; e++) {
            et[e] = vt[col[e]];
            vm[v] |= (uint8_t)(1u << et[e]);
        }
    float *bp = malloc(E * 4), *vs = malloc(V * 4), *lbp = malloc(E * 4), *lvs = malloc(V * 4);
    uint32_t *bprim = malloc(E * 4), *balias = malloc(E * 4), *bmeta = malloc(E * 4);
    uint32_t *lprim = malloc(E * 4), *lalias = malloc(E * 4);
    orc_alias_build(V, rp, col, w, et, bp, bprim, balias, bmeta, vs);
    orc_alias_build_local(V, rp, w, lbp, lprim, lalias, lvs);
    /* the file-order CSR of the raw edges, mirrored and sorted */
    uint64_t *rp2 = calloc(V + 1, 8);
    uint32_t *col2 = malloc(2 * n * 4);
    float *w2 = malloc(2 * n * 4), *wr = malloc(n * 4);
    for (uint64_t i = 0; i < n; i++) wr[i] = 1.0f + (float)(i % 4);
    orc_csr_from_edges(V, src, dst, wr, n, 1, 1, rp2, col2, w2);

    orc_graph g = {V, E, rp, col, w, et, bp, bprim, balias, bmeta, vs};
    const uint32_t walkers = 3000, L = 24, stride = 200;
    uint32_t *paths = malloc((uint64_t)walkers * stride * 4), *lens = malloc(walkers * 4);
    const uint32_t starts[3] = {1, 5, 77};
    const uint32_t scheme_len[2] = {2, 3};
    const uint8_t scheme_mask[16] = {3, 28, 0, 0, 0, 0, 0, 0, 5, 10, 31, 0, 0, 0, 0, 0};
    int bad = 0;
    for (int threads = 1; threads <= 2; threads++) {
        for (int kind = 0; kind < 4; kind++) {
            for (int weighted = 0; weighted < 2; weighted++) {
                orc_job j;
                memset(&j, 0, sizeof(j));
                j.kind = kind;
                j.use_alias = weighted;
                j.walker_begin = 11;
                j.walker_num = walkers;
                j.walk_length = kind == ORC_PPR ? 150 : L;
                j.path_stride = stride;
                j.seed = 7 + kind;
                j.ppr_ext = 1.0f - 0.15f;
                j.start_mode = kind == ORC_PPR ? ORC_START_LIST : (weighted ? ORC_START_UNIFORM : ORC_START_EQUAL);
                j.n_starts = 3;
                j.starts = starts;
                j.n_schemes = 2;
                j.scheme_len = scheme_len;
                j.scheme_mask = scheme_mask;
                j.vmask = vm;
                j.walker_stride = weighted ? 3 : 1;
                orc_node2vec_constants(weighted ? 2.0f : 0.5f, weighted ? 0.5f : 2.0f, &j.n2v_upper, &j.n2v_lower,
                                       &j.n2v_inv_p, &j.n2v_inv_q, &j.n2v_overflow);
                orc_stats st;
                orc_walk(&g, &j, paths, lens, threads, &st);
                bad |= check_paths(rp, col, paths, lens, walkers, stride);
            }
        }
    }
#ifdef WITH_PROGRAM
    {
        float up, lo, ip, iq, ov;
        orc_node2vec_constants(0.5f, 2.0f, &up, &lo, &ip, &iq, &ov);
        const float params[8] = {(float)L, up, lo, ip, iq, ov, 1.0f, 0.0f};
        const void *arrays[8] = {0};
        orc_program_walk(V, rp, col, w, et, 1, arrays, params, lbp, lprim, lalias, lvs, 5, 0, walkers, L, L + 4,
                         ORC_START_EQUAL, 0, 0, 1, paths, lens);
        bad |= check_paths(rp, col, paths, lens, walkers, L + 4);
    }
#endif
    free(src); free(dst); free(rp); free(col); free(w); free(vt); free(et); free(vm);
    free(bp); free(vs); free(lbp); free(lvs); free(bprim); free(balias); free(bmeta); free(lprim); free(lalias);
    free(rp2); free(col2); free(w2); free(wr); free(paths); free(lens);
    printf("sanitized oracle run: %s (E = %llu)\n", bad ? "bad paths" : "ok", (unsigned long long)E);
    return bad ? 1 : 0;
}
