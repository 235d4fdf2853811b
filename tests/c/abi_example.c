/* The C ABI from plain C (C11, no C++ runtime): load the karate text edge
 * list, build alias tables and hash sets, run node2vec, dump the PathSet text
 * format, then run a batch of two jobs and read one back.  Every call's status
 * is checked; the first failure prints kkw_last_error() and exits non-zero.
 * usage: abi_example <karate.txt> <out_file> */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kkw.h"

#define CHECK(call)                                                                   \
    do {                                                                              \
        int rc_ = (call);                                                             \
        if (rc_ != KKW_OK) {                                                          \
            fprintf(stderr, "%s -> %d: %s\n", #call, rc_, kkw_last_error());          \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    kkw_engine *eng = NULL;
    kkw_engine_opts o;
    memset(&o, 0, sizeof o);
    CHECK(kkw_engine_create(&o, &eng));
    CHECK(kkw_graph_load_file(eng, 34, argv[1], /*text*/ 1, /*weighted*/ 1, /*undirected*/ 0, /*sort_adj*/ 1));
    uint32_t v = 0;
    uint64_t e = 0;
    CHECK(kkw_graph_info(eng, &v, &e));
    if (v != 34 || e != 156) {
        fprintf(stderr, "karate: %u vertices, %llu entries\n", v, (unsigned long long)e);
        return 3;
    }
    CHECK(kkw_prepare(eng));

    kkw_job job;
    memset(&job, 0, sizeof job);
    job.kind = KKW_NODE2VEC;
    job.use_weights = 1;
    job.walker_num = 34;
    job.walk_length = 10;
    job.p = 2.0f;
    job.q = 0.5f;
    job.seed = 5;
    kkw_paths out;
    CHECK(kkw_run(eng, &job, 1, KKW_MODE_PERSISTENT, KKW_OUT_HOST, 0, &out));
    if (out.walker_num != 34 || out.steps != 34 * 10) return 4;
    CHECK(kkw_paths_dump(&out, argv[2], "w", 1));
    CHECK(kkw_paths_free(&out));

    /* a resident batch: two jobs, launched twice, same paths each time */
    kkw_job jobs[2];
    memcpy(&jobs[0], &job, sizeof job);
    memset(&jobs[1], 0, sizeof job);
    jobs[1].kind = KKW_DEEPWALK;
    jobs[1].use_weights = 1;
    jobs[1].walker_num = 100;
    jobs[1].walk_length = 7;
    jobs[1].seed = 9;
    kkw_batch *b = NULL;
    CHECK(kkw_batch_create(eng, jobs, 2, KKW_MODE_PERSISTENT, KKW_BATCH_GRAPH, &b));
    double ms = 0;
    kkw_paths p1, p2;
    CHECK(kkw_batch_launch(b));
    CHECK(kkw_batch_sync(b, &ms));
    CHECK(kkw_batch_paths(b, 1, KKW_OUT_HOST, &p1));
    CHECK(kkw_batch_launch(b));
    CHECK(kkw_batch_sync(b, &ms));
    CHECK(kkw_batch_paths(b, 1, KKW_OUT_HOST, &p2));
    if (p1.walker_num != 100 || memcmp(p1.lengths, p2.lengths, 100 * sizeof(uint32_t)) != 0) return 5;
    for (uint32_t i = 0; i < 100; i++)
        if (memcmp(p1.vertices + (size_t)i * p1.stride, p2.vertices + (size_t)i * p2.stride,
                   (p1.lengths[i] & 0x7fffffffu) * sizeof(uint32_t)) != 0)
            return 6;
    CHECK(kkw_paths_free(&p1));
    CHECK(kkw_paths_free(&p2));
    CHECK(kkw_batch_destroy(b));

    /* errors come back as status codes */
    job.kind = 99;
    if (kkw_run(eng, &job, 1, KKW_MODE_PLAIN, KKW_OUT_HOST, 0, &out) != KKW_E_INVALID) return 7;
    if (strlen(kkw_last_error()) == 0) return 8;
    CHECK(kkw_engine_destroy(eng));
    printf("abi example ok\n");
    return 0;
}
