// engine.hip -- the C ABI of libkkw (include/kkw.h): engine state, graph
// loading, job validation, resident batches, kkw_run, path output, the
// autotuner and the serving loop.
#include "engine_state.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include <sys/prctl.h>

using namespace kkw;

namespace {

uint32_t ppr_default_cap(float t)
{
    // P(walk longer than cap) = (1-t)^cap <= 1e-12
    if (!(t > 0.0f)) return 0;
    double c = std::ceil(27.7 / (double)t);
    return (uint32_t)std::min(c, 1048576.0);
}

}  // namespace

// --------------------------------------------------------------------------
struct ServeState;


// Host copy of a validated job plus its device-side arrays.
struct PreparedJob {
    kkw_job src{};
    DevJob dj{};
    uint32_t stride = 0;
    std::vector<uint32_t> starts;
    std::vector<uint32_t> scheme_len;
    std::vector<uint8_t> scheme_mask;
};

static void node2vec_constants(float p, float q, DevJob &d)
{
    // src/apps/node2vec.hpp:126-127 (double max/min stored as real_t),
    // 136 (overflow), 182-191 (float 1/p, 1/q).
    double iq = 1.0 / (double)q, ip = 1.0 / (double)p;
    double up = std::max(1.0, iq);
    double lo = std::min(ip, std::min(1.0, iq));
    d.upper = (float)up;
    d.lower = (float)lo;
    d.inv_p = 1.0f / p;
    d.inv_q = 1.0f / q;
    d.overflow = (float)(1.0 / (double)p - (double)d.upper);
    d.acc_lo = std::min(1.0f, d.inv_q);
    d.acc_hi = std::max(1.0f, d.inv_q);
}

// line_rows: the job runs in a persistent batch, whose kernels write rows of
// whole 128-byte lines from LDS when the row stride is a multiple of 32
// entries; rows are padded to that when the padding stays within a quarter.
// may_build = false: build no table (no device allocation); a weighted job
// then needs the alias tables resident, while uniform draws without uniform
// buckets read col[] directly and node2vec without hash sets searches the
// sorted lists -- the same candidates and answers either way.
static PreparedJob prepare_job(kkw_engine *eng, const kkw_job &j, bool line_rows = false, bool may_build = true)
{
    GraphStore &gs = eng->gs;
    if (!gs.row_ptr) throw Error(KKW_E_NOGRAPH, "no graph loaded");
    PreparedJob pj;
    pj.src = j;
    DevJob &d = pj.dj;
    d = DevJob{};
    if (j.kind < KKW_DEEPWALK || j.kind > KKW_METAPATH) throw Error(KKW_E_INVALID, "unknown job kind");
    d.kind = j.kind;
    d.use_alias = j.use_weights ? 1 : 0;
    d.walker_begin = j.walker_begin;
    d.walker_num = j.walker_num;
    if ((uint64_t)j.walker_begin + j.walker_num > 0xFFFFFFFFull)
        throw Error(KKW_E_INVALID, "walker ids must fit in 32 bits");
    d.seed = j.seed;
    uint32_t L = j.walk_length;
    if (j.kind == KKW_PPR) {
        if (!(j.terminate_prob >= 0.0f && j.terminate_prob <= 1.0f))
            throw Error(KKW_E_INVALID, "terminate_prob must be in [0,1]");
        d.ppr_ext = 1.0f - j.terminate_prob;  // ppr.hpp:36
        if (L == 0) L = ppr_default_cap(j.terminate_prob);
        if (L == 0) throw Error(KKW_E_INVALID, "PPR with terminate_prob 0 needs a walk_length cap");
    }
    if (L > (1u << 24)) throw Error(KKW_E_INVALID, "walk_length above 2^24");
    d.walk_length = L;
    pj.stride = (L + 1 + 3) & ~3u;   // rows of whole 16-byte groups (path write-out)
    const uint64_t line_stride = ((uint64_t)L + 1 + 31) & ~(uint64_t)31;
    if (line_rows && L + 1 >= 32 && (line_stride - (L + 1)) * 4 <= (uint64_t)L + 1)
        pj.stride = (uint32_t)line_stride;
    d.stride = pj.stride;
    if (j.kind == KKW_NODE2VEC) {
        if (!(j.p > 0.0f && std::isfinite(j.p) && j.q > 0.0f && std::isfinite(j.q)))
            throw Error(KKW_E_INVALID, "node2vec p and q must be positive and finite");
        if (!gs.sorted)
            throw Error(KKW_E_INVALID, "node2vec needs adjacency lists sorted by neighbour "
                                       "(load with sort_adj=1, as node2vec.hpp:122-125 sorts)");
        node2vec_constants(j.p, j.q, d);
        // neighbour hash sets for the adjacency test, unless the caller turned
        // them off (kkw_graph_adj_index(eng, 0): binary search of sorted lists)
        if (may_build) auto_adj_index(eng);
    }
    if (!j.use_weights && !gs.ubucket && may_build) {   // uniform draws read uniform buckets
        eng->graph_version++;
        build_uniform_buckets(gs, eng->stream);
    }
    if (j.use_weights) {
        if (!gs.w) throw Error(KKW_E_INVALID, "use_weights on a graph without weights");
        if (!gs.bucket && !may_build)
            throw Error(KKW_E_NOGRAPH, "the alias tables are not resident (call kkw_prepare first)");
        if (!gs.bucket) {
            eng->graph_version++;
            double t0 = now_ms();
            build_alias(gs, eng->stream);
            eng->timing.alias_ms = now_ms() - t0;
        }
    }
    d.start_mode = j.start_mode;
    if (j.start_mode == KKW_START_LIST) {
        if (j.n_starts == 0 || !j.starts) throw Error(KKW_E_INVALID, "start list is empty");
        pj.starts.assign(j.starts, j.starts + j.n_starts);
        for (uint32_t v : pj.starts)
            if (v >= gs.V) throw Error(KKW_E_INVALID, "start vertex >= v_num");
        d.n_starts = j.n_starts;
    } else if (j.start_mode != KKW_START_EQUAL && j.start_mode != KKW_START_UNIFORM) {
        throw Error(KKW_E_INVALID, "unknown start mode");
    }
    if (j.kind == KKW_METAPATH) {
        if (!gs.vmask) throw Error(KKW_E_INVALID, "metapath needs edge types (kkw_graph_set_edge_types)");
        if (j.n_schemes == 0 || !j.scheme_len || !j.scheme_mask)
            throw Error(KKW_E_INVALID, "metapath needs at least one scheme");
        pj.scheme_len.assign(j.scheme_len, j.scheme_len + j.n_schemes);
        for (uint32_t l : pj.scheme_len)
            if (l < 1 || l > 8) throw Error(KKW_E_INVALID, "scheme length must be in [1,8]");
        pj.scheme_mask.assign(j.scheme_mask, j.scheme_mask + (size_t)j.n_schemes * 8);
        d.n_schemes = j.n_schemes;
    }
    return pj;
}

// Entry points that replace or free device graph storage refuse while the
// serving loop runs: its flights read the tables between admission and
// completion without holding the engine lock (tables the loop only adds, on
// first use by a query, are never freed while it runs).
static void refuse_while_serving(const kkw_engine *eng)
{
    if (eng->serve)
        throw Error(KKW_E_BUSY, "the graph and its tables cannot change while serving (call kkw_serve_stop first)");
}

// --------------------------------------------------------------------------
struct kkw_batch {
    kkw_engine *eng = nullptr;
    hipStream_t stream = nullptr;   // stream the batch runs on (engine's or a serving stream)
    int mode = KKW_MODE_PLAIN;
    bool stats = false;
    bool use_graph = false;
    bool line_rows = false;        // persistent: some job's rows are whole 128-byte lines
    std::vector<PreparedJob> jobs;
    std::vector<DevBuf> bufs;
    DevJob *d_jobs = nullptr;
    unsigned long long *d_queue = nullptr;
    DevStats *d_stats = nullptr;
    uint64_t n_chunks = 0;
    int block = 256, blocks_per_cu = 8, chunk = 64;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool launched = false;
    double host_t0 = 0.0;
    uint32_t graph_V = 0;          // graph the jobs were validated against
    uint64_t graph_E = 0;
    uint64_t captured_version = 0; // graph version baked into gexec

    ~kkw_batch()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    void *alloc(size_t bytes)
    {
        bufs.emplace_back(bytes);
        return bufs.back().p;
    }
};

static void batch_setup(kkw_batch *b, const kkw_job *jobs, int n_jobs)
{
    kkw_engine *eng = b->eng;
    if (n_jobs < 0 || (n_jobs > 0 && !jobs)) throw Error(KKW_E_INVALID, "bad job array");
    for (int i = 0; i < n_jobs; i++)
        b->jobs.push_back(prepare_job(eng, jobs[i], b->mode == KKW_MODE_PERSISTENT));
    for (auto &pj : b->jobs) b->line_rows = b->line_rows || pj.stride % 32 == 0;
    b->graph_V = eng->gs.V;
    b->graph_E = eng->gs.E;
    KKW_HIP(hipEventCreate(&b->e0));
    KKW_HIP(hipEventCreate(&b->e1));
    b->d_queue = (unsigned long long *)b->alloc(sizeof(unsigned long long) * 2);
    b->d_stats = (DevStats *)b->alloc(sizeof(DevStats));
    uint64_t chunk_acc = 0;
    for (auto &pj : b->jobs) {
        DevJob &d = pj.dj;
        size_t pbytes = (size_t)d.walker_num * pj.stride * sizeof(uint32_t);
        d.paths = (uint32_t *)b->alloc(pbytes);
        d.lens = (uint32_t *)b->alloc((size_t)d.walker_num * sizeof(uint32_t));
        if (!pj.starts.empty()) {
            uint32_t *p = (uint32_t *)b->alloc(pj.starts.size() * 4);
            KKW_HIP(hipMemcpyAsync(p, pj.starts.data(), pj.starts.size() * 4, hipMemcpyHostToDevice, b->stream));
            d.starts = p;
        }
        if (!pj.scheme_len.empty()) {
            uint32_t *p = (uint32_t *)b->alloc(pj.scheme_len.size() * 4);
            KKW_HIP(hipMemcpyAsync(p, pj.scheme_len.data(), pj.scheme_len.size() * 4, hipMemcpyHostToDevice,
                                   b->stream));
            d.scheme_len = p;
            uint8_t *m = (uint8_t *)b->alloc(pj.scheme_mask.size());
            KKW_HIP(hipMemcpyAsync(m, pj.scheme_mask.data(), pj.scheme_mask.size(), hipMemcpyHostToDevice,
                                   b->stream));
            d.scheme_mask = m;
        }
        d.chunk_begin = chunk_acc;
        chunk_acc += (d.walker_num + b->chunk - 1) / b->chunk;
    }
    b->n_chunks = chunk_acc;
    if (!b->jobs.empty()) {
        b->d_jobs = (DevJob *)b->alloc(sizeof(DevJob) * b->jobs.size());
        std::vector<DevJob> h;
        for (auto &pj : b->jobs) h.push_back(pj.dj);
        KKW_HIP(hipMemcpyAsync(b->d_jobs, h.data(), sizeof(DevJob) * h.size(), hipMemcpyHostToDevice, b->stream));
    }
    KKW_HIP(hipStreamSynchronize(b->stream));
}

// recompute chunk boundaries after a chunk-size change
static void batch_rechunk(kkw_batch *b)
{
    uint64_t acc = 0;
    std::vector<DevJob> h;
    for (auto &pj : b->jobs) {
        pj.dj.chunk_begin = acc;
        acc += (pj.dj.walker_num + b->chunk - 1) / b->chunk;
        h.push_back(pj.dj);
    }
    b->n_chunks = acc;
    if (!h.empty()) {
        KKW_HIP(hipMemcpyAsync(b->d_jobs, h.data(), sizeof(DevJob) * h.size(), hipMemcpyHostToDevice, b->stream));
        KKW_HIP(hipStreamSynchronize(b->stream));
    }
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
}

// The walk launches of one set of jobs: one plain launch per job, or one
// persistent launch over the work queue (d_queue, cleared here by a fill
// kernel) with the job passed by value (one job) or read from d_jobs.
struct WalkLaunch {
    int mode = KKW_MODE_PLAIN;
    int block = 256, blocks_per_cu = 8, chunk = 64;
    uint64_t n_chunks = 0;
    bool line_rows = false;
    bool stats = false;
    DevStats *d_stats = nullptr;
    unsigned long long *d_queue = nullptr;
    DevJob *d_jobs = nullptr;
};

static void enqueue_walk(const kkw_engine *eng, const std::vector<PreparedJob> &jobs, const WalkLaunch &w,
                         hipStream_t s)
{
    DevGraph g = eng->gs.view();
    if (w.stats) launch_fill_u64((unsigned long long *)w.d_stats, 0ull, 8, s);
    if (w.mode == KKW_MODE_PLAIN) {
        for (auto &pj : jobs) launch_walk_plain(g, pj.dj, w.stats, w.d_stats, w.block, s);
    } else {
        if (w.n_chunks == 0) return;
        launch_fill_u64(w.d_queue, 0ull, 1, s);
        uint64_t want = (uint64_t)eng->cu_count * w.blocks_per_cu;
        uint64_t waves_needed = w.n_chunks;  // one chunk per wave keeps every wave busy
        uint64_t blocks_needed = (waves_needed * 64 + w.block - 1) / w.block;
        int grid = (int)std::max<uint64_t>(1, std::min(want, blocks_needed));
        if (jobs.size() == 1)
            launch_walk_persistent_single(g, jobs[0].dj, w.n_chunks, (uint32_t)w.chunk, w.d_queue, w.stats,
                                          w.d_stats, grid, w.block, eng->cu_count, s);
        else
            launch_walk_persistent_mixed(g, w.d_jobs, (int)jobs.size(), w.n_chunks, (uint32_t)w.chunk, w.d_queue,
                                         w.stats, w.d_stats, grid, w.block, w.line_rows, s);
    }
    KKW_HIP(hipGetLastError());
}

static void batch_enqueue(kkw_batch *b)
{
    WalkLaunch w;
    w.mode = b->mode;
    w.block = b->block;
    w.blocks_per_cu = b->blocks_per_cu;
    w.chunk = b->chunk;
    w.n_chunks = b->n_chunks;
    w.line_rows = b->line_rows;
    w.stats = b->stats;
    w.d_stats = b->d_stats;
    w.d_queue = b->d_queue;
    w.d_jobs = b->d_jobs;
    enqueue_walk(b->eng, b->jobs, w, b->stream);
}

static void batch_launch(kkw_batch *b)
{
    TraceRange tr("walk_launch");
    kkw_engine *eng = b->eng;
    if (!eng->gs.row_ptr || eng->gs.V != b->graph_V || eng->gs.E != b->graph_E)
        throw Error(KKW_E_INVALID, "the graph changed since this batch was created; create a new batch");
    for (auto &pj : b->jobs)   // tables a job needs may have been dropped since
        if ((pj.dj.use_alias && !eng->gs.bucket) || (!pj.dj.use_alias && !eng->gs.ubucket) ||
            (pj.dj.kind == KKW_METAPATH && !eng->gs.vmask))
            throw Error(KKW_E_INVALID, "the graph's alias tables or edge types changed; create a new batch");
    if (b->gexec && b->captured_version != eng->graph_version) {   // stale pointers in the capture
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    b->host_t0 = now_ms();
    KKW_HIP(hipEventRecord(b->e0, b->stream));
    if (b->use_graph && b->mode == KKW_MODE_PERSISTENT) {
        if (!b->gexec) {
            b->captured_version = eng->graph_version;
            KKW_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            try {
                batch_enqueue(b);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            KKW_HIP(hipStreamEndCapture(b->stream, &b->graph));
            KKW_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        KKW_HIP(hipGraphLaunch(b->gexec, b->stream));
    } else {
        batch_enqueue(b);
    }
    KKW_HIP(hipEventRecord(b->e1, b->stream));
    b->launched = true;
}

static double batch_sync(kkw_batch *b)
{
    if (!b->launched) throw Error(KKW_E_INVALID, "batch was not launched");
    TraceRange tr("walk_wait");
    KKW_HIP(hipEventSynchronize(b->e1));
    float ms = 0.0f;
    KKW_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
    kkw_timing &t = b->eng->timing;
    t.kernel_ms = ms;
    t.total_ms = now_ms() - b->host_t0;
    if (b->stats) {
        DevStats h{};
        KKW_HIP(hipMemcpy(&h, b->d_stats, sizeof(h), hipMemcpyDeviceToHost));
        t.steps = h.steps;
        t.trials = h.trials;
        t.searches = h.searches;
        t.search_probes = h.probes;
        t.row_loads = h.row_loads;
        t.search_lines = h.search_lines;
        t.ref_queries = h.ref_queries;
        t.ref_query_lines = h.ref_lines;
    }
    return ms;
}

static void fill_paths_host(const PreparedJob &pj, hipStream_t s, kkw_paths *out)
{
    TraceRange tr("paths_to_host");
    const DevJob &d = pj.dj;
    memset(out, 0, sizeof(*out));
    out->walker_begin = d.walker_begin;
    out->walker_num = d.walker_num;
    out->stride = pj.stride;
    size_t n = (size_t)d.walker_num * pj.stride;
    out->vertices = (uint32_t *)malloc(n ? n * 4 : 4);
    out->lengths = (uint32_t *)malloc(d.walker_num ? (size_t)d.walker_num * 4 : 4);
    if (!out->vertices || !out->lengths) {
        free(out->vertices);
        free(out->lengths);
        throw Error(KKW_E_NOMEM, "host allocation for paths failed");
    }
    if (n) KKW_HIP(hipMemcpyAsync(out->vertices, d.paths, n * 4, hipMemcpyDeviceToHost, s));
    if (d.walker_num) KKW_HIP(hipMemcpyAsync(out->lengths, d.lens, (size_t)d.walker_num * 4, hipMemcpyDeviceToHost, s));
    KKW_HIP(hipStreamSynchronize(s));
    uint64_t steps = 0, trunc = 0;
    for (uint32_t i = 0; i < d.walker_num; i++) {
        uint32_t l = out->lengths[i];
        steps += (l & 0x7fffffffu) - 1;
        trunc += l >> 31;
    }
    out->steps = steps;
    out->truncated = trunc;
    out->on_device = 0;
    out->priv = (void *)1;  // host buffers owned by the caller (free with kkw_paths_free)
}

static void fill_paths_device(const PreparedJob &pj, hipStream_t s, kkw_paths *out)
{
    const DevJob &d = pj.dj;
    memset(out, 0, sizeof(*out));
    out->walker_begin = d.walker_begin;
    out->walker_num = d.walker_num;
    out->stride = pj.stride;
    out->vertices = d.paths;
    out->lengths = d.lens;
    out->on_device = 1;
    std::vector<uint32_t> lens(d.walker_num);
    if (d.walker_num) {
        KKW_HIP(hipMemcpyAsync(lens.data(), d.lens, (size_t)d.walker_num * 4, hipMemcpyDeviceToHost, s));
        KKW_HIP(hipStreamSynchronize(s));
    }
    uint64_t steps = 0, trunc = 0;
    for (uint32_t l : lens) {
        steps += (l & 0x7fffffffu) - 1;
        trunc += l >> 31;
    }
    out->steps = steps;
    out->truncated = trunc;
    out->priv = nullptr;  // owned by the batch
}

// ==========================================================================
// Serving
// ==========================================================================
// Continuous batching on a resident graph.  kkw_submit validates a query and
// queues it; one worker thread admits queued queries into "flights" while slots
// (walkers in flight) and memory (a device arena sized at start) allow, and
// launches each flight as one persistent-kernel launch on one of four streams,
// so flights admitted while others run overlap on the device.  A flight's
// paths are packed on the device (row-major rows under a PPR length cap are
// mostly empty) and copied back asynchronously into pinned staging memory in
// two phases (lengths + offsets, then the payload of exactly the packed size).
// No hipMalloc/hipFree happens per query.

// First-fit sub-allocator over [0, cap) in 256-byte units.
struct SubAlloc {
    size_t cap = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    static size_t up(size_t n) { return (n + 255) & ~(size_t)255; }
    void init(size_t c)
    {
        cap = c & ~(size_t)255;
        free_.clear();
        if (cap) free_[0] = cap;
    }
    bool alloc(size_t n, size_t &off)
    {
        n = up(n ? n : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                off = it->first;
                size_t rem = it->second - n;
                free_.erase(it);
                if (rem) free_[off + n] = rem;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = up(n ? n : 1);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t largest() const
    {
        size_t m = 0;
        for (auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }
};

// Pinned staging memory shared by the serving loop and the results it hands
// out: a served result's paths stay in the flight's pinned block (no copy),
// and the block returns to the arena when the last result using it is freed
// -- after kkw_serve_stop and kkw_engine_destroy too.
struct PinArena {
    std::mutex mu;
    SubAlloc a;
    char *base = nullptr;       // host address
    char *dbase = nullptr;      // the same memory as the device addresses it
    ~PinArena()
    {
        if (base) (void)hipHostFree(base);
    }
    bool alloc(size_t n, size_t &off)
    {
        std::lock_guard<std::mutex> lk(mu);
        return a.alloc(n, off);
    }
    void release(size_t off, size_t n)
    {
        std::lock_guard<std::mutex> lk(mu);
        a.release(off, n);
    }
    size_t largest()
    {
        std::lock_guard<std::mutex> lk(mu);
        return a.largest();
    }
    bool in_use_by_results()
    {
        std::lock_guard<std::mutex> lk(mu);
        return a.largest() < a.cap;
    }
};

struct PinLease {
    std::shared_ptr<PinArena> arena;
    size_t off = 0, bytes = 0;
    PinLease(std::shared_ptr<PinArena> a, size_t o, size_t b) : arena(std::move(a)), off(o), bytes(b) {}
    ~PinLease() { arena->release(off, bytes); }
};

// kkw_paths.priv of a served result
struct ServedPaths {
    static constexpr uint64_t kMagic = 0x6b6b7753657276ull;
    uint64_t magic = kMagic;
    std::shared_ptr<PinLease> lease;
};

struct Request {
    uint64_t ticket = 0;
    PreparedJob pj;
    double t_submit = 0.0, t_done = 0.0;
    int status = 1;  // 1 pending, 0 done, <0 error
    bool cancelled = false;
    bool taken = false;     // leaves the pending queue in this admission pass
    uint32_t graph_V = 0;   // vertex count of the graph the query was checked against
    kkw_paths result{};
};

// Layout of one flight inside the device arena (byte offsets from its base)
// and inside the pinned staging arena.
// Rounds of one flight queued on its stream at once: the next round is
// already there when one ends (no host latency between rounds), and another
// stream's work waits behind at most this many rounds of a long flight.
constexpr uint32_t kRoundsAhead = 3;

struct Flight {
    std::vector<std::shared_ptr<Request>> reqs;
    double t_admit = 0, t_enqueued = 0, t_phase1 = 0, t_phase2 = 0, t_retired = 0;   // flight trace
    int sidx = 0;
    hipEvent_t ev = nullptr;       // recorded after walk, pack (into pinned memory) and D2H(lens, offs)
    size_t dev_off = 0, dev_bytes = 0;
    size_t pin_off = 0, pin_bytes = 0;
    uint64_t walkers = 0;
    uint64_t total = 0;            // packed vertices
    std::vector<uint64_t> row0;    // first flight row of each request
    size_t o_lens = 0, o_offs = 0; // device offsets of lens_all and offs (adjacent)
    size_t p_down = 0;             // pinned offsets: lens + offs, packed payload
    size_t p_payload = 0;
    // device offsets the later phases use
    size_t o_jobs = 0, o_pjobs = 0, o_queue = 0, o_cnt = 0, o_scan = 0, scan_bytes = 0;
    uint64_t chunks = 0;           // work items of the first round
    // serving rounds (kkw_serve_opts.round_steps > 0): walkers run at most
    // round_steps steps per launch; the walk is relaunched on the suspended
    // ones until none is left, then the rows are packed
    bool rounds = false;
    bool packing = false;          // pack and copy-back enqueued: ev says when the flight is done
    uint32_t round_steps = 0;      // steps per walker per launch (the option, or more for very long walks)
    uint32_t n_rounds = 0;         // at most this many launches: ceil(longest path / round_steps)
    uint32_t next_round = 0;       // rounds enqueued
    uint32_t done_rounds = 0;      // rounds seen complete (their counts read)
    hipEvent_t round_ev[kRoundsAhead] = {};   // round r records round_ev[r % kRoundsAhead]
    DevGraph g{};                  // the graph view the flight was launched on (fixed while serving)
    roctx_range_id_t trace_id = 0; // marker-trace range from launch to retirement (0: none)
    int round_grid = 0;            // blocks of every round's launch
    uint64_t rec_cap = 0;          // slots of a suspended-walker list: walkers + waves x ROUND_SLOTS
    size_t o_list[2] = {0, 0};     // suspended-walker records, one list per round parity
    size_t o_rq = 0;               // per-round work-queue words (zeroed by the upload)
    size_t o_rcnt = 0;             // per round [n_rounds + 1]: slot counts, then (ended waves, listed walkers) words
    size_t p_rcnt = 0;             // pinned: round r's listed count lands at p_rcnt + 8 (r + 1)
};

struct ServeState {
    kkw_engine *eng = nullptr;
    kkw_serve_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    hipStream_t streams[4] = {nullptr, nullptr, nullptr, nullptr};
    char *dev_base = nullptr;
    SubAlloc dev;
    std::shared_ptr<PinArena> pin;
    // KKW_SERVE_TRACE=<file>: one CSV line per flight, written at kkw_serve_stop
    // (a measurement aid: where a query's latency goes)
    std::string trace_path;
    std::vector<std::string> trace;
};

static size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }

// Blocks of a serving launch over `chunks` chunks of 64 walkers: one lane per
// walker up to eight 256-thread blocks per CU.
static int round_grid(const kkw_engine *eng, uint64_t chunks)
{
    const uint64_t blocks_needed = (chunks * 64 + 255) / 256;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)eng->cu_count * 8, blocks_needed));
}

// Device bytes of a flight walked in rounds beyond its requests' shares: the
// slots its launches' waves may leave unused in each of the two lists.
static size_t rounds_flight_bytes(const kkw_engine *eng)
{
    return 2 * (size_t)eng->cu_count * 8 * 4 * ROUND_SLOTS * ROUND_REC_PLANES * 16 + 4096;
}

// Length class of a query for flight grouping: 0 when a walker takes 16
// steps or fewer on average (its cap, or a PPR walk's mean length 1 / t),
// else 1.
static int length_class(const PreparedJob &pj)
{
    double e = pj.dj.walk_length;
    if (pj.dj.kind == KKW_PPR && pj.dj.ppr_ext < 1.0f) e = std::min(e, 1.0 / (1.0 - (double)pj.dj.ppr_ext));
    return e <= 16.0 ? 0 : 1;
}

// Device bytes one request adds to a flight: rows, lengths, counts, offsets,
// starts and scheme tables, and with rounds its share of the two
// suspended-walker record lists and of the per-round counters.
static size_t request_dev_bytes(const PreparedJob &pj, bool rounds)
{
    const size_t n = pj.dj.walker_num;
    return al256(n * pj.stride * 4) + al256(n * 4) + al256(n * 8) * 2 + al256(pj.starts.size() * 4) +
           al256(pj.scheme_len.size() * 4) + al256(pj.scheme_mask.size()) + 1024 +
           (rounds ? 2 * al256(n * ROUND_REC_PLANES * 16) + 2 * 1024 * 3 * 8 : 0);
}

// Pinned bytes one request adds to a flight: its share of the upload mirror,
// lengths and offsets, and the packed payload at its worst (rows x stride).
static size_t request_pin_bytes(const PreparedJob &pj)
{
    const size_t n = pj.dj.walker_num;
    return al256(n * pj.stride * 4) + al256(n * 12 + 8) + al256(pj.starts.size() * 4) +
           al256(pj.scheme_len.size() * 4) + al256(pj.scheme_mask.size()) + 1024;
}

static void finish_request(ServeState *S, const std::shared_ptr<Request> &r, int status)
{
    std::lock_guard<std::mutex> lk(S->mu);
    r->status = status;
    r->t_done = now_ms();
    S->cv_done.notify_all();
}

static void enqueue_pack(ServeState *S, Flight &f);
static bool advance_rounds(ServeState *S, Flight &f);

// Build and launch one flight.  Returns false (and leaves reqs pending) when
// the device or pinned arena has no room for them yet; *pin_short tells which.  One stream-ordered
// sequence: upload the job tables, walk, pack the rows, their lengths and
// offsets straight into the flight's pinned block, record the event; the host
// touches the flight again only when it has completed (or, walked in rounds,
// to queue its next rounds).
static bool launch_flight(ServeState *S, Flight &f, bool *pin_short)
{
    *pin_short = false;
    kkw_engine *eng = S->eng;
    hipStream_t st = S->streams[f.sidx];
    const int nj = (int)f.reqs.size();
    uint64_t N = 0;
    for (auto &r : f.reqs) {
        f.row0.push_back(N);
        N += r->pj.dj.walker_num;
    }
    f.walkers = N;
    const size_t scan_bytes = pack_scan_bytes(N ? N : 1);
    // ---- device: [upload: jobs | packjobs | per-job small arrays][queue][lens | offs][cnt][scan][rows...]
    size_t o = 0;
    const size_t o_jobs = o;
    o += al256(sizeof(DevJob) * nj);
    const size_t o_pjobs = o;
    o += al256(sizeof(PackJob) * nj);
    std::vector<size_t> o_starts(nj), o_slen(nj), o_smask(nj);
    for (int j = 0; j < nj; j++) {
        const PreparedJob &pj = f.reqs[j]->pj;
        o_starts[j] = o;
        o += al256(pj.starts.size() * 4);
        o_slen[j] = o;
        o += al256(pj.scheme_len.size() * 4);
        o_smask[j] = o;
        o += al256(pj.scheme_mask.size());
    }
    const size_t o_queue = o;   // the walk's work counter, uploaded as zero
    o += 256;
    f.rounds = S->opts.round_steps != 0;
    if (f.rounds) {
        // every walker ends within ceil(longest path / round_steps) rounds; a
        // flight takes at most kMaxRounds launches (its rounds lengthen for
        // longer walks)
        constexpr uint64_t kMaxRounds = 1024;
        uint64_t longest = 0;
        for (auto &r : f.reqs) longest = std::max<uint64_t>(longest, r->pj.dj.walk_length);
        f.round_steps = (uint32_t)std::max<uint64_t>(S->opts.round_steps, (longest + kMaxRounds - 1) / kMaxRounds);
        f.n_rounds = (uint32_t)std::max<uint64_t>(1, (longest + f.round_steps - 1) / f.round_steps);
        uint64_t chunks = 0;
        for (auto &r : f.reqs) chunks += (r->pj.dj.walker_num + 63) / 64;
        f.round_grid = round_grid(eng, chunks);
        f.rec_cap = N + (uint64_t)f.round_grid * 4 * ROUND_SLOTS;
        f.o_rq = o;             // uploaded as zero, like the counts after it
        o += al256((size_t)f.n_rounds * 8);
        f.o_rcnt = o;
        o += al256(2 * ((size_t)f.n_rounds + 1) * 8);
    }
    const size_t upload_bytes = o;
    f.o_lens = o;
    const size_t lens_bytes = ((N * 4 + 7) & ~(size_t)7);
    f.o_offs = o + lens_bytes;
    const size_t down_bytes = lens_bytes + (N + 1) * 8;
    o += al256(down_bytes);
    const size_t o_cnt = o;
    o += al256(N * 8);
    const size_t o_scan = o;
    o += al256(scan_bytes);
    std::vector<size_t> o_rows(nj);
    size_t packed_max = 0;
    for (int j = 0; j < nj; j++) {
        const PreparedJob &pj = f.reqs[j]->pj;
        o_rows[j] = o;
        o += al256((size_t)pj.dj.walker_num * pj.stride * 4);
        packed_max += (size_t)pj.dj.walker_num * pj.stride;
    }
    if (f.rounds) {
        f.o_list[0] = o;
        o += al256(f.rec_cap * ROUND_REC_PLANES * 16);
        f.o_list[1] = o;
        o += al256(f.rec_cap * ROUND_REC_PLANES * 16);
    }
    f.dev_bytes = o;
    // ---- pinned: [upload mirror][lens | offs][packed payload, at most rows x stride]
    f.p_down = upload_bytes;
    f.p_payload = upload_bytes + al256(down_bytes);
    f.pin_bytes = f.p_payload + al256(packed_max * 4);
    if (!S->dev.alloc(f.dev_bytes, f.dev_off)) return false;
    if (!S->pin->alloc(f.pin_bytes, f.pin_off)) {
        S->dev.release(f.dev_off, f.dev_bytes);
        *pin_short = true;
        return false;
    }
    char *D = S->dev_base + f.dev_off;
    char *H = S->pin->base + f.pin_off;
    // ---- fill the upload mirror in pinned memory
    std::vector<DevJob> djs(nj);
    std::vector<PackJob> pjs(nj);
    uint64_t chunk_acc = 0;
    const uint32_t chunk = 64;
    for (int j = 0; j < nj; j++) {
        const PreparedJob &pj = f.reqs[j]->pj;
        DevJob d = pj.dj;
        d.paths = (uint32_t *)(D + o_rows[j]);
        d.lens = (uint32_t *)(D + f.o_lens) + f.row0[j];
        if (!pj.starts.empty()) {
            memcpy(H + o_starts[j], pj.starts.data(), pj.starts.size() * 4);
            d.starts = (const uint32_t *)(D + o_starts[j]);
        }
        if (!pj.scheme_len.empty()) {
            memcpy(H + o_slen[j], pj.scheme_len.data(), pj.scheme_len.size() * 4);
            memcpy(H + o_smask[j], pj.scheme_mask.data(), pj.scheme_mask.size());
            d.scheme_len = (const uint32_t *)(D + o_slen[j]);
            d.scheme_mask = (const uint8_t *)(D + o_smask[j]);
        }
        d.chunk_begin = chunk_acc;
        chunk_acc += (d.walker_num + chunk - 1) / chunk;
        djs[j] = d;
        pjs[j] = PackJob{f.row0[j], d.walker_num, pj.stride, d.paths};
    }
    memcpy(H + o_jobs, djs.data(), sizeof(DevJob) * nj);
    memcpy(H + o_pjobs, pjs.data(), sizeof(PackJob) * nj);
    memset(H + o_queue, 0, sizeof(unsigned long long));
    if (f.rounds) {
        memset(H + f.o_rq, 0, (size_t)f.n_rounds * 8);
        memset(H + f.o_rcnt, 0, 2 * ((size_t)f.n_rounds + 1) * 8);
        f.p_rcnt = f.o_rcnt;    // slots of the upload mirror, free once the upload has been read
    }
    f.o_jobs = o_jobs;
    f.o_pjobs = o_pjobs;
    f.o_queue = o_queue;
    f.o_cnt = o_cnt;
    f.o_scan = o_scan;
    f.scan_bytes = scan_bytes;
    f.chunks = chunk_acc;
    // ---- enqueue: upload, walk, pack into pinned memory, lengths + offsets
    DevGraph g;
    {
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        g = eng->gs.view();
    }
    for (auto &r : f.reqs)   // what each query was admitted with is still there
        if ((r->pj.dj.use_alias && !g.bucket) || (r->pj.dj.kind == KKW_METAPATH && !g.vmask) || !g.row_ptr ||
            g.V != r->graph_V)
            throw Error(KKW_E_NOGRAPH, "the graph or the tables a query needs are gone");
    KKW_HIP(hipMemcpyAsync(D, H, upload_bytes, hipMemcpyHostToDevice, st));
    KKW_HIP(hipEventCreateWithFlags(&f.ev, hipEventDisableTiming));
    if (f.rounds && chunk_acc) {
        for (auto &e : f.round_ev) KKW_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        f.g = g;
        advance_rounds(S, f);   // the first rounds
        return true;
    }
    if (chunk_acc) {
        const int grid = round_grid(eng, chunk_acc);
        // serving rows stay 16-byte groups (short PPR walks; full occupancy)
        launch_walk_persistent_mixed(g, (const DevJob *)(D + o_jobs), nj, chunk_acc, chunk,
                                     (unsigned long long *)(D + o_queue), false, nullptr, grid, 256, false, st);
    }
    enqueue_pack(S, f);
    return true;
}

// After the walk (its last round): pack the rows, their lengths and offsets
// straight into the flight's pinned block, record the event.
static void enqueue_pack(ServeState *S, Flight &f)
{
    hipStream_t st = S->streams[f.sidx];
    char *D = S->dev_base + f.dev_off;
    char *H = S->pin->base + f.pin_off;
    char *HD = S->pin->dbase + f.pin_off;   // the same pinned block as the kernels address it
    const uint64_t N = f.walkers;
    const int nj = (int)f.reqs.size();
    const size_t lens_bytes = ((N * 4 + 7) & ~(size_t)7);
    if (f.chunks) {
        // the kernels write lengths and offsets straight into the pinned block
        launch_pack((const PackJob *)(D + f.o_pjobs), nj, (const uint32_t *)(D + f.o_lens), (uint64_t *)(D + f.o_cnt),
                    (uint64_t *)(D + f.o_offs), N, (uint32_t *)(HD + f.p_payload), (uint32_t *)(HD + f.p_down),
                    (uint64_t *)(HD + f.p_down + lens_bytes), D + f.o_scan, f.scan_bytes, st);
        KKW_HIP(hipGetLastError());
    } else {
        launch_fill_u64((unsigned long long *)(D + f.o_offs), 0ull, 1, st);
        KKW_HIP(hipGetLastError());
        KKW_HIP(hipMemcpyAsync(H + f.p_down, D + f.o_lens, lens_bytes + (N + 1) * 8, hipMemcpyDeviceToHost, st));
    }
    KKW_HIP(hipEventRecord(f.ev, st));
    f.packing = true;
}

// Round r of a flight walked in rounds: continue the walkers round r - 1
// listed (all of them in round 0), reading their count on the device, list
// the ones it suspends; the launch's last wave writes that count to the
// round's pinned slot, and the round's event says when it is there.
static void enqueue_round(ServeState *S, Flight &f, uint32_t r)
{
    hipStream_t st = S->streams[f.sidx];
    char *D = S->dev_base + f.dev_off;
    char *HD = S->pin->dbase + f.pin_off;   // the pinned block as the kernels address it
    const uint32_t chunk = 64;
    unsigned long long *cnt = (unsigned long long *)(D + f.o_rcnt);
    unsigned long long *live = cnt + f.n_rounds + 1;
    RoundArgs R;
    R.round_steps = f.round_steps;
    R.cap = f.rec_cap;
    R.in = r == 0 ? nullptr : (const uint4 *)(D + f.o_list[r & 1]);
    R.n_in = cnt + r;
    R.out = (uint4 *)(D + f.o_list[(r + 1) & 1]);
    R.n_out = cnt + r + 1;
    R.n_live = live + r + 1;
    R.h_live = (unsigned long long *)(HD + f.p_rcnt) + r + 1;
    launch_walk_rounds(f.g, (const DevJob *)(D + f.o_jobs), (int)f.reqs.size(), f.chunks, chunk,
                       (unsigned long long *)(D + f.o_rq) + r, R, f.round_grid, 256, st);
    KKW_HIP(hipGetLastError());
    KKW_HIP(hipEventRecord(f.round_ev[r % kRoundsAhead], st));
}

// A flight walked in rounds: note the rounds that have completed, pack once
// one leaves no walker (rounds still queued behind it find nothing to do),
// and keep up to kRoundsAhead rounds queued on the flight's stream.  A walker
// at its step cap is never suspended, so the n_rounds-th round ends every
// walker; a count left after it is a device error.  Returns true when
// anything moved.
static bool advance_rounds(ServeState *S, Flight &f)
{
    const char *H = S->pin->base + f.pin_off;
    bool moved = false;
    while (f.done_rounds < f.next_round &&
           hipEventQuery(f.round_ev[f.done_rounds % kRoundsAhead]) == hipSuccess) {
        const uint64_t left = *(volatile const uint64_t *)(H + f.p_rcnt + 8 * ((size_t)f.done_rounds + 1));
        f.done_rounds++;
        moved = true;
        if (left == 0) {
            enqueue_pack(S, f);
            return true;
        }
        if (f.done_rounds == f.n_rounds) throw Error(KKW_E_DEVICE, "serving rounds left walkers unfinished");
    }
    while (f.next_round < f.n_rounds && f.next_round - f.done_rounds < kRoundsAhead) {
        enqueue_round(S, f, f.next_round++);
        moved = true;
    }
    return moved;
}

static void destroy_events(Flight &f)
{
    if (f.ev) (void)hipEventDestroy(f.ev);
    f.ev = nullptr;
    for (auto &e : f.round_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

// The flight has completed: hand each request its paths, in place in the
// flight's pinned block (shared by the requests through one lease).
static void flight_retire(ServeState *S, Flight &f)
{
    const char *H = S->pin->base + f.pin_off;
    const uint64_t N = f.walkers;
    const size_t lens_bytes = ((N * 4 + 7) & ~(size_t)7);
    uint32_t *lens = (uint32_t *)(H + f.p_down);
    uint64_t *offs = (uint64_t *)(H + f.p_down + lens_bytes);
    uint32_t *payload = (uint32_t *)(H + f.p_payload);
    f.total = offs[N];
    // keep only lengths, offsets and the payload actually written; the upload
    // mirror and the unused worst-case tail go back to the arena now
    const size_t used = f.p_payload + al256(f.total * 4);
    if (f.pin_bytes > used) S->pin->release(f.pin_off + used, f.pin_bytes - used);
    if (f.p_down) S->pin->release(f.pin_off, f.p_down);
    auto lease = std::make_shared<PinLease>(S->pin, f.pin_off + f.p_down, used - f.p_down);
    f.pin_bytes = 0;   // owned by the lease from here on
    for (size_t j = 0; j < f.reqs.size(); j++) {
        auto &r = f.reqs[j];
        const DevJob &d = r->pj.dj;
        const uint64_t r0 = f.row0[j], n = d.walker_num;
        kkw_paths &p = r->result;
        memset(&p, 0, sizeof(p));
        p.walker_begin = d.walker_begin;
        p.walker_num = (uint32_t)n;
        p.stride = r->pj.stride;
        p.vertices = payload;            // rows at payload + offsets[k]
        p.lengths = lens + r0;
        p.offsets = offs + r0;
        uint64_t trunc = 0;
        for (uint64_t k = 0; k < n; k++) trunc += lens[r0 + k] >> 31;
        p.steps = offs[r0 + n] - offs[r0] - n;
        p.truncated = trunc;
        auto *sp = new ServedPaths();
        sp->lease = lease;
        p.priv = sp;
        finish_request(S, r, 0);
    }
    S->dev.release(f.dev_off, f.dev_bytes);
    f.dev_bytes = 0;
    destroy_events(f);
}

// Copy finished results that still point into pinned memory to heap memory
// and drop their leases (callers collect results in their own order, so
// pinned memory held by uncollected results must not starve admission).
// Returns how many results were moved.
static int spill_uncollected(ServeState *S)
{
    int moved = 0;
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto &kv : S->all) {
        Request &r = *kv.second;
        kkw_paths &p = r.result;
        if (r.status != 0 || !p.priv || p.priv == (void *)1) continue;
        const uint64_t n = p.walker_num, base = p.offsets[0], cnt = p.offsets[n] - base;
        uint32_t *lens = (uint32_t *)malloc(n ? n * 4 : 4);
        uint64_t *offs = (uint64_t *)malloc((n + 1) * 8);
        uint32_t *verts = (uint32_t *)malloc(cnt ? cnt * 4 : 4);
        if (!lens || !offs || !verts) {
            free(lens);
            free(offs);
            free(verts);
            break;   // leave the rest where they are
        }
        memcpy(lens, p.lengths, n * 4);
        for (uint64_t k = 0; k <= n; k++) offs[k] = p.offsets[k] - base;
        if (cnt) memcpy(verts, p.vertices + base, cnt * 4);
        delete static_cast<ServedPaths *>(p.priv);
        p.lengths = lens;
        p.offsets = offs;
        p.vertices = verts;
        p.priv = (void *)1;
        moved++;
    }
    return moved;
}

// Give back what a failed flight still holds (arena blocks, event).
static void release_flight(ServeState *S, Flight &f)
{
    if (f.dev_bytes && S->dev_base) {
        S->dev.release(f.dev_off, f.dev_bytes);
        f.dev_bytes = 0;
    }
    if (f.pin_bytes) {
        S->pin->release(f.pin_off, f.pin_bytes);
        f.pin_bytes = 0;
    }
    destroy_events(f);
}

static void serve_loop(ServeState *S)
{
    // The loop polls flight events with 10-microsecond sleeps; under Linux's
    // default 50-microsecond timer slack each sleep lasted ~60 us.  With 1 us
    // of slack: PPR query p50 0.22 -> 0.18 ms, p99 at 2000 qps 0.50 -> 0.44 ms.
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
    roctxNameOsThread("kkw serving loop");
    kkw_engine *eng = S->eng;
    (void)hipSetDevice(eng->device);
    std::list<Flight> flights;
    bool stream_busy[4] = {false, false, false, false};
    uint64_t walkers_in_flight = 0;
    double last_active = now_ms();   // last time a flight ran or a query was admitted
    for (;;) {
        if (!flights.empty()) last_active = now_ms();
        bool progressed = false;
        // 1. advance flights: queue their next rounds, retire the ones whose
        //    copies have landed
        for (auto it = flights.begin(); it != flights.end();) {
            Flight &f = *it;
            try {
                if (f.rounds && !f.packing) {   // rounds still to see end (or to enqueue)
                    if (advance_rounds(S, f)) progressed = true;
                    if (!f.packing) {
                        ++it;
                        continue;
                    }
                }
                if (hipEventQuery(f.ev) != hipSuccess) {
                    ++it;
                    continue;
                }
                progressed = true;
                f.t_phase1 = f.t_phase2 = now_ms();
                flight_retire(S, f);
                f.t_retired = now_ms();
                if (!S->trace_path.empty()) {
                    double first_submit = f.t_admit;
                    for (auto &r : f.reqs) first_submit = std::min(first_submit, r->t_submit);
                    char line[256];
                    snprintf(line, sizeof line, "%.4f,%.4f,%.4f,%.4f,%.4f,%.4f,%zu,%llu,%llu", first_submit,
                             f.t_admit, f.t_enqueued, f.t_phase1, f.t_phase2, f.t_retired, f.reqs.size(),
                             (unsigned long long)f.walkers, (unsigned long long)f.total);
                    S->trace.push_back(line);
                }
            } catch (const std::exception &ex) {
                progressed = true;
                const int code = dynamic_cast<const Error *>(&ex) ? static_cast<const Error &>(ex).code : KKW_E_NOMEM;
                (void)hipStreamSynchronize(S->streams[f.sidx]);   // nothing of it may still run
                for (auto &r : f.reqs)
                    if (r->status == 1) finish_request(S, r, code);
                release_flight(S, f);
            }
            if (f.trace_id) roctxRangeStop(f.trace_id);
            stream_busy[f.sidx] = false;
            walkers_in_flight -= f.walkers;
            it = flights.erase(it);
        }
        // 2. admit queued requests into a new flight on a free stream
        int sidx = -1;
        for (int k = 0; k < 4; k++)
            if (!stream_busy[k]) {
                sidx = k;
                break;
            }
        std::vector<std::shared_ptr<Request>> admit;
        bool pin_blocked = false;   // the head request waits for pinned memory
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop && S->pending.empty() && flights.empty()) break;
            if (sidx >= 0) {
                size_t room = S->dev.largest();
                const size_t flight_extra = S->opts.round_steps ? rounds_flight_bytes(eng) : 0;
                size_t used = 4096 + flight_extra;
                const size_t pin_room = S->pin->largest();
                size_t pin_used = 4096;
                uint64_t walkers = walkers_in_flight;
                // one flight takes at most a quarter of the walker budget (a
                // larger request goes alone): behind a backlog the four streams
                // then hold four flights, and one's packing overlaps the others'
                // walks, where one flight holding the whole budget would leave
                // the device idle while it packs
                const uint64_t flight_cap = std::max<uint64_t>(1, S->opts.max_walkers_in_flight / 4);
                uint64_t flight_walkers = 0;
                // A flight holds queries of one length class (the head's): short
                // walks do not wait for long ones that happen to share their
                // flight.  Queries of the other class keep their place in the
                // queue for a flight of their own.  Taken queries are marked and
                // removed in one pass; the scan looks at most kScan queries.
                constexpr size_t kScan = 4096;
                int cls = -1;
                size_t looked = 0;
                bool taken_any = false;
                for (auto &r : S->pending) {
                    if (++looked > kScan) break;
                    if (r->cancelled) {
                        r->taken = true;
                        taken_any = true;
                        r->status = KKW_E_CANCELLED;
                        r->t_done = now_ms();
                        S->cv_done.notify_all();
                        continue;
                    }
                    size_t need = request_dev_bytes(r->pj, S->opts.round_steps != 0) + 2 * pack_scan_bytes(r->pj.dj.walker_num + 1);
                    if (need + 4096 + flight_extra > S->dev.cap || request_pin_bytes(r->pj) + 4096 > S->pin->a.cap) {
                        r->taken = true;   // can never fit the arenas
                        taken_any = true;
                        r->status = KKW_E_NOMEM;
                        r->t_done = now_ms();
                        S->cv_done.notify_all();
                        continue;
                    }
                    const int c = length_class(r->pj);
                    if (cls >= 0 && c != cls) continue;
                    const size_t pneed = request_pin_bytes(r->pj);
                    bool fits_mem = used + need <= room;
                    bool fits_pin = pin_used + pneed <= pin_room;   // worst-case payloads of the whole flight
                    bool fits_slots = walkers + r->pj.dj.walker_num <= S->opts.max_walkers_in_flight ||
                                      (walkers == 0 && admit.empty());
                    const bool fits_flight = flight_walkers + r->pj.dj.walker_num <= flight_cap || admit.empty();
                    if (!fits_mem || !fits_pin || !fits_slots || !fits_flight || admit.size() >= 1024) {
                        pin_blocked = admit.empty() && !fits_pin;
                        break;
                    }
                    r->taken = true;
                    taken_any = true;
                    cls = c;
                    admit.push_back(r);
                    used += need;
                    pin_used += pneed;
                    walkers += r->pj.dj.walker_num;
                    flight_walkers += r->pj.dj.walker_num;
                }
                if (taken_any)
                    S->pending.erase(std::remove_if(S->pending.begin(), S->pending.end(),
                                                    [](const std::shared_ptr<Request> &q) { return q->taken; }),
                                     S->pending.end());
            }
            if (admit.empty() && flights.empty() && !pin_blocked) {
                if (S->stop && S->pending.empty()) break;
                // idle: with idle_poll_us, keep looking for new queries that
                // long after the last activity (an arrival is then admitted
                // microseconds later instead of after a condition-variable
                // wake-up), then sleep until kkw_submit signals
                if (S->opts.idle_poll_us && now_ms() - last_active < 1e-3 * S->opts.idle_poll_us) {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(2));
                    continue;
                }
                S->cv_work.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
        }
        if (admit.empty() && pin_blocked) {
            // pinned memory is held by finished results: move uncollected ones
            // to heap memory; otherwise wait for flights or callers to free some
            if (spill_uncollected(S) == 0) {
                if (flights.empty() && S->pin->in_use_by_results()) {
                    std::lock_guard<std::mutex> lk(S->mu);
                    if (S->stop) {   // stopping: callers still hold results, nothing more can start
                        for (auto &r : S->pending) {
                            r->status = KKW_E_CANCELLED;
                            r->t_done = now_ms();
                        }
                        S->pending.clear();
                        S->cv_done.notify_all();
                    }
                }
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            continue;
        }
        if (!admit.empty()) {
            progressed = true;
            flights.emplace_back();
            Flight &f = flights.back();
            f.sidx = sidx;
            f.reqs = admit;
            f.t_admit = now_ms();
            bool ok = false, pin_short = false;
            try {
                ok = launch_flight(S, f, &pin_short);
                f.t_enqueued = now_ms();
            } catch (const Error &e) {
                (void)hipStreamSynchronize(S->streams[sidx]);   // nothing of it may still run
                for (auto &r : f.reqs) finish_request(S, r, e.code);
                release_flight(S, f);
                flights.pop_back();
                continue;
            }
            if (!ok) {
                // no room after all: put them back at the head of the queue
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    for (auto it = admit.rbegin(); it != admit.rend(); ++it) {
                        (*it)->taken = false;
                        S->pending.push_front(*it);
                    }
                }
                flights.pop_back();
                if (!(pin_short && spill_uncollected(S) > 0)) progressed = false;   // wait for room
            } else {
                f.trace_id = roctxRangeStartA("serve_flight");
                stream_busy[sidx] = true;
                walkers_in_flight += f.walkers;
            }
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
}

// ==========================================================================
// No-allocation runs (kkw_run_noalloc)
// ==========================================================================
// Fixed device scratch of kkw_run_noalloc: the work-queue counter, the job
// table, start lists and scheme tables.  Each GPU holds its own copy of the
// array until the process ends.  Host code must never use the array's name as
// an address: the device address of a GPU's copy comes from
// hipGetSymbolAddress, taken once per device at engine creation.
static __device__ __align__(256) unsigned char scratch[KKW_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_init_mu;
std::map<int, void *> g_scratch_addr;                  // device ordinal -> address of its copy
std::map<int, std::unique_ptr<std::mutex>> g_scratch_mu;   // serialises calls per device
}  // namespace

// The current device's copy of `scratch` (the caller has made eng->device current).
static void *scratch_address(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_init_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        throw Error(KKW_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_addr[device] = p;
    g_scratch_mu[device] = std::make_unique<std::mutex>();
    return p;
}

static std::mutex &scratch_mutex(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_init_mu);
    return *g_scratch_mu.at(device);
}

static int run_noalloc(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode, uint32_t *const *paths,
                       const uint32_t *strides, uint32_t *const *lengths)
{
    if (!eng->noalloc_scratch) throw Error(KKW_E_DEVICE, "the engine has no scratch address");
    TraceRange tr("run_noalloc");
    std::lock_guard<std::mutex> slk(scratch_mutex(eng->device));
    std::vector<PreparedJob> pjs;
    for (int i = 0; i < n_jobs; i++) pjs.push_back(prepare_job(eng, jobs[i], false, /*may_build=*/false));
    // the caller's buffers replace the batch's
    for (int i = 0; i < n_jobs; i++) {
        DevJob &d = pjs[i].dj;
        const uint64_t need = (uint64_t)d.walk_length + 1;
        if (strides[i] % 4 != 0 || strides[i] < need)
            throw Error(KKW_E_INVALID, "job " + std::to_string(i) + ": stride must be a multiple of 4 and at least " +
                                           std::to_string(need));
        if (d.walker_num && (!paths[i] || !lengths[i]))
            throw Error(KKW_E_INVALID, "job " + std::to_string(i) + ": null path or length buffer");
        if ((uintptr_t)paths[i] % 16 != 0 || (uintptr_t)lengths[i] % 4 != 0)
            throw Error(KKW_E_INVALID, "job " + std::to_string(i) + ": paths must be 16-byte aligned, lengths 4-byte");
        pjs[i].stride = strides[i];
        d.stride = strides[i];
        d.paths = paths[i];
        d.lens = lengths[i];
    }
    WalkLaunch w;
    w.mode = mode;
    // layout in the scratch array: [0, 256) queue counter, then 256-byte
    // aligned job table, start lists and scheme tables
    size_t off = 256;
    auto take = [&](size_t n) {
        size_t o = off;
        off = al256(off + n);
        return o;
    };
    const size_t jobs_off = (mode == KKW_MODE_PERSISTENT && n_jobs > 1) ? take(sizeof(DevJob) * n_jobs) : 0;
    std::vector<size_t> starts_off(n_jobs, 0), len_off(n_jobs, 0), mask_off(n_jobs, 0);
    for (int i = 0; i < n_jobs; i++) {
        if (!pjs[i].starts.empty()) starts_off[i] = take(pjs[i].starts.size() * 4);
        if (!pjs[i].scheme_len.empty()) {
            len_off[i] = take(pjs[i].scheme_len.size() * 4);
            mask_off[i] = take(pjs[i].scheme_mask.size());
        }
    }
    if (off > KKW_NOALLOC_SCRATCH_BYTES)
        throw Error(KKW_E_NOMEM, "these jobs need " + std::to_string(off) + " bytes of scratch; the fixed array holds " +
                                     std::to_string((size_t)KKW_NOALLOC_SCRATCH_BYTES));
    char *base = (char *)eng->noalloc_scratch;
    hipStream_t st = eng->stream;
    for (int i = 0; i < n_jobs; i++) {
        PreparedJob &pj = pjs[i];
        if (!pj.starts.empty()) {
            KKW_HIP(hipMemcpyAsync(base + starts_off[i], pj.starts.data(), pj.starts.size() * 4,
                                   hipMemcpyHostToDevice, st));
            pj.dj.starts = (const uint32_t *)(base + starts_off[i]);
        }
        if (!pj.scheme_len.empty()) {
            KKW_HIP(hipMemcpyAsync(base + len_off[i], pj.scheme_len.data(), pj.scheme_len.size() * 4,
                                   hipMemcpyHostToDevice, st));
            KKW_HIP(hipMemcpyAsync(base + mask_off[i], pj.scheme_mask.data(), pj.scheme_mask.size(),
                                   hipMemcpyHostToDevice, st));
            pj.dj.scheme_len = (const uint32_t *)(base + len_off[i]);
            pj.dj.scheme_mask = (const uint8_t *)(base + mask_off[i]);
        }
    }
    uint64_t acc = 0;
    for (auto &pj : pjs) {
        pj.dj.chunk_begin = acc;
        acc += (pj.dj.walker_num + w.chunk - 1) / w.chunk;
        w.line_rows = w.line_rows || (mode == KKW_MODE_PERSISTENT && pj.stride % 32 == 0);
    }
    w.n_chunks = acc;
    w.d_queue = (unsigned long long *)base;
    std::vector<DevJob> table;
    if (jobs_off) {
        for (auto &pj : pjs) table.push_back(pj.dj);
        KKW_HIP(hipMemcpyAsync(base + jobs_off, table.data(), sizeof(DevJob) * table.size(), hipMemcpyHostToDevice, st));
        w.d_jobs = (DevJob *)(base + jobs_off);
    }
    const double t0 = now_ms();
    KKW_HIP(hipEventRecord(eng->noalloc_ev[0], st));
    enqueue_walk(eng, pjs, w, st);
    KKW_HIP(hipEventRecord(eng->noalloc_ev[1], st));
    KKW_HIP(hipStreamSynchronize(st));   // also keeps the host sources of the copies alive until read
    float ms = 0.0f;
    KKW_HIP(hipEventElapsedTime(&ms, eng->noalloc_ev[0], eng->noalloc_ev[1]));
    eng->timing.kernel_ms = ms;
    eng->timing.total_ms = now_ms() - t0;
    return KKW_OK;
}

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char *kkw_last_error(void) { return g_last_error.c_str(); }

int kkw_abi_version(void) { return KKW_ABI_VERSION; }

int kkw_engine_create(const kkw_engine_opts *opts, kkw_engine **out)
{
    return guarded([&]() -> int {
        if (!out) return fail(KKW_E_INVALID, "out is null");
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            return fail(KKW_E_DEVICE, "no HIP device visible (libkkw needs a GPU; there is no CPU fallback)");
        auto eng = std::make_unique<kkw_engine>();
        eng->device = opts ? opts->device : 0;
        eng->verbose = opts ? opts->verbose : 0;
        eng->hbm_budget = opts ? opts->hbm_budget : 0;
        if (eng->device < 0 || eng->device >= ndev) return fail(KKW_E_INVALID, "device ordinal out of range");
        KKW_HIP(hipSetDevice(eng->device));
        KKW_HIP(hipDeviceGetAttribute(&eng->cu_count, hipDeviceAttributeMultiprocessorCount, eng->device));
        KKW_HIP(hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking));
        KKW_HIP(hipStreamCreateWithFlags(&eng->aux_stream, hipStreamNonBlocking));
        roctxNameHipStream("kkw engine", eng->stream);
        roctxNameHipStream("kkw aux", eng->aux_stream);
        // kkw_run_noalloc's scratch: this device's copy of the fixed array, and
        // its events, set up here so that the call itself creates nothing
        eng->noalloc_scratch = scratch_address(eng->device);
        KKW_HIP(hipEventCreate(&eng->noalloc_ev[0]));
        KKW_HIP(hipEventCreate(&eng->noalloc_ev[1]));
        pool_attach(eng->device);
        eng->pool_attached = true;
        *out = eng.release();
        return KKW_OK;
    });
}

int kkw_engine_destroy(kkw_engine *eng)
{
    if (!eng) return KKW_OK;
    kkw_serve_stop(eng);
    (void)hipSetDevice(eng->device);
    eng->gs.release();
    if (eng->stream) (void)hipStreamDestroy(eng->stream);
    if (eng->aux_stream) (void)hipStreamDestroy(eng->aux_stream);
    for (hipEvent_t e : eng->noalloc_ev)
        if (e) (void)hipEventDestroy(e);
    if (eng->pool_attached) pool_detach(eng->device);
    delete eng;
    return KKW_OK;
}

int kkw_graph_from_edges(kkw_engine *eng, uint32_t v_num, const void *edges, uint64_t n_edges, int weighted,
                         int make_undirected, int sort_adj)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        if (n_edges && !edges) return fail(KKW_E_INVALID, "edges is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        double t0 = now_ms();
        std::vector<uint32_t> src(n_edges), dst(n_edges);
        std::vector<float> w(weighted ? n_edges : 0);
        const size_t rec = weighted ? 12 : 8;  // Edge<real_t> / Edge<EmptyData>, type.hpp:43-85
        const uint8_t *p = (const uint8_t *)edges;
        for (uint64_t i = 0; i < n_edges; i++) {
            memcpy(&src[i], p + i * rec, 4);
            memcpy(&dst[i], p + i * rec + 4, 4);
            if (weighted) memcpy(&w[i], p + i * rec + 8, 4);
        }
        eng->graph_version++;
        build_from_edges(eng->gs, v_num, src.data(), dst.data(), weighted ? w.data() : nullptr, nullptr,
                         n_edges, make_undirected != 0, sort_adj != 0, eng->stream);
        eng->timing.graph_ms = now_ms() - t0;
        return KKW_OK;
    });
}

int kkw_graph_from_arrays(kkw_engine *eng, uint32_t v_num, const uint32_t *src, const uint32_t *dst,
                          const float *w, const uint8_t *etype, uint64_t n_edges, int make_undirected,
                          int sort_adj)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        if (n_edges && (!src || !dst)) return fail(KKW_E_INVALID, "src/dst is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        double t0 = now_ms();
        eng->graph_version++;
        build_from_edges(eng->gs, v_num, src, dst, w, etype, n_edges, make_undirected != 0, sort_adj != 0,
                         eng->stream);
        eng->timing.graph_ms = now_ms() - t0;
        return KKW_OK;
    });
}

int kkw_graph_sort_adj(kkw_engine *eng)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        eng->graph_version++;
        sort_adjacency(eng->gs, eng->stream);
        return KKW_OK;
    });
}

int kkw_graph_load_file(kkw_engine *eng, uint32_t v_num, const char *path, int text, int weighted,
                        int make_undirected, int sort_adj)
{
    return guarded([&]() -> int {
        if (!eng || !path) return fail(KKW_E_INVALID, "engine or path is null");
        FILE *f = fopen(path, text ? "r" : "rb");
        if (!f) return fail(KKW_E_IO, std::string("cannot open ") + path);
        std::vector<uint8_t> buf;
        const size_t rec = weighted ? 12 : 8;
        uint64_t n = 0;
        if (!text) {
            // read_graph (include/storage.hpp:35-55), one partition: the whole file
            fseek(f, 0, SEEK_END);
            long sz = ftell(f);
            fseek(f, 0, SEEK_SET);
            n = (uint64_t)sz / rec;
            buf.resize(n * rec);
            if (n && fread(buf.data(), rec, n, f) != n) {
                fclose(f);
                return fail(KKW_E_IO, "short read");
            }
        } else {
            // read_edgelist / read_edge_txt (storage.hpp:110-157): "%u %u[ %f]"
            unsigned a, b;
            float x = 0.0f;
            for (;;) {
                int r = weighted ? fscanf(f, "%u %u %f", &a, &b, &x) : fscanf(f, "%u %u", &a, &b);
                if (r == EOF) break;
                if (r != (weighted ? 3 : 2)) {
                    // skip the rest of a malformed line
                    int c;
                    while ((c = fgetc(f)) != EOF && c != '\n') {
                    }
                    if (c == EOF) break;
                    continue;
                }
                size_t o = buf.size();
                buf.resize(o + rec);
                memcpy(&buf[o], &a, 4);
                memcpy(&buf[o + 4], &b, 4);
                if (weighted) memcpy(&buf[o + 8], &x, 4);
                n++;
            }
        }
        fclose(f);
        return kkw_graph_from_edges(eng, v_num, buf.data(), n, weighted, make_undirected, sort_adj);
    });
}

int kkw_graph_from_csr(kkw_engine *eng, uint32_t v_num, const uint64_t *row_ptr, const uint32_t *col,
                       const float *w)
{
    return guarded([&]() -> int {
        if (!eng || !row_ptr || (!col && row_ptr[v_num])) return fail(KKW_E_INVALID, "null argument");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        double t0 = now_ms();
        for (uint32_t v = 0; v < v_num; v++)
            if (row_ptr[v + 1] < row_ptr[v]) return fail(KKW_E_INVALID, "row_ptr must be non-decreasing");
        eng->graph_version++;
        build_from_csr(eng->gs, v_num, row_ptr, col, w, eng->stream);
        eng->timing.graph_ms = now_ms() - t0;
        return KKW_OK;
    });
}

int kkw_graph_rmat(kkw_engine *eng, int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                   int weighted)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
            return fail(KKW_E_INVALID, "R-MAT probabilities must be non-negative with a+b+c <= 1");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        double t0 = now_ms();
        auto thr = [](double x) -> uint32_t {
            double v = std::floor(x * 4294967296.0);
            return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
        };
        eng->graph_version++;
        build_rmat(eng->gs, scale, edge_factor, thr(a), thr(a + b), thr(a + b + c), seed, weighted != 0,
                   eng->stream);
        eng->timing.graph_ms = now_ms() - t0;
        return KKW_OK;
    });
}

int kkw_graph_set_edge_types(kkw_engine *eng, const uint8_t *etype)
{
    return guarded([&]() -> int {
        if (!eng || !etype) return fail(KKW_E_INVALID, "null argument");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        eng->graph_version++;
        set_edge_types(eng->gs, etype, eng->stream);
        return KKW_OK;
    });
}

int kkw_graph_type_by_dst(kkw_engine *eng, int ntypes, uint64_t seed)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        eng->graph_version++;
        type_by_dst(eng->gs, ntypes, seed, eng->stream);
        return KKW_OK;
    });
}

int kkw_graph_info(kkw_engine *eng, uint32_t *v_num, uint64_t *e_num)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        if (!eng->gs.row_ptr) return fail(KKW_E_NOGRAPH, "no graph loaded");
        if (v_num) *v_num = eng->gs.V;
        if (e_num) *e_num = eng->gs.E;
        return KKW_OK;
    });
}

int kkw_graph_download(kkw_engine *eng, uint64_t *row_ptr, uint32_t *col, float *w, uint8_t *etype)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        GraphStore &gs = eng->gs;
        if (!gs.row_ptr) return fail(KKW_E_NOGRAPH, "no graph loaded");
        KKW_HIP(hipSetDevice(eng->device));
        if (row_ptr) KKW_HIP(hipMemcpy(row_ptr, gs.row_ptr, ((size_t)gs.V + 1) * 8, hipMemcpyDeviceToHost));
        if (col && gs.E) KKW_HIP(hipMemcpy(col, gs.col, gs.E * 4, hipMemcpyDeviceToHost));
        if (w) {
            if (!gs.w) return fail(KKW_E_INVALID, "graph has no weights");
            if (gs.E) KKW_HIP(hipMemcpy(w, gs.w, gs.E * 4, hipMemcpyDeviceToHost));
        }
        if (etype) {
            if (!gs.etype) return fail(KKW_E_INVALID, "graph has no edge types");
            if (gs.E) KKW_HIP(hipMemcpy(etype, gs.etype, gs.E, hipMemcpyDeviceToHost));
        }
        return KKW_OK;
    });
}

int kkw_prepare(kkw_engine *eng)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        if (!eng->gs.row_ptr) return fail(KKW_E_NOGRAPH, "no graph loaded");
        TraceRange tr("prepare");
        eng->graph_version++;
        const double t0 = now_ms();
        if (eng->gs.w && eng->adj_index_auto && !eng->gs.adjset) {
            // the two tables are independent: the hash sets build on the second
            // stream from a helper thread (they write gs.adjset only, the alias
            // build gs.bucket and gs.vsum) while the alias tables build here
            eng->graph_version++;
            std::exception_ptr set_err;
            double index_ms = 0.0;
            std::thread sets([&]() {
                try {
                    KKW_HIP(hipSetDevice(eng->device));
                    const double t1 = now_ms();
                    build_adj_index(eng->gs, eng->aux_stream);
                    index_ms = now_ms() - t1;
                } catch (...) {
                    set_err = std::current_exception();
                }
            });
            std::exception_ptr alias_err;
            try {
                build_alias(eng->gs, eng->stream);
            } catch (...) {
                alias_err = std::current_exception();
            }
            eng->timing.alias_ms = now_ms() - t0;
            sets.join();
            if (alias_err) std::rethrow_exception(alias_err);
            if (set_err) {
                try {
                    std::rethrow_exception(set_err);
                } catch (const Error &e) {   // as auto_adj_index: out of memory leaves the binary search
                    if (e.code != KKW_E_NOMEM) throw;
                    eng->adj_index_auto = false;
                }
            } else {
                eng->timing.index_ms = index_ms;
            }
        } else {
            if (eng->gs.w) {
                build_alias(eng->gs, eng->stream);
                eng->timing.alias_ms = now_ms() - t0;
            }
            auto_adj_index(eng);
        }
        eng->timing.prepare_ms = now_ms() - t0;
        return KKW_OK;
    });
}

int kkw_graph_adj_index(kkw_engine *eng, int enable)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        refuse_while_serving(eng);
        KKW_HIP(hipSetDevice(eng->device));
        eng->adj_index_auto = enable != 0;
        eng->graph_version++;
        if (!enable) {
            pool_free(eng->gs.adjset);
            eng->gs.adjset = nullptr;
            return KKW_OK;
        }
        if (!eng->gs.row_ptr) return fail(KKW_E_NOGRAPH, "no graph loaded");
        double t0 = now_ms();
        build_adj_index(eng->gs, eng->stream);
        eng->timing.index_ms = now_ms() - t0;
        return KKW_OK;
    });
}

int kkw_alias_download(kkw_engine *eng, float *p, uint32_t *primary, uint32_t *alias, uint32_t *meta, float *vsum)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        GraphStore &gs = eng->gs;
        if (!gs.bucket) return fail(KKW_E_NOGRAPH, "alias tables not prepared");
        KKW_HIP(hipSetDevice(eng->device));
        std::vector<Bucket> h(gs.E);
        if (gs.E) KKW_HIP(hipMemcpy(h.data(), gs.bucket, gs.E * sizeof(Bucket), hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < gs.E; i++) {
            if (p) p[i] = h[i].p;
            if (primary) primary[i] = h[i].primary;
            if (alias) alias[i] = h[i].alias;
            if (meta) meta[i] = h[i].meta & 0xFFFFu;   // the edge types (bits 16-31: the candidates' type masks)
        }
        if (vsum) KKW_HIP(hipMemcpy(vsum, gs.vsum, (size_t)gs.V * 4, hipMemcpyDeviceToHost));
        return KKW_OK;
    });
}

int kkw_batch_create(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode, int batch_flags, kkw_batch **out)
{
    return guarded([&]() -> int {
        if (!eng || !out) return fail(KKW_E_INVALID, "null argument");
        *out = nullptr;
        if (mode != KKW_MODE_PLAIN && mode != KKW_MODE_PERSISTENT) return fail(KKW_E_INVALID, "unknown mode");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        auto b = std::make_unique<kkw_batch>();
        b->eng = eng;
        b->stream = eng->stream;
        b->mode = mode;
        b->stats = (batch_flags & KKW_BATCH_STATS) != 0;
        b->use_graph = (batch_flags & KKW_BATCH_GRAPH) != 0 && !b->stats;
        batch_setup(b.get(), jobs, n_jobs);
        *out = b.release();
        return KKW_OK;
    });
}

int kkw_batch_set_shape(kkw_batch *b, int block, int blocks_per_cu, int chunk)
{
    return guarded([&]() -> int {
        if (!b) return fail(KKW_E_INVALID, "batch is null");
        if (block < 64 || block > 256 || block % 64) return fail(KKW_E_INVALID, "block must be 64..256, multiple of 64");
        if (blocks_per_cu < 1 || blocks_per_cu > 64) return fail(KKW_E_INVALID, "blocks_per_cu must be 1..64");
        if (chunk < 1 || chunk > 65536) return fail(KKW_E_INVALID, "chunk must be 1..65536");
        std::lock_guard<std::recursive_mutex> lk(b->eng->mu);
        KKW_HIP(hipSetDevice(b->eng->device));
        bool rechunk = chunk != b->chunk;
        bool reshape = rechunk || block != b->block || blocks_per_cu != b->blocks_per_cu;
        b->block = block;
        b->blocks_per_cu = blocks_per_cu;
        b->chunk = chunk;
        if (rechunk) batch_rechunk(b);
        else if (reshape && b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        return KKW_OK;
    });
}

int kkw_batch_launch(kkw_batch *b)
{
    return guarded([&]() -> int {
        if (!b) return fail(KKW_E_INVALID, "batch is null");
        std::lock_guard<std::recursive_mutex> lk(b->eng->mu);
        KKW_HIP(hipSetDevice(b->eng->device));
        batch_launch(b);
        return KKW_OK;
    });
}

int kkw_batch_sync(kkw_batch *b, double *kernel_ms)
{
    return guarded([&]() -> int {
        if (!b) return fail(KKW_E_INVALID, "batch is null");
        std::lock_guard<std::recursive_mutex> lk(b->eng->mu);
        KKW_HIP(hipSetDevice(b->eng->device));
        double ms = batch_sync(b);
        if (kernel_ms) *kernel_ms = ms;
        return KKW_OK;
    });
}

int kkw_batch_paths(kkw_batch *b, int job, int out_where, kkw_paths *out)
{
    return guarded([&]() -> int {
        if (!b || !out) return fail(KKW_E_INVALID, "null argument");
        if (job < 0 || job >= (int)b->jobs.size()) return fail(KKW_E_INVALID, "job index out of range");
        std::lock_guard<std::recursive_mutex> lk(b->eng->mu);
        KKW_HIP(hipSetDevice(b->eng->device));
        KKW_HIP(hipStreamSynchronize(b->stream));
        if (out_where == KKW_OUT_DEVICE) fill_paths_device(b->jobs[job], b->stream, out);
        else fill_paths_host(b->jobs[job], b->stream, out);
        return KKW_OK;
    });
}

int kkw_batch_destroy(kkw_batch *b)
{
    if (!b) return KKW_OK;
    std::lock_guard<std::recursive_mutex> lk(b->eng->mu);
    (void)hipSetDevice(b->eng->device);
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return KKW_OK;
}

int kkw_run(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode, int out_where, int flags, kkw_paths *out)
{
    return guarded([&]() -> int {
        if (!eng || (n_jobs > 0 && (!jobs || !out))) return fail(KKW_E_INVALID, "null argument");
        if (out_where == KKW_OUT_DEVICE) return fail(KKW_E_INVALID, "kkw_run returns host paths; use a batch for device paths");
        kkw_batch *b = nullptr;
        int rc = kkw_batch_create(eng, jobs, n_jobs, mode, (flags & KKW_FLAG_STATS) ? KKW_BATCH_STATS : 0, &b);
        if (rc) return rc;
        std::unique_ptr<kkw_batch, int (*)(kkw_batch *)> hold(b, kkw_batch_destroy);
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        batch_launch(b);
        batch_sync(b);
        for (int i = 0; i < n_jobs; i++) fill_paths_host(b->jobs[i], b->stream, &out[i]);
        return KKW_OK;
    });
}

uint32_t kkw_uniform_start(uint64_t seed, uint32_t walker, uint32_t v_num)
{
    if (v_num == 0) return 0;
    const u32x4 r = draw(seed, walker, 0, 0, TAG_INIT);   // the kernels' start_vertex()
    return bounded(r.z, r.w, v_num);
}

int kkw_run_noalloc(kkw_engine *eng, const kkw_job *jobs, int n_jobs, int mode, uint32_t *const *paths,
                    const uint32_t *strides, uint32_t *const *lengths)
{
    return guarded([&]() -> int {
        if (!eng || n_jobs < 0 || (n_jobs > 0 && (!jobs || !paths || !strides || !lengths)))
            return fail(KKW_E_INVALID, "null argument");
        if (mode != KKW_MODE_PLAIN && mode != KKW_MODE_PERSISTENT) return fail(KKW_E_INVALID, "unknown mode");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        return run_noalloc(eng, jobs, n_jobs, mode, paths, strides, lengths);
    });
}

int kkw_run_epochs(kkw_engine *eng, const kkw_job *job, uint32_t epoch_walkers, int mode,
                   kkw_epoch_consumer consumer, void *user)
{
    return guarded([&]() -> int {
        if (!eng || !job || !consumer) return fail(KKW_E_INVALID, "null argument");
        if (epoch_walkers == 0) return fail(KKW_E_INVALID, "epoch_walkers must be positive");
        if (mode != KKW_MODE_PLAIN && mode != KKW_MODE_PERSISTENT) return fail(KKW_E_INVALID, "unknown mode");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        const uint32_t total = job->walker_num;
        const uint32_t W = std::min(epoch_walkers, std::max(total, 1u));
        kkw_job first = *job;
        first.walker_num = std::min(W, total);
        PreparedJob proto = prepare_job(eng, first, mode == KKW_MODE_PERSISTENT);   // tables built here
        // two epochs' buffers (ping-pong), one work-queue word
        DevBuf paths0((size_t)W * proto.stride * 4), paths1((size_t)W * proto.stride * 4);
        DevBuf lens0((size_t)W * 4), lens1((size_t)W * 4), queue(256);
        uint32_t *paths[2] = {(uint32_t *)paths0.p, (uint32_t *)paths1.p};
        uint32_t *lens[2] = {(uint32_t *)lens0.p, (uint32_t *)lens1.p};
        struct Res {
            hipStream_t cs = nullptr;
            hipEvent_t walked[2] = {nullptr, nullptr}, consumed[2] = {nullptr, nullptr}, k0[2] = {nullptr, nullptr};
            ~Res()
            {
                if (cs) {
                    (void)hipStreamSynchronize(cs);
                    (void)hipStreamDestroy(cs);
                }
                for (int k = 0; k < 2; k++) {
                    if (walked[k]) (void)hipEventDestroy(walked[k]);
                    if (consumed[k]) (void)hipEventDestroy(consumed[k]);
                    if (k0[k]) (void)hipEventDestroy(k0[k]);
                }
            }
        } r;
        KKW_HIP(hipStreamCreateWithFlags(&r.cs, hipStreamNonBlocking));
        for (int k = 0; k < 2; k++) {
            KKW_HIP(hipEventCreate(&r.walked[k]));
            KKW_HIP(hipEventCreate(&r.k0[k]));
            KKW_HIP(hipEventCreateWithFlags(&r.consumed[k], hipEventDisableTiming));
        }
        hipStream_t st = eng->stream;
        const uint32_t n_epochs = total ? (total + W - 1) / W : 0;
        std::vector<PreparedJob> ej(1, proto);
        // upload the start list / schemes once; every epoch's DevJob points at them
        std::vector<DevBuf> keep;
        if (!proto.starts.empty()) {
            keep.emplace_back(proto.starts.size() * 4);
            KKW_HIP(hipMemcpyAsync(keep.back().p, proto.starts.data(), proto.starts.size() * 4,
                                   hipMemcpyHostToDevice, st));
            ej[0].dj.starts = (const uint32_t *)keep.back().p;
        }
        if (!proto.scheme_len.empty()) {
            keep.emplace_back(proto.scheme_len.size() * 4);
            KKW_HIP(hipMemcpyAsync(keep.back().p, proto.scheme_len.data(), proto.scheme_len.size() * 4,
                                   hipMemcpyHostToDevice, st));
            ej[0].dj.scheme_len = (const uint32_t *)keep.back().p;
            keep.emplace_back(proto.scheme_mask.size());
            KKW_HIP(hipMemcpyAsync(keep.back().p, proto.scheme_mask.data(), proto.scheme_mask.size(),
                                   hipMemcpyHostToDevice, st));
            ej[0].dj.scheme_mask = (const uint8_t *)keep.back().p;
        }
        auto enqueue_epoch = [&](uint32_t e) {
            const int b = e & 1;
            if (e >= 2) KKW_HIP(hipStreamWaitEvent(st, r.consumed[b], 0));   // buffer b is free again
            DevJob &d = ej[0].dj;
            d.walker_begin = job->walker_begin + e * W;
            d.walker_num = std::min(W, total - e * W);
            d.paths = paths[b];
            d.lens = lens[b];
            d.chunk_begin = 0;
            WalkLaunch w;
            w.mode = mode;
            w.n_chunks = (d.walker_num + w.chunk - 1) / w.chunk;
            w.d_queue = (unsigned long long *)queue.p;
            KKW_HIP(hipEventRecord(r.k0[b], st));
            enqueue_walk(eng, ej, w, st);
            KKW_HIP(hipEventRecord(r.walked[b], st));
        };
        const double t_start = now_ms();
        double kernel_ms = 0.0;
        int rc = KKW_OK;
        if (n_epochs) enqueue_epoch(0);
        for (uint32_t e = 0; e < n_epochs; e++) {
            const int b = e & 1;
            const uint32_t begin = job->walker_begin + e * W, n = std::min(W, total - e * W);
            if (e + 1 < n_epochs) enqueue_epoch(e + 1);   // walks while epoch e is consumed
            {
                TraceRange tw("epoch_walk_wait");
                KKW_HIP(hipEventSynchronize(r.walked[b]));
            }
            float ms = 0.0f;
            KKW_HIP(hipEventElapsedTime(&ms, r.k0[b], r.walked[b]));
            kernel_ms += ms;
            KKW_HIP(hipStreamWaitEvent(r.cs, r.walked[b], 0));
            kkw_paths p;
            memset(&p, 0, sizeof(p));
            p.walker_begin = begin;
            p.walker_num = n;
            p.stride = proto.stride;
            p.on_device = 1;
            p.vertices = paths[b];
            p.lengths = lens[b];
            int c;
            {
                TraceRange tc("epoch_consumer");
                c = consumer(&p, (void *)r.cs, user);
            }
            KKW_HIP(hipEventRecord(r.consumed[b], r.cs));
            if (c != 0) {
                rc = fail(KKW_E_CANCELLED, "the epoch consumer stopped the run");
                break;
            }
        }
        KKW_HIP(hipStreamSynchronize(st));
        KKW_HIP(hipStreamSynchronize(r.cs));
        eng->timing.kernel_ms = kernel_ms;
        eng->timing.total_ms = now_ms() - t_start;
        return rc;
    });
}

int kkw_paths_download(const kkw_paths *src, kkw_paths *out)
{
    return guarded([&]() -> int {
        if (!src || !out) return fail(KKW_E_INVALID, "null argument");
        if (!src->on_device) return fail(KKW_E_INVALID, "paths are already on the host");
        if (src->offsets) return fail(KKW_E_INVALID, "packed device paths are not supported");
        const size_t n = src->walker_num, cells = n * src->stride;
        memset(out, 0, sizeof(*out));
        out->walker_begin = src->walker_begin;
        out->walker_num = src->walker_num;
        out->stride = src->stride;
        out->vertices = (uint32_t *)malloc(cells ? cells * 4 : 4);
        out->lengths = (uint32_t *)malloc(n ? n * 4 : 4);
        if (!out->vertices || !out->lengths) {
            free(out->vertices);
            free(out->lengths);
            memset(out, 0, sizeof(*out));
            return fail(KKW_E_NOMEM, "host allocation for paths failed");
        }
        out->priv = (void *)1;
        if (cells) KKW_HIP(hipMemcpy(out->vertices, src->vertices, cells * 4, hipMemcpyDeviceToHost));
        if (n) KKW_HIP(hipMemcpy(out->lengths, src->lengths, n * 4, hipMemcpyDeviceToHost));
        uint64_t steps = 0, trunc = 0;
        for (size_t i = 0; i < n; i++) {
            steps += (out->lengths[i] & 0x7fffffffu) - 1;
            trunc += out->lengths[i] >> 31;
        }
        out->steps = steps;
        out->truncated = trunc;
        return KKW_OK;
    });
}

int kkw_paths_free(kkw_paths *p)
{
    if (!p) return KKW_OK;
    if (p->priv && !p->on_device) {
        if (p->priv == (void *)1) {   // host buffers of kkw_run / kkw_batch_paths / kkw_program_run
            free(p->vertices);
            free(p->lengths);
            free(p->offsets);
        } else {                      // a served result: drop its lease on the pinned block
            auto *sp = static_cast<ServedPaths *>(p->priv);
            if (sp->magic == ServedPaths::kMagic) {
                sp->magic = 0;
                delete sp;
            }
        }
    }
    memset(p, 0, sizeof(*p));
    return KKW_OK;
}

int kkw_paths_dump(const kkw_paths *p, const char *file, const char *mode, int with_head_info)
{
    return guarded([&]() -> int {
        if (!p || !file || !mode) return fail(KKW_E_INVALID, "null argument");
        if (p->on_device) return fail(KKW_E_INVALID, "paths are on the device");
        TraceRange tr("paths_dump");
        FILE *f = fopen(file, mode);
        if (!f) return fail(KKW_E_IO, std::string("cannot open ") + file);
        // PathSet::dump, include/path.hpp:85-109
        for (uint32_t i = 0; i < p->walker_num; i++) {
            uint32_t len = p->lengths[i] & 0x7fffffffu;
            if (with_head_info) fprintf(f, "%u %u", p->walker_begin + i, len);
            const uint32_t *row = p->offsets ? p->vertices + p->offsets[i] : p->vertices + (size_t)i * p->stride;
            for (uint32_t k = 0; k < len; k++) fprintf(f, " %u", row[k]);
            fprintf(f, "\n");
        }
        fclose(f);
        return KKW_OK;
    });
}

int kkw_last_timing(kkw_engine *eng, kkw_timing *t)
{
    if (!eng || !t) return fail(KKW_E_INVALID, "null argument");
    *t = eng->timing;
    return KKW_OK;
}

int kkw_diag_stream_copy(kkw_engine *eng, uint64_t bytes, int iters, double *ms_per_iter)
{
    return guarded([&]() -> int {
        if (!eng || iters < 1 || bytes < 16) return fail(KKW_E_INVALID, "bad argument");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        bytes &= ~(uint64_t)15;
        DevBuf a(bytes), b(bytes);
        KKW_HIP(hipMemsetAsync(a.p, 1, bytes, eng->stream));
        hipEvent_t e0, e1;
        KKW_HIP(hipEventCreate(&e0));
        KKW_HIP(hipEventCreate(&e1));
        launch_stream_copy(a.p, b.p, bytes, eng->cu_count, eng->stream);   // warm
        KKW_HIP(hipEventRecord(e0, eng->stream));
        for (int i = 0; i < iters; i++) launch_stream_copy(a.p, b.p, bytes, eng->cu_count, eng->stream);
        KKW_HIP(hipEventRecord(e1, eng->stream));
        KKW_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        KKW_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms_per_iter) *ms_per_iter = ms / iters;
        return KKW_OK;
    });
}

int kkw_diag_random_gather(kkw_engine *eng, uint64_t bytes, int reads, int iters, double *lines_per_s)
{
    return guarded([&]() -> int {
        if (!eng || iters < 1 || reads < 1 || bytes < 128) return fail(KKW_E_INVALID, "bad argument");
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        bytes &= ~(uint64_t)127;
        DevBuf a(bytes), sink(256);
        KKW_HIP(hipMemsetAsync(a.p, 1, bytes, eng->stream));
        const uint32_t threads = (uint32_t)eng->cu_count * 32u * 256u;
        hipEvent_t e0, e1;
        KKW_HIP(hipEventCreate(&e0));
        KKW_HIP(hipEventCreate(&e1));
        launch_random_gather(a.p, bytes, threads, (uint32_t)reads, 1, (uint32_t *)sink.p, eng->stream);   // warm
        KKW_HIP(hipEventRecord(e0, eng->stream));
        for (int i = 0; i < iters; i++)
            launch_random_gather(a.p, bytes, threads, (uint32_t)reads, 2 + (uint64_t)i, (uint32_t *)sink.p, eng->stream);
        KKW_HIP(hipEventRecord(e1, eng->stream));
        KKW_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        KKW_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        KKW_HIP(hipGetLastError());
        if (lines_per_s) *lines_per_s = (double)threads * (double)reads * iters / (ms * 1e-3);
        return KKW_OK;
    });
}

int kkw_autotune(kkw_batch *b, const char *cache_file, int *block, int *blocks_per_cu, int *chunk)
{
    return guarded([&]() -> int {
        if (!b) return fail(KKW_E_INVALID, "batch is null");
        if (b->mode != KKW_MODE_PERSISTENT) return fail(KKW_E_INVALID, "autotune applies to persistent batches");
        std::string key;
        {
            char buf[256];
            uint64_t walkers = 0;
            for (auto &pj : b->jobs) walkers += pj.dj.walker_num;
            snprintf(buf, sizeof buf, "V%u_E%llu_jobs%zu_k%d_w%llu", b->eng->gs.V, (unsigned long long)b->eng->gs.E,
                     b->jobs.size(), b->jobs.empty() ? -1 : b->jobs[0].dj.kind, (unsigned long long)walkers);
            key = buf;
        }
        if (cache_file) {
            FILE *f = fopen(cache_file, "r");
            if (f) {
                char k[256];
                int bl, bp, ch;
                while (fscanf(f, "%255s %d %d %d", k, &bl, &bp, &ch) == 4) {
                    if (key == k) {
                        fclose(f);
                        int rc = kkw_batch_set_shape(b, bl, bp, ch);
                        if (rc) return rc;
                        if (block) *block = bl;
                        if (blocks_per_cu) *blocks_per_cu = bp;
                        if (chunk) *chunk = ch;
                        return KKW_OK;
                    }
                }
                fclose(f);
            }
        }
        // one launch per shape, then the four fastest again, three launches
        // each, ranked by their median (a single launch is too noisy to rank
        // shapes within a few percent of each other)
        const int blocks[] = {64, 128, 256};
        const int bpcs[] = {4, 8, 16, 32};
        const int chunks[] = {16, 64, 256};
        struct Cand {
            double ms;
            int bl, bp, ch;
        };
        std::vector<Cand> cands;
        auto time_shape = [&](int bl, int bp, int ch) -> double {
            int rc = kkw_batch_set_shape(b, bl, bp, ch);
            if (rc) throw Error(rc, kkw_last_error());
            std::lock_guard<std::recursive_mutex> lk(b->eng->mu);
            batch_launch(b);
            return batch_sync(b);
        };
        for (int bl : blocks)
            for (int bp : bpcs) {
                if (bl * bp > 2048) continue;  // at most 32 waves per CU
                for (int ch : chunks) cands.push_back({time_shape(bl, bp, ch), bl, bp, ch});
            }
        std::sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.ms < y.ms; });
        if (cands.size() > 4) cands.resize(4);
        for (Cand &c : cands) {
            double t[3];
            for (double &x : t) x = time_shape(c.bl, c.bp, c.ch);
            std::sort(t, t + 3);
            c.ms = t[1];
        }
        std::sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.ms < y.ms; });
        const int bb = cands[0].bl, bbp = cands[0].bp, bc = cands[0].ch;
        int rc = kkw_batch_set_shape(b, bb, bbp, bc);
        if (rc) return rc;
        if (cache_file) {
            FILE *f = fopen(cache_file, "a");
            if (f) {
                fprintf(f, "%s %d %d %d\n", key.c_str(), bb, bbp, bc);
                fclose(f);
            }
        }
        if (block) *block = bb;
        if (blocks_per_cu) *blocks_per_cu = bbp;
        if (chunk) *chunk = bc;
        return KKW_OK;
    });
}

int kkw_serve_start(kkw_engine *eng, const kkw_serve_opts *opts)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        std::lock_guard<std::recursive_mutex> elk(eng->mu);
        if (eng->serve) return fail(KKW_E_INVALID, "serving already started");
        if (!eng->gs.row_ptr) return fail(KKW_E_NOGRAPH, "no graph loaded");
        auto S = std::make_unique<ServeState>();
        S->eng = eng;
        if (const char *tp = getenv("KKW_SERVE_TRACE")) S->trace_path = tp;
        if (opts) S->opts = *opts;
        if (S->opts.max_walkers_in_flight == 0) S->opts.max_walkers_in_flight = 1u << 22;
        if (S->opts.queue_capacity == 0) S->opts.queue_capacity = 1u << 16;
        KKW_HIP(hipSetDevice(eng->device));
        for (auto &s : S->streams) KKW_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        {
            static const char *const names[4] = {"kkw serve 0", "kkw serve 1", "kkw serve 2", "kkw serve 3"};
            for (int k = 0; k < 4; k++) roctxNameHipStream(names[k], S->streams[k]);
        }
        // device arena: the admission budget (hbm_budget, else half the free HBM, at most 64 GiB)
        // (the graph store's kept blocks count as free: they go back to the
        // runtime when the arena needs the room)
        size_t free_b = 0, total_b = 0;
        KKW_HIP(hipMemGetInfo(&free_b, &total_b));
        free_b += pool_kept_bytes(eng->device);
        size_t cap = eng->hbm_budget ? (size_t)eng->hbm_budget : std::min(free_b / 2, (size_t)64 << 30);
        if (cap > free_b) cap = free_b / 2;
        if (hipMalloc((void **)&S->dev_base, cap ? cap : 256) != hipSuccess) {
            (void)hipGetLastError();
            pool_trim(eng->device);
            KKW_HIP(hipMalloc((void **)&S->dev_base, cap ? cap : 256));
        }
        S->dev.init(cap);
        // pinned staging: flights pack their paths into it and results are
        // handed out in place (released by kkw_paths_free), so it is sized for
        // several flights' worst case (rows x stride)
        size_t pin_cap = std::min(cap, (size_t)4 << 30);
        S->pin = std::make_shared<PinArena>();
        KKW_HIP(hipHostMalloc((void **)&S->pin->base, pin_cap ? pin_cap : 256, hipHostMallocDefault));
        void *dp = nullptr;
        KKW_HIP(hipHostGetDevicePointer(&dp, S->pin->base, 0));
        S->pin->dbase = (char *)dp;
        S->pin->a.init(pin_cap);
        ServeState *raw = S.get();
        S->worker = std::thread(serve_loop, raw);
        eng->serve = std::move(S);
        return KKW_OK;
    });
}

int kkw_submit(kkw_engine *eng, const kkw_job *job, uint64_t *ticket)
{
    return guarded([&]() -> int {
        if (!eng || !job || !ticket) return fail(KKW_E_INVALID, "null argument");
        // the engine lock is held until the request is queued, so that
        // kkw_serve_stop cannot release the serving state in between
        std::lock_guard<std::recursive_mutex> elk(eng->mu);
        ServeState *S = eng->serve.get();
        if (!S) return fail(KKW_E_INVALID, "serving not started");
        auto r = std::make_shared<Request>();
        KKW_HIP(hipSetDevice(eng->device));
        r->pj = prepare_job(eng, *job);
        r->graph_V = eng->gs.V;
        // prepare_job keeps host copies; pointers inside src must not be used later
        r->pj.src.starts = r->pj.starts.empty() ? nullptr : r->pj.starts.data();
        r->pj.src.scheme_len = r->pj.scheme_len.empty() ? nullptr : r->pj.scheme_len.data();
        r->pj.src.scheme_mask = r->pj.scheme_mask.empty() ? nullptr : r->pj.scheme_mask.data();
        r->t_submit = now_ms();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->stop) return fail(KKW_E_INVALID, "serving is stopping");
            if (S->pending.size() >= S->opts.queue_capacity) return fail(KKW_E_BUSY, "serving queue is full");
            r->ticket = S->next_ticket++;
            S->pending.push_back(r);
            S->all[r->ticket] = r;
            *ticket = r->ticket;
        }
        S->cv_work.notify_one();
        return KKW_OK;
    });
}

int kkw_poll(kkw_engine *eng, uint64_t ticket, int *done)
{
    return guarded([&]() -> int {
        if (!eng || !done) return fail(KKW_E_INVALID, "null argument");
        ServeState *S = eng->serve.get();
        if (!S) return fail(KKW_E_INVALID, "serving not started");
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->all.find(ticket);
        if (it == S->all.end()) return fail(KKW_E_INVALID, "unknown ticket");
        *done = it->second->status != 1;
        return KKW_OK;
    });
}

int kkw_wait(kkw_engine *eng, uint64_t ticket, double timeout_s, kkw_paths *out, double *latency_ms)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        ServeState *S = eng->serve.get();
        if (!S) return fail(KKW_E_INVALID, "serving not started");
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->all.find(ticket);
        if (it == S->all.end()) return fail(KKW_E_INVALID, "unknown ticket");
        auto r = it->second;
        auto pred = [&] { return r->status != 1; };
        if (timeout_s < 0) S->cv_done.wait(lk, pred);
        else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
            return fail(KKW_E_TIMEOUT, "request not finished in time");
        S->all.erase(ticket);
        if (latency_ms) *latency_ms = r->t_done - r->t_submit;
        if (r->status != 0) {
            kkw_paths_free(&r->result);
            return fail(r->status, r->status == KKW_E_CANCELLED ? "request cancelled" : "request failed");
        }
        if (out) *out = r->result;
        else kkw_paths_free(&r->result);
        return KKW_OK;
    });
}

int kkw_cancel(kkw_engine *eng, uint64_t ticket)
{
    return guarded([&]() -> int {
        if (!eng) return fail(KKW_E_INVALID, "engine is null");
        ServeState *S = eng->serve.get();
        if (!S) return fail(KKW_E_INVALID, "serving not started");
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->all.find(ticket);
        if (it == S->all.end()) return fail(KKW_E_INVALID, "unknown ticket");
        it->second->cancelled = true;   // honoured while the request is still pending
        return KKW_OK;
    });
}

int kkw_serve_stop(kkw_engine *eng)
{
    if (!eng || !eng->serve) return KKW_OK;
    ServeState *S = eng->serve.get();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    if (!S->trace_path.empty()) {
        if (FILE *f = fopen(S->trace_path.c_str(), "w")) {
            fprintf(f, "first_submit_ms,admit_ms,enqueued_ms,phase1_seen_ms,phase2_seen_ms,retired_ms,requests,"
                       "walkers,packed_vertices\n");
            for (auto &l : S->trace) fprintf(f, "%s\n", l.c_str());
            fclose(f);
        }
    }
    for (auto &kv : S->all) kkw_paths_free(&kv.second->result);
    for (auto &s : S->streams)
        if (s) (void)hipStreamDestroy(s);
    if (S->dev_base) (void)hipFree(S->dev_base);
    S->pin.reset();   // freed once the last served result using it is freed
    std::lock_guard<std::recursive_mutex> elk(eng->mu);   // graph changes are allowed again from here
    eng->serve.reset();
    return KKW_OK;
}

}  // extern "C"
