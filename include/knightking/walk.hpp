// walk.hpp -- WalkEngine and its configuration classes, the C++ boundary of
// KnightKing (include/walk.hpp:31-437, include/graph.hpp:408-530), on top of
// the C ABI of libkkw (kkw.h).  Link with -lkkw.
//
// What carries over unchanged: the class and function names, the config
// fields and setters, WalkConfig's output file / consumer / epoch rate, the
// binary and text graph formats, PathSet.
//
// What differs: the transition functions run on the GPU, so random_walk()
// executes the device descriptor that node2vec(), deepwalk(), ppr() and
// metapath() attach to the transition config.  A config built from arbitrary
// std::function callbacks carries no descriptor and random_walk() throws
// std::logic_error for it (closures cannot run on the device, and there is no
// CPU fallback); such a transition is written instead as a WalkProgram --
// the same functions as device code (kkw_program.h) -- and run by the
// random_walk(WalkerConfig*, WalkProgram*, WalkConfig*) overload.  Walks are reproducible: each random_walk() call draws from
// the Philox stream keyed by the engine seed (set_seed, default 1) plus the
// call's index on this engine.
#pragma once

#include <assert.h>
#include <stdio.h>
#include <string.h>
#include <sys/mman.h>

#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../kkw.h"
#include "path.hpp"
#include "type.hpp"
#include "util.hpp"

template <typename edge_data_t>
struct AdjUnit {
    vertex_id_t neighbour;
    edge_data_t data;
};

template <>
struct AdjUnit<EmptyData> {
    union {
        vertex_id_t neighbour;
        EmptyData data;
    };
};

template <typename edge_data_t>
struct AdjList {
    AdjUnit<edge_data_t> *begin;
    AdjUnit<edge_data_t> *end;
};

// The CSR as KnightKing's graph engine exposes it (include/graph.hpp:77-96):
// one AdjList {begin, end} per vertex over an array of AdjUnits.
template <typename edge_data_t>
struct EdgeContainer {
    AdjList<edge_data_t> *adj_lists;
    AdjUnit<edge_data_t> *adj_units;
    EdgeContainer() : adj_lists(nullptr), adj_units(nullptr) {}
    EdgeContainer(const EdgeContainer &) = delete;
    EdgeContainer &operator=(const EdgeContainer &) = delete;
    ~EdgeContainer()
    {
        delete[] adj_lists;
        delete[] adj_units;
    }
};

template <typename walker_data_t>
struct Walker {
    walker_id_t id;
    step_t step;
    walker_data_t data;
};

template <>
struct Walker<EmptyData> {
    uint32_t id;
    union {
        step_t step;
        EmptyData data;
    };
};

template <typename query_data_t>
struct stateQuery {
    vertex_id_t src_v;
    walker_id_t walker_idx;
    query_data_t data;
};

template <typename response_data_t>
struct stateResponse {
    walker_id_t walker_idx;
    response_data_t data;
};

namespace knightking_detail {

inline void check(int rc)
{
    if (rc != KKW_OK) throw std::runtime_error(std::string("kkw: ") + kkw_last_error());
}

// Static weight and edge type of an edge record, per edge data type, and
// how the host CSR mirror fills one from the device's weight and type.
template <typename T>
struct edge_traits {
    static const bool weighted = false, typed = false;
    static float weight(const T &) { return 1.0f; }
    static int meta(const T &) { return 0; }
    static void set(T &, float, int) {}
};
template <>
struct edge_traits<real_t> {
    static const bool weighted = true, typed = false;
    static float weight(const real_t &x) { return x; }
    static int meta(const real_t &) { return 0; }
    static void set(real_t &x, float w, int) { x = w; }
};

// A device job descriptor attached to a transition config by an app function.
struct DeviceDescriptor {
    bool present = false;
    kkw_job job{};
    std::vector<uint32_t> starts;
    std::vector<uint32_t> scheme_len;
    std::vector<uint8_t> scheme_mask;
};

struct EqualDist {
    vertex_id_t v_num;
    vertex_id_t operator()(walker_id_t w) const { return w % v_num; }
};
// get_uniform_dist_func: the start vertex the device draws for walker w in
// the engine's next random_walk call (seed = engine seed + calls so far).
struct UniformDist {
    vertex_id_t v_num;
    const uint64_t *seed;
    const uint64_t *calls;
    vertex_id_t operator()(walker_id_t w) const { return kkw_uniform_start(*seed + *calls, w, v_num); }
};

}  // namespace knightking_detail

class WalkConfig {
public:
    bool output_file_flag;
    std::string output_path_prefix;
    bool print_with_head_info;
    bool output_consumer_flag;
    std::function<void(PathSet *)> output_consumer_func;
    bool device_consumer_flag;
    std::function<void(const kkw_paths &)> device_consumer_func;
    bool device_stream_consumer_flag;
    std::function<void(const kkw_paths &, void *)> device_stream_consumer_func;
    double rate;

    WalkConfig()
        : output_file_flag(false), print_with_head_info(false), output_consumer_flag(false),
          output_consumer_func(nullptr), device_consumer_flag(false), device_consumer_func(nullptr),
          device_stream_consumer_flag(false), device_stream_consumer_func(nullptr), rate(1)
    {
    }
    void set_output_file(const char *prefix, bool with_head_info = false)
    {
        assert(prefix != nullptr);
        output_file_flag = true;
        output_path_prefix = prefix;
        print_with_head_info = with_head_info;
    }
    void set_output_consumer(std::function<void(PathSet *)> f)
    {
        assert(f != nullptr);
        output_consumer_flag = true;
        output_consumer_func = f;
    }
    // No counterpart in the reference: each epoch's paths handed over where
    // they were written, in HBM (on_device = 1, rows of `stride` entries,
    // lengths with bit 31 marking a cap), for a consumer on the GPU such as a
    // skip-gram trainer.  The pointers are valid during the call only.
    void set_device_output_consumer(std::function<void(const kkw_paths &)> f)
    {
        assert(f != nullptr);
        device_consumer_flag = true;
        device_consumer_func = f;
    }
    // No counterpart in the reference: each epoch's device paths with a HIP
    // stream (hipStream_t as void *) ordered after the epoch's walk, for a
    // consumer that enqueues GPU work there instead of finishing it in the
    // call; the next epoch walks meanwhile (kkw_run_epochs).  The paths stay
    // valid for the work enqueued on the stream.
    void set_device_stream_consumer(std::function<void(const kkw_paths &, void *)> f)
    {
        assert(f != nullptr);
        device_stream_consumer_flag = true;
        device_stream_consumer_func = f;
    }
    void set_walk_rate(double r)
    {
        assert(r > 0 && r <= 1);
        rate = r;
    }
};

template <typename edge_data_t, typename walker_data_t>
class WalkerConfig {
public:
    walker_id_t walker_num;
    std::function<vertex_id_t(walker_id_t)> walker_init_dist_func;
    std::function<void(Walker<walker_data_t> &, vertex_id_t)> walker_init_state_func;
    std::function<void(Walker<walker_data_t> &, vertex_id_t, AdjUnit<edge_data_t> *)> walker_update_state_func;

    WalkerConfig() : walker_num(0) {}
    WalkerConfig(walker_id_t n,
                 std::function<void(Walker<walker_data_t> &, vertex_id_t)> init_state = nullptr,
                 std::function<void(Walker<walker_data_t> &, vertex_id_t, AdjUnit<edge_data_t> *)> update = nullptr,
                 std::function<vertex_id_t(walker_id_t)> init_dist = nullptr)
    {
        set_walkers(n, init_state, update, init_dist);
    }
    void set_walkers(walker_id_t n,
                     std::function<void(Walker<walker_data_t> &, vertex_id_t)> init_state = nullptr,
                     std::function<void(Walker<walker_data_t> &, vertex_id_t, AdjUnit<edge_data_t> *)> update = nullptr,
                     std::function<vertex_id_t(walker_id_t)> init_dist = nullptr)
    {
        walker_num = n;
        walker_init_state_func = init_state;
        walker_update_state_func = update;
        walker_init_dist_func = init_dist;
    }
};

template <typename edge_data_t, typename walker_data_t>
class TransitionConfig {
public:
    std::function<real_t(Walker<walker_data_t> &, vertex_id_t)> extension_comp_func;
    std::function<real_t(vertex_id_t, AdjUnit<edge_data_t> *)> static_comp_func;
    std::function<real_t(Walker<walker_data_t> &, vertex_id_t, AdjUnit<edge_data_t> *)> dynamic_comp_func;
    std::function<real_t(vertex_id_t, AdjList<edge_data_t> *)> dcomp_upperbound_func;
    std::function<real_t(vertex_id_t, AdjList<edge_data_t> *)> dcomp_lowerbound_func;
    std::function<void(Walker<walker_data_t> &, vertex_id_t, AdjList<edge_data_t> *, real_t &, vertex_id_t &)>
        outlier_upperbound_func;
    std::function<AdjUnit<edge_data_t> *(Walker<walker_data_t> &, vertex_id_t, AdjList<edge_data_t> *, vertex_id_t)>
        outlier_search_func;
    knightking_detail::DeviceDescriptor device;

    TransitionConfig() {}
    TransitionConfig(decltype(extension_comp_func) ext, decltype(static_comp_func) st = nullptr,
                     decltype(dynamic_comp_func) dyn = nullptr, decltype(dcomp_upperbound_func) ub = nullptr,
                     decltype(dcomp_lowerbound_func) lb = nullptr, decltype(outlier_upperbound_func) oub = nullptr,
                     decltype(outlier_search_func) os = nullptr)
    {
        set_transition(ext, st, dyn, ub, lb, oub, os);
    }
    void set_transition(decltype(extension_comp_func) ext, decltype(static_comp_func) st = nullptr,
                        decltype(dynamic_comp_func) dyn = nullptr, decltype(dcomp_upperbound_func) ub = nullptr,
                        decltype(dcomp_lowerbound_func) lb = nullptr, decltype(outlier_upperbound_func) oub = nullptr,
                        decltype(outlier_search_func) os = nullptr)
    {
        // validation of include/walk.hpp:251-261
        if (dyn != nullptr) {
            assert(ub != nullptr);
        } else {
            assert(ub == nullptr && lb == nullptr && oub == nullptr && os == nullptr);
        }
        assert((oub == nullptr && os == nullptr) || (oub != nullptr && os != nullptr));
        extension_comp_func = ext;
        static_comp_func = st;
        dynamic_comp_func = dyn;
        dcomp_upperbound_func = ub;
        dcomp_lowerbound_func = lb;
        outlier_upperbound_func = oub;
        outlier_search_func = os;
    }
};

template <typename edge_data_t, typename walker_data_t, typename query_data_t, typename response_data_t>
class SecondOrderTransitionConfig {
public:
    std::function<real_t(Walker<walker_data_t> &, vertex_id_t)> extension_comp_func;
    std::function<real_t(vertex_id_t, AdjUnit<edge_data_t> *)> static_comp_func;
    std::function<void(Walker<walker_data_t> &, walker_id_t, vertex_id_t, AdjUnit<edge_data_t> *)> post_query_func;
    std::function<void(vertex_id_t, stateQuery<query_data_t> &, AdjList<edge_data_t> *)> respond_query_func;
    std::function<real_t(Walker<walker_data_t> &, stateResponse<response_data_t> &, vertex_id_t,
                         AdjUnit<edge_data_t> *)>
        dynamic_comp_func;
    std::function<real_t(vertex_id_t, AdjList<edge_data_t> *)> dcomp_upperbound_func;
    std::function<real_t(vertex_id_t, AdjList<edge_data_t> *)> dcomp_lowerbound_func;
    std::function<void(Walker<walker_data_t> &, vertex_id_t, AdjList<edge_data_t> *, real_t &, vertex_id_t &)>
        outlier_upperbound_func;
    std::function<AdjUnit<edge_data_t> *(Walker<walker_data_t> &, vertex_id_t, AdjList<edge_data_t> *, vertex_id_t)>
        outlier_search_func;
    knightking_detail::DeviceDescriptor device;

    SecondOrderTransitionConfig() {}
    SecondOrderTransitionConfig(decltype(extension_comp_func) ext, decltype(static_comp_func) st,
                                decltype(post_query_func) pq, decltype(respond_query_func) rq,
                                decltype(dynamic_comp_func) dyn, decltype(dcomp_upperbound_func) ub,
                                decltype(dcomp_lowerbound_func) lb = nullptr,
                                decltype(outlier_upperbound_func) oub = nullptr,
                                decltype(outlier_search_func) os = nullptr)
    {
        set_transition(ext, st, pq, rq, dyn, ub, lb, oub, os);
    }
    void set_transition(decltype(extension_comp_func) ext, decltype(static_comp_func) st,
                        decltype(post_query_func) pq, decltype(respond_query_func) rq,
                        decltype(dynamic_comp_func) dyn, decltype(dcomp_upperbound_func) ub,
                        decltype(dcomp_lowerbound_func) lb = nullptr, decltype(outlier_upperbound_func) oub = nullptr,
                        decltype(outlier_search_func) os = nullptr)
    {
        // validation of include/walk.hpp:339-351
        if (dyn != nullptr) {
            assert(ub != nullptr);
        } else {
            assert(ub == nullptr && lb == nullptr && oub == nullptr && os == nullptr);
        }
        assert(pq != nullptr && rq != nullptr);
        assert((oub == nullptr && os == nullptr) || (oub != nullptr && os != nullptr));
        extension_comp_func = ext;
        static_comp_func = st;
        post_query_func = pq;
        respond_query_func = rq;
        dynamic_comp_func = dyn;
        dcomp_upperbound_func = ub;
        dcomp_lowerbound_func = lb;
        outlier_upperbound_func = oub;
        outlier_search_func = os;
    }
};

// A first-order transition with arbitrary logic, as device code: the
// counterpart of TransitionConfig plus WalkerConfig's state closures
// (include/walk.hpp:147-271).  `source` defines the functions kkw_program.h
// lists; the data a closure would capture by reference goes in arrays
// (kkw_ctx.array[k]) and param (kkw_ctx.param).
struct WalkProgram {
    std::string source;
    std::string options;                         // extra compiler options, e.g. "-DTAGS=4"
    std::vector<std::vector<unsigned char> > arrays;
    float param[8];
    bool use_weights;                            // static component: edge weights (else uniform)
    bool second_order;                           // re-test the extension on every retry
    bool persistent;                             // lanes refilled from a work queue (same paths)
    step_t max_length;                           // path capacity in steps

    WalkProgram() : use_weights(false), second_order(false), persistent(true), max_length(0)
    {
        for (int k = 0; k < 8; k++) param[k] = 0.0f;
    }
    explicit WalkProgram(const std::string &src) : WalkProgram() { source = src; }
    template <typename T>
    void set_array(int slot, const T *data, size_t n)
    {
        if (slot < 0 || slot >= 8) throw std::invalid_argument("WalkProgram: array slot must be 0-7");
        if ((size_t)slot >= arrays.size()) arrays.resize(slot + 1);
        const unsigned char *b = reinterpret_cast<const unsigned char *>(data);
        arrays[slot].assign(b, b + n * sizeof(T));
    }
};

template <typename edge_data_t, typename walker_data_t>
class WalkEngine {
    kkw_engine *eng_ = nullptr;
    vertex_id_t v_num_ = 0;
    edge_id_t e_num_ = 0;
    uint64_t seed_ = 1;
    uint64_t calls_ = 0;
    // host mirror of the device CSR, made on first use of csr / degrees
    std::unique_ptr<EdgeContainer<edge_data_t> > mirror_;
    std::vector<vertex_id_t> out_deg_, in_deg_;
    uint64_t mirror_sig_ = 0;
    std::recursive_mutex mirror_mu_;

public:
    typedef knightking_detail::edge_traits<edge_data_t> traits;

    // KnightKing's public graph members (include/graph.hpp:152-156): `csr`
    // (csr->adj_lists[v].begin / end over AdjUnits) and the degree arrays.
    // They read a host mirror of the device CSR, downloaded on first use.
    // The mirror may be edited in place as KnightKing apps do (node2vec sorts
    // every list through process_vertices): the next walk uploads it when its
    // contents changed.
    class CsrHandle {
        WalkEngine *g_;

    public:
        explicit CsrHandle(WalkEngine *g) : g_(g) {}
        EdgeContainer<edge_data_t> *operator->() const { return g_->host_csr(); }
        EdgeContainer<edge_data_t> &operator*() const { return *g_->host_csr(); }
        operator EdgeContainer<edge_data_t> *() const { return g_->host_csr(); }
    };
    class DegreeHandle {
        WalkEngine *g_;
        bool out_;

    public:
        DegreeHandle(WalkEngine *g, bool out) : g_(g), out_(out) {}
        vertex_id_t &operator[](size_t v) const { return g_->host_degree(out_)[v]; }
        operator vertex_id_t *() const { return g_->host_degree(out_); }
    };
    CsrHandle csr;
    DegreeHandle vertex_out_degree;
    DegreeHandle vertex_in_degree;

    WalkEngine() : csr(this), vertex_out_degree(this, true), vertex_in_degree(this, false)
    {
        kkw_engine_opts o{};
        knightking_detail::check(kkw_engine_create(&o, &eng_));
    }
    ~WalkEngine() { kkw_engine_destroy(eng_); }
    WalkEngine(const WalkEngine &) = delete;
    WalkEngine &operator=(const WalkEngine &) = delete;

    kkw_engine *native() { return eng_; }
    void set_concurrency(int) {}          // one GPU; kept for source compatibility
    void set_seed(uint64_t s) { seed_ = s; calls_ = 0; }
    vertex_id_t get_vertex_num() { return v_num_; }
    edge_id_t get_edge_num() { return e_num_; }
    int get_worker_num() { return 1; }

    // One process drives the whole graph: the local partition is every vertex
    // (include/graph.hpp:164-201).
    vertex_id_t get_local_vertex_begin() { return 0; }
    vertex_id_t get_local_vertex_end() { return v_num_; }
    vertex_id_t get_vertex_begin(partition_id_t p) { return p == 0 ? 0 : v_num_; }
    vertex_id_t get_vertex_end(partition_id_t) { return v_num_; }
    bool is_local_vertex(vertex_id_t v) { return v < v_num_; }
    bool is_valid_edge(Edge<edge_data_t> e) { return e.src < v_num_ && e.dst < v_num_; }

    // Vertex arrays (include/graph.hpp:203-234): zero-filled anonymous
    // mappings, as KnightKing allocates them.
    template <typename T>
    T *alloc_array(size_t num)
    {
        void *p = mmap(NULL, sizeof(T) * (num ? num : 1), PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (p == MAP_FAILED) throw std::bad_alloc();
        return static_cast<T *>(p);
    }
    template <typename T>
    void dealloc_array(T *array, size_t num)
    {
        munmap(array, sizeof(T) * (num ? num : 1));
    }
    template <typename T>
    T *alloc_vertex_array()
    {
        return alloc_array<T>(v_num_);
    }
    template <typename T>
    void dealloc_vertex_array(T *array)
    {
        dealloc_array(array, v_num_);
    }

    // process_vertices (include/graph.hpp:920-941): `process` on every vertex
    // on the host, summed.  Closures may read and reorder csr.
    template <typename reducer_data_t>
    reducer_data_t process_vertices(std::function<reducer_data_t(vertex_id_t)> process)
    {
        (void)host_csr();
        reducer_data_t reducer = 0;
        for (vertex_id_t v = 0; v < v_num_; v++) reducer += process(v);
        return reducer;
    }

    // The generator user closures draw from on the host (walk.hpp:398-401).
    StdRandNumGenerator *get_thread_local_rand_gen()
    {
        static thread_local StdRandNumGenerator gen;
        return &gen;
    }

    // GraphEngine::load_graph (include/graph.hpp:408-530), storage.hpp:35-157 formats.
    void load_graph(vertex_id_t v_num, const char *graph_path, bool load_as_undirected = false,
                    GraphFormat graph_format = GF_Binary)
    {
        std::vector<uint32_t> src, dst;
        std::vector<float> w;
        std::vector<uint8_t> et;
        FILE *f = fopen(graph_path, graph_format == GF_Binary ? "rb" : "r");
        if (!f) throw std::runtime_error(std::string("cannot open ") + graph_path);
        if (graph_format == GF_Binary) {
            Edge<edge_data_t> e;
            while (fread(&e, sizeof(e), 1, f) == 1) push(e, src, dst, w, et);
        } else {
            read_text(f, src, dst, w);
        }
        fclose(f);
        drop_mirror();
        knightking_detail::check(kkw_graph_from_arrays(
            eng_, v_num, src.data(), dst.data(), traits::weighted ? w.data() : nullptr,
            traits::typed ? et.data() : nullptr, src.size(), load_as_undirected ? 1 : 0, 0));
        uint32_t v = 0;
        uint64_t e = 0;
        knightking_detail::check(kkw_graph_info(eng_, &v, &e));
        v_num_ = v;
        e_num_ = e;
    }

    void sort_adjacency()
    {
        sync_mirror();
        knightking_detail::check(kkw_graph_sort_adj(eng_));
        drop_mirror();
    }

    std::function<vertex_id_t(walker_id_t)> get_equal_dist_func()
    {
        return knightking_detail::EqualDist{v_num_};
    }
    std::function<vertex_id_t(walker_id_t)> get_uniform_dist_func()
    {
        return knightking_detail::UniformDist{v_num_, &seed_, &calls_};
    }

    // WalkEngine::random_walk (include/walk.hpp:423-436, epochs 698-878).
    template <typename transition_config_t>
    void random_walk(WalkerConfig<edge_data_t, walker_data_t> *walker_config,
                     transition_config_t *transition_config, WalkConfig *walk_config_param = nullptr)
    {
        WalkConfig local;
        WalkConfig *wc = walk_config_param ? walk_config_param : &local;
        sync_mirror();
        knightking_detail::DeviceDescriptor &dd = transition_config->device;
        if (!dd.present)
            throw std::logic_error("random_walk: this transition config has no device descriptor; "
                                   "arbitrary std::function callbacks cannot run on the GPU "
                                   "(use node2vec, deepwalk, ppr or metapath)");
        kkw_job job = dd.job;
        job.seed = seed_ + calls_++;
        const walker_id_t walker_num = walker_config->walker_num;
        std::vector<uint32_t> host_starts;
        auto &dist = walker_config->walker_init_dist_func;
        if (dist && !dist.template target<knightking_detail::EqualDist>() &&
            !dist.template target<knightking_detail::UniformDist>() && job.start_mode == KKW_START_EQUAL) {
            // a caller-supplied start distribution: evaluate it per walker id on the host
            host_starts.resize(walker_num);
            for (walker_id_t i = 0; i < walker_num; i++) host_starts[i] = dist(i);
            job.start_mode = KKW_START_LIST;
        } else if (dist && dist.template target<knightking_detail::UniformDist>()) {
            job.start_mode = KKW_START_UNIFORM;
        }
        if (job.start_mode == KKW_START_LIST) {
            const std::vector<uint32_t> &s = host_starts.empty() ? dd.starts : host_starts;
            job.starts = s.data();
            job.n_starts = (uint32_t)s.size();
        }
        job.scheme_len = dd.scheme_len.empty() ? nullptr : dd.scheme_len.data();
        job.scheme_mask = dd.scheme_mask.empty() ? nullptr : dd.scheme_mask.data();
        walker_id_t per_iter = (walker_id_t)(walker_num * wc->rate);
        if (per_iter == 0) per_iter = 1;
        if (per_iter > walker_num) per_iter = walker_num;
        if (wc->device_stream_consumer_flag) {
            if (wc->device_consumer_flag)
                throw std::invalid_argument("WalkConfig: set either a device consumer or a device stream consumer");
            run_epochs_streamed(job, walker_num, per_iter, wc);
            return;
        }
        walker_id_t remained = walker_num;
        int iter = 0;
        while (remained != 0) {
            walker_id_t begin = walker_num - remained;
            walker_id_t n = remained < per_iter ? remained : per_iter;
            remained -= n;
            job.walker_begin = begin;
            job.walker_num = n;
            kkw_batch *b = nullptr;
            knightking_detail::check(kkw_batch_create(eng_, &job, 1, KKW_MODE_PERSISTENT, 0, &b));
            struct BatchHolder {
                kkw_batch *b;
                ~BatchHolder() { kkw_batch_destroy(b); }
            } hold{b};
            knightking_detail::check(kkw_batch_launch(b));
            knightking_detail::check(kkw_batch_sync(b, nullptr));
            if (wc->device_consumer_flag) {
                kkw_paths dev{};
                knightking_detail::check(kkw_batch_paths(b, 0, KKW_OUT_DEVICE, &dev));
                wc->device_consumer_func(dev);
            }
            if (wc->output_file_flag || wc->output_consumer_flag) {
                kkw_paths out{};
                knightking_detail::check(kkw_batch_paths(b, 0, KKW_OUT_HOST, &out));
                PathSet *ps = to_pathset(out);
                kkw_paths_free(&out);
                if (wc->output_file_flag) {
                    std::string p = wc->output_path_prefix + ".0";
                    ps->dump(p.c_str(), iter == 0 ? "w" : "a", wc->print_with_head_info);
                }
                if (wc->output_consumer_flag) wc->output_consumer_func(ps);
                else delete ps;
            }
            iter++;
        }
    }

    // random_walk for a WalkProgram: compiled for the GPU on each call, run in
    // the epochs WalkConfig's rate gives, output as for the apps.
    void random_walk(WalkerConfig<edge_data_t, walker_data_t> *walker_config, WalkProgram *program,
                     WalkConfig *walk_config_param = nullptr)
    {
        WalkConfig local;
        WalkConfig *wc = walk_config_param ? walk_config_param : &local;
        if (program->max_length == 0) throw std::invalid_argument("WalkProgram: max_length must be set");
        sync_mirror();
        if (wc->device_consumer_flag)
            throw std::invalid_argument("WalkProgram: paths come back to the host; no device consumer");
        kkw_program *prog = nullptr;
        knightking_detail::check(kkw_program_create(eng_, program->source.c_str(),
                                                    program->options.empty() ? nullptr : program->options.c_str(),
                                                    &prog));
        struct Holder {
            kkw_program *p;
            ~Holder() { kkw_program_destroy(p); }
        } hold{prog};
        for (size_t k = 0; k < program->arrays.size(); k++)
            if (!program->arrays[k].empty())
                knightking_detail::check(
                    kkw_program_set_array(prog, (int)k, program->arrays[k].data(), program->arrays[k].size()));
        kkw_prog_job job;
        memset(&job, 0, sizeof(job));
        job.max_length = program->max_length;
        job.use_weights = program->use_weights ? 1 : 0;
        job.second_order = program->second_order ? 1 : 0;
        job.mode = program->persistent ? KKW_MODE_PERSISTENT : KKW_MODE_PLAIN;
        job.seed = seed_ + calls_++;
        for (int k = 0; k < 8; k++) job.param[k] = program->param[k];
        const walker_id_t walker_num = walker_config->walker_num;
        std::vector<uint32_t> host_starts;
        auto &dist = walker_config->walker_init_dist_func;
        job.start_mode = KKW_START_EQUAL;
        if (dist && dist.template target<knightking_detail::UniformDist>()) {
            job.start_mode = KKW_START_UNIFORM;
        } else if (dist && !dist.template target<knightking_detail::EqualDist>()) {
            host_starts.resize(walker_num);
            for (walker_id_t i = 0; i < walker_num; i++) host_starts[i] = dist(i);
            job.start_mode = KKW_START_LIST;
            job.starts = host_starts.data();
            job.n_starts = (uint32_t)host_starts.size();
        }
        walker_id_t per_iter = (walker_id_t)(walker_num * wc->rate);
        if (per_iter == 0) per_iter = 1;
        if (per_iter > walker_num) per_iter = walker_num;
        walker_id_t remained = walker_num;
        int iter = 0;
        while (remained != 0) {
            walker_id_t n = remained < per_iter ? remained : per_iter;
            job.walker_begin = walker_num - remained;
            job.walker_num = n;
            remained -= n;
            kkw_paths out{};
            knightking_detail::check(kkw_program_run(prog, &job, &out));
            if (wc->output_file_flag || wc->output_consumer_flag) {
                PathSet *ps = to_pathset(out);
                if (wc->output_file_flag) {
                    std::string p = wc->output_path_prefix + ".0";
                    ps->dump(p.c_str(), iter == 0 ? "w" : "a", wc->print_with_head_info);
                }
                if (wc->output_consumer_flag) wc->output_consumer_func(ps);
                else delete ps;
            }
            kkw_paths_free(&out);
            iter++;
        }
    }

private:
    // Epochs through kkw_run_epochs: the stream consumer gets each epoch in
    // HBM while the next one walks; file and host consumers get host copies.
    struct EpochCtx {
        WalkConfig *wc;
        int iter;
        std::exception_ptr err;
    };
    static int epoch_trampoline(const kkw_paths *p, void *stream, void *user)
    {
        EpochCtx *c = static_cast<EpochCtx *>(user);
        try {
            c->wc->device_stream_consumer_func(*p, stream);
            if (c->wc->output_file_flag || c->wc->output_consumer_flag) {
                kkw_paths host{};
                knightking_detail::check(kkw_paths_download(p, &host));
                PathSet *ps = to_pathset(host);
                kkw_paths_free(&host);
                if (c->wc->output_file_flag) {
                    std::string path = c->wc->output_path_prefix + ".0";
                    ps->dump(path.c_str(), c->iter == 0 ? "w" : "a", c->wc->print_with_head_info);
                }
                if (c->wc->output_consumer_flag) c->wc->output_consumer_func(ps);
                else delete ps;
            }
            c->iter++;
            return 0;
        } catch (...) {
            c->err = std::current_exception();
            return 1;
        }
    }
    void run_epochs_streamed(kkw_job job, walker_id_t walker_num, walker_id_t per_iter, WalkConfig *wc)
    {
        job.walker_begin = 0;
        job.walker_num = walker_num;
        EpochCtx ctx{wc, 0, nullptr};
        const int rc = kkw_run_epochs(eng_, &job, per_iter, KKW_MODE_PERSISTENT, epoch_trampoline, &ctx);
        if (ctx.err) std::rethrow_exception(ctx.err);
        knightking_detail::check(rc);
    }

    void drop_mirror()
    {
        std::lock_guard<std::recursive_mutex> lk(mirror_mu_);
        mirror_.reset();
        out_deg_.clear();
        in_deg_.clear();
    }
    EdgeContainer<edge_data_t> *host_csr()
    {
        std::lock_guard<std::recursive_mutex> lk(mirror_mu_);
        if (!mirror_) download_mirror();
        return mirror_.get();
    }
    vertex_id_t *host_degree(bool out)
    {
        std::lock_guard<std::recursive_mutex> lk(mirror_mu_);
        if (!mirror_) download_mirror();
        return out ? out_deg_.data() : in_deg_.data();
    }
    void download_mirror()
    {
        if (v_num_ == 0) throw std::logic_error("no graph loaded");
        std::vector<uint64_t> rp((size_t)v_num_ + 1);
        std::vector<uint32_t> col(e_num_);
        std::vector<float> w(traits::weighted ? e_num_ : 0);
        std::vector<uint8_t> et(traits::typed ? e_num_ : 0);
        knightking_detail::check(kkw_graph_download(eng_, rp.data(), col.data(), traits::weighted ? w.data() : nullptr,
                                                    traits::typed ? et.data() : nullptr));
        std::unique_ptr<EdgeContainer<edge_data_t> > ec(new EdgeContainer<edge_data_t>());
        ec->adj_units = new AdjUnit<edge_data_t>[e_num_ ? e_num_ : 1];
        ec->adj_lists = new AdjList<edge_data_t>[v_num_];
        for (edge_id_t e = 0; e < e_num_; e++) {
            ec->adj_units[e].neighbour = col[e];
            traits::set(unit_data(ec->adj_units[e]), traits::weighted ? w[e] : 1.0f, traits::typed ? et[e] : 0);
        }
        out_deg_.assign(v_num_, 0);
        in_deg_.assign(v_num_, 0);
        for (vertex_id_t v = 0; v < v_num_; v++) {
            ec->adj_lists[v].begin = ec->adj_units + rp[v];
            ec->adj_lists[v].end = ec->adj_units + rp[v + 1];
            out_deg_[v] = (vertex_id_t)(rp[v + 1] - rp[v]);
        }
        for (edge_id_t e = 0; e < e_num_; e++) in_deg_[col[e]]++;
        mirror_ = std::move(ec);
        mirror_sig_ = signature();
    }
    template <typename U>
    static auto unit_data(AdjUnit<U> &u) -> decltype((u.data)) { return u.data; }
    // FNV-1a over every unit's neighbour, weight and type, in CSR order
    uint64_t signature() const
    {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&h](uint32_t x) {
            for (int b = 0; b < 4; b++) {
                h ^= (x >> (8 * b)) & 0xffu;
                h *= 1099511628211ull;
            }
        };
        for (edge_id_t e = 0; e < e_num_; e++) {
            AdjUnit<edge_data_t> &u = mirror_->adj_units[e];
            mix(u.neighbour);
            if (traits::weighted) {
                float x = traits::weight(unit_data(u));
                uint32_t bits;
                memcpy(&bits, &x, 4);
                mix(bits);
            }
            if (traits::typed) mix((uint32_t)traits::meta(unit_data(u)));
        }
        return h;
    }
    // A mirror edited on the host goes back to the device before the next walk.
    void sync_mirror()
    {
        std::lock_guard<std::recursive_mutex> lk(mirror_mu_);
        if (!mirror_) return;
        const uint64_t sig = signature();
        if (sig == mirror_sig_) return;
        std::vector<uint64_t> rp((size_t)v_num_ + 1, 0);
        std::vector<uint32_t> col(e_num_);
        std::vector<float> w(traits::weighted ? e_num_ : 0);
        std::vector<uint8_t> et(traits::typed ? e_num_ : 0);
        edge_id_t o = 0;
        for (vertex_id_t v = 0; v < v_num_; v++) {
            for (AdjUnit<edge_data_t> *p = mirror_->adj_lists[v].begin; p < mirror_->adj_lists[v].end; p++, o++) {
                if (o >= e_num_) throw std::logic_error("csr mirror: lists hold more units than the graph");
                col[o] = p->neighbour;
                if (traits::weighted) w[o] = traits::weight(unit_data(*p));
                if (traits::typed) et[o] = (uint8_t)traits::meta(unit_data(*p));
            }
            rp[v + 1] = o;
        }
        if (o != e_num_) throw std::logic_error("csr mirror: lists hold fewer units than the graph");
        knightking_detail::check(kkw_graph_from_csr(eng_, v_num_, rp.data(), col.data(),
                                                    traits::weighted ? w.data() : nullptr));
        if (traits::typed) knightking_detail::check(kkw_graph_set_edge_types(eng_, et.data()));
        mirror_sig_ = sig;
    }

    template <typename T = edge_data_t>
    void push(const Edge<T> &e, std::vector<uint32_t> &src, std::vector<uint32_t> &dst, std::vector<float> &w,
              std::vector<uint8_t> &et)
    {
        src.push_back(e.src);
        dst.push_back(e.dst);
        if (traits::weighted) w.push_back(traits::weight(e.data));
        if (traits::typed) et.push_back((uint8_t)traits::meta(e.data));
    }
    void push(const Edge<EmptyData> &e, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
              std::vector<float> &, std::vector<uint8_t> &)
    {
        src.push_back(e.src);
        dst.push_back(e.dst);
    }
    void read_text(FILE *f, std::vector<uint32_t> &src, std::vector<uint32_t> &dst, std::vector<float> &w)
    {
        // read_edge_txt (include/storage.hpp:106-121): "%u %u" or "%u %u %f";
        // other edge types cannot be read from text (a run-time error there too)
        if (!std::is_same<edge_data_t, EmptyData>::value && !std::is_same<edge_data_t, real_t>::value)
            throw std::invalid_argument("load_graph: this edge type does not support reading from text");
        unsigned a, b;
        float x;
        for (;;) {
            int r = traits::weighted ? fscanf(f, "%u %u %f", &a, &b, &x) : fscanf(f, "%u %u", &a, &b);
            if (r != (traits::weighted ? 3 : 2)) break;
            src.push_back(a);
            dst.push_back(b);
            if (traits::weighted) w.push_back(x);
        }
    }
    static PathSet *to_pathset(const kkw_paths &p)
    {
        PathSet *ps = new PathSet();
        ps->seg_num = 1;
        ps->path_set = new vertex_id_t *[1];
        ps->walker_id = new walker_id_t *[1];
        ps->path_begin = new vertex_id_t **[1];
        ps->path_end = new vertex_id_t **[1];
        ps->path_length = new step_t *[1];
        ps->path_num = new walker_id_t[1];
        uint64_t total = 0;
        for (uint32_t i = 0; i < p.walker_num; i++) total += p.lengths[i] & 0x7fffffffu;
        ps->path_set[0] = new vertex_id_t[total ? total : 1];
        ps->walker_id[0] = new walker_id_t[p.walker_num ? p.walker_num : 1];
        ps->path_begin[0] = new vertex_id_t *[p.walker_num ? p.walker_num : 1];
        ps->path_end[0] = new vertex_id_t *[p.walker_num ? p.walker_num : 1];
        ps->path_length[0] = new step_t[p.walker_num ? p.walker_num : 1];
        ps->path_num[0] = p.walker_num;
        uint64_t o = 0;
        for (uint32_t i = 0; i < p.walker_num; i++) {
            uint32_t len = p.lengths[i] & 0x7fffffffu;
            const uint32_t *row = p.offsets ? p.vertices + p.offsets[i] : p.vertices + (uint64_t)i * p.stride;
            memcpy(ps->path_set[0] + o, row, len * sizeof(vertex_id_t));
            ps->walker_id[0][i] = p.walker_begin + i;
            ps->path_begin[0][i] = ps->path_set[0] + o;
            ps->path_end[0][i] = ps->path_set[0] + o + len;
            ps->path_length[0][i] = len;
            o += len;
        }
        return ps;
    }
};
