// engine_state.h -- engine state and the helpers every C-ABI translation
// unit of libkkw shares: thread-local last error, exception-to-status
// translation, owned device buffers.
#pragma once

#include "kkw_internal.h"
#include "../../include/kkw.h"

#include <chrono>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>

namespace kkw {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

template <typename F>
inline int guarded(F &&f)
{
    try {
        return f();
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(KKW_E_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(KKW_E_INVALID, e.what());
    }
}

inline double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Device allocation owned by a batch / request (from the device pool, so a
// batch's path buffers are kept for the next batch, and blocks the graph store
// keeps go back to the runtime when a batch needs the room).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n)
    {
        bytes = n;
        p = pool_alloc(n ? n : 1);   // Error(KKW_E_NOMEM) when out of memory
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; }
    ~DevBuf() { pool_free(p); }
};

}  // namespace kkw

struct ServeState;

struct kkw_engine {
    int device = 0;
    int verbose = 0;
    uint64_t hbm_budget = 0;
    int cu_count = 256;
    hipStream_t stream = nullptr;
    hipStream_t aux_stream = nullptr;   // kkw_prepare: the hash sets build here beside the alias tables
    std::recursive_mutex mu;
    kkw::GraphStore gs;
    kkw_timing timing{};
    bool adj_index_auto = true;
    // bumped whenever device graph storage is reallocated; batches compare it
    // before replaying a captured launch
    uint64_t graph_version = 1;
    std::unique_ptr<ServeState> serve;
    // kkw_run_noalloc: this device's copy of the library's fixed scratch array
    // (its device address, from hipGetSymbolAddress) and the two timing events
    void *noalloc_scratch = nullptr;
    hipEvent_t noalloc_ev[2] = {nullptr, nullptr};
    bool pool_attached = false;   // registered with the device pool (kkw::pool_attach)
};

// Neighbour hash sets when the engine builds them on its own (first node2vec
// job, kkw_prepare, a program that tests adjacency).  They only speed up the
// adjacency test: when they do not fit in device memory the walks use binary
// search over the sorted lists, with the same answers, and automatic builds
// stay off until kkw_graph_adj_index(eng, 1).
inline void auto_adj_index(kkw_engine *eng)
{
    if (!eng->adj_index_auto || eng->gs.adjset) return;
    eng->graph_version++;
    const double t0 = kkw::now_ms();
    try {
        kkw::build_adj_index(eng->gs, eng->stream);
    } catch (const kkw::Error &e) {
        if (e.code != KKW_E_NOMEM) throw;
        eng->adj_index_auto = false;
        return;
    }
    eng->timing.index_ms = kkw::now_ms() - t0;
}
