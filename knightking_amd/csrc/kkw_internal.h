// kkw_internal.h -- declarations shared by the translation units of libkkw.
#pragma once

#include "kkw_dev.h"

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdexcept>
#include <string>

namespace kkw {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// A named span of host time on the calling thread (graph builds, table
// builds, walk launches and waits, path output), recorded by `rocprofv3
// --marker-trace` beside the kernels it encloses; without a profiler the two
// calls return at once.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

#define KKW_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::kkw::Error(-4, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

// Device memory of the graph store and of the graph and table builders,
// cached per device.  Freeing a large device buffer can hold the host for up
// to a second of driver time (MI355X: 0.26-1.08 s to free ~36 GiB with the
// device idle, while allocating is under a millisecond), and graph loads and
// table builds free their temporaries, and rebuilt tables their
// predecessors.  A freed block waits for the device to finish its work (as
// hipFree does) and is kept; an allocation takes the smallest kept block of
// its device that holds it without wasting more than half of it (or 1 MiB),
// else asks the runtime, and when the runtime has no memory left every kept
// block of the device goes back to it and the request is tried once more.
// Engines register with their device; when the last one goes, so do the kept
// blocks.  Thread-safe.
void *pool_alloc(size_t bytes);   // on the current device; throws Error(-3) when out of memory
void pool_free(void *p);          // null is ignored
void pool_attach(int device);
void pool_detach(int device);
size_t pool_kept_bytes(int device);
void pool_trim(int device);       // returns every kept block of the device to the runtime

// Device-resident CSR and its derived arrays.  Every pointer is a device
// allocation owned by the store (null when absent).
struct GraphStore {
    uint32_t V = 0;
    uint64_t E = 0;
    uint64_t *row_ptr = nullptr;
    uint32_t *col = nullptr;
    float *w = nullptr;
    uint8_t *etype = nullptr;
    uint8_t *vmask = nullptr;
    Bucket *bucket = nullptr;
    Bucket *ubucket = nullptr;
    Bucket *lbucket = nullptr;   // alias tables holding local edge indices (walk programs)
    float *vsum = nullptr;
    uint32_t *adjset = nullptr;
    bool sorted = false;
    bool symmetric = false;
    bool packable = false;   // every row fits the packed form
    void release();
    DevGraph view() const;
};

// graph.hip
void build_rmat(GraphStore &gs, int scale, int edge_factor, uint32_t thr_a, uint32_t thr_ab,
                uint32_t thr_abc, uint64_t seed, bool weighted, hipStream_t s);
void build_from_edges(GraphStore &gs, uint32_t V, const uint32_t *h_src, const uint32_t *h_dst,
                      const float *h_w, const uint8_t *h_et, uint64_t n, bool undirected, bool sort_adj,
                      hipStream_t s);
void sort_adjacency(GraphStore &gs, hipStream_t s);
void build_from_csr(GraphStore &gs, uint32_t V, const uint64_t *h_row_ptr, const uint32_t *h_col,
                    const float *h_w, hipStream_t s);
void set_edge_types(GraphStore &gs, const uint8_t *h_etype, hipStream_t s);
void type_by_dst(GraphStore &gs, int ntypes, uint64_t seed, hipStream_t s);
// local: buckets hold local edge indices (walk programs) instead of neighbour
// ids and rows; either way the per-vertex weight sums go to vsum
void build_alias(GraphStore &gs, hipStream_t s, bool local = false);
void build_adj_index(GraphStore &gs, hipStream_t s);
void build_uniform_buckets(GraphStore &gs, hipStream_t s);

// one job's rows inside a serving flight (packed output)
struct PackJob {
    uint64_t row0;
    uint32_t n;
    uint32_t stride;
    const uint32_t *paths;
};

// walk.hip
void launch_walk_plain(const DevGraph &g, const DevJob &J, bool stats, DevStats *st, int block,
                       hipStream_t stream);
// cu_count: the device's CUs (the kernel form depends on walkers per CU)
void launch_walk_persistent_single(const DevGraph &g, const DevJob &J, uint64_t n_chunks,
                                   uint32_t chunk, unsigned long long *queue, bool stats,
                                   DevStats *st, int grid, int block, int cu_count, hipStream_t stream);
// lines: some job's rows are a multiple of 32 entries (written as whole lines
// through the LDS line buffer the launch then allocates)
void launch_walk_persistent_mixed(const DevGraph &g, const DevJob *jobs, int n_jobs, uint64_t n_chunks,
                                  uint32_t chunk, unsigned long long *queue, bool stats, DevStats *st,
                                  int grid, int block, bool lines, hipStream_t stream);
// Serving rounds: every walker takes at most round_steps steps in one launch;
// walkers still going are suspended: written to `out` as a record of
// ROUND_REC_PLANES 16-byte words (plane p of record k at out[p * cap + k],
// so consecutive lanes touch consecutive words).  Each wave reserves slots
// ROUND_SLOTS at a time from *n_out and marks the ones it leaves unused as
// holes; *n_live counts the walkers listed and the waves ended (both zero at
// launch), and the launch's last wave copies the walkers listed to *h_live.  The first
// round (in == null) takes the n_chunks chunks of the jobs' walker ranges; a
// later one takes chunks of the *n_in slots of `in` -- a count the previous
// round left on the device, so rounds follow each other on the stream
// without the host.  A list holds at most walkers + waves x ROUND_SLOTS slots.
constexpr uint32_t ROUND_REC_PLANES = 5;
constexpr uint32_t ROUND_SLOTS = 64;
constexpr uint32_t ROUND_HOLE = 0xFFFFFFFFu;   // plane 0 .x (the job) of an unused slot
constexpr uint32_t ROUND_WAVE_SHIFT = 44;      // *n_live: ended waves above, listed walkers below
struct RoundArgs {
    uint32_t round_steps;
    uint64_t cap;                  // slots per plane
    const uint4 *in;
    const unsigned long long *n_in;
    uint4 *out;
    unsigned long long *n_out;
    unsigned long long *n_live;    // waves ended << ROUND_WAVE_SHIFT | walkers listed
    unsigned long long *h_live;    // host memory (device address): the walkers listed, once the launch ends
};
void launch_walk_rounds(const DevGraph &g, const DevJob *jobs, int n_jobs, uint64_t n_chunks, uint32_t chunk,
                        unsigned long long *queue, const RoundArgs &R, int grid, int block, hipStream_t stream);
void launch_fill_u64(unsigned long long *p, unsigned long long v, int n, hipStream_t stream);
void launch_stream_copy(const void *src, void *dst, uint64_t bytes, int cu_count, hipStream_t stream);
void launch_random_gather(const void *buf, uint64_t bytes, uint32_t threads, uint32_t reads, uint64_t seed,
                          uint32_t *sink, hipStream_t stream);
// Packed rows: counts (lengths clamped to their row), offsets (offs[0] = 0,
// then the inclusive scan), rows copied to out at their offsets.  h_lens /
// h_offs, when not null, receive the lengths and offsets too (pinned memory
// the caller reads after the stream: no copy needed).
void launch_pack(const PackJob *jobs, int n_jobs, const uint32_t *lens, uint64_t *cnt, uint64_t *offs, uint64_t n,
                 uint32_t *out, uint32_t *h_lens, uint64_t *h_offs, void *scan_tmp, size_t scan_bytes, hipStream_t s);
size_t pack_scan_bytes(uint64_t n);

}  // namespace kkw
