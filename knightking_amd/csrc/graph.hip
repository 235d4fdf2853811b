// graph.hip -- graph storage on the device: CSR construction (from an edge
// list, from a CSR, or from a synthetic R-MAT), edge types, vertex type masks
// and the static-component alias tables.
//
// CSR semantics follow KnightKing include/graph.hpp: mirroring of undirected
// input (430-443), per-source lists in file order (309-330; one partition keeps
// file order through shuffle_edges, 332-406), optional per-list sort by
// neighbour as node2vec() does (src/apps/node2vec.hpp:122-125).  Order is made
// deterministic on the device with stable radix sorts.
#include "kkw_internal.h"

#include <hipcub/hipcub.hpp>

#include <cstdlib>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace kkw {

// ------------------------------------------------------------ device pool ----
namespace {

struct Pool {
    std::mutex mu;
    std::map<int, std::multimap<size_t, void *>> kept;       // device -> (bytes -> block)
    std::unordered_map<void *, std::pair<int, size_t>> owned;  // block -> (device, bytes), kept or in use
    std::map<int, int> engines;                               // device -> registered engines
};

Pool &pool()
{
    static Pool *p = new Pool;   // never destroyed: engines may outlive static destruction order
    return *p;
}

// kept blocks of `device` back to the runtime (pool lock held)
void trim_locked(Pool &P, int device)
{
    auto it = P.kept.find(device);
    if (it == P.kept.end()) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    for (auto &kv : it->second) {
        (void)hipFree(kv.second);
        P.owned.erase(kv.second);
    }
    if (cur != device) (void)hipSetDevice(cur);
    P.kept.erase(it);
}

}  // namespace

void *pool_alloc(size_t bytes)
{
    if (bytes == 0) bytes = 1;
    int dev = 0;
    KKW_HIP(hipGetDevice(&dev));
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto &kept = P.kept[dev];
    auto it = kept.lower_bound(bytes);
    if (it != kept.end() && (it->first - bytes <= bytes || it->first - bytes <= (1u << 20))) {
        void *q = it->second;
        kept.erase(it);
        return q;
    }
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess && !kept.empty()) {
        (void)hipGetLastError();
        trim_locked(P, dev);
        e = hipMalloc(&q, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();   // a later launch check must not see this failure
        throw Error(-3, "hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    }
    P.owned[q] = {dev, bytes};
    return q;
}

void pool_free(void *p)
{
    if (!p) return;
    Pool &P = pool();
    int dev = 0;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.owned.find(p);
        if (it == P.owned.end()) {   // not the pool's: hand it back directly
            (void)hipFree(p);
            return;
        }
        dev = it->second.first;
        bytes = it->second.second;
    }
    // No queued work may still use the block when it is handed out again.  The
    // wait happens outside the lock: a thread waiting here for the device (the
    // two table builds of kkw_prepare run on two threads) must not hold up the
    // other thread's allocations; the block is not in the kept list until the
    // wait is over, so nobody can take it early.
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != dev) (void)hipSetDevice(dev);
    (void)hipDeviceSynchronize();
    if (cur != dev) (void)hipSetDevice(cur);
    std::lock_guard<std::mutex> lk(P.mu);
    P.kept[dev].emplace(bytes, p);
}

void pool_attach(int device)
{
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.engines[device]++;
}

void pool_detach(int device)
{
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    if (--P.engines[device] <= 0) {
        P.engines.erase(device);
        trim_locked(P, device);
    }
}

size_t pool_kept_bytes(int device)
{
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    size_t n = 0;
    auto it = P.kept.find(device);
    if (it != P.kept.end())
        for (auto &kv : it->second) n += kv.first;
    return n;
}

void pool_trim(int device)
{
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    trim_locked(P, device);
}

namespace {

template <typename T>
T *dalloc(size_t n)
{
    return static_cast<T *>(pool_alloc((n ? n : 1) * sizeof(T)));
}

template <typename T>
void dfree(T *&p)
{
    pool_free(p);
    p = nullptr;
}

// Device array owned by a scope: freed on every exit, exceptions included;
// release() hands the allocation over (a table is published only once built).
template <typename T>
struct DArr {
    T *p = nullptr;
    DArr() = default;
    explicit DArr(size_t n) : p(dalloc<T>(n)) {}
    ~DArr() { dfree(p); }
    DArr(const DArr &) = delete;
    DArr &operator=(const DArr &) = delete;
    T *release()
    {
        T *q = p;
        p = nullptr;
        return q;
    }
};

// A graph builder that fails part-way leaves no half-built graph behind.
template <typename F>
void building(GraphStore &gs, F &&f)
{
    try {
        f();
    } catch (...) {
        gs.release();
        throw;
    }
}

struct Temp {
    void *p = nullptr;
    size_t n = 0;
    void ensure(size_t bytes)
    {
        if (bytes <= n) return;
        pool_free(p);
        p = nullptr;
        n = 0;
        p = pool_alloc(bytes ? bytes : 1);
        n = bytes;
    }
    ~Temp() { pool_free(p); }
};

inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

int bits_for(uint64_t v)  // bits needed for values < v
{
    int b = 0;
    while (b < 64 && (1ull << b) < v) b++;
    return b ? b : 1;
}

// ---------------------------------------------------------------- R-MAT ----
// Edge i, levels 4g..4g+3 take the words of Philox(key seed, ctr (i_lo, i_hi,
// g, TAG_RMAT)); level l sets bit (scale-1-l).  Key = min << scale | max.
__global__ void k_rmat_keys(uint64_t n, int scale, uint32_t ta, uint32_t tab, uint32_t tabc,
                            uint64_t seed, uint64_t *keys)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = 0, d = 0;
    for (int g = 0; g * 4 < scale; g++) {
        u32x4 r = philox((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)g, TAG_RMAT, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
        uint32_t w4[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int l = g * 4 + j;
            if (l >= scale) break;
            uint32_t bit = 1u << (scale - 1 - l);
            uint32_t u = w4[j];
            if (u < ta) {
            } else if (u < tab) {
                d |= bit;
            } else if (u < tabc) {
                s |= bit;
            } else {
                s |= bit;
                d |= bit;
            }
        }
    }
    uint32_t a = s < d ? s : d, b = s < d ? d : s;
    keys[i] = ((uint64_t)a << scale) | b;
}

// keep the first of each run of equal keys, drop self loops (min == max)
__global__ void k_unique_flags(const uint64_t *keys, uint64_t n, int scale, uint8_t *flags)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    uint64_t mask = (1ull << scale) - 1;
    bool self = (k >> scale) == (k & mask);
    flags[i] = (!self && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
}

__global__ void k_mirror(const uint64_t *keys, uint64_t u, int scale, uint64_t *out)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= u) return;
    uint64_t k = keys[i];
    uint64_t mask = (1ull << scale) - 1;
    out[2 * i] = k;
    out[2 * i + 1] = ((k & mask) << scale) | (k >> scale);
}

// row_ptr[v] = first position whose source is >= v (keys sorted)
__global__ void k_rowptr_from_sorted(const uint64_t *keys, uint64_t n, int shift, uint32_t V,
                                     uint64_t *row_ptr)
{
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > V) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((keys[mid] >> shift) < v) lo = mid + 1;
        else hi = mid;
    }
    row_ptr[v] = lo;
}

__global__ void k_col_weight_from_keys(const uint64_t *keys, uint64_t n, int scale, uint64_t seed,
                                       uint32_t *col, float *w)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    uint64_t mask = (1ull << scale) - 1;
    uint32_t s = (uint32_t)(k >> scale), d = (uint32_t)(k & mask);
    col[i] = d;
    if (w) {
        uint32_t a = s < d ? s : d, b = s < d ? d : s;
        u32x4 r = philox(a, b, 0, TAG_WEIGHT, (uint32_t)seed, (uint32_t)(seed >> 32));
        w[i] = 1.0f + 4.0f * u01(r.x);
    }
}

// ------------------------------------------------------ edges -> CSR ----
// Entry j of the (possibly mirrored) list: j = 2i is edge i as read, j = 2i+1
// its reverse (graph.hpp:430-443).  Sort key = src (file order) or
// src << vb | dst (sorted lists); value = j, so the stable sort keeps file
// order among equal keys.
__global__ void k_entry_keys(const uint32_t *src, const uint32_t *dst, uint64_t n, int undirected,
                             int sort_adj, int vb, uint32_t V, uint64_t *keys, uint64_t *vals,
                             unsigned long long *bad)
{
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t m = undirected ? 2 * n : n;
    if (j >= m) return;
    uint64_t i = undirected ? (j >> 1) : j;
    bool rev = undirected && (j & 1);
    uint32_t s = rev ? dst[i] : src[i];
    uint32_t d = rev ? src[i] : dst[i];
    if (s >= V || d >= V) {
        atomicAdd(bad, 1ull);
        s = 0;
        d = 0;
    }
    keys[j] = sort_adj ? (((uint64_t)s << vb) | d) : (uint64_t)s;
    vals[j] = j;
}

__global__ void k_gather_entries(const uint64_t *perm, uint64_t m, const uint32_t *src,
                                 const uint32_t *dst, const float *w, const uint8_t *et, int undirected,
                                 uint32_t *col, float *wout, uint8_t *etout, uint32_t *deg_src_out)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    uint64_t j = perm[k];
    uint64_t i = undirected ? (j >> 1) : j;
    bool rev = undirected && (j & 1);
    col[k] = rev ? src[i] : dst[i];
    if (wout) wout[k] = w[i];
    if (etout) etout[k] = et[i];
    deg_src_out[k] = rev ? dst[i] : src[i];
}

// source vertex of every CSR entry
__global__ void k_expand_src(const uint64_t *row_ptr, uint32_t V, uint32_t *srcs)
{
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (uint64_t k = row_ptr[v]; k < row_ptr[v + 1]; k++) srcs[k] = (uint32_t)v;
}

__global__ void k_sort_keys_from_csr(const uint32_t *srcs, const uint32_t *col, uint64_t E, int vb,
                                     uint64_t *keys, uint64_t *vals)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    keys[k] = ((uint64_t)srcs[k] << vb) | col[k];
    vals[k] = k;
}

__global__ void k_permute_csr(const uint64_t *perm, uint64_t E, const uint32_t *col, const float *w,
                              const uint8_t *et, uint32_t *col2, float *w2, uint8_t *et2)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    uint64_t j = perm[k];
    col2[k] = col[j];
    if (w2) w2[k] = w[j];
    if (et2) et2[k] = et[j];
}

__global__ void k_rowptr_from_src(const uint32_t *srcs, uint64_t m, uint32_t V, uint64_t *row_ptr)
{
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > V) return;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (srcs[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    row_ptr[v] = lo;
}

// ------------------------------------------------------------- checks ----
__global__ void k_check_csr(const uint64_t *row_ptr, const uint32_t *col, const float *w, uint32_t V,
                            uint64_t E, unsigned long long *bad, unsigned int *sorted_flag)
{
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint64_t b = row_ptr[v], e = row_ptr[v + 1];
    if (b > e || e > E) {
        atomicAdd(bad, 1ull);
        return;
    }
    for (uint64_t k = b; k < e; k++) {
        if (col[k] >= V) atomicAdd(bad, 1ull);
        if (w && !(w[k] >= 0.0f && w[k] < 3.0e38f)) atomicAdd(bad, 1ull);
        if (k > b && col[k] < col[k - 1]) atomicAnd(sorted_flag, 0u);
    }
}

// every (v,u) has its reverse (u,v): binary search of v in u's sorted list
__global__ void k_check_symmetric(const uint64_t *row_ptr, const uint32_t *col, uint32_t V,
                                  unsigned long long *missing)
{
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (uint64_t k = row_ptr[v]; k < row_ptr[v + 1]; k++) {
        const uint32_t u = col[k];
        uint64_t lo = row_ptr[u], hi = row_ptr[u + 1];
        while (lo < hi) {
            uint64_t mid = (lo + hi) >> 1;
            if (col[mid] < (uint32_t)v) lo = mid + 1;
            else hi = mid;
        }
        if (lo == row_ptr[u + 1] || col[lo] != (uint32_t)v) {
            atomicAdd(missing, 1ull);
            return;
        }
    }
}

// Neighbour hash sets (layout in kkw_dev.h) are built in three tiers by list
// length, every table written by exactly one builder: lists of up
// to ADJ_LDS_MAX_DEG entries one wave each in LDS (k_adjset_small), lists of
// up to ADJ_MID_MAX_DEG one workgroup each in LDS (k_adjset_mid) -- each
// table's lines leave once, instead of every insert fetching and dirtying a
// random line of HBM -- and longer lists in segments of ADJ_SEG slots, one
// workgroup per segment in LDS (k_adjset_seg), the few keys whose probe runs
// past a segment's end inserted afterwards with global atomics
// (k_adjset_overflow).  Empty lists and lists of at most ADJ_TINY_MAX_DEG
// entries take one lane each (k_adjset_tiny).
constexpr uint32_t ADJ_TINY_MAX_DEG = 7;     // table of at most 4 * 7 + 4 = 32 slots
constexpr uint32_t ADJ_LDS_MAX_DEG = 255;    // table of at most 4 * 255 + 4 = 1024 slots
constexpr uint32_t ADJ_MID_MAX_DEG = 4095;   // table of at most 16384 slots (64 KiB of LDS)
constexpr uint32_t ADJ_SEG = 16384;          // slots of one LDS segment of a longer list's table
constexpr uint64_t ADJ_OVERFLOW_CAP = 1u << 20;   // keys past a segment's end kept for k_adjset_overflow

__device__ __forceinline__ void adj_insert(uint32_t *tab, uint32_t size, uint32_t u)
{
    uint32_t slot = adj_slot(u, size);
    for (uint32_t n = 0; n < size; n++) {   // deg < size entries: a free slot exists
        const uint32_t old = atomicCAS(tab + slot, ADJ_EMPTY, u);
        if (old == ADJ_EMPTY || old == u) return;
        slot = slot + 1 == size ? 0 : slot + 1;
    }
}

// Keys of long lists whose probe ran past the end of their LDS segment
// (k_adjset_seg), inserted with global atomics from their home slot on: the
// probe passes the written segments and takes the first empty slot.  Filling
// an empty slot never shortens another key's probe, so every key stays
// reachable from its home slot.
__global__ void k_adjset_overflow(const uint2 *__restrict__ ovf, const unsigned int *n_ovf, uint32_t cap,
                                  const uint64_t *__restrict__ row_ptr, uint32_t *__restrict__ t)
{
    const uint32_t n = *n_ovf < cap ? *n_ovf : cap;   // beyond cap: k_adjset_insert_long
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint2 kv = ovf[i];   // {vertex, key}
        const uint64_t b = row_ptr[kv.x];
        adj_insert(t + adj_table_start(b, kv.x), adj_table_size(row_ptr[kv.x + 1] - b), kv.y);
    }
}

// Every key of the lists longer than ADJ_MID_MAX_DEG inserted with global
// atomics, one lane per edge (source by binary search over row_ptr): the
// fallback when more keys ran past their segments than the overflow buffer
// holds.  A key already in its table is found on its probe and left alone.
__global__ void k_adjset_insert_long(const uint64_t *row_ptr, const uint32_t *col, uint32_t V, uint64_t E,
                                     uint32_t *t)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    uint32_t lo = 0, hi = V - 1;   // source: last v with row_ptr[v] <= e
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo + 1) >> 1);
        if (row_ptr[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    const uint64_t b = row_ptr[lo], deg = row_ptr[lo + 1] - b;
    if (deg <= ADJ_MID_MAX_DEG) return;
    adj_insert(t + adj_table_start(b, lo), adj_table_size(deg), col[e]);
}

// -------------------------------------------------------------- types ----
__global__ void k_vertex_types(uint32_t V, int ntypes, uint64_t seed, uint8_t *vtype)
{
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    u32x4 r = philox((uint32_t)v, 0, 0, TAG_VTYPE, (uint32_t)seed, (uint32_t)(seed >> 32));
    vtype[v] = (uint8_t)(((uint64_t)r.x * (uint32_t)ntypes) >> 32);
}

__global__ void k_etype_by_dst(const uint32_t *col, uint64_t E, const uint8_t *vtype, uint8_t *etype)
{
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    etype[k] = vtype[col[k]];
}

// vertex_masks[v] = OR of 1 << meta(e) over out-edges (src/apps/metapath.hpp:77-88).
// One wave per vertex: lanes stride the list and the lane masks meet by a
// butterfly of shuffles (one lane per vertex would leave a hub's list of 10^5
// edges to a single lane).  Types of 8 and above are counted in *bad.
__global__ void __launch_bounds__(256) k_vmask(const uint64_t *row_ptr, const uint8_t *etype, uint32_t V,
                                               uint8_t *vmask, unsigned long long *bad)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    // v is the same for all lanes of a wave: the shuffles see every lane
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); v < V; v += waves) {
        const uint64_t b = row_ptr[v], e = row_ptr[v + 1];
        uint32_t m = 0, nbad = 0;
        for (uint64_t k = b + lane; k < e; k += 64) {
            const uint32_t t = etype[k];
            nbad += t >= 8 ? 1u : 0u;
            m |= 1u << (t & 7u);
        }
        for (int off = 32; off > 0; off >>= 1) {
            m |= (uint32_t)__shfl_xor((int)m, off, 64);
            nbad += (uint32_t)__shfl_xor((int)nbad, off, 64);
        }
        if (lane == 0) {
            vmask[v] = (uint8_t)m;
            if (nbad) atomicAdd(bad, (unsigned long long)nbad);
        }
    }
}

// -------------------------------------------------------------- alias ----
// include/walk.hpp:538-652, one vertex per lane.  sum: sequential float sum
// (591-595); ave = sum / d (600); edges split in order into small (< ave) and
// large FIFO queues (601-615); each small is paired with the next large:
// bucket {p = prob[s] / ave, primary s, alias l}, prob[l] -= ave - prob[s], l
// re-queued by its new value (617-634); leftovers get p = 1 and no alias
// (635-647).  Buckets are stored in creation order.  An element sits in at most
// one queue and its mass changes only when it is popped as the large partner,
// so each queue entry carries its current mass with its index.
//
// This kernel takes the lists shorter than AW_MIN (k_alias_wave the rest).
// Such a list never has more than AS_RING - 1 live queue entries, so each
// lane keeps its two queues as rings in LDS (mass and local index apart,
// rows padded against bank conflicts) and the pairing makes no global memory
// round trip.  Buckets leave as AliasRec records (local indices), for
// k_finish_buckets.
constexpr uint32_t AS_RING = 32;

// A bucket as the alias kernels leave it: the coin's threshold p, the local
// indices of its primary and alias edges, and the vertex whose list it
// belongs to.  Records are dense (16 bytes, where the finished bucket takes
// 32): the alias kernels write and k_finish_buckets reads half the bytes that
// whole buckets cost.
struct __align__(16) AliasRec {
    float p;
    uint32_t primary, alias, v;
};
static_assert(sizeof(AliasRec) == 16, "alias record is 16 bytes");
constexpr uint32_t AW_MIN = 32;   // lists of AW_MIN or more entries: k_alias_wave
static_assert(AW_MIN <= AS_RING, "a short list's live queue entries must fit its ring");

__global__ void __launch_bounds__(64) k_alias(const uint64_t *__restrict__ row_ptr, const float *__restrict__ w,
                                              uint32_t V, AliasRec *__restrict__ out, float *__restrict__ vsum)
{
    __shared__ float sp[64][AS_RING + 1], lp[64][AS_RING + 1];   // masses of the small / large rings
    __shared__ uint8_t si[64][AS_RING + 4], li[64][AS_RING + 4];  // their local edge indices
    const uint32_t lane = threadIdx.x;
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint64_t b = row_ptr[v];
    const uint32_t d = (uint32_t)(row_ptr[v + 1] - b);
    if (d == 0) {
        vsum[v] = 0.0f;
        return;
    }
    if (d >= AW_MIN) return;   // k_alias_wave's
    const float *wv = w + b;
    float sum = 0.0f;
    for (uint32_t e = 0; e < d; e++) sum += wv[e];
    vsum[v] = sum;
    const float ave = sum / (float)d;
    constexpr uint32_t M = AS_RING - 1;
    uint32_t st = 0, lt = 0;
    for (uint32_t e = 0; e < d; e++) {
        const float p = wv[e];
        if (p < ave) {
            sp[lane][st & M] = p;
            si[lane][st & M] = (uint8_t)e;
            st++;
        } else {
            lp[lane][lt & M] = p;
            li[lane][lt & M] = (uint8_t)e;
            lt++;
        }
    }
    AliasRec *o = out + b;
    AliasRec bk;
    bk.v = (uint32_t)v;
    uint32_t sh = 0, lh = 0;
    while (sh != st) {
        const float ps = sp[lane][sh & M];
        const uint32_t se = si[lane][sh & M];
        sh++;
        bk.primary = se;
        if (lh != lt) {
            float pl = lp[lane][lh & M];
            const uint32_t le = li[lane][lh & M];
            lh++;
            bk.p = ps / ave;
            bk.alias = le;
            const float diff = ave - ps;
            pl -= diff;
            if (pl < ave) {
                sp[lane][st & M] = pl;
                si[lane][st & M] = (uint8_t)le;
                st++;
            } else {
                lp[lane][lt & M] = pl;
                li[lane][lt & M] = (uint8_t)le;
                lt++;
            }
        } else {
            bk.p = 1.0f;
            bk.alias = se;
        }
        *o++ = bk;
    }
    while (lh != lt) {
        const uint32_t le = li[lane][lh & M];
        lh++;
        bk.p = 1.0f;
        bk.primary = le;
        bk.alias = le;
        *o++ = bk;
    }
}


// ------------------------------------------------ alias, longer lists ----
// The same construction for lists of AW_MIN or more entries, one wave per
// vertex.  The float sum stays sequential (lane 0, in edge order, from weights
// staged in LDS); the split into small/large queues keeps edge order (ballot +
// prefix popcount).  The pairing runs in batches of up to 64: while the queues
// hold k entries ahead of both heads, the next k pairings take small entries
// S[sh..sh+k) and large entries L[lh..lh+k) -- each pairing pops a different
// large entry, so the k pairings are independent -- and their re-queued large
// entries go to the tails in pairing order (ballot again).  This is exactly the
// sequential FIFO run k steps at a time.  Queue windows of AW_WIN entries live
// in LDS; lists longer than that also keep their queues in global memory and
// refill the windows from there.  Lists of more than AW_POOL_MIN entries come
// first in the list and take a whole block each: its one working wave has the
// block's four windows as one of AW_POOL_WIN entries, so it refills a quarter
// as often, and these lists start before the shorter ones.  The longest list
// bounds the launch: on R-MAT-24 (largest list 406,605 entries) its wave
// spends 3.1 ms on the sequential sum, 2.2 ms queueing and 4.7 ms pairing
// (6,357 batches, 198 refills) of a 10.1-10.3 ms launch (10.8 ms with one
// wave and 512-entry windows for every list); when the table builds run
// together in kkw_prepare, the neighbour-set kernels fill the CUs this tail
// leaves idle.  The pairing works on local edge indices
// and reads global memory only to refill its windows (all loads before one
// wait), so a batch costs its LDS round trip: each bucket leaves as an
// AliasRec (local indices and the list's vertex), and k_finish_buckets then
// gives every bucket its neighbour ids, rows and edge types in one
// edge-parallel pass (finish_bucket; the tables of walk programs keep the
// indices and add the ids).
constexpr uint32_t AW_WIN = 512;
constexpr int AW_WAVES = 4;
constexpr uint32_t AW_POOL_WIN = AW_WIN * AW_WAVES;
constexpr uint64_t AW_POOL_MIN = 4096;
constexpr uint32_t AW_SUM_CHUNK = 256;
constexpr uint32_t AW_Q_CHUNKS = 16;         // chunks of 64 weights per load group when queueing

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A bucket built with local edge indices gets its edge types (meta bits 0-15)
// and, for the engine's tables, its candidates' neighbour ids in place of the
// indices, their type masks (meta bits 16-31, when the graph has edge types)
// and their packed rows (nrow, when the graph packs rows); tables of local
// indices (walk programs) keep them and take the ids in the row fields.  Done
// as each bucket is built, so every bucket is written once.
__device__ __forceinline__ void finish_bucket(Bucket &bk, uint64_t b, const uint32_t *__restrict__ col,
                                              const uint8_t *__restrict__ etype, bool to_ids,
                                              const uint64_t *__restrict__ nrow, const uint8_t *__restrict__ vmask)
{
    const uint32_t pe = bk.primary, ae = bk.alias;
    bk.meta = etype ? ((uint32_t)etype[b + pe] | ((uint32_t)etype[b + ae] << 8)) : 0u;
    if (to_ids) {
        bk.primary = col[b + pe];
        bk.alias = col[b + ae];
        if (etype && vmask) bk.meta |= ((uint32_t)vmask[bk.primary] << 16) | ((uint32_t)vmask[bk.alias] << 24);
        if (nrow) {
            bk.prow = nrow[b + pe];
            bk.arow = nrow[b + ae];
        }
    } else {
        bk.prow = col[b + pe];
        bk.arow = col[b + ae];
    }
}

// list[0 .. n_pool): the lists of more than AW_POOL_MIN entries, one block each
__global__ void __launch_bounds__(64 * AW_WAVES) k_alias_wave(const uint32_t *__restrict__ list, uint32_t n_list,
                                                            uint32_t n_pool,
                                                            const uint64_t *__restrict__ row_ptr,
                                                            const float *__restrict__ w,
                                                            AliasRec *__restrict__ out, float *__restrict__ vsum,
                                                            uint2 *__restrict__ qs, uint2 *__restrict__ ql)
{
    __shared__ uint2 win_s[AW_POOL_WIN], win_l[AW_POOL_WIN];
    __shared__ float stage[AW_WAVES][AW_SUM_CHUNK];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool pooled = blockIdx.x < n_pool;
    if (pooled && wave != 0) return;   // the block's windows serve its first wave
    const uint32_t idx = pooled ? blockIdx.x : n_pool + (blockIdx.x - n_pool) * AW_WAVES + wave;
    if (idx >= n_list) return;   // whole wave; nothing below synchronises across waves
    const uint32_t WIN = pooled ? AW_POOL_WIN : AW_WIN;
    const uint32_t v = list[idx];
    const uint64_t b = row_ptr[v];
    const uint32_t d = (uint32_t)(row_ptr[v + 1] - b);
    const float *wv = w + b;
    uint2 *S = qs + b, *L = ql + b;
    uint2 *sS = win_s + (pooled ? 0u : wave * AW_WIN), *sL = win_l + (pooled ? 0u : wave * AW_WIN);
    float *sw = stage[wave];
    AliasRec *o = out + b;
    const bool spill = d > WIN;   // queues also kept in global memory
    const uint64_t lt_mask = (1ull << lane) - 1ull;

    // 1. sequential float sum in edge order (walk.hpp:591-595)
    float nx[AW_SUM_CHUNK / 64];
#pragma unroll
    for (uint32_t j = 0; j < AW_SUM_CHUNK / 64; j++) nx[j] = j * 64 + lane < d ? wv[j * 64 + lane] : 0.0f;
    float sum = 0.0f;
    for (uint32_t base = 0; base < d; base += AW_SUM_CHUNK) {
#pragma unroll
        for (uint32_t j = 0; j < AW_SUM_CHUNK / 64; j++) sw[j * 64 + lane] = nx[j];
        wave_sync();
        const uint32_t nb = base + AW_SUM_CHUNK;
#pragma unroll
        for (uint32_t j = 0; j < AW_SUM_CHUNK / 64; j++) nx[j] = nb + j * 64 + lane < d ? wv[nb + j * 64 + lane] : 0.0f;
        if (lane == 0) {
            const uint32_t n = d - base < AW_SUM_CHUNK ? d - base : AW_SUM_CHUNK;
            for (uint32_t i = 0; i < n; i++) sum += sw[i];
        }
        wave_sync();
    }
    sum = __shfl(sum, 0);
    if (lane == 0) vsum[v] = sum;
    const float ave = sum / (float)d;

    // 2. small / large queues in edge order (walk.hpp:601-615); the weights
    // of AW_Q_CHUNKS chunks of 64 are loaded together (one wait per group)
    uint32_t st = 0, lt = 0;
    for (uint32_t gbase = 0; gbase < d; gbase += 64 * AW_Q_CHUNKS) {
        float pc[AW_Q_CHUNKS];
#pragma unroll
        for (uint32_t j = 0; j < AW_Q_CHUNKS; j++) {
            const uint32_t e = gbase + j * 64 + lane;
            pc[j] = e < d ? wv[e] : 0.0f;
        }
#pragma unroll
        for (uint32_t j = 0; j < AW_Q_CHUNKS; j++) {   // chunks past the end queue nothing
            const uint32_t base = gbase + j * 64;
            const uint32_t e = base + lane;
            const bool valid = e < d;
            const float p = pc[j];
            const bool small = valid && p < ave, large = valid && !(p < ave);
            const uint64_t ms = __ballot(small), ml = __ballot(large);
            const uint2 ent = make_uint2(e, __float_as_uint(p));
            if (small) {
                const uint32_t pos = st + (uint32_t)__popcll(ms & lt_mask);
                if (spill) S[pos] = ent;
                if (pos < WIN) sS[pos] = ent;
            }
            if (large) {
                const uint32_t pos = lt + (uint32_t)__popcll(ml & lt_mask);
                if (spill) L[pos] = ent;
                if (pos < WIN) sL[pos] = ent;
            }
            st += (uint32_t)__popcll(ms);
            lt += (uint32_t)__popcll(ml);
        }
    }
    wave_sync();

    // 3. FIFO pairing, k steps at a time (walk.hpp:617-634)
    uint32_t sh = 0, lh = 0, ws = 0, wl = 0, ob = 0;
    while (sh != st) {
        const bool paired = lh != lt;
        const uint32_t av_s = (st < ws + WIN ? st : ws + WIN) - sh;
        const uint32_t av_l = paired ? (lt < wl + WIN ? lt : wl + WIN) - lh : 64u;
        if ((av_s < 64 && st > ws + WIN) || (av_l < 64 && paired && lt > wl + WIN)) {
            // refill both windows from the global queues at the current heads
            // (which hold every entry of a list that spills), AW_WIN entries
            // at a time: all loads first, then the LDS writes (one wait each)
            for (uint32_t c = 0; c < WIN; c += AW_WIN) {
                uint2 rs[AW_WIN / 64], rl[AW_WIN / 64];
#pragma unroll
                for (uint32_t j = 0; j < AW_WIN / 64; j++) {
                    const uint32_t i = c + j * 64 + lane;
                    rs[j] = sh + i < st ? S[sh + i] : make_uint2(0, 0);
                    rl[j] = lh + i < lt ? L[lh + i] : make_uint2(0, 0);
                }
#pragma unroll
                for (uint32_t j = 0; j < AW_WIN / 64; j++) {
                    const uint32_t i = c + j * 64 + lane;
                    if (sh + i < st) sS[i] = rs[j];
                    if (lh + i < lt) sL[i] = rl[j];
                }
            }
            ws = sh;
            wl = lh;
            wave_sync();
            continue;
        }
        uint32_t k = av_s < av_l ? av_s : av_l;
        k = k < 64 ? k : 64;
        const bool act = lane < k;
        AliasRec bk;
        bk.v = v;
        bool to_s = false, to_l = false;
        uint2 ne = make_uint2(0, 0);
        if (act) {
            const uint2 se = sS[sh - ws + lane];
            const float ps = __uint_as_float(se.y);
            bk.primary = se.x;
            if (paired) {
                const uint2 le = sL[lh - wl + lane];
                float pl = __uint_as_float(le.y);
                bk.p = ps / ave;
                bk.alias = le.x;
                const float diff = ave - ps;
                pl -= diff;
                ne = make_uint2(le.x, __float_as_uint(pl));
                to_s = pl < ave;
                to_l = !to_s;
            } else {
                bk.p = 1.0f;
                bk.alias = se.x;
            }
            o[ob + lane] = bk;
        }
        const uint64_t ms = __ballot(to_s), ml = __ballot(to_l);
        if (to_s) {
            const uint32_t pos = st + (uint32_t)__popcll(ms & lt_mask);
            if (spill) S[pos] = ne;
            if (pos < ws + WIN) sS[pos - ws] = ne;
        }
        if (to_l) {
            const uint32_t pos = lt + (uint32_t)__popcll(ml & lt_mask);
            if (spill) L[pos] = ne;
            if (pos < wl + WIN) sL[pos - wl] = ne;
        }
        st += (uint32_t)__popcll(ms);
        lt += (uint32_t)__popcll(ml);
        sh += k;
        if (paired) lh += k;
        ob += k;
        wave_sync();
    }
    // 4. leftover large entries (walk.hpp:635-647), in queue order
    for (uint32_t i = lane; lh + i < lt; i += 64) {
        const uint32_t pos = lh + i;
        const uint32_t e = pos < wl + WIN ? sL[pos - wl].x : L[pos].x;
        AliasRec bk;
        bk.p = 1.0f;
        bk.primary = e;
        bk.alias = e;
        bk.v = v;
        o[ob + i] = bk;
    }
}

typedef uint32_t rows_u32x4 __attribute__((ext_vector_type(4), aligned(4)));

// Every bucket as the alias kernels left it (an AliasRec: local indices and
// the list's vertex) gets its neighbour ids, rows and edge types
// (finish_bucket), one lane per bucket.
__global__ void k_finish_buckets(const AliasRec *__restrict__ rec, Bucket *__restrict__ out, uint64_t E,
                                 const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col,
                                 const uint8_t *__restrict__ etype, bool to_ids, const uint64_t *__restrict__ nrow,
                                 const uint8_t *__restrict__ vmask)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= E) return;
    const AliasRec r = rec[k];
    Bucket bk;
    bk.p = r.p;
    bk.primary = r.primary;
    bk.alias = r.alias;
    bk.prow = 0;
    bk.arow = 0;
    finish_bucket(bk, row_ptr[r.v], col, etype, to_ids, nrow, vmask);
    out[k] = bk;
}

// Neighbour hash sets of empty lists and lists of at most ADJ_TINY_MAX_DEG
// entries: one lane per vertex.  The lane inserts its keys in list order into
// its own slice of LDS (linear probing from each key's home slot, as the
// probes expect; no other lane writes there) and writes the table out in
// 16-byte groups.  Consecutive vertices' tables are consecutive in memory, so
// a wave's row reads and table writes are nearly contiguous.  (With one wave
// per vertex, each such list was a chain of dependent row, key and LDS round
// trips: R-MAT-24's 15 M lists under 32 entries, 6.8 M of them empty, held
// that kernel at 3.9 ms.)
__global__ void __launch_bounds__(256) k_adjset_tiny(const uint64_t *__restrict__ row_ptr,
                                                     const uint32_t *__restrict__ col, uint32_t V,
                                                     uint32_t *__restrict__ t)
{
    constexpr uint32_t SLICE = 4 * ADJ_TINY_MAX_DEG + 4 + 1;   // odd stride: lanes hit different banks
    __shared__ uint32_t tabs[256 * SLICE];
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint64_t b = row_ptr[v], e = row_ptr[v + 1];
    uint4 *gt = reinterpret_cast<uint4 *>(t + adj_table_start(b, v));
    if (e == b) {   // an empty list's table: four empty slots
        gt[0] = make_uint4(ADJ_EMPTY, ADJ_EMPTY, ADJ_EMPTY, ADJ_EMPTY);
        return;
    }
    if (e - b > ADJ_TINY_MAX_DEG) return;
    const uint32_t d = (uint32_t)(e - b), size = adj_table_size(d);
    uint32_t keys[ADJ_TINY_MAX_DEG];
#pragma unroll
    for (uint32_t k = 0; k < ADJ_TINY_MAX_DEG; k++) keys[k] = k < d ? col[b + k] : 0u;
    uint32_t *lt = tabs + threadIdx.x * SLICE;
    for (uint32_t i = 0; i < size; i++) lt[i] = ADJ_EMPTY;
#pragma unroll
    for (uint32_t k = 0; k < ADJ_TINY_MAX_DEG; k++) {
        if (k < d) {   // d < size: the probe finds a free slot
            uint32_t slot = adj_slot(keys[k], size);
            while (lt[slot] != ADJ_EMPTY && lt[slot] != keys[k]) slot = slot + 1 == size ? 0 : slot + 1;
            lt[slot] = keys[k];
        }
    }
    for (uint32_t g = 0; g < size / 4; g++) gt[g] = make_uint4(lt[4 * g], lt[4 * g + 1], lt[4 * g + 2], lt[4 * g + 3]);
}

// Neighbour hash sets of lists with ADJ_TINY_MAX_DEG + 1 .. ADJ_LDS_MAX_DEG
// entries (list ids from k_list_range, *n of them): one wave per list, lists
// strided over the grid's waves, builds the table in LDS with LDS atomics and
// writes it out whole.
__global__ void __launch_bounds__(256) k_adjset_small(const uint32_t *__restrict__ list, const unsigned int *n,
                                                      const uint64_t *__restrict__ row_ptr,
                                                      const uint32_t *__restrict__ col, uint32_t *__restrict__ t)
{
    __shared__ uint32_t tabs[4][4 * ADJ_LDS_MAX_DEG + 4];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t *lt = tabs[wave];
    const uint32_t count = *n;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + wave; i < count; i += waves) {   // wave-uniform
        const uint32_t v = list[i];
        const uint64_t b = row_ptr[v], e = row_ptr[v + 1];
        const uint32_t size = adj_table_size(e - b);
        for (uint32_t j = lane; j < size; j += 64) lt[j] = ADJ_EMPTY;
        wave_sync();
        for (uint64_t k = b + lane; k < e; k += 64) adj_insert(lt, size, col[k]);
        wave_sync();
        uint32_t *gt = t + adj_table_start(b, v);
        for (uint32_t j = lane; j < size; j += 64) gt[j] = lt[j];
        wave_sync();   // the next list reuses the LDS table
    }
}

// Lists of ADJ_LDS_MAX_DEG + 1 .. MAX_DEG entries: one workgroup per list
// (list ids from k_list_range, *n of them), the table built with LDS atomics
// and written out in 16-byte groups (tables start 16-byte aligned and hold a
// multiple of four slots).  A 128 KiB tier (lists up to 8191 entries, 1024
// lanes) cut the atomics kernel from 10.2 to 8.5 ms but made every later
// rebuild in the process take 0.2-0.5 s of host time, so tables stop at 64 KiB.
template <uint32_t MAX_DEG, uint32_t BLOCK>
__global__ void __launch_bounds__(BLOCK) k_adjset_mid(const uint32_t *__restrict__ list, const unsigned int *n,
                                                      const uint64_t *__restrict__ row_ptr,
                                                      const uint32_t *__restrict__ col, uint32_t *__restrict__ t)
{
    __shared__ uint4 tab4[(4 * MAX_DEG + 4) / 4];
    uint32_t *tab = reinterpret_cast<uint32_t *>(tab4);
    const uint32_t count = *n;
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {   // workgroup-uniform
        const uint32_t v = list[i];
        const uint64_t b = row_ptr[v], e = row_ptr[v + 1];
        const uint32_t size = adj_table_size(e - b);
        for (uint32_t k = threadIdx.x; k < size; k += blockDim.x) tab[k] = ADJ_EMPTY;
        __syncthreads();
        for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x) adj_insert(tab, size, col[k]);
        __syncthreads();
        uint4 *gt = reinterpret_cast<uint4 *>(t + adj_table_start(b, v));
        for (uint32_t k = threadIdx.x; k < size / 4; k += blockDim.x) gt[k] = tab4[k];
        __syncthreads();   // the next list reuses the LDS table
    }
}

// Tables of lists longer than ADJ_MID_MAX_DEG, one workgroup per (list,
// segment) item: the workgroup reads the whole list, inserts the keys whose
// home slot lies in its segment with LDS atomics (linear probing inside the
// segment) and writes the segment out whole in 16-byte groups (segments start
// at multiples of ADJ_SEG slots of 16-byte aligned tables, and table sizes
// are multiples of four).  A key whose probe runs past the segment's end goes
// to ovf (cap entries) for k_adjset_overflow.
__global__ void __launch_bounds__(256) k_adjset_seg(const uint2 *__restrict__ items,
                                                    const uint32_t *__restrict__ n_items_d,
                                                    const uint32_t *__restrict__ list,
                                                    const uint64_t *__restrict__ row_ptr,
                                                    const uint32_t *__restrict__ col, uint32_t *__restrict__ t,
                                                    uint2 *__restrict__ ovf, uint32_t cap, unsigned int *n_ovf)
{
    __shared__ uint4 tab4[ADJ_SEG / 4];
    uint32_t *tab = reinterpret_cast<uint32_t *>(tab4);
    const uint32_t n_items = *n_items_d;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {   // workgroup-uniform
        const uint2 item = items[it];   // {index into list, segment}
        const uint32_t v = list[item.x];
        const uint64_t b = row_ptr[v], e = row_ptr[v + 1];
        const uint32_t size = adj_table_size(e - b);
        const uint32_t s0 = item.y * ADJ_SEG;
        const uint32_t w = size - s0 < ADJ_SEG ? size - s0 : ADJ_SEG;
        for (uint32_t k = threadIdx.x; k < w; k += blockDim.x) tab[k] = ADJ_EMPTY;
        __syncthreads();
        for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
            const uint32_t u = col[k];
            const uint32_t h = adj_slot(u, size);
            if (h < s0 || h - s0 >= w) continue;   // homed in another segment
            for (uint32_t slot = h - s0;;) {
                const uint32_t old = atomicCAS(tab + slot, ADJ_EMPTY, u);
                if (old == ADJ_EMPTY || old == u) break;
                if (++slot == w) {   // counted always, kept while there is room
                    const unsigned int pos = atomicAdd(n_ovf, 1u);
                    if (pos < cap) ovf[pos] = make_uint2(v, u);
                    break;
                }
            }
        }
        __syncthreads();
        uint4 *gt = reinterpret_cast<uint4 *>(t + adj_table_start(b, v) + s0);
        for (uint32_t k = threadIdx.x; k < w / 4; k += blockDim.x) gt[k] = tab4[k];
        __syncthreads();   // the next item reuses the LDS segment
    }
}

// Work items of k_adjset_seg, made on the device: segs[i] = the number of
// ADJ_SEG-slot segments of list[i]'s table (segs[n] = 0, for the scan), then,
// with seg_off its exclusive prefix sum, item (i, g) for every segment g.  (A
// host loop over the degrees had the first build of a process wait 5.4 ms for
// its first pageable device-to-host copy.)
__global__ void k_list_segments(const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ row_ptr,
                                uint32_t *__restrict__ segs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint64_t d = row_ptr[list[i] + 1] - row_ptr[list[i]];
        segs[i] = (uint32_t)(((uint64_t)adj_table_size(d) + ADJ_SEG - 1) / ADJ_SEG);
    } else if (i == n) {
        segs[n] = 0;
    }
}

__global__ void k_segment_items(const uint32_t *__restrict__ segs, const uint32_t *__restrict__ seg_off, uint32_t n,
                                uint2 *__restrict__ items)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        for (uint32_t g = 0; g < segs[i]; g++) items[seg_off[i] + g] = make_uint2(i, g);
}

// ids of the vertices whose lists hold lo .. hi entries, in no particular order.
// One atomic per workgroup of LIST_BLOCK lanes (same-address atomics serialise
// in L2): picks are counted by wave ballots, the waves' counts offset in LDS.
constexpr uint32_t LIST_BLOCK = 1024;

__global__ void __launch_bounds__(LIST_BLOCK) k_list_range(const uint64_t *row_ptr, uint32_t V, uint64_t lo,
                                                           uint64_t hi, uint32_t *out, unsigned int *n)
{
    __shared__ unsigned int wave_count[LIST_BLOCK / 64], block_base;
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    if (v < V) {
        const uint64_t d = row_ptr[v + 1] - row_ptr[v];
        take = d >= lo && d <= hi;
    }
    const uint64_t m = __ballot(take);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_count[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int total = 0;
        for (uint32_t k = 0; k < blockDim.x / 64; k++) {
            const unsigned int c = wave_count[k];
            wave_count[k] = total;   // now the wave's offset in the block
            total += c;
        }
        block_base = total ? atomicAdd(n, total) : 0u;
    }
    __syncthreads();
    if (take) out[block_base + wave_count[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)v;
}

// Packed row of every edge's neighbour, nrow[k] = row of col[k], for the
// alias kernels to copy into buckets (each bucket's candidates are neighbours
// of its list: E random row reads instead of two per bucket).
__device__ __forceinline__ uint64_t row_of(const uint64_t *row_ptr, uint32_t v)
{
    const rows_u32x4 q = *reinterpret_cast<const rows_u32x4 *>(row_ptr + v);   // row_ptr[v], row_ptr[v + 1]
    return pack_row((uint64_t)q.x | ((uint64_t)q.y << 32), (uint64_t)q.z | ((uint64_t)q.w << 32));
}

__global__ void k_neighbour_rows(const uint32_t *__restrict__ col, uint64_t E, const uint64_t *__restrict__ row_ptr,
                                 uint64_t *__restrict__ nrow)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < E) nrow[k] = row_of(row_ptr, col[k]);
}

// uniform candidates as buckets: p = 1, alias = primary = col[e]
__global__ void k_uniform_buckets(const uint32_t *col, const uint8_t *etype, const uint8_t *vmask,
                                  const uint64_t *row_ptr, uint64_t E, int rows, Bucket *out)
{
    uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t u = col[e];
    Bucket b;
    b.p = 1.0f;
    b.primary = u;
    b.alias = u;
    b.meta = etype ? (uint32_t)etype[e] * 0x101u : 0u;
    if (etype && vmask) b.meta |= (uint32_t)vmask[u] * 0x01010000u;   // the candidate's type mask, twice
    b.prow = rows ? row_of(row_ptr, u) : 0;
    b.arow = b.prow;
    out[e] = b;
}

// Maximum degree: a bounded grid strides over the vertices, each block
// reduces with shuffles and LDS, and only block maxima meet in one atomic
// (same-address atomics serialise in L2).  Launched with 256 threads a block.
__global__ void __launch_bounds__(256) k_max_degree(const uint64_t *row_ptr, uint32_t V, unsigned long long *out)
{
    __shared__ unsigned long long wmax[4];
    unsigned long long d = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long x = row_ptr[v + 1] - row_ptr[v];
        d = x > d ? x : d;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(d, off, 64);
        d = o > d ? o : d;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmax[0];
        for (int k = 1; k < 4; k++) m = wmax[k] > m ? wmax[k] : m;
        if (m) atomicMax(out, m);
    }
}

unsigned long long read_counter(unsigned long long *d, hipStream_t s)
{
    unsigned long long h = 0;
    KKW_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    KKW_HIP(hipStreamSynchronize(s));
    return h;
}

void check_graph(GraphStore &gs, hipStream_t s)
{
    DArr<unsigned long long> bad(1);
    DArr<unsigned int> sorted(1);
    KKW_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), s));
    unsigned int one = 1;
    KKW_HIP(hipMemcpyAsync(sorted.p, &one, sizeof(one), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_check_csr, dim3(grid_for(gs.V, 256)), dim3(256), 0, s, gs.row_ptr, gs.col,
                       gs.w, gs.V, gs.E, bad.p, sorted.p);
    KKW_HIP(hipGetLastError());
    unsigned long long nbad = read_counter(bad.p, s);
    unsigned int hs = 0;
    KKW_HIP(hipMemcpy(&hs, sorted.p, sizeof(hs), hipMemcpyDeviceToHost));
    if (nbad) throw Error(-1, "graph has " + std::to_string(nbad) +
                                  " invalid entries (vertex id >= v_num, bad row offsets, or a "
                                  "weight that is negative or not finite)");
    gs.sorted = hs != 0;
}

// rows pack into 40 + 24 bits when E < 2^40 and every degree < 2^24
void compute_packable(GraphStore &gs, hipStream_t s)
{
    gs.packable = false;
    if (gs.V == 0 || gs.E >= (1ull << ROW_DEG_SHIFT)) return;
    DArr<unsigned long long> mx(1);
    KKW_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_max_degree, dim3(std::min<uint64_t>(grid_for(gs.V, 256), 2048)), dim3(256), 0, s, gs.row_ptr,
                       gs.V, mx.p);
    KKW_HIP(hipGetLastError());
    gs.packable = read_counter(mx.p, s) < (1ull << (64 - ROW_DEG_SHIFT));
}

void detect_symmetric(GraphStore &gs, hipStream_t s)
{
    gs.symmetric = false;
    if (!gs.sorted || gs.V == 0) return;
    DArr<unsigned long long> miss(1);
    KKW_HIP(hipMemsetAsync(miss.p, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_check_symmetric, dim3(grid_for(gs.V, 256)), dim3(256), 0, s, gs.row_ptr, gs.col, gs.V,
                       miss.p);
    KKW_HIP(hipGetLastError());
    gs.symmetric = read_counter(miss.p, s) == 0;
}

}  // namespace

static void rebuild_vmask(GraphStore &gs, hipStream_t s);

void GraphStore::release()
{
    dfree(row_ptr);
    dfree(col);
    dfree(w);
    dfree(etype);
    dfree(vmask);
    dfree(bucket);
    dfree(ubucket);
    dfree(lbucket);
    dfree(vsum);
    dfree(adjset);
    V = 0;
    E = 0;
    sorted = false;
    symmetric = false;
    packable = false;
}

DevGraph GraphStore::view() const
{
    DevGraph g;
    g.V = V;
    g.E = E;
    g.row_ptr = row_ptr;
    g.col = col;
    g.w = w;
    g.etype = etype;
    g.vmask = vmask;
    g.bucket = bucket;
    g.ubucket = ubucket;
    g.packed_rows = packable ? 1u : 0u;
    g.vsum = vsum;
    g.symmetric = symmetric ? 1u : 0u;
    g.adjset = adjset;
    return g;
}

void build_rmat(GraphStore &gs, int scale, int edge_factor, uint32_t thr_a, uint32_t thr_ab,
                uint32_t thr_abc, uint64_t seed, bool weighted, hipStream_t s)
{
    if (scale < 1 || scale > 30 || edge_factor < 1)
        throw Error(-1, "rmat: scale must be in [1,30] and edge_factor >= 1");
    TraceRange tr("graph_rmat");
    gs.release();
    building(gs, [&] {
        const uint32_t V = 1u << scale;
        const uint64_t n = (uint64_t)edge_factor << scale;
        const int kb = 2 * scale;
        DArr<uint64_t> k0(n);
        Temp tmp;
        uint64_t u = 0;
        {
            DArr<uint64_t> k1(n);
            hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, n, scale, thr_a, thr_ab,
                               thr_abc, seed, k0.p);
            KKW_HIP(hipGetLastError());
            size_t tb = 0;
            KKW_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, k0.p, k1.p, n, 0, kb, s));
            tmp.ensure(tb);
            KKW_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, k0.p, k1.p, n, 0, kb, s));
            DArr<uint8_t> flags(n);
            hipLaunchKernelGGL(k_unique_flags, dim3(grid_for(n, 256)), dim3(256), 0, s, k1.p, n, scale, flags.p);
            KKW_HIP(hipGetLastError());
            DArr<unsigned long long> nsel(1);
            tb = 0;
            KKW_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, k1.p, flags.p, k0.p, nsel.p, (int64_t)n, s));
            tmp.ensure(tb);
            KKW_HIP(hipcub::DeviceSelect::Flagged(tmp.p, tb, k1.p, flags.p, k0.p, nsel.p, (int64_t)n, s));
            u = read_counter(nsel.p, s);
        }
        const uint64_t m = 2 * u;
        DArr<uint64_t> m1(m);
        {
            DArr<uint64_t> m0(m);
            hipLaunchKernelGGL(k_mirror, dim3(grid_for(u, 256)), dim3(256), 0, s, k0.p, u, scale, m0.p);
            KKW_HIP(hipGetLastError());
            size_t tb = 0;
            KKW_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, m0.p, m1.p, m, 0, kb, s));
            tmp.ensure(tb);
            KKW_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, m0.p, m1.p, m, 0, kb, s));
            KKW_HIP(hipStreamSynchronize(s));
        }
        dfree(k0.p);
        gs.V = V;
        gs.E = m;
        gs.row_ptr = dalloc<uint64_t>((uint64_t)V + 1);
        gs.col = dalloc<uint32_t>(m + 4);
        gs.w = weighted ? dalloc<float>(m) : nullptr;
        hipLaunchKernelGGL(k_rowptr_from_sorted, dim3(grid_for((uint64_t)V + 1, 256)), dim3(256), 0, s, m1.p, m,
                           scale, V, gs.row_ptr);
        hipLaunchKernelGGL(k_col_weight_from_keys, dim3(grid_for(m, 256)), dim3(256), 0, s, m1.p, m, scale, seed,
                           gs.col, gs.w);
        KKW_HIP(hipGetLastError());
        KKW_HIP(hipStreamSynchronize(s));
        gs.sorted = true;
        gs.symmetric = true;   // every edge was mirrored
        compute_packable(gs, s);
    });
}

void build_from_edges(GraphStore &gs, uint32_t V, const uint32_t *h_src, const uint32_t *h_dst,
                      const float *h_w, const uint8_t *h_et, uint64_t n, bool undirected, bool sort_adj,
                      hipStream_t s)
{
    if (V == 0) throw Error(-1, "v_num must be positive");
    TraceRange tr("graph_from_edges");
    gs.release();
    building(gs, [&] {
        const uint64_t m = undirected ? 2 * n : n;
        DArr<uint32_t> src(n), dst(n);
        DArr<float> w;
        DArr<uint8_t> et;
        if (h_w) w.p = dalloc<float>(n);
        if (h_et) et.p = dalloc<uint8_t>(n);
        KKW_HIP(hipMemcpyAsync(src.p, h_src, n * 4, hipMemcpyHostToDevice, s));
        KKW_HIP(hipMemcpyAsync(dst.p, h_dst, n * 4, hipMemcpyHostToDevice, s));
        if (w.p) KKW_HIP(hipMemcpyAsync(w.p, h_w, n * 4, hipMemcpyHostToDevice, s));
        if (et.p) KKW_HIP(hipMemcpyAsync(et.p, h_et, n, hipMemcpyHostToDevice, s));
        const int vb = bits_for(V);
        DArr<uint64_t> v1(m);
        {
            DArr<uint64_t> k0(m), k1(m), v0(m);
            DArr<unsigned long long> bad(1);
            KKW_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), s));
            if (m) {
                hipLaunchKernelGGL(k_entry_keys, dim3(grid_for(m, 256)), dim3(256), 0, s, src.p, dst.p, n,
                                   undirected ? 1 : 0, sort_adj ? 1 : 0, vb, V, k0.p, v0.p, bad.p);
                KKW_HIP(hipGetLastError());
            }
            if (read_counter(bad.p, s)) throw Error(-1, "edge list holds a vertex id >= v_num");
            Temp tmp;
            if (m) {
                size_t tb = 0;
                const int eb = sort_adj ? 2 * vb : vb;
                KKW_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0.p, k1.p, v0.p, v1.p, m, 0, eb, s));
                tmp.ensure(tb);
                KKW_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.p, k1.p, v0.p, v1.p, m, 0, eb, s));
            }
            KKW_HIP(hipStreamSynchronize(s));
        }
        gs.V = V;
        gs.E = m;
        gs.row_ptr = dalloc<uint64_t>((uint64_t)V + 1);
        gs.col = dalloc<uint32_t>(m + 4);
        gs.w = w.p ? dalloc<float>(m) : nullptr;
        gs.etype = et.p ? dalloc<uint8_t>(m) : nullptr;
        DArr<uint32_t> srcs(m);
        if (m) {
            hipLaunchKernelGGL(k_gather_entries, dim3(grid_for(m, 256)), dim3(256), 0, s, v1.p, m, src.p, dst.p,
                               w.p, et.p, undirected ? 1 : 0, gs.col, gs.w, gs.etype, srcs.p);
            KKW_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_rowptr_from_src, dim3(grid_for((uint64_t)V + 1, 256)), dim3(256), 0, s, srcs.p, m,
                           V, gs.row_ptr);
        KKW_HIP(hipGetLastError());
        KKW_HIP(hipStreamSynchronize(s));
        check_graph(gs, s);
        if (undirected) gs.symmetric = true;   // every edge was mirrored
        else detect_symmetric(gs, s);
        compute_packable(gs, s);
        if (gs.etype) rebuild_vmask(gs, s);
    });
}

void build_from_csr(GraphStore &gs, uint32_t V, const uint64_t *h_row_ptr, const uint32_t *h_col,
                    const float *h_w, hipStream_t s)
{
    if (V == 0) throw Error(-1, "v_num must be positive");
    if (h_row_ptr[0] != 0) throw Error(-1, "row_ptr[0] must be 0");
    TraceRange tr("graph_from_csr");
    gs.release();
    building(gs, [&] {
        const uint64_t E = h_row_ptr[V];
        gs.V = V;
        gs.E = E;
        gs.row_ptr = dalloc<uint64_t>((uint64_t)V + 1);
        gs.col = dalloc<uint32_t>(E + 4);
        gs.w = h_w ? dalloc<float>(E) : nullptr;
        KKW_HIP(hipMemcpyAsync(gs.row_ptr, h_row_ptr, ((uint64_t)V + 1) * 8, hipMemcpyHostToDevice, s));
        if (E) KKW_HIP(hipMemcpyAsync(gs.col, h_col, E * 4, hipMemcpyHostToDevice, s));
        if (gs.w && E) KKW_HIP(hipMemcpyAsync(gs.w, h_w, E * 4, hipMemcpyHostToDevice, s));
        KKW_HIP(hipStreamSynchronize(s));
        check_graph(gs, s);
        detect_symmetric(gs, s);
        compute_packable(gs, s);
    });
}

static void rebuild_vmask(GraphStore &gs, hipStream_t s)
{
    dfree(gs.vmask);
    DArr<uint8_t> vm(gs.V);
    DArr<unsigned long long> bad(1);
    KKW_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), s));
    const uint32_t vm_blocks = (uint32_t)std::min<uint64_t>(((uint64_t)gs.V + 3) / 4, 1u << 16);   // 4 waves each
    hipLaunchKernelGGL(k_vmask, dim3(std::max(vm_blocks, 1u)), dim3(256), 0, s, gs.row_ptr, gs.etype, gs.V,
                       vm.p, bad.p);
    KKW_HIP(hipGetLastError());
    if (read_counter(bad.p, s))
        throw Error(-1, "edge types must be < 8 (one bit each in a u8 mask, as metapath.hpp:30)");
    gs.vmask = vm.release();
}

// tables whose entries carry edge types (bucket meta) or depend on the list
// order: rebuilt on next use
static void drop_derived(GraphStore &gs)
{
    dfree(gs.bucket);
    dfree(gs.ubucket);
    dfree(gs.lbucket);
    dfree(gs.vsum);
}

void set_edge_types(GraphStore &gs, const uint8_t *h_etype, hipStream_t s)
{
    if (!gs.row_ptr) throw Error(-2, "no graph loaded");
    DArr<uint8_t> et(gs.E);
    if (gs.E) KKW_HIP(hipMemcpyAsync(et.p, h_etype, gs.E, hipMemcpyHostToDevice, s));
    KKW_HIP(hipStreamSynchronize(s));
    uint8_t *old = gs.etype;
    gs.etype = et.p;
    try {
        rebuild_vmask(gs, s);
    } catch (...) {
        gs.etype = old;   // keep the previous types (and masks) on bad input
        if (old) rebuild_vmask(gs, s);
        throw;
    }
    et.release();
    dfree(old);
    drop_derived(gs);
}

void type_by_dst(GraphStore &gs, int ntypes, uint64_t seed, hipStream_t s)
{
    if (!gs.row_ptr) throw Error(-2, "no graph loaded");
    if (ntypes < 1 || ntypes > 8) throw Error(-1, "ntypes must be in [1,8]");
    DArr<uint8_t> vt(gs.V), et(gs.E);
    hipLaunchKernelGGL(k_vertex_types, dim3(grid_for(gs.V, 256)), dim3(256), 0, s, gs.V, ntypes, seed, vt.p);
    if (gs.E)
        hipLaunchKernelGGL(k_etype_by_dst, dim3(grid_for(gs.E, 256)), dim3(256), 0, s, gs.col, gs.E, vt.p,
                           et.p);
    KKW_HIP(hipGetLastError());
    KKW_HIP(hipStreamSynchronize(s));
    dfree(gs.etype);
    gs.etype = et.release();
    rebuild_vmask(gs, s);   // types < 8 by construction
    drop_derived(gs);
}

// Per-list stable sort by neighbour id, in place (node2vec.hpp:122-125 sorts
// every list before walking).  Derived tables are dropped.
void sort_adjacency(GraphStore &gs, hipStream_t s)
{
    if (!gs.row_ptr) throw Error(-2, "no graph loaded");
    if (gs.sorted) return;
    TraceRange tr("sort_adjacency");
    const uint64_t E = gs.E;
    if (E == 0) {
        gs.sorted = true;
        return;
    }
    const int vb = bits_for(gs.V);
    DArr<uint64_t> v1(E);
    {
        DArr<uint32_t> srcs(E);
        DArr<uint64_t> k0(E), k1(E), v0(E);
        hipLaunchKernelGGL(k_expand_src, dim3(grid_for(gs.V, 256)), dim3(256), 0, s, gs.row_ptr, gs.V, srcs.p);
        hipLaunchKernelGGL(k_sort_keys_from_csr, dim3(grid_for(E, 256)), dim3(256), 0, s, srcs.p, gs.col, E, vb,
                           k0.p, v0.p);
        KKW_HIP(hipGetLastError());
        Temp tmp;
        size_t tb = 0;
        KKW_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0.p, k1.p, v0.p, v1.p, E, 0, 2 * vb, s));
        tmp.ensure(tb);
        KKW_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.p, k1.p, v0.p, v1.p, E, 0, 2 * vb, s));
        KKW_HIP(hipStreamSynchronize(s));
    }
    DArr<uint32_t> col2(E + 4);
    DArr<float> w2;
    DArr<uint8_t> et2;
    if (gs.w) w2.p = dalloc<float>(E);
    if (gs.etype) et2.p = dalloc<uint8_t>(E);
    hipLaunchKernelGGL(k_permute_csr, dim3(grid_for(E, 256)), dim3(256), 0, s, v1.p, E, gs.col, gs.w, gs.etype,
                       col2.p, w2.p, et2.p);
    KKW_HIP(hipGetLastError());
    KKW_HIP(hipStreamSynchronize(s));
    // the permuted lists replace the old ones only once complete
    dfree(gs.col);
    gs.col = col2.release();
    if (w2.p) {
        dfree(gs.w);
        gs.w = w2.release();
    }
    if (et2.p) {
        dfree(gs.etype);
        gs.etype = et2.release();
    }
    drop_derived(gs);   // the hash sets hold sets: sorting leaves them valid
    gs.sorted = true;
    detect_symmetric(gs, s);
}

void build_adj_index(GraphStore &gs, hipStream_t s)
{
    if (!gs.row_ptr) throw Error(-2, "no graph loaded");
    if (gs.adjset) return;
    TraceRange tr("neighbour_sets");
    const uint64_t n = adj_table_start(gs.E, gs.V) + 4;   // every table ends by slot 4 (E + V)
    DArr<uint32_t> t(n);
    DArr<uint32_t> small, mid, big;    // freed after the stream is synchronised below
    DArr<unsigned int> n_small, n_mid, n_big, n_ovf;
    DArr<uint2> items_d, ovf;
    if (gs.V)   // empty lists and lists up to ADJ_TINY_MAX_DEG entries, one lane each
        hipLaunchKernelGGL(k_adjset_tiny, dim3(grid_for(gs.V, 256)), dim3(256), 0, s, gs.row_ptr, gs.col, gs.V, t.p);
    if (gs.E) {
        // lists up to ADJ_LDS_MAX_DEG entries: one wave each, the grid striding over them
        small.p = dalloc<uint32_t>(gs.V);
        n_small.p = dalloc<unsigned int>(1);
        KKW_HIP(hipMemsetAsync(n_small.p, 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_list_range, dim3(grid_for(gs.V, LIST_BLOCK)), dim3(LIST_BLOCK), 0, s, gs.row_ptr, gs.V,
                           (uint64_t)ADJ_TINY_MAX_DEG + 1, (uint64_t)ADJ_LDS_MAX_DEG, small.p, n_small.p);
        hipLaunchKernelGGL(k_adjset_small, dim3((uint32_t)std::min<uint64_t>(((uint64_t)gs.V + 3) / 4, 1u << 16)),
                           dim3(256), 0, s, small.p, n_small.p, gs.row_ptr, gs.col, t.p);
        mid.p = dalloc<uint32_t>(gs.V);
        n_mid.p = dalloc<unsigned int>(1);
        KKW_HIP(hipMemsetAsync(n_mid.p, 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_list_range, dim3(grid_for(gs.V, LIST_BLOCK)), dim3(LIST_BLOCK), 0, s, gs.row_ptr, gs.V,
                           (uint64_t)ADJ_LDS_MAX_DEG + 1, (uint64_t)ADJ_MID_MAX_DEG, mid.p, n_mid.p);
        // two 64 KiB workgroups fit a CU's LDS; the grid strides over the lists
        hipLaunchKernelGGL((k_adjset_mid<ADJ_MID_MAX_DEG, 256>), dim3(std::min<uint64_t>(gs.V, 4096)), dim3(256), 0,
                           s, mid.p, n_mid.p, gs.row_ptr, gs.col, t.p);
        // longer lists: LDS segments, then the keys that ran past a segment
        big.p = dalloc<uint32_t>(gs.V);
        n_big.p = dalloc<unsigned int>(1);
        KKW_HIP(hipMemsetAsync(n_big.p, 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_list_range, dim3(grid_for(gs.V, LIST_BLOCK)), dim3(LIST_BLOCK), 0, s, gs.row_ptr, gs.V,
                           (uint64_t)ADJ_MID_MAX_DEG + 1, ~0ull, big.p, n_big.p);
        unsigned int nb = 0;
        KKW_HIP(hipMemcpyAsync(&nb, n_big.p, sizeof(nb), hipMemcpyDeviceToHost, s));
        KKW_HIP(hipStreamSynchronize(s));
        if (nb) {
            // the segments of every long list as work items, counted and listed on the device
            DArr<uint32_t> segs(nb + 1), seg_off(nb + 1);
            hipLaunchKernelGGL(k_list_segments, dim3(grid_for(nb + 1, 256)), dim3(256), 0, s, big.p, nb, gs.row_ptr,
                               segs.p);
            size_t scan_bytes = 0;
            KKW_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, segs.p, seg_off.p, nb + 1, s));
            DArr<unsigned char> scan_tmp(scan_bytes ? scan_bytes : 1);
            KKW_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_bytes, segs.p, seg_off.p, nb + 1, s));
            // at most (4 E + 4 nb) / ADJ_SEG segments plus one partial segment per list
            const uint64_t max_items = (4 * gs.E + 4 * (uint64_t)nb + ADJ_SEG - 1) / ADJ_SEG + nb;
            items_d.p = dalloc<uint2>(max_items);
            hipLaunchKernelGGL(k_segment_items, dim3(grid_for(nb, 256)), dim3(256), 0, s, segs.p, seg_off.p, nb,
                               items_d.p);
            // a few keys per table run past a segment; KKW_ADJ_OVERFLOW_CAP (tests) shrinks the buffer
            uint64_t cap = std::min<uint64_t>(gs.E, ADJ_OVERFLOW_CAP);
            if (const char *env = getenv("KKW_ADJ_OVERFLOW_CAP"))
                cap = std::min<uint64_t>(cap, strtoull(env, nullptr, 10));
            ovf.p = dalloc<uint2>(cap ? cap : 1);
            n_ovf.p = dalloc<unsigned int>(1);
            KKW_HIP(hipMemsetAsync(n_ovf.p, 0, sizeof(unsigned int), s));
            // two 64 KiB workgroups fit a CU's LDS; the grid strides over the items (their count is seg_off[nb])
            hipLaunchKernelGGL(k_adjset_seg, dim3((unsigned)std::min<uint64_t>(max_items, 8192)), dim3(256), 0, s,
                               items_d.p, seg_off.p + nb, big.p, gs.row_ptr, gs.col, t.p, ovf.p, (uint32_t)cap,
                               n_ovf.p);
            hipLaunchKernelGGL(k_adjset_overflow, dim3(1024), dim3(256), 0, s, ovf.p, n_ovf.p, (uint32_t)cap,
                               gs.row_ptr, t.p);
            unsigned int no = 0;
            KKW_HIP(hipMemcpyAsync(&no, n_ovf.p, sizeof(no), hipMemcpyDeviceToHost, s));
            KKW_HIP(hipStreamSynchronize(s));   // also before the segment and scan buffers go
            if (no > cap)   // the buffer could not hold them all: insert every key of the long lists
                hipLaunchKernelGGL(k_adjset_insert_long, dim3(grid_for(gs.E, 256)), dim3(256), 0, s, gs.row_ptr,
                                   gs.col, gs.V, gs.E, t.p);
        }
    }
    KKW_HIP(hipGetLastError());
    KKW_HIP(hipStreamSynchronize(s));
    gs.adjset = t.release();
}

void build_uniform_buckets(GraphStore &gs, hipStream_t s)
{
    if (!gs.row_ptr) throw Error(-2, "no graph loaded");
    if (gs.ubucket) return;
    TraceRange tr("uniform_buckets");
    DArr<Bucket> t(gs.E);
    if (gs.E)
        hipLaunchKernelGGL(k_uniform_buckets, dim3(grid_for(gs.E, 256)), dim3(256), 0, s, gs.col, gs.etype, gs.vmask,
                           gs.row_ptr, gs.E, gs.packable ? 1 : 0, t.p);
    KKW_HIP(hipGetLastError());
    KKW_HIP(hipStreamSynchronize(s));
    gs.ubucket = t.release();
}

void build_alias(GraphStore &gs, hipStream_t s, bool local)
{
    if (!gs.row_ptr) throw Error(-2, "no graph loaded");
    if (!gs.w) throw Error(-2, "graph has no weights: alias tables need a weighted graph");
    TraceRange tr(local ? "alias_tables_local" : "alias_tables");
    Bucket *&target = local ? gs.lbucket : gs.bucket;
    dfree(target);
    DArr<Bucket> t(gs.E);
    DArr<AliasRec> rec(gs.E);   // freed (kept by the device pool) when the build ends
    // both tables give the same sums; keep the array (and its address) if present
    DArr<float> vs;
    float *vsum = gs.vsum;
    if (!vsum) vsum = vs.p = dalloc<float>(gs.V);
    {
        DArr<uint2> qs(gs.E), ql(gs.E);
        // the engine's tables carry their candidates' rows when the graph packs them
        DArr<uint64_t> nrow;
        if (!local && gs.packable && gs.E) {
            nrow.p = dalloc<uint64_t>(gs.E);
            hipLaunchKernelGGL(k_neighbour_rows, dim3(grid_for(gs.E, 256)), dim3(256), 0, s, gs.col, gs.E, gs.row_ptr,
                               nrow.p);
        }
        // lists of more than AW_POOL_MIN entries: one block each, listed
        // first; the rest of AW_MIN or more: one wave each; shorter lists:
        // one lane each
        DArr<uint32_t> big(gs.V);
        DArr<unsigned int> n_big_d(2);
        KKW_HIP(hipMemsetAsync(n_big_d.p, 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_list_range, dim3(grid_for(gs.V, LIST_BLOCK)), dim3(LIST_BLOCK), 0, s, gs.row_ptr, gs.V,
                           AW_POOL_MIN + 1, ~0ull, big.p, n_big_d.p);
        KKW_HIP(hipMemcpyAsync(n_big_d.p + 1, n_big_d.p, sizeof(unsigned int), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_list_range, dim3(grid_for(gs.V, LIST_BLOCK)), dim3(LIST_BLOCK), 0, s, gs.row_ptr, gs.V,
                           (uint64_t)AW_MIN, AW_POOL_MIN, big.p, n_big_d.p);
        unsigned int counts[2] = {0, 0};   // all listed, pooled (listed first)
        KKW_HIP(hipMemcpyAsync(counts, n_big_d.p, sizeof(counts), hipMemcpyDeviceToHost, s));
        KKW_HIP(hipStreamSynchronize(s));
        const unsigned int n_big = counts[0], n_pool = counts[1];
        if (n_big)
            hipLaunchKernelGGL(k_alias_wave, dim3(n_pool + grid_for(n_big - n_pool, AW_WAVES)), dim3(64 * AW_WAVES), 0,
                               s, big.p, n_big, n_pool, gs.row_ptr, gs.w, rec.p, vsum, qs.p, ql.p);
        hipLaunchKernelGGL(k_alias, dim3(grid_for(gs.V, 64)), dim3(64), 0, s, gs.row_ptr, gs.w, gs.V, rec.p, vsum);
        if (gs.E)
            hipLaunchKernelGGL(k_finish_buckets, dim3(grid_for(gs.E, 256)), dim3(256), 0, s, rec.p, t.p, gs.E,
                               gs.row_ptr, gs.col, gs.etype, !local, nrow.p, gs.vmask);
        KKW_HIP(hipGetLastError());
        KKW_HIP(hipStreamSynchronize(s));
    }
    target = t.release();
    if (vs.p) gs.vsum = vs.release();
}

}  // namespace kkw
