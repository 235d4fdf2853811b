// program.hip -- walk programs: a user's first-order transition (walker state,
// extension, dynamic component, bounds, outlier pair; include/kkw_program.h)
// compiled at run time with hipRTC for the engine's GPU, together with the
// library's walk loop (program_walk.h).  This is what replaces the generic
// WalkEngine::random_walk(WalkerConfig*, TransitionConfig*, WalkConfig*) of
// include/walk.hpp:423-436, whose std::function closures cannot run on a GPU.
#include "engine_state.h"
#include "program_args.h"

#include <hip/hiprtc.h>

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "program_sources.inc"

using namespace kkw;

static_assert(sizeof(kkw_prog_bucket) == sizeof(Bucket), "program bucket layout");
static_assert(offsetof(kkw_prog_bucket, p) == offsetof(Bucket, p), "program bucket layout");
static_assert(offsetof(kkw_prog_bucket, primary) == offsetof(Bucket, primary), "program bucket layout");
static_assert(offsetof(kkw_prog_bucket, alias) == offsetof(Bucket, alias), "program bucket layout");
static_assert(offsetof(kkw_prog_bucket, prow) == offsetof(Bucket, prow), "program bucket layout");
static_assert(offsetof(kkw_prog_bucket, arow) == offsetof(Bucket, arow), "program bucket layout");
static_assert(KKW_PROG_ROW_DEG_SHIFT == (unsigned)ROW_DEG_SHIFT, "packed row layout");

struct kkw_program {
    kkw_engine *eng = nullptr;
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;     // plain form
    hipFunction_t fn_p = nullptr;   // persistent form
    std::string log;
    bool uses_adjacency = false;   // the source calls kkw_adjacent: give it the hash sets
    bool edge_index = true;        // reads e.index: local-index tables; else the engine's id + row tables
    void *arrays[KKW_PROGRAM_ARRAYS] = {};
    uint64_t array_bytes[KKW_PROGRAM_ARRAYS] = {};
    ~kkw_program()
    {
        for (void *&a : arrays)
            if (a) (void)hipFree(a);
        if (mod) (void)hipModuleUnload(mod);
    }
};

namespace {

#define KKW_RTC(call)                                                                          \
    do {                                                                                       \
        hiprtcResult r_ = (call);                                                              \
        if (r_ != HIPRTC_SUCCESS)                                                              \
            throw Error(KKW_E_DEVICE, std::string(#call) + ": " + hiprtcGetErrorString(r_));  \
    } while (0)

struct RtcProgram {
    hiprtcProgram p = nullptr;
    ~RtcProgram()
    {
        if (p) (void)hiprtcDestroyProgram(&p);
    }
};

std::vector<std::string> split_options(const char *options)
{
    std::vector<std::string> out;
    if (!options) return out;
    std::istringstream is(options);
    std::string tok;
    while (is >> tok) out.push_back(tok);
    return out;
}

// Compile source + walk loop for the device's architecture; returns the code
// object, or throws KKW_E_INVALID with the compiler log.
std::vector<char> compile(const std::string &arch, const char *source, const char *options, std::string &log)
{
    std::string full = "#include \"kkw_program.h\"\n#line 1 \"program\"\n";
    full += source;
    full += "\n#include \"program_walk.h\"\n";
    RtcProgram rp;
    KKW_RTC(hiprtcCreateProgram(&rp.p, full.c_str(), "program.hip", kProgramHeaderCount, kProgramHeaderTexts,
                                kProgramHeaderNames));
    std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17", "-ffp-contract=off",
                                     // no packed fp32 instructions, as for the library's own kernels
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    for (auto &o : split_options(options)) opts.push_back(o);
    std::vector<const char *> argv;
    for (auto &o : opts) argv.push_back(o.c_str());
    const hiprtcResult rc = hiprtcCompileProgram(rp.p, (int)argv.size(), argv.data());
    size_t n = 0;
    KKW_RTC(hiprtcGetProgramLogSize(rp.p, &n));
    log.assign(n, '\0');
    if (n) KKW_RTC(hiprtcGetProgramLog(rp.p, &log[0]));
    while (!log.empty() && log.back() == '\0') log.pop_back();
    if (rc != HIPRTC_SUCCESS) throw Error(KKW_E_INVALID, "walk program does not compile:\n" + log);
    size_t cs = 0;
    KKW_RTC(hiprtcGetCodeSize(rp.p, &cs));
    std::vector<char> code(cs);
    KKW_RTC(hiprtcGetCode(rp.p, code.data()));
    return code;
}

uint32_t round_up4(uint64_t x) { return (uint32_t)((x + 3) & ~(uint64_t)3); }

}  // namespace

int kkw_program_create(kkw_engine *eng, const char *source, const char *options, kkw_program **out)
{
    return guarded([&]() -> int {
        if (!eng || !source || !out) return fail(KKW_E_INVALID, "null argument");
        *out = nullptr;
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        hipDeviceProp_t prop;
        KKW_HIP(hipGetDeviceProperties(&prop, eng->device));
        std::string arch = prop.gcnArchName;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch.resize(colon);
        std::unique_ptr<kkw_program> p(new kkw_program());
        p->eng = eng;
        p->uses_adjacency = strstr(source, "kkw_adjacent") != nullptr;
        std::vector<char> code = compile(arch, source, options, p->log);
        KKW_HIP(hipModuleLoadData(&p->mod, code.data()));
        KKW_HIP(hipModuleGetFunction(&p->fn, p->mod, "kkw_program_walk"));
        KKW_HIP(hipModuleGetFunction(&p->fn_p, p->mod, "kkw_program_walk_persistent"));
        hipDeviceptr_t flag = nullptr;
        size_t flag_bytes = 0;
        KKW_HIP(hipModuleGetGlobal(&flag, &flag_bytes, p->mod, "kkw_prog_edge_index"));
        int edge_index = 1;
        KKW_HIP(hipMemcpyDtoH(&edge_index, flag, sizeof(edge_index)));
        p->edge_index = edge_index != 0;
        *out = p.release();
        return KKW_OK;
    });
}

int kkw_program_destroy(kkw_program *prog)
{
    if (!prog) return KKW_OK;
    {
        std::lock_guard<std::recursive_mutex> lk(prog->eng->mu);
        (void)hipSetDevice(prog->eng->device);
        delete prog;
    }
    return KKW_OK;
}

const char *kkw_program_log(const kkw_program *prog) { return prog ? prog->log.c_str() : ""; }

int kkw_program_set_array(kkw_program *prog, int slot, const void *data, uint64_t bytes)
{
    return guarded([&]() -> int {
        if (!prog || slot < 0 || slot >= KKW_PROGRAM_ARRAYS || (bytes && !data))
            return fail(KKW_E_INVALID, "bad program array slot or data");
        std::lock_guard<std::recursive_mutex> lk(prog->eng->mu);
        KKW_HIP(hipSetDevice(prog->eng->device));
        if (prog->arrays[slot]) {
            KKW_HIP(hipFree(prog->arrays[slot]));
            prog->arrays[slot] = nullptr;
            prog->array_bytes[slot] = 0;
        }
        if (!bytes) return KKW_OK;
        void *d = nullptr;
        if (hipMalloc(&d, bytes) != hipSuccess) return fail(KKW_E_NOMEM, "program array allocation failed");
        prog->arrays[slot] = d;
        prog->array_bytes[slot] = bytes;
        KKW_HIP(hipMemcpy(d, data, bytes, hipMemcpyHostToDevice));
        return KKW_OK;
    });
}

int kkw_program_run(kkw_program *prog, const kkw_prog_job *job, kkw_paths *out)
{
    return guarded([&]() -> int {
        if (!prog || !job || !out) return fail(KKW_E_INVALID, "null argument");
        kkw_engine *eng = prog->eng;
        std::lock_guard<std::recursive_mutex> lk(eng->mu);
        KKW_HIP(hipSetDevice(eng->device));
        GraphStore &gs = eng->gs;
        if (!gs.row_ptr) return fail(KKW_E_NOGRAPH, "no graph loaded");
        if (gs.V == 0) return fail(KKW_E_INVALID, "graph has no vertices");
        if (job->start_mode == KKW_START_LIST) {
            if (!job->starts || job->n_starts == 0) return fail(KKW_E_INVALID, "start list is empty");
            for (uint32_t i = 0; i < job->n_starts; i++)
                if (job->starts[i] >= gs.V) return fail(KKW_E_INVALID, "start vertex out of range");
        } else if (job->start_mode != KKW_START_EQUAL && job->start_mode != KKW_START_UNIFORM) {
            return fail(KKW_E_INVALID, "unknown start mode");
        }
        if (job->use_weights && !gs.w) return fail(KKW_E_INVALID, "use_weights on a graph without weights");
        if ((uint64_t)job->walker_begin + job->walker_num > 0xFFFFFFFFull)
            return fail(KKW_E_INVALID, "walker ids exceed 32 bits");
        if (job->max_length >= 0x7FFFFFF0u) return fail(KKW_E_INVALID, "max_length too large");
        if (job->mode != KKW_MODE_PLAIN && job->mode != KKW_MODE_PERSISTENT)
            return fail(KKW_E_INVALID, "unknown mode");
        if (prog->uses_adjacency) auto_adj_index(eng);
        if (prog->edge_index && job->use_weights && !gs.lbucket) {
            // local-index alias tables; the sums array keeps its address, so
            // batches captured earlier stay valid
            double t0 = now_ms();
            build_alias(gs, eng->stream, true);
            eng->timing.alias_ms = now_ms() - t0;
        }
        if (!prog->edge_index) {
            // the engine's own tables, as its built-in walks build them
            if (job->use_weights && !gs.bucket) {
                eng->graph_version++;
                double t0 = now_ms();
                build_alias(gs, eng->stream);
                eng->timing.alias_ms = now_ms() - t0;
            } else if (!job->use_weights && !gs.ubucket) {
                eng->graph_version++;
                build_uniform_buckets(gs, eng->stream);
            }
        }
        const uint32_t n = job->walker_num;
        const bool persistent = job->mode == KKW_MODE_PERSISTENT;
        // persistent rows are padded to whole 128-byte lines when that adds at
        // most a quarter (the kernel then writes 64-byte half lines from LDS)
        const uint64_t len = (uint64_t)job->max_length + 1;
        const uint64_t line_stride = (len + 31) & ~(uint64_t)31;
        const bool line_rows = persistent && len >= 32 && (line_stride - len) * 4 <= len;
        const uint32_t stride = line_rows ? (uint32_t)line_stride : round_up4(len);
        memset(out, 0, sizeof(*out));
        out->walker_begin = job->walker_begin;
        out->walker_num = n;
        out->stride = stride;
        if (n == 0) {
            out->vertices = (uint32_t *)malloc(4);
            out->lengths = (uint32_t *)malloc(4);
            out->priv = (void *)1;
            return KKW_OK;
        }
        // queue word, then the envelope-violation counter
        DevBuf paths((size_t)n * stride * 4), lens((size_t)n * 4), queue(2 * sizeof(unsigned long long));
        DevBuf starts(job->start_mode == KKW_START_LIST ? (size_t)job->n_starts * 4 : 0);
        if (job->start_mode == KKW_START_LIST)
            KKW_HIP(hipMemcpyAsync(starts.p, job->starts, (size_t)job->n_starts * 4, hipMemcpyHostToDevice,
                                   eng->stream));
        kkw_prog_args A;
        memset(&A, 0, sizeof(A));
        A.c.v_num = gs.V;
        A.c.row_v = KKW_NO_VERTEX;   // each lane keeps its walker's row from here
        A.c.sorted = gs.sorted ? 1u : 0u;
        A.c.e_num = gs.E;
        A.c.adjset = gs.adjset;
        A.c.row_ptr = gs.row_ptr;
        A.c.col = gs.col;
        A.c.weight = gs.w;
        A.c.etype = gs.etype;
        for (int k = 0; k < KKW_PROGRAM_ARRAYS; k++) A.c.array[k] = prog->arrays[k];
        for (int k = 0; k < KKW_PROGRAM_PARAMS; k++) A.c.param[k] = job->param[k];
        if (prog->edge_index) {
            A.lb = job->use_weights ? reinterpret_cast<const kkw_prog_bucket *>(gs.lbucket) : nullptr;
        } else {
            A.lb = reinterpret_cast<const kkw_prog_bucket *>(job->use_weights ? gs.bucket : gs.ubucket);
            A.packed_rows = gs.packable ? 1u : 0u;
        }
        A.vsum = job->use_weights ? gs.vsum : nullptr;
        A.seed = job->seed;
        A.walker_begin = job->walker_begin;
        A.walker_num = n;
        A.max_len = job->max_length;
        A.stride = stride;
        A.start_mode = job->start_mode;
        A.n_starts = job->n_starts;
        A.starts = (const uint32_t *)starts.p;
        A.retest_ext = job->second_order ? 1 : 0;
        A.paths = (uint32_t *)paths.p;
        A.lens = (uint32_t *)lens.p;
        const uint32_t chunk = 64;
        A.chunk = chunk;
        A.n_chunks = ((uint64_t)n + chunk - 1) / chunk;
        A.queue = (unsigned long long *)queue.p;
        A.violations = (unsigned long long *)queue.p + 1;
        A.line_rows = line_rows ? 1u : 0u;
        // persistent: 8 blocks of 256 per CU, fewer when there are fewer work items
        const uint64_t blocks_needed = (A.n_chunks * 64 + KKW_PROG_BLOCK - 1) / KKW_PROG_BLOCK;
        const uint32_t grid = persistent ? (uint32_t)std::min<uint64_t>((uint64_t)eng->cu_count * 8, blocks_needed)
                                         : (n + KKW_PROG_BLOCK - 1) / KKW_PROG_BLOCK;
        KKW_HIP(hipMemsetAsync(queue.p, 0, 2 * sizeof(unsigned long long), eng->stream));
        void *args[] = {&A};
        hipEvent_t e0, e1;
        KKW_HIP(hipEventCreate(&e0));
        KKW_HIP(hipEventCreate(&e1));
        const double t0 = now_ms();
        KKW_HIP(hipEventRecord(e0, eng->stream));
        const unsigned lds = line_rows ? KKW_PROG_BLOCK * (KKW_PROG_HALF_LINE + 1) * 4u : 0u;
        KKW_HIP(hipModuleLaunchKernel(persistent ? prog->fn_p : prog->fn, grid, 1, 1, KKW_PROG_BLOCK, 1, 1, lds,
                                      eng->stream, args, nullptr));
        KKW_HIP(hipEventRecord(e1, eng->stream));
        KKW_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        KKW_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        eng->timing.kernel_ms = ms;
        out->vertices = (uint32_t *)malloc((size_t)n * stride * 4);
        out->lengths = (uint32_t *)malloc((size_t)n * 4);
        if (!out->vertices || !out->lengths) {
            free(out->vertices);
            free(out->lengths);
            memset(out, 0, sizeof(*out));
            return fail(KKW_E_NOMEM, "host allocation for paths failed");
        }
        out->priv = (void *)1;   // host buffers owned by the caller (kkw_paths_free)
        KKW_HIP(hipMemcpy(out->vertices, paths.p, (size_t)n * stride * 4, hipMemcpyDeviceToHost));
        KKW_HIP(hipMemcpy(out->lengths, lens.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        uint64_t steps = 0, trunc = 0;
        for (uint32_t i = 0; i < n; i++) {
            steps += (out->lengths[i] & 0x7fffffffu) - 1;
            trunc += out->lengths[i] >> 31;
        }
        unsigned long long viol = 0;
        KKW_HIP(hipMemcpy(&viol, A.violations, sizeof(viol), hipMemcpyDeviceToHost));
        out->steps = steps;
        out->truncated = trunc;
        eng->timing.steps = steps;
        eng->timing.envelope_violations = viol;
        eng->timing.total_ms = now_ms() - t0;
        return KKW_OK;
    });
}
