// libfgx: context, errors, memory/stream/graph wrappers, fills, ID assignment,
// spatial metadata and scratch sizing.
#include "fgx_internal.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <limits>

namespace fgx {

static thread_local std::string g_last_error;

std::mutex& clean_mutex() {
    static std::mutex m;
    return m;
}
std::unordered_map<const void*, uint32_t>& clean_record() {
    static std::unordered_map<const void*, uint32_t> r;
    return r;
}
std::unordered_map<const void*, uint32_t>& bins_history() {
    static std::unordered_map<const void*, uint32_t> r;
    return r;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int make_var_table(uint32_t nvars, const void* const* in, void* const* out, const uint32_t* bytes, VarTable* t) {
    if (nvars > FGX_MAX_VARS) return fail(FGX_ERR_ARG, "too many variables (max " + std::to_string(FGX_MAX_VARS) + ")");
    std::memset(t, 0, sizeof(*t));
    t->n = nvars;
    for (uint32_t v = 0; v < nvars; ++v) {
        if (!bytes[v]) return fail(FGX_ERR_ARG, "variable with zero element size");
        t->bytes[v] = bytes[v];
        t->in[v] = static_cast<const char*>(in ? in[v] : nullptr);
        t->out[v] = static_cast<char*>(out ? out[v] : nullptr);
    }
    return FGX_OK;
}

// ---------------------------------------------------------------- kernels
__global__ void k_fill_u32(uint32_t* __restrict__ dst, uint32_t value, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}

struct FillArgs {
    uint32_t nvars;
    uint32_t n;
    uint32_t bytes[FGX_MAX_VARS];
    char* ptr[FGX_MAX_VARS];
    uint32_t words[FGX_MAX_VARS][4];  // default value, up to 16 bytes per element
};

// One launch fills every variable (grid.y indexes the variable).
__global__ void k_fill_vars(FillArgs a) {
    const uint32_t v = blockIdx.y;
    const uint32_t b = a.bytes[v];
    char* base = a.ptr[v];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        char* d = base + (size_t)i * b;
        if ((b & 3u) == 0) {
            for (uint32_t w = 0; w < b / 4; ++w) reinterpret_cast<uint32_t*>(d)[w] = a.words[v][w & 3];
        } else {
            const char* src = reinterpret_cast<const char*>(a.words[v]);
            for (uint32_t w = 0; w < b; ++w) d[w] = src[w & 15];
        }
    }
}

__global__ void k_add_u32(uint32_t* p, uint32_t delta) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += delta;
}

struct SetCounts {
    uint32_t n;
    uint32_t* dst[FGX_SET_COUNTS_MAX];
    uint32_t value[FGX_SET_COUNTS_MAX];
    uint32_t* inc;
};

__global__ void k_set_counts(SetCounts a) {
    const uint32_t t = threadIdx.x;
    if (blockIdx.x != 0) return;
    if (t < a.n) *a.dst[t] = a.value[t];
    if (t == 63 && a.inc) *a.inc += 1u;
}

__global__ void k_sub_u32(uint32_t* dst, const uint32_t* a, const uint32_t* b) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t x = *a, y = *b;
        *dst = x > y ? x - y : 0u;
    }
}

__global__ void k_add_u32_n(uint32_t* __restrict__ p, uint32_t n, uint32_t delta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] += delta;
}

// One thread per destination element d = dst_off[0] + j: its range is the last one
// with dst_off <= d (binary search; destination ranges are contiguous and sorted),
// then every variable's element is copied.
__global__ void k_copy_ranges(VarTable vt, const uint32_t* __restrict__ ranges, uint32_t nranges, uint32_t n_total) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t base = ranges[1];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_total; j += stride) {
        const uint32_t d = base + j;
        uint32_t lo = 0, hi = nranges;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ranges[3 * mid + 1] <= d) lo = mid;
            else hi = mid;
        }
        const uint32_t k = d - ranges[3 * lo + 1];
        if (k < ranges[3 * lo + 2]) copy_element(vt, ranges[3 * lo] + k, d);
    }
}

// ---------------------------------------------------------------- batches whose populations change
// Per-simulation counts live on the device; these kernels recompute a packed list's
// per-simulation layout and patch the tables that record it.
constexpr uint32_t kSegBlock = 1024;

struct SegTargets {
    uint32_t n;
    fgx_seg_target t[FGX_SEG_MAX_TARGETS];
};

// Exclusive prefix of v over the block (1024 threads); *sum receives the block total.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_w, uint32_t* sum) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t inc = wave_inclusive_u32(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t k = 0; k < kSegBlock / 64; ++k) {
        const uint32_t x = s_w[k];
        before += k < w ? x : 0u;
        total += x;
    }
    __syncthreads();
    *sum = total;
    return before + inc - v;
}

__device__ __forceinline__ void put_u32(const fgx_seg_target& t, uint32_t rec, int32_t field, uint32_t v) {
    if (field >= 0) *reinterpret_cast<uint32_t*>(static_cast<char*>(t.table) + (size_t)rec * t.stride + field) = v;
}

// One block: c[s] = cnt[s] (+ add[s], stored back) -> off[s], *total, and every target's
// offset / count / first-tile fields (plus the sentinel record nsegs when asked).
__global__ void __launch_bounds__(kSegBlock) k_seg_layout(uint32_t nsegs, uint32_t* __restrict__ cnt,
                                                          const uint32_t* __restrict__ add, uint32_t* __restrict__ off,
                                                          uint32_t* __restrict__ total, SegTargets tg) {
    __shared__ uint32_t s_w[kSegBlock / 64];
    uint32_t elems = 0;                          // running offset before this chunk
    uint32_t tiles[FGX_SEG_MAX_TARGETS];         // running first tile per target
    for (uint32_t k = 0; k < FGX_SEG_MAX_TARGETS; ++k) tiles[k] = 0;
    for (uint32_t base = 0; base < nsegs; base += kSegBlock) {
        const uint32_t s = base + threadIdx.x;
        const bool on = s < nsegs;
        const uint32_t c = on ? cnt[s] + (add ? add[s] : 0u) : 0u;
        uint32_t chunk = 0;
        const uint32_t eo = elems + block_exclusive(c, s_w, &chunk);
        if (on) {
            if (add) cnt[s] = c;
            off[s] = eo;
        }
        for (uint32_t k = 0; k < FGX_SEG_MAX_TARGETS; ++k) {
            if (k >= tg.n) break;
            const fgx_seg_target& t = tg.t[k];
            const uint32_t nt = (t.tile_field >= 0 && t.tile_elems) ? (c + t.tile_elems - 1u) / t.tile_elems : 0u;
            uint32_t tsum = 0;
            const uint32_t to = tiles[k] + block_exclusive(nt, s_w, &tsum);
            if (on) {
                put_u32(t, s, t.off_field, eo);
                put_u32(t, s, t.n_field, c);
                put_u32(t, s, t.tile_field, to);
            }
            tiles[k] += tsum;
        }
        elems += chunk;
    }
    if (threadIdx.x == 0) {
        *total = elems;
        for (uint32_t k = 0; k < tg.n && k < FGX_SEG_MAX_TARGETS; ++k) {
            const fgx_seg_target& t = tg.t[k];
            if (!t.sentinel) continue;
            put_u32(t, nsegs, t.off_field, elems);
            put_u32(t, nsegs, t.n_field, 0u);
            put_u32(t, nsegs, t.tile_field, tiles[k]);
        }
    }
}

// Block s: out[s] = flags equal to `keep` in [off[s], off[s] + cnt[s]).  out may be cnt:
// each block reads its own count before the barrier that precedes its write.
__global__ void __launch_bounds__(256) k_seg_count(const uint32_t* __restrict__ off, const uint32_t* cnt,
                                                   const uint32_t* __restrict__ flags, uint32_t keep, uint32_t* out) {
    __shared__ uint32_t s_w[4];
    const uint32_t s = blockIdx.x;
    const uint32_t o = off[s], n = cnt[s];
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 256u) c += flags[o + i] == keep ? 1u : 0u;
    c = wave_sum_u32(c);
    if ((threadIdx.x & 63u) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) out[s] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// One block: work = [new_off[0..S) | new_total | boff[0..S) | born_total], new_off the
// exclusive scan of cnt + born, boff that of born.
__global__ void __launch_bounds__(kSegBlock) k_seg_append_scan(uint32_t nsegs, const uint32_t* __restrict__ cnt,
                                                               const uint32_t* __restrict__ born,
                                                               uint32_t* __restrict__ work) {
    __shared__ uint32_t s_w[kSegBlock / 64];
    uint32_t e = 0, b = 0;
    for (uint32_t base = 0; base < nsegs; base += kSegBlock) {
        const uint32_t s = base + threadIdx.x;
        const bool on = s < nsegs;
        const uint32_t bs = on ? born[s] : 0u, c = on ? cnt[s] + bs : 0u;
        uint32_t ce = 0, cb = 0;
        const uint32_t eo = e + block_exclusive(c, s_w, &ce);
        const uint32_t bo = b + block_exclusive(bs, s_w, &cb);
        if (on) {
            work[s] = eo;
            work[nsegs + 1 + s] = bo;
        }
        e += ce;
        b += cb;
    }
    if (threadIdx.x == 0) {
        work[nsegs] = e;
        work[2 * nsegs + 1] = b;
    }
}

// Output element q of the new layout: simulation s (last new_off[s] <= q), then either
// its existing element q - new_off[s] or its new agent r = q - new_off[s] - cnt[s]
// (born[boff[s] + r], ID next_id[s] + r).
__global__ void k_seg_append_copy(uint32_t nsegs, const uint32_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                                  const uint32_t* __restrict__ work, uint32_t n_cap_out, VarTable list, VarTable born,
                                  int32_t id_var, const uint32_t* __restrict__ next_id) {
    const uint32_t total = work[nsegs] < n_cap_out ? work[nsegs] : n_cap_out;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
        uint32_t lo = 0, hi = nsegs;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (work[mid] <= q) lo = mid;
            else hi = mid;
        }
        const uint32_t local = q - work[lo];
        const uint32_t c = cnt[lo];
        if (local < c) {
            copy_element(list, off[lo] + local, q);
        } else {
            const uint32_t r = local - c;
            copy_element(born, work[nsegs + 1 + lo] + r, q, id_var);
            if (id_var >= 0) reinterpret_cast<uint32_t*>(born.out[id_var])[q] = next_id[lo] + r;
        }
    }
}

__global__ void k_add_n_u32(uint32_t* __restrict__ p, const uint32_t* __restrict__ d, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] += d[i];
}

__global__ void k_assign_ids(uint32_t* __restrict__ ids, uint32_t n, uint32_t next_id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ids[i] == 0u) ids[i] = next_id + i;
}

}  // namespace fgx

using namespace fgx;

struct fgx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

const char* fgx_last_error(void) { return g_last_error.c_str(); }

int fgx_ctx_create(int device, fgx_ctx** out) {
    if (!out) return fail(FGX_ERR_ARG, "fgx_ctx_create: out is NULL");
    int count = 0;
    FGX_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(FGX_ERR_ARG, "fgx_ctx_create: device " + std::to_string(device) + " out of range");
    FGX_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    FGX_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FGX_ERR_ARG, std::string("fgx_ctx_create: built for gfx950, device is ") + prop.gcnArchName);
    {
        const int rc = noalloc_scratch_init(device);
        if (rc) return rc;
    }
    auto* c = new fgx_ctx();
    c->c.device = device;
    c->c.cu_count = prop.multiProcessorCount;
    c->c.hbm_bytes = prop.totalGlobalMem;
    c->c.xcd_count = 8;
    // fused-layer image space, allocated here so that no later call allocates (calls may be captured)
    if (hipMalloc((void**)&c->c.layers.dev, FGX_LAYER_SPACE_BYTES) != hipSuccess ||
        hipHostMalloc((void**)&c->c.layers.host, FGX_LAYER_SPACE_BYTES, hipHostMallocDefault) != hipSuccess) {
        fgx_ctx_destroy(c);
        return fail(FGX_ERR_HIP, "fgx_ctx_create: layer image space allocation failed");
    }
    *out = c;
    return FGX_OK;
}

int fgx_ctx_destroy(fgx_ctx* ctx) {
    if (ctx) {
        if (ctx->c.layers.dev || ctx->c.layers.host) hipSetDevice(ctx->c.device);
        if (ctx->c.layers.dev) hipFree(ctx->c.layers.dev);
        if (ctx->c.layers.host) hipHostFree(ctx->c.layers.host);
    }
    delete ctx;
    return FGX_OK;
}

int fgx_ctx_set_option(fgx_ctx* ctx, int option, int value) {
    if (!ctx) return fail(FGX_ERR_ARG, "fgx_ctx_set_option: NULL context");
    if (option != FGX_OPT_KEEP_SCRATCH_ZERO) return fail(FGX_ERR_ARG, "fgx_ctx_set_option: unknown option");
    ctx->c.keep_zero = value != 0;
    return FGX_OK;
}

int fgx_device_info(fgx_ctx* ctx, int* cu_count, int* xcd_count, size_t* hbm_bytes) {
    if (!ctx) return fail(FGX_ERR_ARG, "fgx_device_info: NULL context");
    if (cu_count) *cu_count = ctx->c.cu_count;
    if (xcd_count) *xcd_count = ctx->c.xcd_count;
    if (hbm_bytes) *hbm_bytes = ctx->c.hbm_bytes;
    return FGX_OK;
}

int fgx_malloc(fgx_ctx* ctx, size_t bytes, void** out) {
    if (!out) return fail(FGX_ERR_ARG, "fgx_malloc: out is NULL");
    if (ctx) FGX_HIP(hipSetDevice(ctx->c.device));
    *out = nullptr;
    if (!bytes) return FGX_OK;
    FGX_HIP(hipMalloc(out, bytes));
    return FGX_OK;
}

int fgx_free(fgx_ctx* ctx, void* p) {
    (void)ctx;
    if (!p) return FGX_OK;
    {
        // a later allocation at this address starts with unknown contents
        std::lock_guard<std::mutex> lk(clean_mutex());
        clean_record().erase(p);
        bins_history().erase(p);
    }
    FGX_HIP(hipFree(p));
    return FGX_OK;
}

int fgx_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(FGX_ERR_ARG, "fgx_host_alloc: out is NULL");
    FGX_HIP(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return FGX_OK;
}

int fgx_host_free(void* p) {
    if (p) FGX_HIP(hipHostFree(p));
    return FGX_OK;
}

int fgx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (!bytes) return FGX_OK;
    FGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return FGX_OK;
}

int fgx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    if (!bytes) return FGX_OK;
    FGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return FGX_OK;
}

int fgx_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (!bytes) return FGX_OK;
    FGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return FGX_OK;
}

int fgx_fill_u32(uint32_t* dst, uint32_t value, size_t n, void* stream) {
    if (!n) return FGX_OK;
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dst, value, n);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_set_counts(uint32_t n, uint32_t* const* dst, const uint32_t* values, uint32_t* inc, void* stream) {
    if (n > FGX_SET_COUNTS_MAX) return fail(FGX_ERR_ARG, "fgx_set_counts: too many counters");
    if (n && (!dst || !values)) return fail(FGX_ERR_ARG, "fgx_set_counts: NULL arrays");
    SetCounts a;
    std::memset(&a, 0, sizeof(a));
    a.n = n;
    for (uint32_t k = 0; k < n; ++k) {
        if (!dst[k]) return fail(FGX_ERR_ARG, "fgx_set_counts: NULL counter");
        a.dst[k] = dst[k];
        a.value[k] = values[k];
    }
    a.inc = inc;
    if (!n && !inc) return FGX_OK;
    hipLaunchKernelGGL(k_set_counts, dim3(1), dim3(64), 0, (hipStream_t)stream, a);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_add_u32(uint32_t* p, uint32_t delta, void* stream) {
    if (!p) return fail(FGX_ERR_ARG, "fgx_add_u32: NULL pointer");
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, (hipStream_t)stream, p, delta);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_fill_vars(uint32_t nvars, void* const* vars, const uint32_t* var_bytes, const void* defaults,
                  const uint32_t* def_off, uint32_t n, void* stream) {
    if (!nvars || !n) return FGX_OK;
    if (nvars > FGX_MAX_VARS) return fail(FGX_ERR_ARG, "fgx_fill_vars: too many variables");
    FillArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nvars = nvars;
    a.n = n;
    for (uint32_t v = 0; v < nvars; ++v) {
        a.bytes[v] = var_bytes[v];
        a.ptr[v] = static_cast<char*>(vars[v]);
        const uint32_t b = var_bytes[v];
        // Element default: up to 16 bytes are taken verbatim; larger (array) elements
        // repeat their first 16 bytes, which covers the per-element defaults used here
        // (arrays whose elements share one value).
        if (defaults) std::memcpy(a.words[v], static_cast<const char*>(defaults) + def_off[v], std::min<uint32_t>(b, 16));
    }
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n, 256), 1024);
    hipLaunchKernelGGL(k_fill_vars, dim3(blocks, nvars), dim3(256), 0, (hipStream_t)stream, a);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_stream_create(void** out) {
    hipStream_t s;
    FGX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = s;
    return FGX_OK;
}

int fgx_stream_destroy(void* stream) {
    if (stream) FGX_HIP(hipStreamDestroy((hipStream_t)stream));
    return FGX_OK;
}

int fgx_stream_synchronize(void* stream) {
    FGX_HIP(hipStreamSynchronize((hipStream_t)stream));
    return FGX_OK;
}

int fgx_device_synchronize(void) {
    FGX_HIP(hipDeviceSynchronize());
    return FGX_OK;
}

int fgx_event_create(void** out) {
    hipEvent_t e;
    FGX_HIP(hipEventCreate(&e));
    *out = e;
    return FGX_OK;
}

int fgx_event_destroy(void* ev) {
    if (ev) FGX_HIP(hipEventDestroy((hipEvent_t)ev));
    return FGX_OK;
}

int fgx_event_record(void* ev, void* stream) {
    FGX_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return FGX_OK;
}

int fgx_stream_wait_event(void* stream, void* ev) {
    FGX_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0));
    return FGX_OK;
}

int fgx_event_elapsed_ms(void* start, void* stop, float* ms) {
    FGX_HIP(hipEventSynchronize((hipEvent_t)stop));
    FGX_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return FGX_OK;
}

int fgx_graph_begin(void* stream) {
    if (!stream) return fail(FGX_ERR_ARG, "fgx_graph_begin: capture needs a non-default stream");
    FGX_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return FGX_OK;
}

int fgx_graph_end(void* stream, fgx_graph** out) {
    auto* g = new fgx_graph();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
    if (e != hipSuccess) {
        delete g;
        return fail(FGX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g->graph);
        delete g;
        return fail(FGX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    *out = g;
    return FGX_OK;
}

int fgx_graph_launch(fgx_graph* g, void* stream) {
    if (!g) return fail(FGX_ERR_ARG, "fgx_graph_launch: NULL graph");
    FGX_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return FGX_OK;
}

int fgx_graph_destroy(fgx_graph* g) {
    if (!g) return FGX_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return FGX_OK;
}

// Fused layer launch: assemble the image (header + launch records), find it in the
// context's image space or upload it there, then one launch over all functions' blocks.
int fgx_layer_launch(fgx_ctx* ctx, const void* kernel, uint32_t block, const fgx_fn_launch* fns, uint32_t nfns,
                     uint32_t args_bytes, void* stream) {
    if (!ctx || !kernel || !fns) return fail(FGX_ERR_ARG, "fgx_layer_launch: NULL context, kernel or functions");
    if (nfns == 0 || nfns > FGX_LAYER_MAX_FNS)
        return fail(FGX_ERR_ARG, "fgx_layer_launch: 1 to " + std::to_string(FGX_LAYER_MAX_FNS) + " functions per launch");
    if (args_bytes == 0 || args_bytes > FGX_LAYER_MAX_ARGS_BYTES)
        return fail(FGX_ERR_ARG, "fgx_layer_launch: launch record size out of range");
    if (block == 0 || block > 1024 || block % 64 != 0) return fail(FGX_ERR_ARG, "fgx_layer_launch: block must be 64..1024, a multiple of 64");
    static_assert(sizeof(fgx_layer_header) <= FGX_LAYER_ARGS_OFFSET, "layer header must fit before the launch records");
    fgx_layer_header h;
    std::memset(&h, 0, sizeof(h));
    h.nfns = nfns;
    h.stride = (args_bytes + 15u) & ~15u;
    uint32_t lds = 0;
    uint64_t blocks = 0, per_xcd = 0;
    for (uint32_t k = 0; k < nfns; ++k) {
        if (!fns[k].args || fns[k].blocks == 0) return fail(FGX_ERR_ARG, "fgx_layer_launch: function without launch record or blocks");
        h.begin[k] = (uint32_t)blocks;
        h.fn[k] = fns[k].fn;
        blocks += fns[k].blocks;
        per_xcd += (fns[k].blocks + 7u) / 8u;   // the most blocks of function k any XCD runs
        lds = std::max(lds, fns[k].lds_bytes);
    }
    // the kernel spreads each function's blocks over the 8 XCDs (workgroup w on XCD w % 8)
    const uint64_t grid = 8u * per_xcd;
    if (grid >= (1ull << 31)) return fail(FGX_ERR_ARG, "fgx_layer_launch: too many blocks");
    h.begin[nfns] = (uint32_t)blocks;
    const size_t bytes = FGX_LAYER_ARGS_OFFSET + (size_t)nfns * h.stride;
    std::string img(bytes, '\0');
    std::memcpy(&img[0], &h, sizeof(h));
    for (uint32_t k = 0; k < nfns; ++k) std::memcpy(&img[FGX_LAYER_ARGS_OFFSET + (size_t)k * h.stride], fns[k].args, args_bytes);
    LayerSpace& L = ctx->c.layers;
    const char* d_img = nullptr;
    auto it = L.images.find(img);
    if (it != L.images.end()) {
        d_img = it->second;
    } else {
        if (!L.dev || !L.host) return fail(FGX_ERR_STATE, "fgx_layer_launch: context has no layer image space");
        const size_t at = (L.used + 255) & ~(size_t)255;
        if (at + bytes > FGX_LAYER_SPACE_BYTES) return fail(FGX_ERR_CAPACITY, "fgx_layer_launch: layer image space full");
        std::memcpy(L.host + at, img.data(), bytes);
        FGX_HIP(hipMemcpyAsync(L.dev + at, L.host + at, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
        L.used = at + bytes;
        d_img = L.dev + at;
        L.images.emplace(std::move(img), d_img);
    }
    void* arg = const_cast<char*>(d_img);
    void* kargs[] = {&arg};
    FGX_HIP(hipLaunchKernel(kernel, dim3((uint32_t)grid), dim3(block), kargs, lds, (hipStream_t)stream));
    return FGX_OK;
}

int fgx_layer_space_info(fgx_ctx* ctx, uint32_t* images, size_t* bytes) {
    if (!ctx) return fail(FGX_ERR_ARG, "fgx_layer_space_info: NULL context");
    if (images) *images = (uint32_t)ctx->c.layers.images.size();
    if (bytes) *bytes = ctx->c.layers.used;
    return FGX_OK;
}

int fgx_sub_u32(uint32_t* dst, const uint32_t* a, const uint32_t* b, void* stream) {
    if (!dst || !a || !b) return fail(FGX_ERR_ARG, "fgx_sub_u32: NULL pointer");
    hipLaunchKernelGGL(k_sub_u32, dim3(1), dim3(64), 0, (hipStream_t)stream, dst, a, b);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_add_u32_n(uint32_t* p, uint32_t n, uint32_t delta, void* stream) {
    if (!n) return FGX_OK;
    if (!p) return fail(FGX_ERR_ARG, "fgx_add_u32_n: NULL pointer");
    hipLaunchKernelGGL(k_add_u32_n, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, p, n, delta);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_copy_ranges(uint32_t nvars, const void* const* src, void* const* dst, const uint32_t* var_bytes,
                    const uint32_t* d_ranges, uint32_t nranges, uint32_t n_total, void* stream) {
    if (!n_total || !nranges) return FGX_OK;
    if (!d_ranges) return fail(FGX_ERR_ARG, "fgx_copy_ranges: NULL ranges");
    VarTable vt;
    int rc = make_var_table(nvars, src, dst, var_bytes, &vt);
    if (rc) return rc;
    const uint32_t blocks = std::min<uint32_t>(ceil_div(n_total, 256), 8192);
    hipLaunchKernelGGL(k_copy_ranges, dim3(blocks), dim3(256), 0, (hipStream_t)stream, vt, d_ranges, nranges, n_total);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_seg_layout(uint32_t nsegs, uint32_t* d_cnt, const uint32_t* d_add, uint32_t* d_off, uint32_t* d_total,
                   const fgx_seg_target* targets, uint32_t ntargets, void* stream) {
    if (!d_cnt || !d_off || !d_total) return fail(FGX_ERR_ARG, "fgx_seg_layout: NULL count, offset or total array");
    if (ntargets > FGX_SEG_MAX_TARGETS) return fail(FGX_ERR_ARG, "fgx_seg_layout: too many targets");
    if (ntargets && !targets) return fail(FGX_ERR_ARG, "fgx_seg_layout: NULL targets");
    SegTargets tg;
    std::memset(&tg, 0, sizeof(tg));
    tg.n = ntargets;
    for (uint32_t k = 0; k < ntargets; ++k) {
        const fgx_seg_target& t = targets[k];
        if (!t.table || t.stride < 4) return fail(FGX_ERR_ARG, "fgx_seg_layout: target without table or stride");
        for (int32_t f : {t.off_field, t.n_field, t.tile_field})
            if (f >= 0 && ((uint32_t)f + 4u > t.stride || (f & 3))) return fail(FGX_ERR_ARG, "fgx_seg_layout: field outside its record");
        if (t.tile_field >= 0 && !t.tile_elems) return fail(FGX_ERR_ARG, "fgx_seg_layout: tile field without tile size");
        tg.t[k] = t;
    }
    hipLaunchKernelGGL(k_seg_layout, dim3(1), dim3(kSegBlock), 0, (hipStream_t)stream, nsegs, d_cnt, d_add, d_off, d_total, tg);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_seg_count(uint32_t nsegs, const uint32_t* d_off, const uint32_t* d_cnt, const uint32_t* flags, uint32_t keep,
                  uint32_t* d_out, void* stream) {
    if (!nsegs) return FGX_OK;
    if (!d_off || !d_cnt || !flags || !d_out) return fail(FGX_ERR_ARG, "fgx_seg_count: NULL argument");
    hipLaunchKernelGGL(k_seg_count, dim3(nsegs), dim3(256), 0, (hipStream_t)stream, d_off, d_cnt, flags, keep, d_out);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_seg_append(uint32_t nsegs, const uint32_t* d_off, const uint32_t* d_cnt, const uint32_t* d_born, uint32_t n_cap_out,
                   uint32_t nvars, const void* const* list_in, const void* const* born_in, void* const* list_out,
                   const uint32_t* var_bytes, int32_t id_var, uint32_t* d_next_id, uint32_t* d_work, void* stream) {
    if (!nsegs) return FGX_OK;
    if (!d_off || !d_cnt || !d_born || !d_work) return fail(FGX_ERR_ARG, "fgx_seg_append: NULL table");
    if (id_var >= (int32_t)nvars) return fail(FGX_ERR_ARG, "fgx_seg_append: ID variable out of range");
    if (id_var >= 0 && (!d_next_id || var_bytes[id_var] != 4)) return fail(FGX_ERR_ARG, "fgx_seg_append: IDs need next-ID counters and 4-byte IDs");
    VarTable list, born;
    int rc = make_var_table(nvars, list_in, list_out, var_bytes, &list);
    if (rc) return rc;
    rc = make_var_table(nvars, born_in, list_out, var_bytes, &born);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_seg_append_scan, dim3(1), dim3(kSegBlock), 0, st, nsegs, d_cnt, d_born, d_work);
    FGX_HIP(hipGetLastError());
    if (n_cap_out) {
        const uint32_t blocks = std::min<uint32_t>(ceil_div(n_cap_out, 256), 8192);
        hipLaunchKernelGGL(k_seg_append_copy, dim3(blocks), dim3(256), 0, st, nsegs, d_off, d_cnt, d_work, n_cap_out, list,
                           born, id_var, d_next_id);
        FGX_HIP(hipGetLastError());
    }
    if (id_var >= 0) {
        hipLaunchKernelGGL(k_add_n_u32, dim3(ceil_div(nsegs, 256)), dim3(256), 0, st, d_next_id, d_born, nsegs);
        FGX_HIP(hipGetLastError());
    }
    return FGX_OK;
}

static_assert(sizeof(fgx_seg_key) == 80, "fgx_seg_key layout is part of the C ABI");

int fgx_seg_key_init(int kind, int dims, const float* mn, const float* mx, float radius, fgx_seg_key* out) {
    if (!out || !mn || !mx) return fail(FGX_ERR_ARG, "fgx_seg_key_init: NULL argument");
    if (kind != FGX_SEG_MESSAGE_BINS && kind != FGX_SEG_AGENT_SORT) return fail(FGX_ERR_ARG, "fgx_seg_key_init: unknown kind");
    if (dims != 2 && dims != 3) return fail(FGX_ERR_ARG, "fgx_seg_key_init: dims must be 2 or 3");
    out->dims = dims;
    out->radius = radius;
    if (kind == FGX_SEG_MESSAGE_BINS) {
        fgx_spatial_meta m;
        int rc = fgx_spatial_meta_init(dims, mn, mx, radius, &m);
        if (rc) return rc;
        for (int a = 0; a < 3; ++a) {
            out->min[a] = m.min[a];
            out->width[a] = m.width[a];
            out->grid[a] = m.grid[a] ? m.grid[a] : 1u;
        }
        out->keys = m.bins;
        out->key_mask = 0xFFFFFFFFu;
    } else {
        uint32_t bits = 0;
        int rc = fgx_agent_sort_key_bits(dims, mn, mx, radius, &bits);
        if (rc) return rc;
        if (bits > 30) return fail(FGX_ERR_ARG, "fgx_seg_key_init: key range too large");
        for (int a = 0; a < 3; ++a) {
            if (a < dims) {
                out->min[a] = mn[a];
                out->width[a] = mx[a] - mn[a];
                out->grid[a] = out->width[a] != 0.0f ? (uint32_t)std::ceil(out->width[a] / radius) : 1u;
            } else {
                out->min[a] = 0.0f;
                out->width[a] = 1.0f;
                out->grid[a] = 1u;
            }
        }
        out->keys = 1u << bits;
        out->key_mask = out->keys - 1u;
    }
    return FGX_OK;
}

int fgx_segments_pack(fgx_seg_key* segs, uint32_t nsegs, uint32_t* n_total, uint32_t* keys_total, uint32_t* tiles_total) {
    if (nsegs && !segs) return fail(FGX_ERR_ARG, "fgx_segments_pack: NULL segments");
    uint64_t n = 0, k = 0, t = 0;
    for (uint32_t s = 0; s < nsegs; ++s) {
        segs[s].elem_off = (uint32_t)n;
        segs[s].key_off = (uint32_t)k;
        segs[s].tile_off = (uint32_t)t;
        n += segs[s].n;
        k += segs[s].keys;
        t += ceil_div(segs[s].n, 2048u);
        if (n >= (1ull << 31) || k >= (1ull << 31)) return fail(FGX_ERR_ARG, "fgx_segments_pack: batch too large");
    }
    if (n_total) *n_total = (uint32_t)n;
    if (keys_total) *keys_total = (uint32_t)k;
    if (tiles_total) *tiles_total = (uint32_t)t;
    return FGX_OK;
}

int fgx_assign_ids(uint32_t* ids, uint32_t n, uint32_t next_id, void* stream) {
    if (!n) return FGX_OK;
    hipLaunchKernelGGL(k_assign_ids, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, ids, n, next_id);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// MessageSpatial3D.cu:31-52 / MessageSpatial2D.cu:36-54 and numeric.h:12-20.
static bool approx_exactly_divisible(float x, float y) {
    const float eps = std::max(std::fabs(x), std::fabs(y)) * std::numeric_limits<float>::epsilon();
    const float v = std::fmod(x, y);
    return v <= eps || v > y - eps;
}

int fgx_spatial_meta_init_mult(int dims, const float* mn, const float* mx, float radius, float mult,
                               fgx_spatial_meta* m) {
    if (dims != 2 && dims != 3) return fail(FGX_ERR_ARG, "spatial messages are 2-D or 3-D");
    if (!(radius > 0.0f)) return fail(FGX_ERR_ARG, "spatial message radius must be > 0");
    if (!(mult == 1.0f || mult == 0.5f)) return fail(FGX_ERR_ARG, "bin-width multiplier must be 1 or 0.5");
    std::memset(m, 0, sizeof(*m));
    m->dims = dims;
    m->radius = radius;
    m->cell = radius * mult;
    m->span = mult == 1.0f ? 1 : 2;
    m->bins = 1;
    m->wrap_ok = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < dims) {
            if (!(mx[a] > mn[a])) return fail(FGX_ERR_ARG, "spatial message max must exceed min on every axis");
            m->min[a] = mn[a];
            m->max[a] = mx[a];
            m->width[a] = mx[a] - mn[a];
            const float q = m->width[a] / m->cell;
            m->grid[a] = (uint32_t)std::ceil(q);
            const uint64_t bins = (uint64_t)m->bins * m->grid[a];
            if (bins >= (1ull << 31)) return fail(FGX_ERR_ARG, "spatial grid too large");
            m->bins = (uint32_t)bins;
            m->wrap_ok = m->wrap_ok && approx_exactly_divisible(m->width[a], m->cell);
        } else {
            m->grid[a] = 1;
        }
    }
    return FGX_OK;
}

int fgx_spatial_meta_init(int dims, const float* mn, const float* mx, float radius, fgx_spatial_meta* m) {
    return fgx_spatial_meta_init_mult(dims, mn, mx, radius, 1.0f, m);
}

int fgx_agent_sort_key_bits(int dims, const float* mn, const float* mx, float radius, uint32_t* key_bits) {
    uint64_t bins = 1;
    for (int a = 0; a < dims; ++a) {
        const float w = mx[a] - mn[a];
        bins *= w != 0.0f ? (uint64_t)std::ceil(w / radius) : 1ull;
    }
    if (bins >= (1ull << 31)) return fail(FGX_ERR_ARG, "agent sort grid too large");
    *key_bits = (uint32_t)std::floor(std::log2((double)bins)) + 1u;
    return FGX_OK;
}

}  // extern "C"
