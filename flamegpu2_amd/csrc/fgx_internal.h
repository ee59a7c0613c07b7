// Internal helpers shared by the libfgx translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <mutex>
#include <unordered_map>

#include "../../include/fgx.h"

namespace fgx {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define FGX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::fgx::fail(FGX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device space for fused-layer images: immutable once written, found by content.
struct LayerSpace {
    char* dev = nullptr;        // FGX_LAYER_SPACE_BYTES of device memory
    char* host = nullptr;       // pinned staging copy (the source of each image's upload)
    size_t used = 0;
    std::unordered_map<std::string, const char*> images;
};

struct Ctx {
    int device = 0;
    int cu_count = 0;
    int xcd_count = 8;
    size_t hbm_bytes = 0;
    // FGX_OPT_KEEP_SCRATCH_ZERO: bin sorts leave their histogram zeroed
    int keep_zero = 0;
    LayerSpace layers;
};

// Host-side record, per scratch buffer, of the bin count whose histogram region a sort
// left zero (kAllZero after fgx_scratch_clear).  Consulted only with FGX_OPT_KEEP_SCRATCH_ZERO.
constexpr uint32_t kAllZero = 0xFFFFFFFFu;
std::mutex& clean_mutex();
std::unordered_map<const void*, uint32_t>& clean_record();
// Per scratch buffer: the one bin count every sort of it has used so far (kMixedBins once
// two differ).  Consulted only with FGX_OPT_KEEP_SCRATCH_ZERO.
constexpr uint32_t kMixedBins = 0xFFFFFFFEu;
std::unordered_map<const void*, uint32_t>& bins_history();

// ---------------------------------------------------------------- variable tables
// Passed by value in the kernel argument segment, so every field is wave-uniform
// and read through the scalar unit.
struct VarTable {
    uint32_t n;
    uint32_t bytes[FGX_MAX_VARS];
    const char* in[FGX_MAX_VARS];
    char* out[FGX_MAX_VARS];
    // optional 16-byte records written beside the variables (fgx_spatial_build_records):
    // word k of record i is 4-byte variable rec_var[k] of element i, or 0 when it is -1
    char* rec;
    int32_t rec_var[4];
};

int make_var_table(uint32_t nvars, const void* const* in, void* const* out, const uint32_t* bytes, VarTable* t);

// Copy element `src_i` of every variable (except `skip`) to element `dst_i` of its output array.
__device__ __forceinline__ void copy_element(const VarTable& t, uint32_t src_i, uint32_t dst_i, int32_t skip = -1) {
    for (uint32_t v = 0; v < t.n; ++v) {
        if ((int32_t)v == skip) continue;
        const uint32_t b = t.bytes[v];
        const char* s = t.in[v] + (size_t)src_i * b;
        char* d = t.out[v] + (size_t)dst_i * b;
        if (b == 4) {
            *reinterpret_cast<uint32_t*>(d) = *reinterpret_cast<const uint32_t*>(s);
        } else if (b == 8) {
            *reinterpret_cast<uint2*>(d) = *reinterpret_cast<const uint2*>(s);
        } else if ((b & 3u) == 0) {
            for (uint32_t w = 0; w < b; w += 4)
                *reinterpret_cast<uint32_t*>(d + w) = *reinterpret_cast<const uint32_t*>(s + w);
        } else {
            for (uint32_t w = 0; w < b; ++w) d[w] = s[w];
        }
    }
}

// Record dst_i of vt.rec from element src_i of the input variables (no-op without records).
__device__ __forceinline__ void write_record(const VarTable& t, uint32_t src_i, uint32_t dst_i) {
    if (!t.rec) return;
    uint4 r;
    r.x = t.rec_var[0] >= 0 ? reinterpret_cast<const uint32_t*>(t.in[t.rec_var[0]])[src_i] : 0u;
    r.y = t.rec_var[1] >= 0 ? reinterpret_cast<const uint32_t*>(t.in[t.rec_var[1]])[src_i] : 0u;
    r.z = t.rec_var[2] >= 0 ? reinterpret_cast<const uint32_t*>(t.in[t.rec_var[2]])[src_i] : 0u;
    r.w = t.rec_var[3] >= 0 ? reinterpret_cast<const uint32_t*>(t.in[t.rec_var[3]])[src_i] : 0u;
    reinterpret_cast<uint4*>(t.rec)[dst_i] = r;
}

// ---------------------------------------------------------------- XCD-contiguous block order
// Workgroups go to the 8 XCDs round-robin (workgroup w runs on XCD w % 8), each with
// its own L2.  Renumbered, XCD x runs one contiguous range of logical blocks, in order
// (q + (x < r) of them, q = G / 8, r = G % 8, from x*q + min(x, r)), so neighbouring
// elements, bins and scatter targets share an L2.
__device__ __forceinline__ uint32_t xcd_block() {
    constexpr uint32_t kXcds = 8;
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t x = b % kXcds, k = b / kXcds, q = g / kXcds, r = g % kXcds;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------- wave64 helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_u32(uint32_t v) {
    const uint32_t lane = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// ---------------------------------------------------------------- scan / compaction scratch
// Scans and compactions run as passes ordered by kernel boundaries, with no waits
// between blocks.  Their scratch starts with this 256-byte header (unused words kept
// so that the per-tile words after it stay 256-byte aligned), then 8 bytes per tile:
// tile sums, or a compaction's stashed inputs and per-tile counts.
// Looks up, on the current device, the address of the library's own build working space
// (used by fgx_spatial_build_noalloc); called by fgx_ctx_create once per device.
int noalloc_scratch_init(int device);

struct ScratchHeader {
    uint32_t pad[64];
};
static_assert(sizeof(ScratchHeader) == 256, "header keeps the per-tile words 256-B aligned");

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Launch helpers
inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace fgx

struct fgx_ctx {
    fgx::Ctx c;
};
