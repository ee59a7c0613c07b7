// Internal helpers shared by the libfgx translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>

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

struct Ctx {
    int device = 0;
    int cu_count = 0;
    int xcd_count = 8;
    size_t hbm_bytes = 0;
};

// ---------------------------------------------------------------- variable tables
// Passed by value in the kernel argument segment, so every field is wave-uniform
// and read through the scalar unit.
struct VarTable {
    uint32_t n;
    uint32_t bytes[FGX_MAX_VARS];
    const char* in[FGX_MAX_VARS];
    char* out[FGX_MAX_VARS];
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

// ---------------------------------------------------------------- decoupled look-back
// Tile t is block t.  Workgroups are dispatched in index order on each XCD, so
// the lowest unfinished tile is always resident and the chain of waits makes
// progress.  Tile status words pack {flag:2, epoch:30, value:32}; the epoch
// advances when the last block of a launch retires, so words from an older
// launch read as "not yet published" and the status array never needs
// clearing between launches (graph replays included).  The header must be
// zeroed once before first use.  Spins are bounded: a tile that waits too long
// records an error in the header instead of hanging the device.
struct LookbackHeader {
    uint32_t done;
    uint32_t epoch;
    uint32_t error;
    uint32_t pad[61];
};
static_assert(sizeof(LookbackHeader) == 256, "header keeps the status array 256-B aligned");

constexpr uint64_t kFlagAggregate = 1ull;
constexpr uint64_t kFlagInclusive = 2ull;

__device__ __forceinline__ uint64_t pack_status(uint64_t flag, uint32_t epoch, uint32_t value) {
    return (flag << 62) | ((uint64_t)(epoch & 0x3FFFFFFFu) << 32) | value;
}

__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by every thread of the block at kernel start: the tile is blockIdx.x,
// the epoch is read once and broadcast through shared memory.
__device__ __forceinline__ void acquire_tile(LookbackHeader* hdr, uint32_t num_tiles, uint32_t* s_tile, uint32_t* s_epoch) {
    (void)num_tiles;
    if (threadIdx.x == 0) {
        const uint32_t ep = __hip_atomic_load(&hdr->epoch, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        *s_tile = blockIdx.x;
        *s_epoch = ep & 0x3FFFFFFFu;
    }
    __syncthreads();
}

// Called by every thread of the block after its last status access.  The block
// that retires last (all others have read the epoch and finished) advances it.
__device__ __forceinline__ void retire_tile(LookbackHeader* hdr, uint32_t num_tiles) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t d = __hip_atomic_fetch_add(&hdr->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == num_tiles - 1) {
            __hip_atomic_store(&hdr->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&hdr->epoch, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

constexpr uint32_t kMaxSpins = 1u << 22;

// Exclusive prefix of `aggregate` over tiles 0..tile-1.  Executed by wave 0 of
// the block (all 64 lanes, uniform control flow); returns the prefix in every lane.
__device__ __forceinline__ uint32_t lookback_wave0(uint64_t* status, uint32_t tile, uint32_t epoch, uint32_t aggregate,
                                                   LookbackHeader* hdr) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) status_store(&status[0], pack_status(kFlagInclusive, epoch, aggregate));
        return 0;
    }
    if (lane == 0) status_store(&status[tile], pack_status(kFlagAggregate, epoch, aggregate));
    uint32_t excl = 0;
    int64_t top = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (true) {
        const int64_t idx = top - (int64_t)lane;
        uint64_t s;
        uint32_t flag;
        if (idx >= 0) {
            s = status_load(&status[idx]);
            flag = (((s >> 32) & 0x3FFFFFFFu) == epoch) ? (uint32_t)(s >> 62) : 0u;
        } else {
            s = 0;
            flag = (uint32_t)kFlagInclusive;   // before tile 0: inclusive prefix 0
        }
        const uint64_t incl = __ballot(flag == kFlagInclusive);
        const uint64_t valid = __ballot(flag != 0);
        const int j = incl ? (int)__builtin_ctzll(incl) : 63;
        const uint64_t need = (j == 63) ? ~0ull : ((2ull << j) - 1ull);
        if ((valid & need) != need) {
            if (++spins > kMaxSpins) {
                if (lane == 0) __hip_atomic_store(&hdr->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += wave_sum_u32((int)lane <= j ? (uint32_t)s : 0u);
        if (incl) break;
        top -= 64;
    }
    if (lane == 0) status_store(&status[tile], pack_status(kFlagInclusive, epoch, excl + aggregate));
    return excl;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Launch helpers
inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace fgx

struct fgx_ctx {
    fgx::Ctx c;
};
