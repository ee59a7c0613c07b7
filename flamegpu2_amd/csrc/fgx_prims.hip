// libfgx: hot-path primitives for gfx950 (wave64).
//
//   message-list build : bin key + run-aggregated wave64 histogram -> exclusive scan in
//                        passes ordered by kernel boundaries (tile sums, then tile
//                        prefixes) -> bin start table -> index scatter (bins grouped) ->
//                        in-bin order restored by index (large bins sorted by one block
//                        each) -> SoA payload scatter
//   agent spatial sort : same pipeline with the unclamped sort key
//   compaction/births  : kept count per tile, then per-tile prefixes, ballot ranks and
//                        the copy of every variable (two passes, no inter-block waits)
//   reduction          : fixed-shape (deterministic) tree
//
// Nothing here allocates or synchronises with the host.
#include "fgx_internal.h"

#include <algorithm>
#include <cstring>

namespace fgx {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kItems = 8;
constexpr uint32_t kTile = kBlock * kItems;   // 2048 items per scan / compaction tile

__device__ __forceinline__ uint32_t load_count(const uint32_t* d_n, uint32_t n_cap) {
    if (!d_n) return n_cap;
    const uint32_t n = __hip_atomic_load(d_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return n < n_cap ? n : n_cap;
}

// ================================================================ bin keys
struct KeyArgs {
    int32_t mode;        // 0: message bin (clamped, MessageSpatial3DDevice.cuh:646-672)
                         // 1: agent sort key (unclamped, CUDASimulation.cu:372-408)
                         // 2: bucket key, x holds int32 keys (MessageBucket.cu:58-72)
    int32_t dims;
    float min[3];
    float radius;
    float width[3];
    uint32_t grid[3];
    uint32_t key_mask;   // mode 1: (1 << key_bits) - 1
    int32_t key_min;     // mode 2: lowest key
    uint32_t keys;       // mode 2: number of keys
};

__device__ __forceinline__ uint32_t bin_key(const KeyArgs& k, float x, float y, float z) {
    if (k.mode == 2) {
        // keys outside the range were reported by setKey; clamp so they stay in the table
        const int v = __float_as_int(x) - k.key_min;
        return v < 0 ? 0u : ((uint32_t)v >= k.keys ? k.keys - 1u : (uint32_t)v);
    }
    if (k.mode == 0) {
        int c[3];
        const float p[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
            if (a < k.dims) {
                const int g = (int)floorf((p[a] - k.min[a]) / k.radius);
                const int hi = (int)k.grid[a] - 1;
                c[a] = g < 0 ? 0 : (g > hi ? hi : g);
            } else {
                c[a] = 0;
            }
        }
        return (uint32_t)c[2] * k.grid[0] * k.grid[1] + (uint32_t)c[1] * k.grid[0] + (uint32_t)c[0];
    } else {
        const int gx = (int)floorf(((x - k.min[0]) / k.width[0]) * (float)k.grid[0]);
        const int gy = (int)floorf(((y - k.min[1]) / k.width[1]) * (float)k.grid[1]);
        uint32_t b;
        if (k.dims == 3) {
            const int gz = (int)floorf(((z - k.min[2]) / k.width[2]) * (float)k.grid[2]);
            b = (uint32_t)gz * k.grid[0] * k.grid[1] + (uint32_t)gy * k.grid[0] + (uint32_t)gx;
        } else {
            b = (uint32_t)gy * k.grid[0] + (uint32_t)gx;
        }
        return b & k.key_mask;
    }
}

// Histogram with wave-level run aggregation: consecutive lanes with equal keys
// (the common case, as agents are kept in bin order) issue one atomic per run.
// sub[i] is the slot of message i inside its bin (atomic order, fixed later).
// *unsorted is set when some key is smaller than its predecessor's: when it stays 0
// the input is already in bin order and the scatters reduce to a copy.
__global__ void __launch_bounds__(kBlock) k_bin_hist(KeyArgs ka, const uint32_t* d_n, uint32_t n_cap,
                                                     const float* __restrict__ x, const float* __restrict__ y,
                                                     const float* __restrict__ z, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ sub, uint32_t* __restrict__ hist,
                                                     uint32_t* __restrict__ unsorted) {
    __shared__ uint32_t s_down;
    if (threadIdx.x == 0) s_down = 0u;
    __syncthreads();
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t lane = lane_id();
    const uint32_t wave = (xcd_block() * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t base = wave * 64u; base < n; base += nwaves * 64u) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        uint32_t k = 0xFFFFFFFFu;
        if (valid) k = bin_key(ka, x[i], y[i], ka.dims == 3 ? z[i] : 0.0f);
        const uint32_t prev = __shfl_up(k, 1, 64);
        {
            uint32_t before = prev;
            if (lane == 0 && base > 0) before = bin_key(ka, x[base - 1], y[base - 1], ka.dims == 3 ? z[base - 1] : 0.0f);
            const bool down = valid && (lane > 0 || base > 0) && k < before;
            if (__ballot(down) != 0ull && lane == 0) s_down = 1u;
        }
        const bool head = valid && (lane == 0 || k != prev);
        const uint64_t hm = __ballot(head);
        const uint64_t vm = __ballot(valid);
        const uint32_t nvalid = (uint32_t)__builtin_popcountll(vm);
        const uint64_t above = lane == 63 ? 0ull : (hm & (~0ull << (lane + 1)));
        const uint32_t end = above ? (uint32_t)__builtin_ctzll(above) : nvalid;
        uint32_t b = 0;
        if (head) b = atomicAdd(&hist[k], end - lane);
        const uint64_t below = hm & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const uint32_t hl = below ? 63u - (uint32_t)__builtin_clzll(below) : 0u;
        b = __shfl(b, (int)hl, 64);
        if (valid) {
            key[i] = k;
            sub[i] = b + (lane - hl);
        }
    }
    // one flag update per block, skipped once another block has set it
    __syncthreads();
    if (threadIdx.x == 0 && s_down && !__hip_atomic_load(unsorted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicOr(unsorted, 1u);
}

// Segmented variant: block t is tile t of the packed batch; its simulation s (last
// with tile_off <= t, a scalar binary search) gives the geometry and the key offset.
// Tiles never straddle simulations, so run aggregation stays within one simulation.
__global__ void __launch_bounds__(kBlock) k_bin_hist_seg(int32_t mode, const fgx_seg_key* __restrict__ segs,
                                                         uint32_t nsegs, const float* __restrict__ x,
                                                         const float* __restrict__ y, const float* __restrict__ z,
                                                         uint32_t* __restrict__ key, uint32_t* __restrict__ sub,
                                                         uint32_t* __restrict__ hist, uint32_t* __restrict__ unsorted) {
    __shared__ uint32_t s_down;
    if (threadIdx.x == 0) s_down = 0u;
    __syncthreads();
    const uint32_t t = xcd_block();   // tiles in XCD-contiguous order
    uint32_t lo = 0, hi = nsegs;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs[mid].tile_off <= t) lo = mid;
        else hi = mid;
    }
    const fgx_seg_key& sg = segs[lo];
    KeyArgs ka;
    ka.mode = mode;
    ka.dims = sg.dims;
    ka.radius = sg.radius;
    ka.key_mask = sg.key_mask;
    for (int a = 0; a < 3; ++a) {
        ka.min[a] = sg.min[a];
        ka.width[a] = sg.width[a];
        ka.grid[a] = sg.grid[a];
    }
    const uint32_t first = sg.elem_off + (t - sg.tile_off) * kTile;
    const uint32_t end = sg.elem_off + sg.n;
    const uint32_t koff = sg.key_off;
    const uint32_t lane = lane_id();
    for (uint32_t c = threadIdx.x >> 6; c < kTile / 64u; c += kBlock / 64u) {
        const uint32_t i = first + c * 64u + lane;
        const bool valid = i < end;
        if (__ballot(valid) == 0ull) break;
        uint32_t k = 0xFFFFFFFFu;
        if (valid) k = koff + bin_key(ka, x[i], y[i], ka.dims == 3 ? z[i] : 0.0f);
        const uint32_t prev = __shfl_up(k, 1, 64);
        {
            // the predecessor of a simulation's first element belongs to an earlier one,
            // whose keys are all smaller
            const uint32_t c0 = first + c * 64u;
            const bool has_pred = c0 > sg.elem_off;
            uint32_t before = prev;
            if (lane == 0 && has_pred) before = koff + bin_key(ka, x[c0 - 1], y[c0 - 1], ka.dims == 3 ? z[c0 - 1] : 0.0f);
            const bool down = valid && (lane > 0 || has_pred) && k < before;
            if (__ballot(down) != 0ull && lane == 0) s_down = 1u;
        }
        const bool head = valid && (lane == 0 || k != prev);
        const uint64_t hm = __ballot(head);
        const uint64_t vm = __ballot(valid);
        const uint32_t nvalid = (uint32_t)__builtin_popcountll(vm);
        const uint64_t above = lane == 63 ? 0ull : (hm & (~0ull << (lane + 1)));
        const uint32_t run_end = above ? (uint32_t)__builtin_ctzll(above) : nvalid;
        uint32_t b = 0;
        if (head) b = atomicAdd(&hist[k], run_end - lane);
        const uint64_t below = hm & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const uint32_t hl = below ? 63u - (uint32_t)__builtin_clzll(below) : 0u;
        b = __shfl(b, (int)hl, 64);
        if (valid) {
            key[i] = k;
            sub[i] = b + (lane - hl);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_down && !__hip_atomic_load(unsorted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicOr(unsorted, 1u);
}

__global__ void __launch_bounds__(kBlock) k_index_scatter(const uint32_t* d_n, uint32_t n_cap,
                                                          const uint32_t* __restrict__ key, const uint32_t* __restrict__ sub,
                                                          const uint32_t* __restrict__ start, uint32_t* __restrict__ tmp,
                                                          const uint32_t* __restrict__ unsorted) {
    if (!*unsorted) return;   // in bin order already: the permutation is the identity
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += stride)
        tmp[start[key[i]] + sub[i]] = i;
}

// ---------------------------------------------------------------- in-bin order
// After k_index_scatter every bin's members sit in its slots in atomic order.  The
// stable order is ascending input index.  Bins of at most kSmallBin members are put in
// order by k_stable_scatter itself (each slot counts the smaller members of its bin);
// larger bins are ordered first by k_sort_big_bins, one block per bin, into `sorted`:
//   * up to kLdsBin members: the bin is staged in LDS and each member counts the smaller
//     ones (every thread reads the same LDS word at once, a broadcast);
//   * beyond: a bitmap over a window of kMapWords * 32 input indices is set in LDS, and a
//     member's rank is the number of set bits below its own (per-8-word prefix counts
//     plus at most 8 word popcounts), window by window over [min, max] of the bin.
// Work per bin is O(n_b) per window, so a build costs O(n) however the bins are filled.
constexpr uint32_t kSmallBin = 32;
constexpr uint32_t kLdsBin = 2 * kBlock;
constexpr uint32_t kMapWords = 16384;
constexpr uint32_t kMapGroups = kMapWords / 8;

__device__ __forceinline__ void block_min_max(uint32_t& lo, uint32_t& hi, uint32_t* s_red) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_red[w] = lo;
        s_red[kBlock / 64 + w] = hi;
    }
    __syncthreads();
    lo = s_red[0];
    hi = s_red[kBlock / 64];
    for (uint32_t k = 1; k < kBlock / 64; ++k) {
        lo = s_red[k] < lo ? s_red[k] : lo;
        hi = s_red[kBlock / 64 + k] > hi ? s_red[kBlock / 64 + k] : hi;
    }
    __syncthreads();
}

// Exclusive block scan of one value per thread; *total receives the block's sum.
__device__ __forceinline__ uint32_t block_exclusive_u32(uint32_t v, uint32_t* s_red, uint32_t* total) {
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_u32(v);
    if (lane_id() == 63) s_red[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t k = 0; k < kBlock / 64; ++k) {
        off += k < w ? s_red[k] : 0u;
        tot += s_red[k];
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// Block-wide: members tmp[st, en) of one bin (en - st > kSmallBin) to sorted[st, en) in
// ascending order.  Ends with a barrier, so LDS is free again for the next bin.
__device__ void sort_one_bin(uint32_t st, uint32_t en, const uint32_t* __restrict__ tmp, uint32_t* __restrict__ sorted,
                             uint32_t* s_map, uint32_t* s_gpre, uint32_t* s_red) {
    const uint32_t nb = en - st, tid = threadIdx.x;
    if (nb <= kLdsBin) {
        for (uint32_t i = tid; i < nb; i += kBlock) s_map[i] = tmp[st + i];
        __syncthreads();
        const uint32_t e0 = tid < nb ? s_map[tid] : 0xFFFFFFFFu;
        const uint32_t e1 = tid + kBlock < nb ? s_map[tid + kBlock] : 0xFFFFFFFFu;
        uint32_t r0 = 0, r1 = 0;
        for (uint32_t t = 0; t < nb; ++t) {
            const uint32_t v = s_map[t];
            r0 += v < e0 ? 1u : 0u;
            r1 += v < e1 ? 1u : 0u;
        }
        if (tid < nb) sorted[st + r0] = e0;
        if (tid + kBlock < nb) sorted[st + r1] = e1;
        __syncthreads();
        return;
    }
    // Member passes load kUnroll members per thread before using them, so that one block
    // keeps enough loads in flight to stream a large bin.
    constexpr uint32_t kUnroll = 8;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t i0 = 0; i0 < nb; i0 += kBlock * kUnroll) {
        uint32_t v[kUnroll];
#pragma unroll
        for (uint32_t j = 0; j < kUnroll; ++j) {
            const uint32_t i = i0 + j * kBlock + tid;
            v[j] = i < nb ? tmp[st + i] : 0u;
        }
#pragma unroll
        for (uint32_t j = 0; j < kUnroll; ++j) {
            if (i0 + j * kBlock + tid < nb) {
                lo = v[j] < lo ? v[j] : lo;
                hi = v[j] > hi ? v[j] : hi;
            }
        }
    }
    block_min_max(lo, hi, s_red);
    constexpr uint64_t kSpan = (uint64_t)kMapWords * 32u;
    uint32_t running = 0;
    for (uint64_t w0 = lo; w0 <= hi; w0 += kSpan) {
        const uint32_t span = (uint32_t)((uint64_t)hi - w0 + 1u < kSpan ? (uint64_t)hi - w0 + 1u : kSpan);
        const uint32_t words = (span + 31u) >> 5;
        const uint32_t base = (uint32_t)w0;
        for (uint32_t i = tid; i < words; i += kBlock) s_map[i] = 0u;
        __syncthreads();
        for (uint32_t i0 = 0; i0 < nb; i0 += kBlock * kUnroll) {
            uint32_t v[kUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t i = i0 + j * kBlock + tid;
                v[j] = i < nb ? tmp[st + i] : 0u;
            }
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t off = v[j] - base;   // wraps above span for members below the window
                if (i0 + j * kBlock + tid < nb && off < span) atomicOr(&s_map[off >> 5], 1u << (off & 31u));
            }
        }
        __syncthreads();
        // thread t owns groups [8t, 8t + 8) of 8 words each
        uint32_t gs[8];
        uint32_t tsum = 0;
#pragma unroll
        for (uint32_t g = 0; g < 8; ++g) {
            uint32_t c = 0;
            const uint32_t w8 = (tid * 8u + g) * 8u;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) c += w8 + k < words ? (uint32_t)__builtin_popcount(s_map[w8 + k]) : 0u;
            gs[g] = c;
            tsum += c;
        }
        uint32_t total = 0;
        uint32_t run = block_exclusive_u32(tsum, s_red, &total);
#pragma unroll
        for (uint32_t g = 0; g < 8; ++g) {
            s_gpre[tid * 8u + g] = run;
            run += gs[g];
        }
        __syncthreads();
        for (uint32_t i0 = 0; i0 < nb; i0 += kBlock * kUnroll) {
            uint32_t v[kUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t i = i0 + j * kBlock + tid;
                v[j] = i < nb ? tmp[st + i] : 0u;
            }
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t off = v[j] - base;
                if (i0 + j * kBlock + tid >= nb || off >= span) continue;
                const uint32_t wi = off >> 5, gi = wi >> 3;
                uint32_t r = running + s_gpre[gi];
                for (uint32_t x = gi * 8u; x < wi; ++x) r += (uint32_t)__builtin_popcount(s_map[x]);
                r += (uint32_t)__builtin_popcount(s_map[wi] & ((1u << (off & 31u)) - 1u));
                sorted[st + r] = v[j];
            }
        }
        running += total;
        __syncthreads();
    }
}

// Block k handles the bins b = k, k + G, k + 2G, ... (G = gridDim.x), so that the few
// large bins of a small table land on different blocks.  Each block queues its bins with
// more than kSmallBin members in LDS and orders them one after another.
__global__ void __launch_bounds__(kBlock) k_sort_big_bins(const uint32_t* __restrict__ start, uint32_t bins,
                                                          const uint32_t* __restrict__ tmp, uint32_t* __restrict__ sorted,
                                                          const uint32_t* __restrict__ unsorted,
                                                          const uint32_t* __restrict__ tile_big, uint32_t ntiles) {
    __shared__ uint32_t s_map[kMapWords];
    __shared__ uint32_t s_gpre[kMapGroups];
    __shared__ uint32_t s_q[kBlock];
    __shared__ uint32_t s_red[2 * (kBlock / 64)];
    __shared__ uint32_t s_qn;
    if (!*unsorted) return;
    {
        // no tile has a large bin: nothing to order here
        uint32_t any = 0;
        for (uint32_t t = threadIdx.x; t < ntiles; t += kBlock) any |= tile_big[t];
        if (__syncthreads_or(any) == 0) return;
    }
    const uint32_t G = gridDim.x, k = blockIdx.x;
    const uint32_t mine = k < bins ? (bins - k + G - 1u) / G : 0u;
    for (uint32_t j0 = 0; j0 < mine; j0 += kBlock) {
        if (threadIdx.x == 0) s_qn = 0u;
        __syncthreads();
        const uint32_t j = j0 + threadIdx.x;
        if (j < mine) {
            const uint32_t b = j * G + k;
            if (start[b + 1] - start[b] > kSmallBin) s_q[atomicAdd(&s_qn, 1u)] = b;
        }
        __syncthreads();
        const uint32_t qn = s_qn;
        for (uint32_t q = 0; q < qn; ++q) {
            const uint32_t b = s_q[q];
            sort_one_bin(start[b], start[b + 1], tmp, sorted, s_map, s_gpre, s_red);
        }
        __syncthreads();
    }
}

// For output slot s: the message e = tmp[s] sits in bin b = key[e].  In a bin of at most
// kSmallBin members its stable rank is the number of bin members with a smaller input
// index, and every variable of e goes to start[b] + rank; a larger bin was ordered by
// k_sort_big_bins, and slot s takes sorted[s].
__global__ void __launch_bounds__(kBlock) k_stable_scatter(const uint32_t* d_n, uint32_t n_cap,
                                                           const uint32_t* __restrict__ key,
                                                           const uint32_t* __restrict__ start,
                                                           const uint32_t* __restrict__ tmp,
                                                           const uint32_t* __restrict__ sorted, VarTable vt,
                                                           const uint32_t* __restrict__ unsorted) {
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    if (!*unsorted) {
        for (uint32_t s = xcd_block() * blockDim.x + threadIdx.x; s < n; s += stride) {
            copy_element(vt, s, s);
            write_record(vt, s, s);
        }
        return;
    }
    for (uint32_t s = xcd_block() * blockDim.x + threadIdx.x; s < n; s += stride) {
        const uint32_t e = tmp[s];
        const uint32_t b = key[e];
        const uint32_t st = start[b];
        const uint32_t en = start[b + 1];
        if (en - st > kSmallBin) {
            copy_element(vt, sorted[s], s);
            write_record(vt, sorted[s], s);
            continue;
        }
        uint32_t rank = 0;
        for (uint32_t t = st; t < en; ++t) rank += tmp[t] < e ? 1u : 0u;
        copy_element(vt, e, st + rank);
        write_record(vt, e, st + rank);
    }
}

// 16-byte records of four 4-byte variables (fgx_pack_records); NULL sources give 0 words.
__global__ void __launch_bounds__(kBlock) k_pack_records(const uint32_t* d_n, uint32_t n_cap,
                                                         const uint32_t* __restrict__ w0, const uint32_t* __restrict__ w1,
                                                         const uint32_t* __restrict__ w2, const uint32_t* __restrict__ w3,
                                                         uint4* __restrict__ out) {
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint4 r;
        r.x = w0 ? w0[i] : 0u;
        r.y = w1 ? w1[i] : 0u;
        r.z = w2 ? w2[i] : 0u;
        r.w = w3 ? w3[i] : 0u;
        out[i] = r;
    }
}

// ================================================================ kernel-boundary exclusive scan
// For up to kTwoPassTiles tiles: per-tile sums, then every tile adds the sums of the
// tiles before it.  The kernel boundary orders the two passes, where a single-pass
// scan's look-back waits on status words written by other CUs.
constexpr uint32_t kTwoPassTiles = 1024;

// flag_src/flag_dst (bin sorts; NULL otherwise): block 0 moves the histogram's "unsorted"
// word to the word the scatters read, leaving the histogram region all zero behind it.
// tile_big (bin sorts): tile_big[t] = 1 when a count of tile t exceeds kSmallBin, so that
// k_sort_big_bins can tell at once that no bin needs it.
__global__ void __launch_bounds__(kBlock) k_tile_sums(const uint32_t* __restrict__ in, uint32_t n,
                                                      uint32_t* __restrict__ tile_sum, uint32_t* flag_src,
                                                      uint32_t* flag_dst, uint32_t* __restrict__ tile_big) {
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_big[kBlock / 64];
    if (flag_src && blockIdx.x == 0 && threadIdx.x == 0) {
        *flag_dst = *flag_src;
        *flag_src = 0u;
    }
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kItems;
    uint32_t sum = 0;
    bool big = false;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t c = (base + j < n) ? in[base + j] : 0u;
        sum += c;
        big = big || c > kSmallBin;
    }
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t tot = wave_sum_u32(sum);
    const bool wbig = __ballot(big) != 0ull;
    if (lane_id() == 0) {
        s_wave[w] = tot;
        s_big[w] = wbig ? 1u : 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0, b = 0;
        for (uint32_t k = 0; k < kBlock / 64; ++k) {
            t += s_wave[k];
            b |= s_big[k];
        }
        tile_sum[blockIdx.x] = t;
        if (tile_big) tile_big[blockIdx.x] = b;
    }
}

// zero_in: each element is cleared once read (the bin histograms of self-cleaning sorts)
__global__ void __launch_bounds__(kBlock) k_scan_tiles(uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       uint32_t n, const uint32_t* __restrict__ tile_sum,
                                                       uint32_t zero_in) {
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_prefix;
    const uint32_t tile = blockIdx.x;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    if (w == 0) {
        uint32_t p = 0;
        for (uint32_t k = lane; k < tile; k += 64) p += tile_sum[k];
        p = wave_sum_u32(p);
        if (lane == 0) s_prefix = p;
    }
    const uint32_t base = tile * kTile + threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        sum += v[j];
    }
    if (zero_in) {
#pragma unroll
        for (uint32_t j = 0; j < kItems; ++j)
            if (base + j < n && v[j]) in[base + j] = 0u;
    }
    const uint32_t incl = wave_inclusive_u32(sum);
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t k = 0; k < w; ++k) wave_off += s_wave[k];
    uint32_t run = s_prefix + wave_off + incl - sum;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}

// In-place exclusive scan of the `t` tile sums by one block: thread k sums a contiguous
// chunk, the block scans the chunk totals, and each thread rewrites its chunk.
constexpr uint32_t kSumsBlock = 1024;
__global__ void __launch_bounds__(kSumsBlock) k_scan_sums(uint32_t* __restrict__ ts, uint32_t t) {
    __shared__ uint32_t s_wave[kSumsBlock / 64];
    const uint32_t per = (t + kSumsBlock - 1) / kSumsBlock;
    const uint32_t lo = threadIdx.x * per;
    const uint32_t hi = lo + per < t ? lo + per : t;
    uint32_t sum = 0;
    for (uint32_t k = lo; k < hi; ++k) sum += ts[k];
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_u32(sum);
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (uint32_t k = 0; k < w; ++k) run += s_wave[k];
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t v = ts[k];
        ts[k] = run;
        run += v;
    }
}

// The second pass when the tile sums are already scanned: tile t's prefix is ts[t].
__global__ void __launch_bounds__(kBlock) k_scan_tiles_pre(uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                           uint32_t n, const uint32_t* __restrict__ ts,
                                                           uint32_t zero_in) {
    __shared__ uint32_t s_wave[kBlock / 64];
    const uint32_t tile = blockIdx.x;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t base = tile * kTile + threadIdx.x * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0u;
        sum += v[j];
    }
    if (zero_in) {
#pragma unroll
        for (uint32_t j = 0; j < kItems; ++j)
            if (base + j < n && v[j]) in[base + j] = 0u;
    }
    const uint32_t incl = wave_inclusive_u32(sum);
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t k = 0; k < w; ++k) wave_off += s_wave[k];
    uint32_t run = ts[tile] + wave_off + incl - sum;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}

// Exclusive scan of `n` counts into `out`, ordered by kernel boundaries only (no
// inter-block waits): up to kTwoPassTiles tiles, tile sums then a pass in which each
// tile adds up the sums before it; beyond that, tile sums, one block scanning them,
// then a pass that reads each tile's prefix.  `status` holds the tile sums.
// flag_src/flag_dst and zero_in: see k_tile_sums and k_scan_tiles (bin sorts only).
// tile_big: per-tile "some count exceeds kSmallBin" words for bin sorts (see k_tile_sums);
// bin_sort_tile_big() places them after the tile sums in `status`.
static int scan_counts(const uint32_t* in, uint32_t* out, uint32_t n, ScratchHeader* hdr, uint64_t* status, hipStream_t st,
                       uint32_t zero_in = 0u, uint32_t* flag_src = nullptr, uint32_t* flag_dst = nullptr,
                       uint32_t* tile_big = nullptr) {
    (void)hdr;
    const uint32_t tiles = ceil_div(n, kTile);
    if (!tiles) return FGX_OK;
    uint32_t* ts = reinterpret_cast<uint32_t*>(status);   // 8 bytes per tile are reserved; sums use 4
    uint32_t* inw = const_cast<uint32_t*>(in);           // written only when zero_in
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kBlock), 0, st, in, n, ts, flag_src, flag_dst, tile_big);
    FGX_HIP(hipGetLastError());
    if (tiles <= kTwoPassTiles) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kBlock), 0, st, inw, out, n, ts, zero_in);
    } else {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kSumsBlock), 0, st, ts, tiles);
        FGX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_scan_tiles_pre, dim3(tiles), dim3(kBlock), 0, st, inw, out, n, ts, zero_in);
    }
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// ================================================================ compaction / births
// Striped arrangement: item (j, wave, lane) = tile*2048 + j*256 + wave*64 + lane, so loads
// and the compacted stores of a wave are contiguous.  Index order == (j, wave, lane) order.
struct CompactArgs {
    const uint32_t* flags;
    uint32_t keep;
    const uint32_t* d_n;
    uint32_t n_cap;
    uint32_t num_tiles;
    // compaction output
    uint32_t* d_count_out;
    // births (mode 1)
    int32_t births;
    int32_t id_var;
    uint32_t* d_dst_count;
    uint32_t dst_capacity;
    uint32_t* d_next_id;       // may be NULL (no IDs)
    uint32_t* d_id_slots;      // may be NULL; nonzero: ID = next ID + parent index (cleared here)
    uint32_t* d_overflow;
    // item ranges (device-resident, NULL = 0): items are [*d_src_off, n); item k has
    // flag k - *d_keep_prefix, and the first *d_keep_prefix items are kept unconditionally
    const uint32_t* d_src_off;
    const uint32_t* d_keep_prefix;
    // scratch: the inputs as read by the first pass, then the per-tile kept counts
    uint32_t* stash;           // [0] items n, [1] source offset, [2] kept prefix, [3] dst count, [4] next ID,
                               // [5] ID by parent index
    uint32_t* counts;          // num_tiles words
};

// Kept-item masks of tile `tile`, in the striped arrangement, for items [0, n).
__device__ __forceinline__ void compact_masks(const CompactArgs& a, uint32_t tile, uint32_t n, uint32_t pre,
                                              uint64_t (&masks)[kItems]) {
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t i = tile * kTile + j * kBlock + threadIdx.x;
        const bool keep = i < n && (i < pre || !a.flags || ((a.flags[i - pre] != 0u) == (a.keep != 0u)));
        masks[j] = __ballot(keep);
    }
}

// Pass 1: kept items per tile.  Block 0 also records the device-resident inputs, so that
// pass 2 does not read counters that its own last tile rewrites (the output count may
// be the input count).
__global__ void __launch_bounds__(kBlock) k_compact_counts(CompactArgs a) {
    __shared__ uint32_t s_cnt[kBlock / 64];
    const uint32_t n_all = load_count(a.d_n, a.n_cap);
    const uint32_t off0 = a.d_src_off ? __hip_atomic_load(a.d_src_off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    const uint32_t off = off0 < n_all ? off0 : n_all;
    const uint32_t n = n_all - off;
    const uint32_t pre = a.d_keep_prefix ? __hip_atomic_load(a.d_keep_prefix, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.stash[0] = n;
        a.stash[1] = off;
        a.stash[2] = pre;
        a.stash[3] = a.births ? __hip_atomic_load(a.d_dst_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
        a.stash[4] = a.births && a.d_next_id ? __hip_atomic_load(a.d_next_id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
        a.stash[5] = a.births && a.d_id_slots ? __hip_atomic_load(a.d_id_slots, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    }
    uint64_t masks[kItems];
    compact_masks(a, blockIdx.x, n, pre, masks);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) c += (uint32_t)__builtin_popcountll(masks[j]);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) s_cnt[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < kBlock / 64; ++k) t += s_cnt[k];
        a.counts[blockIdx.x] = t;
    }
}

// Pass 2 (after k_scan_sums when `scanned`): each tile's prefix, its items' ranks, and
// the copy of every variable; the last tile writes the output counts.
__global__ void __launch_bounds__(kBlock) k_compact_scatter(CompactArgs a, VarTable vt, uint32_t scanned) {
    __shared__ uint32_t s_prefix;
    __shared__ uint32_t s_cnt[kItems * (kBlock / 64)];
    const uint32_t tile = blockIdx.x;
    const uint32_t n = a.stash[0], off = a.stash[1], pre = a.stash[2], dst0 = a.stash[3], id0 = a.stash[4];
    const bool by_slot = a.stash[5] != 0u;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    uint64_t masks[kItems];
    compact_masks(a, tile, n, pre, masks);
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j)
        if (lane == 0) s_cnt[j * (kBlock / 64) + w] = (uint32_t)__builtin_popcountll(masks[j]);
    __syncthreads();
    if (w == 0) {
        // exclusive scan of the 32 (j, wave) counts in index order; the tile's prefix
        const uint32_t c = lane < kItems * (kBlock / 64) ? s_cnt[lane] : 0u;
        const uint32_t inc = wave_inclusive_u32(c);
        const uint32_t total = __shfl(inc, 63, 64);
        uint32_t p = 0;
        if (scanned) {
            p = a.counts[tile];
        } else {
            for (uint32_t k = lane; k < tile; k += 64) p += a.counts[k];
            p = wave_sum_u32(p);
        }
        if (lane < kItems * (kBlock / 64)) s_cnt[lane] = inc - c;
        if (lane == 0) s_prefix = p;
        if (lane == 0 && tile == a.num_tiles - 1) {
            const uint32_t kept = p + total;
            if (!a.births) {
                *a.d_count_out = kept;
            } else {
                uint32_t nd = dst0 + kept;
                if (nd > a.dst_capacity) {
                    nd = a.dst_capacity;
                    if (a.d_overflow) *a.d_overflow = 1u;
                }
                *a.d_dst_count = nd;
                if (a.d_next_id) *a.d_next_id = id0 + (by_slot ? n : kept);
                if (a.d_id_slots) *a.d_id_slots = 0u;   // pass 1 stashed it; nothing reads it after
            }
        }
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        if (!((masks[j] >> lane) & 1ull)) continue;
        const uint32_t i = tile * kTile + j * kBlock + threadIdx.x;
        const uint32_t r = prefix + s_cnt[j * (kBlock / 64) + w] + (uint32_t)__builtin_popcountll(masks[j] & lanes_below);
        if (!a.births) {
            copy_element(vt, off + i, r);
        } else {
            const uint32_t d = dst0 + r;
            if (d >= a.dst_capacity) continue;
            copy_element(vt, off + i, d, a.id_var);
            if (a.id_var >= 0) reinterpret_cast<uint32_t*>(vt.out[a.id_var])[d] = id0 + (by_slot ? i : r);
        }
    }
}

// ================================================================ deterministic reduction
constexpr uint32_t kReduceBlocks = 256;

template <typename T, typename Acc>
__global__ void __launch_bounds__(kBlock) k_reduce_partial(const T* __restrict__ in, const uint32_t* d_n, uint32_t n_cap,
                                                           Acc* __restrict__ partial) {
    __shared__ Acc s[kBlock];
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t stride = kReduceBlocks * kBlock;
    Acc acc = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) acc += (Acc)in[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

template <typename Acc>
__global__ void __launch_bounds__(kBlock) k_reduce_final(const Acc* __restrict__ partial, Acc* __restrict__ out) {
    __shared__ Acc s[kBlock];
    s[threadIdx.x] = threadIdx.x < kReduceBlocks ? partial[threadIdx.x] : (Acc)0;
    __syncthreads();
    for (uint32_t o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// ================================================================ array messages
// src/flamegpu/simulation/detail/CUDAScatter.cu:540-640 (reorder_array_messages): message i
// goes to array element index[i]; elements nobody wrote are zero; two writes to one
// element, or more messages than elements, are an error (ArrayMessageWriteConflict).
constexpr uint32_t kErrArrayConflict = 18;   // DERR_ARRAY_CONFLICT of the device API

__device__ __forceinline__ void array_report(uint32_t* d_error, uint32_t detail) {
    if (d_error && atomicCAS(&d_error[0], 0u, kErrArrayConflict) == 0u) {
        d_error[1] = detail;
        d_error[2] = 0xFFFFFFFFu;
    }
}

__global__ void __launch_bounds__(kBlock) k_array_zero(VarTable vt, uint32_t length, uint32_t* __restrict__ counts) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = t0; i < length; i += stride) counts[i] = 0u;
    for (uint32_t v = 0; v < vt.n; ++v) {
        const size_t bytes = (size_t)length * vt.bytes[v];
        if ((vt.bytes[v] & 3u) == 0u) {
            uint32_t* o = reinterpret_cast<uint32_t*>(vt.out[v]);
            for (size_t w = t0; w < bytes / 4; w += stride) o[w] = 0u;
        } else {
            for (size_t b = t0; b < bytes; b += stride) vt.out[v][b] = 0;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_array_scatter(const uint32_t* d_n, uint32_t n_cap, const uint32_t* __restrict__ index,
                                                          uint32_t length, VarTable vt, uint32_t* __restrict__ counts,
                                                          uint32_t* d_error) {
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0 && n > length) array_report(d_error, n);
    for (uint32_t i = t0; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t e = index[i];
        if (e >= length) continue;   // reported by setIndex
        copy_element(vt, i, e);
        atomicAdd(&counts[e], 1u);
    }
}

__global__ void __launch_bounds__(kBlock) k_array_check(const uint32_t* __restrict__ counts, uint32_t length, uint32_t* d_error) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < length; i += gridDim.x * blockDim.x)
        if (counts[i] > 1u) array_report(d_error, i);
}

// ================================================================ deterministic statistics
// Sum, minimum and maximum in one pass, then the sum of squared deviations from the
// mean (which needs the finished sum) in a second.  Each block owns a fixed strided
// set of elements and folds it with a fixed tree, so the result does not depend on
// scheduling.  Acc: double for f32/f64 (exact widening), int64 / uint64 for integers.
template <typename Acc>
struct StatLimits;
template <>
struct StatLimits<double> {
    __device__ static double hi() { return __builtin_huge_val(); }
    __device__ static double lo() { return -__builtin_huge_val(); }
};
template <>
struct StatLimits<long long> {
    __device__ static long long hi() { return 0x7FFFFFFFFFFFFFFFll; }
    __device__ static long long lo() { return -0x7FFFFFFFFFFFFFFFll - 1; }
};
template <>
struct StatLimits<unsigned long long> {
    __device__ static unsigned long long hi() { return ~0ull; }
    __device__ static unsigned long long lo() { return 0ull; }
};

template <typename T, typename Acc>
__global__ void __launch_bounds__(kBlock) k_stats_partial(const T* __restrict__ in, const uint32_t* d_n, uint32_t n_cap,
                                                          Acc* __restrict__ part) {
    __shared__ Acc s_sum[kBlock], s_min[kBlock], s_max[kBlock];
    const uint32_t n = load_count(d_n, n_cap);
    const uint32_t stride = kReduceBlocks * kBlock;
    Acc sum = 0, mn = StatLimits<Acc>::hi(), mx = StatLimits<Acc>::lo();
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const Acc v = (Acc)in[i];
        sum += v;
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    s_sum[threadIdx.x] = sum;
    s_min[threadIdx.x] = mn;
    s_max[threadIdx.x] = mx;
    __syncthreads();
    for (uint32_t o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
            const Acc a = s_min[threadIdx.x + o], b = s_max[threadIdx.x + o];
            if (a < s_min[threadIdx.x]) s_min[threadIdx.x] = a;
            if (b > s_max[threadIdx.x]) s_max[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s_sum[0];
        part[kReduceBlocks + blockIdx.x] = s_min[0];
        part[2 * kReduceBlocks + blockIdx.x] = s_max[0];
    }
}

// out[0] sum, out[1] min, out[2] max (as Acc)
template <typename Acc>
__global__ void __launch_bounds__(kBlock) k_stats_final(const Acc* __restrict__ part, Acc* __restrict__ out) {
    static_assert(kReduceBlocks == kBlock, "one partial per thread");
    __shared__ Acc s_sum[kBlock], s_min[kBlock], s_max[kBlock];
    s_sum[threadIdx.x] = part[threadIdx.x];
    s_min[threadIdx.x] = part[kReduceBlocks + threadIdx.x];
    s_max[threadIdx.x] = part[2 * kReduceBlocks + threadIdx.x];
    __syncthreads();
    for (uint32_t o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
            const Acc a = s_min[threadIdx.x + o], b = s_max[threadIdx.x + o];
            if (a < s_min[threadIdx.x]) s_min[threadIdx.x] = a;
            if (b > s_max[threadIdx.x]) s_max[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_sum[0];
        out[1] = s_min[0];
        out[2] = s_max[0];
    }
}

// Σ ((double)x - mean)^2 with mean = sum / n (0 when n == 0), as the reference's
// standard deviation (include/flamegpu/simulation/AgentLoggingConfig.h,
// getAgentVariableStandardDevFunc; AgentLoggingConfig_Reductions.cuh).
template <typename T, typename Acc>
__global__ void __launch_bounds__(kBlock) k_sqdev_partial(const T* __restrict__ in, const uint32_t* d_n, uint32_t n_cap,
                                                          const Acc* __restrict__ stats, double* __restrict__ part) {
    __shared__ double s[kBlock];
    const uint32_t n = load_count(d_n, n_cap);
    const double mean = n ? (double)stats[0] / (double)n : 0.0;
    const uint32_t stride = kReduceBlocks * kBlock;
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const double d = (double)in[i] - mean;
        acc += d * d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

template <typename T, typename Acc>
static int run_stats(const void* in, const uint32_t* d_n, uint32_t n_cap, void* d_out, char* scratch, hipStream_t st) {
    Acc* part = reinterpret_cast<Acc*>(scratch);
    double* dpart = reinterpret_cast<double*>(scratch + 3 * kReduceBlocks * 8);
    hipLaunchKernelGGL((k_stats_partial<T, Acc>), dim3(kReduceBlocks), dim3(kBlock), 0, st, static_cast<const T*>(in), d_n,
                       n_cap, part);
    hipLaunchKernelGGL((k_stats_final<Acc>), dim3(1), dim3(kBlock), 0, st, part, static_cast<Acc*>(d_out));
    hipLaunchKernelGGL((k_sqdev_partial<T, Acc>), dim3(kReduceBlocks), dim3(kBlock), 0, st, static_cast<const T*>(in), d_n,
                       n_cap, static_cast<const Acc*>(d_out), dpart);
    hipLaunchKernelGGL((k_reduce_final<double>), dim3(1), dim3(kBlock), 0, st, dpart, static_cast<double*>(d_out) + 3);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

// ================================================================ scratch layouts
struct BuildLayout {
    size_t hist, key, sub, tmp, hdr, status, total;
};

static BuildLayout build_layout(uint32_t n, uint32_t bins) {
    BuildLayout L;
    size_t off = 0;
    // counts [0, bins], "unsorted" flag at [bins + 1], the flag as the scatters read it at [bins + 2]
    L.hist = off; off += align256((size_t)(bins + 3) * 4);
    L.key = off; off += align256((size_t)std::max(n, 1u) * 4);
    L.sub = off; off += align256((size_t)std::max(n, 1u) * 4);
    L.tmp = off; off += align256((size_t)std::max(n, 1u) * 4);
    L.hdr = off; off += sizeof(ScratchHeader);
    L.status = off; off += align256((size_t)ceil_div(bins + 1, kTile) * 8);
    L.total = off;
    return L;
}

static size_t tile_scratch_bytes(uint32_t n) {
    return sizeof(ScratchHeader) + align256((size_t)std::max(ceil_div(n, kTile), 1u) * 8);
}

// The per-tile large-bin words of a bin sort: the second half of the 8 bytes per tile that
// `status` reserves (the tile sums use the first half).
static uint32_t* bin_sort_tile_big(uint64_t* status, uint32_t bins) {
    return reinterpret_cast<uint32_t*>(status) + ceil_div(bins + 1, kTile);
}

// k_sort_big_bins: at most two blocks per CU (73 KB of LDS each), at most one per bin.
static uint32_t big_bin_grid(const Ctx* c, uint32_t bins) {
    const uint32_t cap = (uint32_t)std::max(c ? c->cu_count : 256, 1) * 2u;
    return std::max(1u, std::min(bins, cap));
}

// ---------------------------------------------------------------- library-owned build space
// fgx_spatial_build_noalloc works in this array instead of a caller buffer.  Each GPU holds
// its own copy for the life of the process.  In host code the array's name is not a device
// address: its address on a device comes from hipGetSymbolAddress (noalloc_scratch_init,
// once per device, that device current) and only that address is used.
static __device__ __align__(256) char noalloc_scratch[FGX_NOALLOC_BUILD_BYTES];
constexpr int kMaxDevices = 64;
static void* g_noalloc_ptr[kMaxDevices];
static std::mutex g_noalloc_mutex;

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(FGX_ERR_ARG, "noalloc_scratch_init: device out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mutex);
    if (g_noalloc_ptr[device]) return FGX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || !p)
        return fail(FGX_ERR_HIP, std::string("hipGetSymbolAddress(noalloc_scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = p;
    return FGX_OK;
}

static uint32_t stream_grid(const Ctx* c, uint32_t n) {
    // Enough waves to cover HBM latency on every CU; grid-stride beyond that.
    const uint32_t cap = (uint32_t)std::max(c ? c->cu_count : 256, 1) * 8u;
    return std::max(1u, std::min(ceil_div(n, kBlock), cap));
}

}  // namespace fgx

using namespace fgx;

static int run_scan(const uint32_t* in, uint32_t* out, uint32_t n, char* scratch, void* stream) {
    if (!n) return FGX_OK;
    auto* hdr = reinterpret_cast<ScratchHeader*>(scratch);
    auto* status = reinterpret_cast<uint64_t*>(scratch + sizeof(ScratchHeader));
    return scan_counts(in, out, n, hdr, status, (hipStream_t)stream);
}

// Shared by the message build and the agent sort.  shared_space: the working space is the
// library's own per-GPU array (fgx_spatial_build_noalloc), which every such call shares:
// its histogram is always cleared and no clean record is kept for it.
static int run_bin_sort(fgx_ctx* ctx, const KeyArgs& ka, uint32_t bins, const uint32_t* d_n, uint32_t n_cap,
                        const float* x, const float* y, const float* z, const VarTable& vt, uint32_t* start_out,
                        void* scratch, size_t scratch_bytes, void* stream, bool shared_space = false) {
    const BuildLayout L = build_layout(n_cap, bins);
    if (!scratch || scratch_bytes < L.total) return fail(FGX_ERR_CAPACITY, "scratch too small for bin sort");
    char* s = static_cast<char*>(scratch);
    auto* hist = reinterpret_cast<uint32_t*>(s + L.hist);
    auto* key = reinterpret_cast<uint32_t*>(s + L.key);
    auto* sub = reinterpret_cast<uint32_t*>(s + L.sub);
    auto* tmp = reinterpret_cast<uint32_t*>(s + L.tmp);
    hipStream_t st = (hipStream_t)stream;
    uint32_t* unsorted = hist + bins + 1;
    uint32_t* unsorted_read = hist + bins + 2;
    // With FGX_OPT_KEEP_SCRATCH_ZERO every sort leaves [0, bins + 1] zero (the scan clears
    // the counts, k_tile_sums moves the flag), so a buffer last sorted with this bin count
    // needs no clearing launch.  The record describes the buffer after the work already
    // enqueued eagerly.  A captured graph replays later, in an order the record cannot
    // follow: under capture the clearing is skipped only for a buffer that every earlier
    // sort used with this same bin count (each of them left the region zero, and the first
    // one cleared it), since another bin count moves the key arrays over the histogram.
    // A buffer that no sort has used yet still clears under capture: a later sort of the
    // same graph may use it with another bin count, and would leave this region dirty for
    // the next replay.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    FGX_HIP(hipStreamIsCapturing((hipStream_t)stream, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    bool clean = false;
    const bool record = ctx && ctx->c.keep_zero && !shared_space;
    if (record) {
        std::lock_guard<std::mutex> lk(clean_mutex());
        auto it = clean_record().find(scratch);
        auto hit = bins_history().find(scratch);
        const bool single = hit != bins_history().end() && hit->second == bins;
        clean = capturing ? single : (it != clean_record().end() && (it->second == bins || it->second == kAllZero));
        if (it != clean_record().end()) clean_record().erase(it);   // re-recorded once every launch is enqueued
        if (hit == bins_history().end()) bins_history()[scratch] = bins;
        else if (hit->second != bins) hit->second = kMixedBins;
    }
    int rc = FGX_OK;
    if (!clean) {
        rc = fgx_fill_u32(hist, 0u, (size_t)bins + 2, stream);
        if (rc) return rc;
    }
    const uint32_t g = stream_grid(ctx ? &ctx->c : nullptr, n_cap);
    if (n_cap) {
        hipLaunchKernelGGL(k_bin_hist, dim3(g), dim3(kBlock), 0, st, ka, d_n, n_cap, x, y, z, key, sub, hist, unsorted);
        FGX_HIP(hipGetLastError());
    }
    // scan hist -> start (bins+1 entries, start[bins] = n)
    auto* status = reinterpret_cast<uint64_t*>(s + L.status);
    {
        auto* hdr = reinterpret_cast<ScratchHeader*>(s + L.hdr);
        const uint32_t zero = ctx && ctx->c.keep_zero ? 1u : 0u;
        rc = scan_counts(hist, start_out, bins + 1, hdr, status, st, zero, unsorted, unsorted_read,
                         bin_sort_tile_big(status, bins));
        if (rc) return rc;
    }
    if (n_cap) {
        hipLaunchKernelGGL(k_index_scatter, dim3(g), dim3(kBlock), 0, st, d_n, n_cap, key, sub, start_out, tmp,
                           unsorted_read);
        FGX_HIP(hipGetLastError());
        // sub is free once the index scatter has run: it holds the ordered large bins
        hipLaunchKernelGGL(k_sort_big_bins, dim3(big_bin_grid(ctx ? &ctx->c : nullptr, bins)), dim3(kBlock), 0, st,
                           start_out, bins, tmp, sub, unsorted_read, bin_sort_tile_big(status, bins),
                           ceil_div(bins + 1, kTile));
        FGX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_stable_scatter, dim3(g), dim3(kBlock), 0, st, d_n, n_cap, key, start_out, tmp, sub, vt,
                           unsorted_read);
        FGX_HIP(hipGetLastError());
    }
    if (record && !capturing) {
        std::lock_guard<std::mutex> lk(clean_mutex());
        clean_record()[scratch] = bins;
    }
    return FGX_OK;
}

// Segmented: the keys come from k_bin_hist_seg; scan and scatters run over the packed batch.
static int run_bin_sort_seg(fgx_ctx* ctx, int32_t mode, const fgx_seg_key* d_segs, uint32_t nsegs, uint32_t tiles,
                            uint32_t n_total, const uint32_t* d_n_total, uint32_t keys_total, const float* x,
                            const float* y, const float* z,
                            const VarTable& vt, uint32_t* start_out, void* scratch, size_t scratch_bytes, void* stream) {
    const BuildLayout L = build_layout(n_total, keys_total);
    if (!scratch || scratch_bytes < L.total) return fail(FGX_ERR_CAPACITY, "scratch too small for segmented bin sort");
    char* s = static_cast<char*>(scratch);
    auto* hist = reinterpret_cast<uint32_t*>(s + L.hist);
    auto* key = reinterpret_cast<uint32_t*>(s + L.key);
    auto* sub = reinterpret_cast<uint32_t*>(s + L.sub);
    auto* tmp = reinterpret_cast<uint32_t*>(s + L.tmp);
    hipStream_t st = (hipStream_t)stream;
    uint32_t* unsorted = hist + keys_total + 1;
    int rc = fgx_fill_u32(hist, 0u, (size_t)keys_total + 2, stream);
    if (rc) return rc;
    if (n_total && tiles) {
        hipLaunchKernelGGL(k_bin_hist_seg, dim3(tiles), dim3(kBlock), 0, st, mode, d_segs, nsegs, x, y, z, key, sub, hist,
                           unsorted);
        FGX_HIP(hipGetLastError());
    }
    auto* status = reinterpret_cast<uint64_t*>(s + L.status);
    {
        auto* hdr = reinterpret_cast<ScratchHeader*>(s + L.hdr);
        rc = scan_counts(hist, start_out, keys_total + 1, hdr, status, st, 0u, nullptr, nullptr,
                         bin_sort_tile_big(status, keys_total));
        if (rc) return rc;
    }
    if (n_total) {
        const uint32_t g = stream_grid(ctx ? &ctx->c : nullptr, n_total);
        hipLaunchKernelGGL(k_index_scatter, dim3(g), dim3(kBlock), 0, st, d_n_total, n_total, key, sub, start_out, tmp,
                           unsorted);
        FGX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sort_big_bins, dim3(big_bin_grid(ctx ? &ctx->c : nullptr, keys_total)), dim3(kBlock), 0, st,
                           start_out, keys_total, tmp, sub, unsorted, bin_sort_tile_big(status, keys_total),
                           ceil_div(keys_total + 1, kTile));
        FGX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_stable_scatter, dim3(g), dim3(kBlock), 0, st, d_n_total, n_total, key, start_out, tmp, sub, vt,
                           unsorted);
        FGX_HIP(hipGetLastError());
    }
    return FGX_OK;
}

extern "C" {

int fgx_spatial_build_seg(fgx_ctx* ctx, const fgx_seg_key* d_segs, uint32_t nsegs, uint32_t tiles_total,
                          uint32_t n_total, const uint32_t* d_n_total, uint32_t keys_total, const float* x,
                          const float* y, const float* z,
                          uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                          uint32_t* pbm, void* scratch, size_t scratch_bytes, void* stream) {
    if (!pbm) return fail(FGX_ERR_ARG, "fgx_spatial_build_seg: NULL pbm");
    if (nsegs && !d_segs) return fail(FGX_ERR_ARG, "fgx_spatial_build_seg: NULL segment table");
    if (n_total && (!x || !y)) return fail(FGX_ERR_ARG, "fgx_spatial_build_seg: missing location array");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    return run_bin_sort_seg(ctx, 0, d_segs, nsegs, tiles_total, n_total, d_n_total, keys_total, x, y, z, vt, pbm, scratch,
                            scratch_bytes, stream);
}

int fgx_agent_sort_seg(fgx_ctx* ctx, const fgx_seg_key* d_segs, uint32_t nsegs, uint32_t tiles_total,
                       uint32_t n_total, const uint32_t* d_n_total, uint32_t keys_total, const float* x,
                       const float* y, const float* z,
                       uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                       void* scratch, size_t scratch_bytes, void* stream) {
    if (nsegs && !d_segs) return fail(FGX_ERR_ARG, "fgx_agent_sort_seg: NULL segment table");
    if (n_total && (!x || !y)) return fail(FGX_ERR_ARG, "fgx_agent_sort_seg: missing location array");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    const BuildLayout L = build_layout(n_total, keys_total);
    if (!scratch || scratch_bytes < L.total + align256((size_t)(keys_total + 1) * 4))
        return fail(FGX_ERR_CAPACITY, "fgx_agent_sort_seg: scratch too small");
    auto* start = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) + L.total);
    return run_bin_sort_seg(ctx, 1, d_segs, nsegs, tiles_total, n_total, d_n_total, keys_total, x, y, z, vt, start, scratch,
                            L.total, stream);
}

int fgx_scratch_bytes(int op, uint32_t n, uint32_t bins, size_t* bytes) {
    if (!bytes) return fail(FGX_ERR_ARG, "fgx_scratch_bytes: bytes is NULL");
    switch (op) {
        case FGX_OP_SPATIAL_BUILD:
            *bytes = build_layout(n, bins).total;
            return FGX_OK;
        case FGX_OP_AGENT_SORT:
            // start table of the sort keys lives in scratch too
            *bytes = build_layout(n, bins).total + align256((size_t)(bins + 1) * 4);
            return FGX_OK;
        case FGX_OP_COMPACT:
        case FGX_OP_BIRTHS:
        case FGX_OP_SCAN:
            *bytes = tile_scratch_bytes(n) + align256(kReduceBlocks * 8);
            return FGX_OK;
        case FGX_OP_STATS:
            *bytes = align256(4 * kReduceBlocks * 8);
            return FGX_OK;
        default:
            return fail(FGX_ERR_ARG, "fgx_scratch_bytes: unknown op");
    }
}

int fgx_scratch_clear(void* scratch, size_t bytes, void* stream) {
    if (!bytes) return FGX_OK;
    if (bytes & 3u) return fail(FGX_ERR_ARG, "fgx_scratch_clear: size must be a multiple of 4");
    const int rc = fgx_fill_u32(static_cast<uint32_t*>(scratch), 0u, bytes / 4, stream);
    if (rc == FGX_OK) {
        std::lock_guard<std::mutex> lk(clean_mutex());
        clean_record()[scratch] = kAllZero;
        bins_history().erase(scratch);   // the next sort is this buffer's first (it clears under capture)
    }
    return rc;
}

static int spatial_build(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                         const float* x, const float* y, const float* z,
                         uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                         const int32_t* rec_vars, void* records,
                         uint32_t* pbm, void* scratch, size_t scratch_bytes, void* stream) {
    if (!meta || !pbm) return fail(FGX_ERR_ARG, "fgx_spatial_build: NULL meta or pbm");
    if (meta->dims != 2 && meta->dims != 3) return fail(FGX_ERR_ARG, "fgx_spatial_build: dims must be 2 or 3");
    if (n_cap && (!x || !y || (meta->dims == 3 && !z)))
        return fail(FGX_ERR_ARG, "fgx_spatial_build: missing location array");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    if (records) {
        if (!rec_vars || (reinterpret_cast<uintptr_t>(records) & 15u))
            return fail(FGX_ERR_ARG, "fgx_spatial_build_records: NULL rec_vars, or records not 16-byte aligned");
        for (int k = 0; k < 4; ++k) {
            if (rec_vars[k] >= (int32_t)nvars || (rec_vars[k] >= 0 && var_bytes[rec_vars[k]] != 4))
                return fail(FGX_ERR_ARG, "fgx_spatial_build_records: a record word names no 4-byte variable");
            vt.rec_var[k] = rec_vars[k] < 0 ? -1 : rec_vars[k];
        }
        vt.rec = static_cast<char*>(records);
    }
    KeyArgs ka;
    std::memset(&ka, 0, sizeof(ka));
    ka.mode = 0;
    ka.dims = meta->dims;
    ka.radius = meta->cell > 0.0f ? meta->cell : meta->radius;   // bin width
    for (int a = 0; a < 3; ++a) {
        ka.min[a] = meta->min[a];
        ka.width[a] = meta->width[a];
        ka.grid[a] = meta->grid[a] ? meta->grid[a] : 1u;
    }
    return run_bin_sort(ctx, ka, meta->bins, d_n, n_cap, x, y, z, vt, pbm, scratch, scratch_bytes, stream);
}

int fgx_spatial_build(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                      const float* x, const float* y, const float* z,
                      uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                      uint32_t* pbm, void* scratch, size_t scratch_bytes, void* stream) {
    return spatial_build(ctx, meta, d_n, n_cap, x, y, z, nvars, vars_in, vars_out, var_bytes, nullptr, nullptr, pbm,
                         scratch, scratch_bytes, stream);
}

int fgx_spatial_build_records(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                              const float* x, const float* y, const float* z, uint32_t nvars,
                              const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                              const int32_t* rec_vars, void* records, uint32_t* pbm, void* scratch,
                              size_t scratch_bytes, void* stream) {
    if (!records) return fail(FGX_ERR_ARG, "fgx_spatial_build_records: NULL records");
    return spatial_build(ctx, meta, d_n, n_cap, x, y, z, nvars, vars_in, vars_out, var_bytes, rec_vars, records, pbm,
                         scratch, scratch_bytes, stream);
}

int fgx_spatial_build_noalloc(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                              const float* x, const float* y, const float* z,
                              uint32_t nvars, const void* const* vars_in, void* const* vars_out,
                              const uint32_t* var_bytes, uint32_t* pbm, void* stream) {
    static_assert((size_t)FGX_NOALLOC_MAX_BINS + 3 <= 0xFFFFFFFFull, "bin limit fits the layout arithmetic");
    if (!ctx) return fail(FGX_ERR_ARG, "fgx_spatial_build_noalloc: NULL context");
    if (!meta || !pbm) return fail(FGX_ERR_ARG, "fgx_spatial_build_noalloc: NULL meta or pbm");
    if (n_cap > FGX_NOALLOC_MAX_MESSAGES || meta->bins > FGX_NOALLOC_MAX_BINS)
        return fail(FGX_ERR_CAPACITY, "fgx_spatial_build_noalloc: " + std::to_string(n_cap) + " messages / " +
                                          std::to_string(meta->bins) + " bins exceed the limits of " +
                                          std::to_string(FGX_NOALLOC_MAX_MESSAGES) + " / " +
                                          std::to_string(FGX_NOALLOC_MAX_BINS) + "; use fgx_spatial_build");
    const int dev = ctx->c.device;
    void* space = dev >= 0 && dev < kMaxDevices ? g_noalloc_ptr[dev] : nullptr;
    if (!space) return fail(FGX_ERR_HIP, "fgx_spatial_build_noalloc: no build space on this device");
    if (build_layout(n_cap, meta->bins).total > FGX_NOALLOC_BUILD_BYTES)
        return fail(FGX_ERR_CAPACITY, "fgx_spatial_build_noalloc: working space too small");
    if (meta->dims != 2 && meta->dims != 3) return fail(FGX_ERR_ARG, "fgx_spatial_build_noalloc: dims must be 2 or 3");
    if (n_cap && (!x || !y || (meta->dims == 3 && !z)))
        return fail(FGX_ERR_ARG, "fgx_spatial_build_noalloc: missing location array");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    KeyArgs ka;
    std::memset(&ka, 0, sizeof(ka));
    ka.mode = 0;
    ka.dims = meta->dims;
    ka.radius = meta->cell > 0.0f ? meta->cell : meta->radius;
    for (int a = 0; a < 3; ++a) {
        ka.min[a] = meta->min[a];
        ka.width[a] = meta->width[a];
        ka.grid[a] = meta->grid[a] ? meta->grid[a] : 1u;
    }
    return run_bin_sort(ctx, ka, meta->bins, d_n, n_cap, x, y, z, vt, pbm, space, FGX_NOALLOC_BUILD_BYTES, stream,
                        true);
}

int fgx_bucket_build(fgx_ctx* ctx, int32_t key_min, uint32_t keys, const uint32_t* d_n, uint32_t n_cap, const int32_t* key,
                     uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                     uint32_t* pbm, void* scratch, size_t scratch_bytes, void* stream) {
    if (!pbm || !keys) return fail(FGX_ERR_ARG, "fgx_bucket_build: NULL bucket table or no keys");
    if (n_cap && !key) return fail(FGX_ERR_ARG, "fgx_bucket_build: missing key array");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    KeyArgs ka;
    std::memset(&ka, 0, sizeof(ka));
    ka.mode = 2;
    ka.dims = 1;
    ka.key_min = key_min;
    ka.keys = keys;
    const float* kf = reinterpret_cast<const float*>(key);
    return run_bin_sort(ctx, ka, keys, d_n, n_cap, kf, kf, nullptr, vt, pbm, scratch, scratch_bytes, stream);
}

int fgx_array_build(fgx_ctx* ctx, uint32_t length, const uint32_t* d_n, uint32_t n_cap, const uint32_t* index,
                    uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                    uint32_t* counts, uint32_t* d_error, void* stream) {
    if (!length || !counts) return fail(FGX_ERR_ARG, "fgx_array_build: zero length or NULL counts");
    if (n_cap && !index) return fail(FGX_ERR_ARG, "fgx_array_build: missing index array");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t gz = stream_grid(ctx ? &ctx->c : nullptr, length);
    hipLaunchKernelGGL(k_array_zero, dim3(gz), dim3(kBlock), 0, st, vt, length, counts);
    FGX_HIP(hipGetLastError());
    if (n_cap) {
        const uint32_t g = stream_grid(ctx ? &ctx->c : nullptr, n_cap);
        hipLaunchKernelGGL(k_array_scatter, dim3(g), dim3(kBlock), 0, st, d_n, n_cap, index, length, vt, counts, d_error);
        FGX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_array_check, dim3(gz), dim3(kBlock), 0, st, counts, length, d_error);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_agent_sort(fgx_ctx* ctx, int dims, const float* mn, const float* mx, float radius,
                   const uint32_t* d_n, uint32_t n_cap, const float* x, const float* y, const float* z,
                   uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                   void* scratch, size_t scratch_bytes, void* stream) {
    if (dims != 2 && dims != 3) return fail(FGX_ERR_ARG, "fgx_agent_sort: dims must be 2 or 3");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    uint32_t bits = 0;
    rc = fgx_agent_sort_key_bits(dims, mn, mx, radius, &bits);
    if (rc) return rc;
    if (bits > 30) return fail(FGX_ERR_ARG, "fgx_agent_sort: key range too large");
    const uint32_t bins = 1u << bits;
    KeyArgs ka;
    std::memset(&ka, 0, sizeof(ka));
    ka.mode = 1;
    ka.dims = dims;
    ka.radius = radius;
    ka.key_mask = bins - 1u;
    for (int a = 0; a < 3; ++a) {
        if (a < dims) {
            ka.min[a] = mn[a];
            ka.width[a] = mx[a] - mn[a];
            ka.grid[a] = ka.width[a] != 0.0f ? (uint32_t)std::ceil(ka.width[a] / radius) : 1u;
        } else {
            ka.grid[a] = 1u;
            ka.width[a] = 1.0f;
        }
    }
    const BuildLayout L = build_layout(n_cap, bins);
    if (!scratch || scratch_bytes < L.total + align256((size_t)(bins + 1) * 4))
        return fail(FGX_ERR_CAPACITY, "fgx_agent_sort: scratch too small");
    auto* start = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) + L.total);
    return run_bin_sort(ctx, ka, bins, d_n, n_cap, x, y, z, vt, start, scratch, L.total, stream);
}

int fgx_pack_records(const uint32_t* d_n, uint32_t n_cap, const void* w0, const void* w1, const void* w2,
                     const void* w3, void* out, void* stream) {
    if (!n_cap) return FGX_OK;
    if (!out || (reinterpret_cast<uintptr_t>(out) & 15u)) return fail(FGX_ERR_ARG, "fgx_pack_records: out must be 16-byte aligned");
    hipLaunchKernelGGL(k_pack_records, dim3(stream_grid(nullptr, n_cap)), dim3(kBlock), 0, (hipStream_t)stream, d_n, n_cap,
                       static_cast<const uint32_t*>(w0), static_cast<const uint32_t*>(w1),
                       static_cast<const uint32_t*>(w2), static_cast<const uint32_t*>(w3), static_cast<uint4*>(out));
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_exclusive_scan_u32(fgx_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                           void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!scratch || scratch_bytes < tile_scratch_bytes(n)) return fail(FGX_ERR_CAPACITY, "fgx_exclusive_scan_u32: scratch too small");
    return run_scan(in, out, n, static_cast<char*>(scratch), stream);
}

// Stable compaction in passes ordered by kernel boundaries: kept counts per tile, a
// block-level scan of them when there are more than kTwoPassTiles tiles (otherwise each
// tile adds up the counts before it), then ranks and copies.
static int launch_compact(CompactArgs a, const VarTable& vt, void* scratch, size_t scratch_bytes, void* stream) {
    if (!scratch || scratch_bytes < tile_scratch_bytes(a.n_cap)) return fail(FGX_ERR_CAPACITY, "compaction scratch too small");
    uint32_t* words = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) + sizeof(ScratchHeader));
    a.stash = words;          // 8 words
    a.counts = words + 8;     // num_tiles words: the region holds 8 bytes per tile, at least 256
    a.num_tiles = std::max(ceil_div(a.n_cap, kTile), 1u);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_compact_counts, dim3(a.num_tiles), dim3(kBlock), 0, st, a);
    FGX_HIP(hipGetLastError());
    const uint32_t scanned = a.num_tiles > kTwoPassTiles ? 1u : 0u;
    if (scanned) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kSumsBlock), 0, st, a.counts, a.num_tiles);
        FGX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_compact_scatter, dim3(a.num_tiles), dim3(kBlock), 0, st, a, vt, scanned);
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

int fgx_compact(fgx_ctx* ctx, const uint32_t* flags, uint32_t keep, const uint32_t* d_n, uint32_t n_cap,
                uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                uint32_t* d_count_out, void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!d_count_out) return fail(FGX_ERR_ARG, "fgx_compact: d_count_out is NULL");
    if (n_cap && !flags) return fail(FGX_ERR_ARG, "fgx_compact: flags is NULL");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    CompactArgs a;
    std::memset(&a, 0, sizeof(a));
    a.flags = flags;
    a.keep = keep;
    a.d_n = d_n;
    a.n_cap = n_cap;
    a.d_count_out = d_count_out;
    a.births = 0;
    a.id_var = -1;
    return launch_compact(a, vt, scratch, scratch_bytes, stream);
}

int fgx_append_births(fgx_ctx* ctx, const uint32_t* flags, const uint32_t* d_n_parents, uint32_t n_parents_cap,
                      uint32_t nvars, const void* const* newbuf, void* const* dst, const uint32_t* var_bytes,
                      int32_t id_var, uint32_t* d_dst_count, uint32_t dst_capacity, uint32_t* d_next_id,
                      uint32_t* d_id_slots, uint32_t* d_overflow, void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!d_dst_count || !d_next_id) return fail(FGX_ERR_ARG, "fgx_append_births: NULL device counters");
    if (id_var >= (int32_t)nvars) return fail(FGX_ERR_ARG, "fgx_append_births: id_var out of range");
    if (id_var >= 0 && var_bytes[id_var] != 4) return fail(FGX_ERR_ARG, "fgx_append_births: id variable must be 32-bit");
    VarTable vt;
    int rc = make_var_table(nvars, newbuf, dst, var_bytes, &vt);
    if (rc) return rc;
    // The ID variable is produced (next_id + rank), not copied from newbuf.
    CompactArgs a;
    std::memset(&a, 0, sizeof(a));
    a.flags = flags;
    a.keep = 1;
    a.d_n = d_n_parents;
    a.n_cap = n_parents_cap;
    a.births = 1;
    a.id_var = id_var;
    a.d_dst_count = d_dst_count;
    a.dst_capacity = dst_capacity;
    a.d_next_id = d_next_id;
    a.d_id_slots = d_id_slots;
    a.d_overflow = d_overflow;
    return launch_compact(a, vt, scratch, scratch_bytes, stream);
}

int fgx_partition(fgx_ctx* ctx, const uint32_t* flags, const uint32_t* d_n, uint32_t n_cap, uint32_t nvars,
                  const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes, uint32_t* d_first_count,
                  uint32_t* d_work, void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!d_first_count || !d_work) return fail(FGX_ERR_ARG, "fgx_partition: NULL device counters");
    if (n_cap && !flags) return fail(FGX_ERR_ARG, "fgx_partition: flags is NULL");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    // pass 1: flagged items to the front, their count to d_first_count
    CompactArgs a;
    std::memset(&a, 0, sizeof(a));
    a.flags = flags;
    a.keep = 1;
    a.d_n = d_n;
    a.n_cap = n_cap;
    a.d_count_out = d_first_count;
    a.id_var = -1;
    rc = launch_compact(a, vt, scratch, scratch_bytes, stream);
    if (rc) return rc;
    // pass 2: the others appended after them (same scratch: the epoch advanced)
    rc = fgx_memcpy_d2d(d_work, d_first_count, 4, stream);
    if (rc) return rc;
    CompactArgs b;
    std::memset(&b, 0, sizeof(b));
    b.flags = flags;
    b.keep = 0;
    b.d_n = d_n;
    b.n_cap = n_cap;
    b.births = 1;
    b.id_var = -1;
    b.d_dst_count = d_work;
    b.dst_capacity = 0xFFFFFFFFu;
    return launch_compact(b, vt, scratch, scratch_bytes, stream);
}

int fgx_compact_prefix(fgx_ctx* ctx, const uint32_t* flags, uint32_t keep, const uint32_t* d_prefix, const uint32_t* d_n,
                       uint32_t n_cap, uint32_t nvars, const void* const* vars_in, void* const* vars_out,
                       const uint32_t* var_bytes, uint32_t* d_count_out, void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!d_count_out || !d_prefix) return fail(FGX_ERR_ARG, "fgx_compact_prefix: NULL device counters");
    if (n_cap && !flags) return fail(FGX_ERR_ARG, "fgx_compact_prefix: flags is NULL");
    VarTable vt;
    int rc = make_var_table(nvars, vars_in, vars_out, var_bytes, &vt);
    if (rc) return rc;
    CompactArgs a;
    std::memset(&a, 0, sizeof(a));
    a.flags = flags;
    a.keep = keep;
    a.d_n = d_n;
    a.n_cap = n_cap;
    a.d_count_out = d_count_out;
    a.id_var = -1;
    a.d_keep_prefix = d_prefix;
    return launch_compact(a, vt, scratch, scratch_bytes, stream);
}

int fgx_append_range(fgx_ctx* ctx, const uint32_t* d_src_off, const uint32_t* d_n, uint32_t n_cap, uint32_t nvars,
                     const void* const* src, void* const* dst, const uint32_t* var_bytes, uint32_t* d_dst_count,
                     uint32_t dst_capacity, uint32_t* d_overflow, void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!d_dst_count || !d_src_off) return fail(FGX_ERR_ARG, "fgx_append_range: NULL device counters");
    VarTable vt;
    int rc = make_var_table(nvars, src, dst, var_bytes, &vt);
    if (rc) return rc;
    CompactArgs a;
    std::memset(&a, 0, sizeof(a));
    a.keep = 1;
    a.d_n = d_n;
    a.n_cap = n_cap;
    a.births = 1;
    a.id_var = -1;
    a.d_dst_count = d_dst_count;
    a.dst_capacity = dst_capacity;
    a.d_overflow = d_overflow;
    a.d_src_off = d_src_off;
    return launch_compact(a, vt, scratch, scratch_bytes, stream);
}

int fgx_reduce_stats(fgx_ctx* ctx, int type, const void* in, const uint32_t* d_n, uint32_t n_cap, void* d_out,
                     void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    if (!d_out) return fail(FGX_ERR_ARG, "fgx_reduce_stats: NULL output");
    if (n_cap && !in) return fail(FGX_ERR_ARG, "fgx_reduce_stats: NULL input");
    if (!scratch || scratch_bytes < align256(4 * kReduceBlocks * 8))
        return fail(FGX_ERR_CAPACITY, "fgx_reduce_stats: scratch too small");
    char* sc = static_cast<char*>(scratch);
    hipStream_t st = (hipStream_t)stream;
    switch (type) {
        case 0: return run_stats<float, double>(in, d_n, n_cap, d_out, sc, st);
        case 1: return run_stats<double, double>(in, d_n, n_cap, d_out, sc, st);
        case 2: return run_stats<int32_t, long long>(in, d_n, n_cap, d_out, sc, st);
        case 3: return run_stats<uint32_t, unsigned long long>(in, d_n, n_cap, d_out, sc, st);
        case 4: return run_stats<long long, long long>(in, d_n, n_cap, d_out, sc, st);
        case 5: return run_stats<unsigned long long, unsigned long long>(in, d_n, n_cap, d_out, sc, st);
        default: return fail(FGX_ERR_ARG, "fgx_reduce_stats: unknown type");
    }
}

int fgx_reduce_sum(fgx_ctx* ctx, int type, const void* in, const uint32_t* d_n, uint32_t n_cap,
                   void* d_out, void* scratch, size_t scratch_bytes, void* stream) {
    (void)ctx;
    const size_t need = tile_scratch_bytes(n_cap) + align256(kReduceBlocks * 8);
    if (!scratch || scratch_bytes < need) return fail(FGX_ERR_CAPACITY, "fgx_reduce_sum: scratch too small");
    char* part = static_cast<char*>(scratch) + tile_scratch_bytes(n_cap);
    hipStream_t st = (hipStream_t)stream;
    switch (type) {
        case 0:
            hipLaunchKernelGGL((k_reduce_partial<float, double>), dim3(kReduceBlocks), dim3(kBlock), 0, st,
                               static_cast<const float*>(in), d_n, n_cap, reinterpret_cast<double*>(part));
            hipLaunchKernelGGL((k_reduce_final<double>), dim3(1), dim3(kBlock), 0, st, reinterpret_cast<double*>(part),
                               static_cast<double*>(d_out));
            break;
        case 1:
            hipLaunchKernelGGL((k_reduce_partial<double, double>), dim3(kReduceBlocks), dim3(kBlock), 0, st,
                               static_cast<const double*>(in), d_n, n_cap, reinterpret_cast<double*>(part));
            hipLaunchKernelGGL((k_reduce_final<double>), dim3(1), dim3(kBlock), 0, st, reinterpret_cast<double*>(part),
                               static_cast<double*>(d_out));
            break;
        case 2:
            hipLaunchKernelGGL((k_reduce_partial<int32_t, long long>), dim3(kReduceBlocks), dim3(kBlock), 0, st,
                               static_cast<const int32_t*>(in), d_n, n_cap, reinterpret_cast<long long*>(part));
            hipLaunchKernelGGL((k_reduce_final<long long>), dim3(1), dim3(kBlock), 0, st,
                               reinterpret_cast<long long*>(part), static_cast<long long*>(d_out));
            break;
        case 3:
            hipLaunchKernelGGL((k_reduce_partial<uint32_t, unsigned long long>), dim3(kReduceBlocks), dim3(kBlock), 0, st,
                               static_cast<const uint32_t*>(in), d_n, n_cap, reinterpret_cast<unsigned long long*>(part));
            hipLaunchKernelGGL((k_reduce_final<unsigned long long>), dim3(1), dim3(kBlock), 0, st,
                               reinterpret_cast<unsigned long long*>(part), static_cast<unsigned long long*>(d_out));
            break;
        case 4:
            hipLaunchKernelGGL((k_reduce_partial<long long, long long>), dim3(kReduceBlocks), dim3(kBlock), 0, st,
                               static_cast<const long long*>(in), d_n, n_cap, reinterpret_cast<long long*>(part));
            hipLaunchKernelGGL((k_reduce_final<long long>), dim3(1), dim3(kBlock), 0, st,
                               reinterpret_cast<long long*>(part), static_cast<long long*>(d_out));
            break;
        case 5:
            hipLaunchKernelGGL((k_reduce_partial<unsigned long long, unsigned long long>), dim3(kReduceBlocks), dim3(kBlock), 0, st,
                               static_cast<const unsigned long long*>(in), d_n, n_cap,
                               reinterpret_cast<unsigned long long*>(part));
            hipLaunchKernelGGL((k_reduce_final<unsigned long long>), dim3(1), dim3(kBlock), 0, st,
                               reinterpret_cast<unsigned long long*>(part), static_cast<unsigned long long*>(d_out));
            break;
        default:
            return fail(FGX_ERR_ARG, "fgx_reduce_sum: unknown type");
    }
    FGX_HIP(hipGetLastError());
    return FGX_OK;
}

}  // extern "C"
