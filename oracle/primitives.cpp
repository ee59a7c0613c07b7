// Serial CPU oracle: hot-path primitives.  TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

// ---------------------------------------------------------------------------
// Philox4x32-10, restated from Salmon, Moraes, Dror, Shaw, "Parallel Random
// Numbers: As Easy as 1, 2, 3" (SC'11), section 5: round function
//   (hi0,lo0) = M0*c0, (hi1,lo1) = M1*c2,
//   c' = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0),
// with the Weyl key schedule k += (W0, W1) between rounds.
// ---------------------------------------------------------------------------
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

extern "C" void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

// Agent stream: counter = (draw block, function stream, step, agent id),
// key = (seed lo, seed hi).  Draw k is word k&3 of block k>>2.
extern "C" uint32_t orc_rng_u32(uint64_t seed, uint32_t id, uint32_t step, uint32_t stream, uint32_t k) {
    const uint32_t ctr[4] = {k >> 2, stream, step, id};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return o[k & 3];
}

// include/flamegpu/runtime/random/AgentRandom.cuh:76-84: keep the low 24 bits, scale by 2^-24.
extern "C" float orc_rng_uniform(uint64_t seed, uint32_t id, uint32_t step, uint32_t stream, uint32_t k) {
    const uint32_t v = orc_rng_u32(seed, id, step, stream, k);
    return (float)(v & 0xFFFFFFu) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Grid metadata: src/flamegpu/runtime/messaging/MessageSpatial3D.cu:31-52
// (width = max-min, grid = ceil(width/radius) in FP32) and
// include/flamegpu/detail/numeric.h:12-20 (approxExactlyDivisible).
// ---------------------------------------------------------------------------
static bool approx_exactly_divisible(float x, float y) {
    const float eps = std::max(std::fabs(x), std::fabs(y)) * std::numeric_limits<float>::epsilon();
    const float v = std::fmod(x, y);
    return v <= eps || v > y - eps;
}

// With mult = 1 (orc_spatial_meta_init) the grid is the reference's: MessageSpatial3D.cu:31-52.
// mult = 1/2 halves the bin width (an exact FP32 scaling), so the grid is ceil(width / (r/2))
// and the searched neighbourhood widens to 2 bins either side.
extern "C" int orc_spatial_meta_init_mult(int dims, const float* mn, const float* mx, float radius, float mult,
                                          orc_spatial_meta* m) {
    if (dims != 2 && dims != 3) return -1;
    if (!(mult == 1.0f || mult == 0.5f)) return -1;
    std::memset(m, 0, sizeof(*m));
    m->dims = dims;
    m->radius = radius;
    m->cell = radius * mult;
    m->span = mult == 1.0f ? 1 : 2;
    m->bins = 1;
    m->wrap_ok = 1;
    for (int a = 0; a < 3; ++a) {
        if (a < dims) {
            m->min[a] = mn[a];
            m->max[a] = mx[a];
            m->width[a] = mx[a] - mn[a];
            const float q = m->width[a] / m->cell;
            m->grid[a] = (uint32_t)std::ceil(q);
            m->bins *= m->grid[a];
            m->wrap_ok = m->wrap_ok && approx_exactly_divisible(m->width[a], m->cell);
        } else {
            m->grid[a] = 1;
        }
    }
    return 0;
}

extern "C" int orc_spatial_meta_init(int dims, const float* mn, const float* mx, float radius, orc_spatial_meta* m) {
    return orc_spatial_meta_init_mult(dims, mn, mx, radius, 1.0f, m);
}

// getGridPosition3D/2D (MessageSpatial3DDevice.cuh:646-659): floor((p-min)/r), clamped to [0, grid-1]
// (r is the bin width, the radius when the multiplier is 1).
static inline int grid_pos(const orc_spatial_meta* m, int a, float p) {
    const int g = (int)std::floor((p - m->min[a]) / m->cell);
    const int hi = (int)m->grid[a] - 1;
    return g < 0 ? 0 : (g > hi ? hi : g);
}
// getHash3D/2D (MessageSpatial3DDevice.cuh:660-672): x re-clamped, x fastest.
static inline uint32_t hash_cell(const orc_spatial_meta* m, int cx, int cy, int cz) {
    const int hx = (int)m->grid[0] - 1;
    const uint32_t ux = (uint32_t)(cx < 0 ? 0 : (cx >= hx ? hx : cx));
    return (uint32_t)cz * m->grid[0] * m->grid[1] + (uint32_t)cy * m->grid[0] + ux;
}

extern "C" uint32_t orc_spatial_hash(const orc_spatial_meta* m, float x, float y, float z) {
    const int cx = grid_pos(m, 0, x);
    const int cy = grid_pos(m, 1, y);
    const int cz = m->dims == 3 ? grid_pos(m, 2, z) : 0;
    return hash_cell(m, cx, cy, cz);
}

// buildIndex (MessageSpatial3D.cu:113-146): histogram -> exclusive scan -> reorder.
// The reorder here is stable (input order within a bin).
extern "C" int orc_spatial_build(const orc_spatial_meta* m, uint32_t n,
                                 const float* x, const float* y, const float* z,
                                 uint32_t nvars, const void* const* vars_in, const uint32_t* var_bytes,
                                 void* const* vars_out, uint32_t* pbm, uint32_t* perm) {
    std::vector<uint32_t> key(n);
    std::vector<uint32_t> hist(m->bins + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        key[i] = orc_spatial_hash(m, x[i], y[i], m->dims == 3 ? z[i] : 0.0f);
        hist[key[i]]++;
    }
    uint32_t run = 0;
    for (uint32_t b = 0; b <= m->bins; ++b) {
        pbm[b] = run;
        run += hist[b];
    }
    std::vector<uint32_t> fill(m->bins, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t dst = pbm[key[i]] + fill[key[i]]++;
        if (perm) perm[dst] = i;
        for (uint32_t v = 0; v < nvars; ++v) {
            const uint32_t sz = var_bytes[v];
            std::memcpy(static_cast<char*>(vars_out[v]) + (size_t)dst * sz,
                        static_cast<const char*>(vars_in[v]) + (size_t)i * sz, sz);
        }
    }
    return 0;
}

// Neighbour visit order, S = m->span (1 in the reference).
//  Filter 3-D (MessageSpatial3DDevice.cuh:52-59,693-719): strips along x, dy outer, dz inner,
//    strip = [PBM[hash(cx-S,..)], PBM[hash(cx+S,..)+1]), out-of-range y/z strips skipped.
//  WrapFilter 3-D (:264-279,727-749): (2S+1)^3 single cells, dx outer, dy, dz inner, modulo grid.
//  Filter 2-D (MessageSpatial2DDevice.cuh:51-53,646-672): 2S+1 strips, dy ascending.
//  WrapFilter 2-D (:253-260,679-698): (2S+1)^2 cells, dx outer, dy inner.
template <typename Visit>
static void visit_candidates(const orc_spatial_meta* m, const uint32_t* pbm, int mode,
                             float x, float y, float z, Visit&& visit) {
    const int cx = grid_pos(m, 0, x);
    const int cy = grid_pos(m, 1, y);
    const int cz = m->dims == 3 ? grid_pos(m, 2, z) : 0;
    const int gx = (int)m->grid[0], gy = (int)m->grid[1], gz = (int)m->grid[2];
    const int S = m->span;
    if (mode == 0) {
        if (m->dims == 3) {
            for (int dy = -S; dy <= S; ++dy) {
                for (int dz = -S; dz <= S; ++dz) {
                    const int ay = cy + dy, az = cz + dz;
                    if (ay < 0 || az < 0 || ay >= gy || az >= gz) continue;
                    const uint32_t s = pbm[hash_cell(m, cx - S, ay, az)];
                    const uint32_t e = pbm[hash_cell(m, cx + S, ay, az) + 1];
                    for (uint32_t i = s; i < e; ++i) visit(i);
                }
            }
        } else {
            for (int dy = -S; dy <= S; ++dy) {
                const int ay = cy + dy;
                if (ay < 0 || ay >= gy) continue;
                const uint32_t s = pbm[hash_cell(m, cx - S, ay, 0)];
                const uint32_t e = pbm[hash_cell(m, cx + S, ay, 0) + 1];
                for (uint32_t i = s; i < e; ++i) visit(i);
            }
        }
    } else {
        // modulo of a possibly negative cell (S may exceed the grid on tiny grids)
        auto wrapc = [](int c, int g) { return ((c % g) + g) % g; };
        if (m->dims == 3) {
            for (int dx = -S; dx <= S; ++dx)
                for (int dy = -S; dy <= S; ++dy)
                    for (int dz = -S; dz <= S; ++dz) {
                        const uint32_t h = hash_cell(m, wrapc(cx + dx, gx), wrapc(cy + dy, gy), wrapc(cz + dz, gz));
                        for (uint32_t i = pbm[h]; i < pbm[h + 1]; ++i) visit(i);
                    }
        } else {
            for (int dx = -S; dx <= S; ++dx)
                for (int dy = -S; dy <= S; ++dy) {
                    const uint32_t h = hash_cell(m, wrapc(cx + dx, gx), wrapc(cy + dy, gy), 0);
                    for (uint32_t i = pbm[h]; i < pbm[h + 1]; ++i) visit(i);
                }
        }
    }
}

extern "C" uint32_t orc_spatial_candidates(const orc_spatial_meta* m, const uint32_t* pbm, int mode,
                                           float x, float y, float z, uint32_t* out, uint32_t max_out) {
    uint32_t c = 0;
    visit_candidates(m, pbm, mode, x, y, z, [&](uint32_t i) {
        if (c < max_out) out[c] = i;
        ++c;
    });
    return c;
}

extern "C" double orc_spatial_mean_candidates(const orc_spatial_meta* m, const uint32_t* pbm, int mode) {
    const uint32_t n = pbm[m->bins];
    if (!n) return 0.0;
    // Each message in bin b issues a query from b; the candidate count only depends on b.
    double total = 0.0;
    for (uint32_t b = 0; b < m->bins; ++b) {
        const uint32_t nb = pbm[b + 1] - pbm[b];
        if (!nb) continue;
        const uint32_t bx = b % m->grid[0];
        const uint32_t by = (b / m->grid[0]) % m->grid[1];
        const uint32_t bz = b / (m->grid[0] * m->grid[1]);
        // A query point at the centre of bin b lands in bin b.
        const float qx = m->min[0] + ((float)bx + 0.5f) * m->cell;
        const float qy = m->min[1] + ((float)by + 0.5f) * m->cell;
        const float qz = m->dims == 3 ? m->min[2] + ((float)bz + 0.5f) * m->cell : 0.0f;
        uint64_t c = 0;
        visit_candidates(m, pbm, mode, qx, qy, qz, [&](uint32_t) { ++c; });
        total += (double)nb * (double)c;
    }
    return total / (double)n;
}

// scatter_generic (CUDAScatter.cu:68-88): copy every variable of items with flag==keep
// to their exclusive-scan position.
extern "C" uint32_t orc_compact(uint32_t n, const uint32_t* flags, uint32_t keep,
                                uint32_t nvars, const void* const* in, const uint32_t* var_bytes, void* const* out) {
    uint32_t pos = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if ((flags[i] != 0) != (keep != 0)) continue;
        for (uint32_t v = 0; v < nvars; ++v) {
            const uint32_t sz = var_bytes[v];
            std::memcpy(static_cast<char*>(out[v]) + (size_t)pos * sz,
                        static_cast<const char*>(in[v]) + (size_t)i * sz, sz);
        }
        ++pos;
    }
    return pos;
}

// scatterNew (CUDAAgentStateList.cu:189-250): flagged new agents appended after the
// existing agents in parent order.  ID = next_id + rank (replaces atomicInc, DeviceAPI.cuh:552).
extern "C" uint32_t orc_births_ex(uint32_t n_parents, const uint32_t* flags,
                                  uint32_t nvars, const void* const* newbuf, const uint32_t* var_bytes,
                                  void* const* dst, uint32_t dst_count, uint32_t id_var, uint32_t* next_id,
                                  int32_t by_parent_index) {
    uint32_t rank = 0;
    for (uint32_t i = 0; i < n_parents; ++i) {
        if (!flags[i]) continue;
        const uint32_t d = dst_count + rank;
        for (uint32_t v = 0; v < nvars; ++v) {
            const uint32_t sz = var_bytes[v];
            if (v == id_var) {
                const uint32_t id = *next_id + (by_parent_index ? i : rank);
                std::memcpy(static_cast<char*>(dst[v]) + (size_t)d * sz, &id, sizeof(id));
            } else {
                std::memcpy(static_cast<char*>(dst[v]) + (size_t)d * sz,
                            static_cast<const char*>(newbuf[v]) + (size_t)i * sz, sz);
            }
        }
        ++rank;
    }
    *next_id += by_parent_index ? n_parents : rank;
    return dst_count + rank;
}

extern "C" uint32_t orc_births(uint32_t n_parents, const uint32_t* flags,
                               uint32_t nvars, const void* const* newbuf, const uint32_t* var_bytes,
                               void* const* dst, uint32_t dst_count, uint32_t id_var, uint32_t* next_id) {
    return orc_births_ex(n_parents, flags, nvars, newbuf, var_bytes, dst, dst_count, id_var, next_id, 0);
}

// calculateSpatialHash (CUDASimulation.cu:372-408) with the grid of spatialSortAgent_async
// (:518-525) and the radix sort's end bit floor(log2(bins))+1 (:571-572).
extern "C" void orc_agent_sort_keys(int dims, const float* mn, const float* mx, float radius, uint32_t n,
                                    const float* x, const float* y, const float* z, uint32_t* keys, uint32_t* key_bits) {
    float w[3] = {0, 0, 0};
    uint32_t g[3] = {1, 1, 1};
    for (int a = 0; a < dims; ++a) {
        w[a] = mx[a] - mn[a];
        g[a] = w[a] != 0.0f ? (uint32_t)std::ceil(w[a] / radius) : 1u;
    }
    const uint32_t bins = g[0] * g[1] * g[2];
    const int max_bit = (int)std::floor(std::log2((double)bins)) + 1;
    const uint32_t mask = max_bit >= 32 ? 0xFFFFFFFFu : ((1u << max_bit) - 1u);
    if (key_bits) *key_bits = (uint32_t)max_bit;
    for (uint32_t i = 0; i < n; ++i) {
        const int gx = (int)std::floor(((x[i] - mn[0]) / w[0]) * (float)g[0]);
        const int gy = (int)std::floor(((y[i] - mn[1]) / w[1]) * (float)g[1]);
        uint32_t b;
        if (dims == 3) {
            const int gz = (int)std::floor(((z[i] - mn[2]) / w[2]) * (float)g[2]);
            b = (uint32_t)gz * g[0] * g[1] + (uint32_t)gy * g[0] + (uint32_t)gx;
        } else {
            b = (uint32_t)gy * g[0] + (uint32_t)gx;
        }
        keys[i] = b & mask;
    }
}

extern "C" void orc_stable_sort_perm(uint32_t n, const uint32_t* keys, uint32_t* perm) {
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
}
