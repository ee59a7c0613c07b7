// Device runtime: DeviceAPI, message input iterators / output writers, AgentRandom,
// DeviceEnvironment and the agent-function kernel.
//
// Same surface as include/flamegpu/runtime/DeviceAPI.cuh:163-557,
// include/flamegpu/runtime/messaging/*/Message*Device.cuh and
// include/flamegpu/runtime/random/AgentRandom.cuh; the variable tables are
// compile-time hashed names resolved against kernel-argument bindings
// (flamegpu/detail/core.h) instead of the curve shared-memory hash table.
#pragma once

#include <cmath>
#include <limits>

#include "flamegpu/model.h"

namespace flamegpu {

namespace detail {

template <typename T, unsigned N>
__device__ __forceinline__ T load_elem(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, const char (&name)[N], uint32_t idx,
                                       uint32_t len, uint32_t elem, uint32_t* d_error, uint32_t err_unknown,
                                       uint32_t err_type) {
    const uint32_t h = name_hash(name, N);
    const Found f = find_binding(tab, mul, h);
    if (!f.ptr) {
        report_error(d_error, err_unknown, h, idx);
        return T();
    }
    if (f.bytes != sizeof(T) * len) {
        report_error(d_error, err_type, h, idx);
        return T();
    }
    if (elem >= len) {
        report_error(d_error, DERR_ARRAY_INDEX, h, idx);
        return T();
    }
    return reinterpret_cast<const T*>(f.ptr)[(size_t)idx * len + elem];
}

template <typename T, unsigned N>
__device__ __forceinline__ void store_elem(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, const char (&name)[N], uint32_t idx,
                                           uint32_t len, uint32_t elem, T v, uint32_t* d_error, uint32_t err_unknown,
                                           uint32_t err_type) {
    const uint32_t h = name_hash(name, N);
    const Found f = find_binding(tab, mul, h);
    if (!f.ptr) {
        report_error(d_error, err_unknown, h, idx);
        return;
    }
    if (f.bytes != sizeof(T) * len) {
        report_error(d_error, err_type, h, idx);
        return;
    }
    if (elem >= len) {
        report_error(d_error, DERR_ARRAY_INDEX, h, idx);
        return;
    }
    reinterpret_cast<T*>(f.ptr)[(size_t)idx * len + elem] = v;
}

__device__ __forceinline__ uint32_t load_count(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

}  // namespace detail

// ------------------------------------------------------------------ RNG
// include/flamegpu/runtime/random/AgentRandom.cuh:17-136, with the per-thread curand
// state replaced by counter-based Philox4x32-10 keyed by (seed, agent ID, step,
// function stream, draw block): results do not depend on thread placement,
// population order or co-scheduled work, and resume exactly.
class AgentRandom {
 public:
    __device__ __forceinline__ AgentRandom(const detail::FnArgs& a, uint32_t idx) : a_(a), idx_(idx) {}
    template <typename T>
    __device__ __forceinline__ T uniform() const;
    template <typename T>
    __device__ __forceinline__ T uniform(T min, T max) const;
    template <typename T>
    __device__ __forceinline__ T normal() const;
    template <typename T>
    __device__ __forceinline__ T logNormal(T mean, T stddev) const;
    __device__ __forceinline__ unsigned int poisson(double mean = 1.0) const;

    __device__ __forceinline__ uint32_t next_u32() const {
        const uint32_t blk = counter_ >> 2;
        if (blk != cached_) {
            uint32_t c[4] = {blk, a_.stream_id, a_.d_step ? *a_.d_step : 0u, a_.ids ? a_.ids[idx_] : idx_};
            detail::philox4x32_10(c, (uint32_t)a_.seed, (uint32_t)(a_.seed >> 32));
            r0_ = c[0];
            r1_ = c[1];
            r2_ = c[2];
            r3_ = c[3];
            cached_ = blk;
        }
        const uint32_t k = counter_++ & 3u;
        return k == 0 ? r0_ : (k == 1 ? r1_ : (k == 2 ? r2_ : r3_));
    }

 private:
    const detail::FnArgs& a_;
    uint32_t idx_;
    mutable uint32_t counter_ = 0;
    mutable uint32_t cached_ = 0xFFFFFFFFu;
    mutable uint32_t r0_ = 0, r1_ = 0, r2_ = 0, r3_ = 0;
};

// AgentRandom.cuh:76-84: low 24 bits of one draw scaled by 2^-24 -> [0, 1).
template <>
__device__ __forceinline__ float AgentRandom::uniform<float>() const {
    return (float)(next_u32() & 0xFFFFFFu) * (1.0f / 16777216.0f);
}
// AgentRandom.cuh:86-96 (Doornik's conversion of two draws).
template <>
__device__ __forceinline__ double AgentRandom::uniform<double>() const {
    const uint32_t r1 = next_u32();
    const uint32_t r2 = next_u32();
    constexpr double INVM32 = 2.32830643653869628906e-010;
    constexpr double INVM52 = 2.22044604925031308085e-016;
    return (static_cast<int>(r1) * INVM32 + (0.5 + INVM52 / 2) + static_cast<int>(r2 & 0x000FFFFF) * INVM52);
}
// AgentRandom.cuh:120-136: integer range [min, max].
template <typename T>
__device__ __forceinline__ T AgentRandom::uniform(T min, T max) const {
    static_assert(std::is_integral<T>::value, "AgentRandom::uniform(min, max) is for integer types");
    if (min > max) detail::report_error(a_.d_error, detail::DERR_RANDOM_RANGE, 0, idx_);
    return static_cast<T>(min + (1 + max - min) * uniform<float>());
}
// Box-Muller on two 24-bit uniforms mapped to (0, 1].
template <>
__device__ __forceinline__ float AgentRandom::normal<float>() const {
    const float u1 = (float)((next_u32() & 0xFFFFFFu) + 1u) * (1.0f / 16777216.0f);
    const float u2 = (float)(next_u32() & 0xFFFFFFu) * (1.0f / 16777216.0f);
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}
template <>
__device__ __forceinline__ double AgentRandom::normal<double>() const {
    const double u1 = 1.0 - uniform<double>();
    const double u2 = uniform<double>();
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}
template <>
__device__ __forceinline__ float AgentRandom::logNormal<float>(float mean, float stddev) const {
    return expf(mean + stddev * normal<float>());
}
template <>
__device__ __forceinline__ double AgentRandom::logNormal<double>(double mean, double stddev) const {
    return exp(mean + stddev * normal<double>());
}
// Knuth's multiplication method for small means, normal approximation above 64.
__device__ __forceinline__ unsigned int AgentRandom::poisson(double mean) const {
    if (mean <= 0.0) return 0u;
    if (mean > 64.0) {
        const double v = mean + sqrt(mean) * normal<double>();
        return v < 0.0 ? 0u : (unsigned int)(v + 0.5);
    }
    const double L = exp(-mean);
    unsigned int k = 0;
    double p = 1.0;
    do {
        ++k;
        p *= uniform<double>();
    } while (p > L && k < 1024u);
    return k - 1;
}

// ------------------------------------------------------------------ environment
// include/flamegpu/runtime/environment/DeviceEnvironment.cuh
class DeviceEnvironment {
 public:
    __device__ __forceinline__ explicit DeviceEnvironment(const detail::FnArgs& a) : a_(a) {}
    template <typename T, unsigned N>
    __device__ __forceinline__ T getProperty(const char (&name)[N]) const {
        return get<T>(name, 1, 0);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ T getProperty(const char (&name)[N], unsigned int i) const {
        return get<T>(name, L, i);
    }
    template <unsigned N>
    __device__ __forceinline__ bool containsProperty(const char (&name)[N]) const {
        const uint32_t h = detail::name_hash(name, N);
        return a_.env_b[detail::table_slot(h, a_.env_mul, detail::ENV_SLOT_BITS)].hash == h;
    }

 private:
    template <typename T, unsigned N>
    __device__ __forceinline__ T get(const char (&name)[N], uint32_t len, uint32_t i) const {
        const uint32_t h = detail::name_hash(name, N);
        const detail::EnvBinding& b = a_.env_b[detail::table_slot(h, a_.env_mul, detail::ENV_SLOT_BITS)];
        if (b.hash != h) {
            detail::report_error(a_.d_error, detail::DERR_UNKNOWN_ENV, h, 0);
            return T();
        }
        if (b.bytes != sizeof(T) * len || i >= len) {
            detail::report_error(a_.d_error, detail::DERR_ENV_TYPE, h, 0);
            return T();
        }
        return reinterpret_cast<const T*>(a_.env + b.offset)[i];
    }
    const detail::FnArgs& a_;
};

// ------------------------------------------------------------------ message inputs
class MessageNone::In {
 public:
    __device__ __forceinline__ explicit In(const detail::FnArgs&) {}
};
class MessageNone::Out {
 public:
    __device__ __forceinline__ Out(const detail::FnArgs&, uint32_t) {}
};

namespace detail {
// Shared access to the bound input message list.
struct MsgReader {
    const FnArgs& a;
    template <typename T, unsigned N>
    __device__ __forceinline__ T get(const char (&name)[N], uint32_t idx) const {
        const uint32_t h = name_hash(name, N);
        const Found f = find_binding(a.msg_in, a.msg_in_mul, h);
        if (!f.ptr) {
            report_error(a.d_error, DERR_UNKNOWN_MSG_VAR, h, idx);
            return T();
        }
        if (f.bytes != sizeof(T)) {
            report_error(a.d_error, DERR_MSG_VAR_TYPE, h, idx);
            return T();
        }
        return reinterpret_cast<const T*>(f.ptr)[idx];
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ T get(const char (&name)[N], uint32_t idx, uint32_t i) const {
        return load_elem<T>(a.msg_in, a.msg_in_mul, name, idx, L, i, a.d_error, DERR_UNKNOWN_MSG_VAR,
                                          DERR_MSG_VAR_TYPE);
    }
};
}  // namespace detail

// include/flamegpu/runtime/messaging/MessageBruteForce/MessageBruteForceDevice.cuh:19-172
class MessageBruteForce::In {
 public:
    __device__ __forceinline__ explicit In(const detail::FnArgs& a) : r_{a}, len_(a.d_msg_in_count ? detail::load_count(a.d_msg_in_count) : 0u) {}
    class Message {
     public:
        __device__ __forceinline__ Message(const detail::MsgReader& r, uint32_t i) : r_(r), i_(i) {}
        template <typename T, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
            return r_.template get<T>(name, i_);
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
            return r_.template get<T, L>(name, i_, e);
        }
        __device__ __forceinline__ size_type getIndex() const { return i_; }

     private:
        const detail::MsgReader& r_;
        uint32_t i_;
    };
    class iterator {
     public:
        __device__ __forceinline__ iterator(const detail::MsgReader& r, uint32_t i) : r_(r), i_(i) {}
        __device__ __forceinline__ iterator& operator++() {
            ++i_;
            return *this;
        }
        __device__ __forceinline__ bool operator!=(const iterator& o) const { return i_ != o.i_; }
        __device__ __forceinline__ bool operator==(const iterator& o) const { return i_ == o.i_; }
        __device__ __forceinline__ Message operator*() const { return Message(r_, i_); }

     private:
        const detail::MsgReader& r_;
        uint32_t i_;
    };
    __device__ __forceinline__ iterator begin() const { return iterator(r_, 0); }
    __device__ __forceinline__ iterator end() const { return iterator(r_, len_); }
    __device__ __forceinline__ size_type size() const { return len_; }
    __device__ __forceinline__ Message at(size_type i) const { return Message(r_, i); }

 private:
    detail::MsgReader r_;
    uint32_t len_;
};

// include/flamegpu/runtime/messaging/MessageBruteForce/MessageBruteForceDevice.cuh:259-294
class MessageBruteForce::Out {
 public:
    __device__ __forceinline__ Out(const detail::FnArgs& a, uint32_t idx) : a_(a), idx_(idx) {}
    template <typename T, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], T value) const {
        if (!a_.msg_out_flags) {
            detail::report_error(a_.d_error, detail::DERR_MSG_OUTPUT_DISABLED, 0, idx_);
            return;
        }
        detail::store_elem<T>(a_.msg_out, a_.msg_out_mul, name, idx_, 1, 0, value, a_.d_error,
                                                    detail::DERR_UNKNOWN_MSG_VAR, detail::DERR_MSG_VAR_TYPE);
        if (a_.msg_out_optional) a_.msg_out_flags[idx_] = 1u;
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], unsigned int e, T value) const {
        if (!a_.msg_out_flags) {
            detail::report_error(a_.d_error, detail::DERR_MSG_OUTPUT_DISABLED, 0, idx_);
            return;
        }
        detail::store_elem<T>(a_.msg_out, a_.msg_out_mul, name, idx_, L, e, value, a_.d_error,
                                                    detail::DERR_UNKNOWN_MSG_VAR, detail::DERR_MSG_VAR_TYPE);
        if (a_.msg_out_optional) a_.msg_out_flags[idx_] = 1u;
    }

 protected:
    const detail::FnArgs& a_;
    uint32_t idx_;
};

namespace detail {
// getGridPosition / getHash (MessageSpatial3DDevice.cuh:646-672).
__device__ __forceinline__ int grid_pos(const SpatialDev& m, int a, float p) {
    const int g = (int)floorf((p - m.min[a]) / m.radius);
    const int hi = (int)m.grid[a] - 1;
    return g < 0 ? 0 : (g > hi ? hi : g);
}
__device__ __forceinline__ uint32_t cell_hash(const SpatialDev& m, int x, int y, int z) {
    const int hx = (int)m.grid[0] - 1;
    const uint32_t ux = (uint32_t)(x < 0 ? 0 : (x >= hx ? hx : x));
    return (uint32_t)z * m.grid[0] * m.grid[1] + (uint32_t)y * m.grid[0] + ux;
}
// getVirtualX (MessageSpatial3DDevice.cuh:373-397): x2 - (x21/|x21|)*width.  On the
// taken branch |x21| > width/2 > 0, so x21/|x21| is exactly +-1 and the product is
// exactly +-width: copysignf gives the identical value without a division.
__device__ __forceinline__ float virtual_coord(float x2, float x1, float width) {
    const float x21 = x2 - x1;
    return fabsf(x21) > width / 2.0f ? x2 - copysignf(width, x21) : x2;
}
// (c + d) mod g for |d| <= 1 and 0 <= c < g, without an integer division.
__device__ __forceinline__ int wrap_cell(int c, int d, int g) {
    int a = c + d;
    a = a < 0 ? a + g : a;
    return a >= g ? a - g : a;
}
}  // namespace detail

// Spatial message with its search origin, shared by the 2-D and 3-D filters.
// Two sources for the same visit order: the block's LDS copy of the neighbourhood
// (agent_function_kernel_staged) when the search origin's cell is covered by it,
// otherwise the message list in HBM.
template <int DIMS, bool WRAP>
class SpatialFilter {
 public:
    __device__ __forceinline__ SpatialFilter(const detail::FnArgs& a, float x, float y, float z, const uint32_t* smem,
                                             const detail::StageHeader* sh)
        : a_(a), m_(a.in_meta) {
        loc_[0] = x;
        loc_[1] = y;
        loc_[2] = z;
        cell_[0] = detail::grid_pos(m_, 0, x);
        cell_[1] = detail::grid_pos(m_, 1, y);
        cell_[2] = DIMS == 3 ? detail::grid_pos(m_, 2, z) : 0;
        // Staged reads only when every active lane of the wave is covered, so the
        // choice between LDS and HBM is a scalar branch, never both paths.
        int covered = 0;
        if (smem) {
            for (uint32_t k = 0; k < sh->ngroups; ++k) {
                const detail::StageGroup& g = sh->g[k];
                if (g.cy == cell_[1] && g.cz == cell_[2] && cell_[0] >= g.xlo && cell_[0] <= g.xhi) {
                    covered = 1;
                    ebase_ = g.ebase;
                    xbase_ = g.xbase;
                    width_ = (uint32_t)g.width;
                }
            }
        }
        staged_ = __builtin_amdgcn_readfirstlane(__all(covered) ? 1 : 0) != 0;
        smem_ = smem;
    }
    class Message {
     public:
        __device__ __forceinline__ Message(const SpatialFilter& f, uint32_t g, uint32_t l) : f_(f), g_(g), l_(l) {}
        template <typename T, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
            const uint32_t h = detail::name_hash(name, N);
            const detail::VarBinding& b = f_.a_.msg_in[detail::table_slot(h, f_.a_.msg_in_mul, detail::VAR_SLOT_BITS)];
            if (b.hash != h) {
                detail::report_error(f_.a_.d_error, detail::DERR_UNKNOWN_MSG_VAR, h, g_);
                return T();
            }
            if (b.bytes != sizeof(T)) {
                detail::report_error(f_.a_.d_error, detail::DERR_MSG_VAR_TYPE, h, g_);
                return T();
            }
            if constexpr (sizeof(T) == 4) {
                if (f_.staged_ && b.lds != detail::NOT_STAGED) {
                    const uint32_t w = f_.smem_[2 * (f_.a_.stage_max_entries + 1) + (uint32_t)b.lds * f_.a_.stage_cap + l_];
                    T v;
                    __builtin_memcpy(&v, &w, 4);
                    return v;
                }
            }
            return reinterpret_cast<const T*>(b.ptr)[g_];
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
            return detail::load_elem<T>(f_.a_.msg_in, f_.a_.msg_in_mul, name, g_, L, e, f_.a_.d_error,
                                        detail::DERR_UNKNOWN_MSG_VAR, detail::DERR_MSG_VAR_TYPE);
        }
        __device__ __forceinline__ float getVirtualX() const { return getVirtualX(f_.loc_[0]); }
        __device__ __forceinline__ float getVirtualY() const { return getVirtualY(f_.loc_[1]); }
        __device__ __forceinline__ float getVirtualZ() const { return getVirtualZ(f_.loc_[2]); }
        __device__ __forceinline__ float getVirtualX(float x1) const {
            return detail::virtual_coord(getVariable<float>("x"), x1, f_.m_.width[0]);
        }
        __device__ __forceinline__ float getVirtualY(float y1) const {
            return detail::virtual_coord(getVariable<float>("y"), y1, f_.m_.width[1]);
        }
        __device__ __forceinline__ float getVirtualZ(float z1) const {
            return detail::virtual_coord(getVariable<float>("z"), z1, f_.m_.width[2]);
        }
        __device__ __forceinline__ size_type getIndex() const { return g_; }

     private:
        const SpatialFilter& f_;
        uint32_t g_, l_;
    };
    // Visit order of the reference iterators:
    //  Filter 3-D (MessageSpatial3DDevice.cuh:52-59,693-719): 9 x-strips, dy outer, dz inner;
    //  WrapFilter 3-D (:264-279,727-749): 27 cells, dx outer, dy, dz inner, modulo grid;
    //  Filter 2-D (MessageSpatial2DDevice.cuh:646-672): 3 strips; WrapFilter 2-D (:679-698): 9 cells.
    static constexpr int kUnits = WRAP ? (DIMS == 3 ? 27 : 9) : (DIMS == 3 ? 9 : 3);
    class iterator {
     public:
        __device__ __forceinline__ iterator(const SpatialFilter& f, int unit) : f_(f), unit_(unit), g_(0), l_(0), lend_(0) {}
        __device__ __forceinline__ void advance_unit() {
            const detail::SpatialDev& m = f_.m_;
            while (++unit_ < kUnits) {
                int dx, dy, dz;
                if (WRAP) {
                    if (DIMS == 3) {
                        dx = unit_ / 9 - 1;
                        dy = (unit_ / 3) % 3 - 1;
                        dz = unit_ % 3 - 1;
                    } else {
                        dx = unit_ / 3 - 1;
                        dy = unit_ % 3 - 1;
                        dz = 0;
                    }
                } else {
                    dx = 0;
                    if (DIMS == 3) {
                        dy = unit_ / 3 - 1;
                        dz = unit_ % 3 - 1;
                    } else {
                        dy = unit_ - 1;
                        dz = 0;
                    }
                    const int ay = f_.cell_[1] + dy, az = f_.cell_[2] + dz;
                    if (ay < 0 || ay >= (int)m.grid[1] || az < 0 || az >= (int)m.grid[2]) continue;
                }
                const int hx = (int)m.grid[0] - 1;
                const int xs = WRAP ? f_.cell_[0] + dx : (f_.cell_[0] - 1 < 0 ? 0 : f_.cell_[0] - 1);
                const int xe = WRAP ? xs : (f_.cell_[0] + 1 > hx ? hx : f_.cell_[0] + 1);
                if (f_.staged_) {
                    const uint32_t j = DIMS == 3 ? (uint32_t)((dy + 1) * 3 + (dz + 1)) : (uint32_t)(dy + 1);
                    const uint32_t row = f_.ebase_ + j * f_.width_;
                    const uint32_t es = row + (uint32_t)(xs - f_.xbase_);
                    const uint32_t ee = row + (uint32_t)(xe - f_.xbase_);
                    l_ = f_.smem_[es];
                    lend_ = f_.smem_[ee + 1];
                    g_ = f_.smem_[f_.a_.stage_max_entries + 1 + es];
                } else {
                    uint32_t s, e;
                    if (WRAP) {
                        const int ax = detail::wrap_cell(f_.cell_[0], dx, (int)m.grid[0]);
                        const int ay = detail::wrap_cell(f_.cell_[1], dy, (int)m.grid[1]);
                        const int az = detail::wrap_cell(f_.cell_[2], dz, (int)m.grid[2]);
                        s = detail::cell_hash(m, ax, ay, az);
                        e = s + 1;
                    } else {
                        const int ay = f_.cell_[1] + dy, az = f_.cell_[2] + dz;
                        s = detail::cell_hash(m, xs, ay, az);
                        e = detail::cell_hash(m, xe, ay, az) + 1;
                    }
                    g_ = m.pbm[s];
                    l_ = g_;
                    lend_ = m.pbm[e];
                }
                if (l_ < lend_) return;
            }
        }
        __device__ __forceinline__ iterator& operator++() {
            ++g_;
            if (++l_ >= lend_) advance_unit();
            return *this;
        }
        __device__ __forceinline__ bool operator!=(const iterator&) const { return unit_ < kUnits; }
        __device__ __forceinline__ Message operator*() const { return Message(f_, g_, l_); }

     private:
        const SpatialFilter& f_;
        int unit_;
        uint32_t g_, l_, lend_;
    };
    __device__ __forceinline__ iterator begin() const {
        iterator it(*this, -1);
        it.advance_unit();
        return it;
    }
    __device__ __forceinline__ iterator end() const { return iterator(*this, kUnits); }

 private:
    const detail::FnArgs& a_;
    const detail::SpatialDev& m_;
    float loc_[3];
    int cell_[3];
    const uint32_t* smem_;
    bool staged_;
    uint32_t ebase_ = 0;
    int xbase_ = 0;
    uint32_t width_ = 0;
};

// include/flamegpu/runtime/messaging/MessageSpatial3D/MessageSpatial3DDevice.cuh:503-558
class MessageSpatial3D::In {
 public:
    static constexpr int dims = 3;
    typedef SpatialFilter<3, false> Filter;
    typedef SpatialFilter<3, true> WrapFilter;
    __device__ __forceinline__ explicit In(const detail::FnArgs& a) : a_(a) {}
    __device__ __forceinline__ void enable_stage(const uint32_t* smem, const detail::StageHeader* sh) {
        smem_ = smem;
        sh_ = sh;
    }
    __device__ __forceinline__ Filter operator()(float x, float y, float z) const { return Filter(a_, x, y, z, smem_, sh_); }
    __device__ __forceinline__ WrapFilter wrap(float x, float y, float z) const {
        const detail::SpatialDev& m = a_.in_meta;
        if (x > m.max[0] || y > m.max[1] || z > m.max[2] || x < m.min[0] || y < m.min[1] || z < m.min[2]) {
            detail::report_error(a_.d_error, detail::DERR_WRAP_OUT_OF_BOUNDS, 0, 0);
            return WrapFilter(a_, m.min[0], m.min[1], m.min[2], nullptr, nullptr);
        }
        if (!m.wrap_ok) detail::report_error(a_.d_error, detail::DERR_WRAP_NOT_FACTOR, 0, 0);
        return WrapFilter(a_, x, y, z, smem_, sh_);
    }
    __device__ __forceinline__ float radius() const { return a_.in_meta.radius; }

 private:
    const detail::FnArgs& a_;
    const uint32_t* smem_ = nullptr;
    const detail::StageHeader* sh_ = nullptr;
};

// include/flamegpu/runtime/messaging/MessageSpatial2D/MessageSpatial2DDevice.cuh
class MessageSpatial2D::In {
 public:
    static constexpr int dims = 2;
    typedef SpatialFilter<2, false> Filter;
    typedef SpatialFilter<2, true> WrapFilter;
    __device__ __forceinline__ explicit In(const detail::FnArgs& a) : a_(a) {}
    __device__ __forceinline__ void enable_stage(const uint32_t* smem, const detail::StageHeader* sh) {
        smem_ = smem;
        sh_ = sh;
    }
    __device__ __forceinline__ Filter operator()(float x, float y) const { return Filter(a_, x, y, 0.0f, smem_, sh_); }
    __device__ __forceinline__ WrapFilter wrap(float x, float y) const {
        const detail::SpatialDev& m = a_.in_meta;
        if (x > m.max[0] || y > m.max[1] || x < m.min[0] || y < m.min[1]) {
            detail::report_error(a_.d_error, detail::DERR_WRAP_OUT_OF_BOUNDS, 0, 0);
            return WrapFilter(a_, m.min[0], m.min[1], 0.0f, nullptr, nullptr);
        }
        if (!m.wrap_ok) detail::report_error(a_.d_error, detail::DERR_WRAP_NOT_FACTOR, 0, 0);
        return WrapFilter(a_, x, y, 0.0f, smem_, sh_);
    }
    __device__ __forceinline__ float radius() const { return a_.in_meta.radius; }

 private:
    const detail::FnArgs& a_;
    const uint32_t* smem_ = nullptr;
    const detail::StageHeader* sh_ = nullptr;
};

class MessageSpatial2D::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ Out(const detail::FnArgs& a, uint32_t idx) : MessageBruteForce::Out(a, idx) {}
    __device__ __forceinline__ void setLocation(float x, float y) const {
        setVariable<float>("x", x);
        setVariable<float>("y", y);
    }
};

// MessageSpatial3DDevice.cuh:674-684
class MessageSpatial3D::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ Out(const detail::FnArgs& a, uint32_t idx) : MessageBruteForce::Out(a, idx) {}
    __device__ __forceinline__ void setLocation(float x, float y, float z) const {
        setVariable<float>("x", x);
        setVariable<float>("y", y);
        setVariable<float>("z", z);
    }
};

// ------------------------------------------------------------------ DeviceAPI
template <typename MsgIn, typename MsgOut>
class DeviceAPI {
 public:
    // include/flamegpu/runtime/DeviceAPI.cuh:187-260 (AgentOut)
    class AgentOut {
     public:
        __device__ __forceinline__ AgentOut(const detail::FnArgs& a, uint32_t idx) : a_(a), idx_(idx) {}
        template <typename T, unsigned N>
        __device__ __forceinline__ void setVariable(const char (&name)[N], T value) const {
            if (!a_.birth_flags) {
                detail::report_error(a_.d_error, detail::DERR_AGENT_OUTPUT_DISABLED, 0, idx_);
                return;
            }
            if (name[0] == '_') {
                detail::report_error(a_.d_error, detail::DERR_RESERVED_NAME, 0, idx_);
                return;
            }
            detail::store_elem<T>(a_.new_agent, a_.new_mul, name, idx_, 1, 0, value, a_.d_error,
                                                          detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
            a_.birth_flags[idx_] = 1u;
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ void setVariable(const char (&name)[N], unsigned int e, T value) const {
            if (!a_.birth_flags) {
                detail::report_error(a_.d_error, detail::DERR_AGENT_OUTPUT_DISABLED, 0, idx_);
                return;
            }
            if (name[0] == '_') {
                detail::report_error(a_.d_error, detail::DERR_RESERVED_NAME, 0, idx_);
                return;
            }
            detail::store_elem<T>(a_.new_agent, a_.new_mul, name, idx_, L, e, value, a_.d_error,
                                                          detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
            a_.birth_flags[idx_] = 1u;
        }
        // New-agent IDs are nextID + the parent's rank among this launch's births, which is
        // only known after the launch; the ID can be read from the new agent next step.
        __device__ __forceinline__ id_t getID() const {
            detail::report_error(a_.d_error, detail::DERR_AGENT_OUT_ID, 0, idx_);
            return ID_NOT_SET;
        }

     private:
        const detail::FnArgs& a_;
        uint32_t idx_;
    };

    __device__ __forceinline__ DeviceAPI(const detail::FnArgs& a, uint32_t idx)
        : message_in(a), message_out(a, idx), agent_out(a, idx), random(a, idx), environment(a), a_(a), idx_(idx) {}

    template <typename T, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
        return detail::load_elem<T>(a_.agent, a_.agent_mul, name, idx_, 1, 0, a_.d_error,
                                                            detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
        return detail::load_elem<T>(a_.agent, a_.agent_mul, name, idx_, L, e, a_.d_error,
                                                            detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], T value) const {
        if (name[0] == '_') {
            detail::report_error(a_.d_error, detail::DERR_RESERVED_NAME, 0, idx_);
            return;
        }
        detail::store_elem<T>(a_.agent, a_.agent_mul, name, idx_, 1, 0, value, a_.d_error,
                                                      detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], unsigned int e, T value) const {
        if (name[0] == '_') {
            detail::report_error(a_.d_error, detail::DERR_RESERVED_NAME, 0, idx_);
            return;
        }
        detail::store_elem<T>(a_.agent, a_.agent_mul, name, idx_, L, e, value, a_.d_error,
                                                      detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
    }
    __device__ __forceinline__ id_t getID() const { return a_.ids ? a_.ids[idx_] : ID_NOT_SET; }
    __device__ __forceinline__ unsigned int getStepCounter() const { return a_.d_step ? *a_.d_step : 0u; }
    __device__ __forceinline__ unsigned int getThreadIndex() const { return idx_; }

    typename MsgIn::In message_in;
    typename MsgOut::Out message_out;
    AgentOut agent_out;
    AgentRandom random;
    DeviceEnvironment environment;

 private:
    const detail::FnArgs& a_;
    uint32_t idx_;
};

// ------------------------------------------------------------------ the agent-function kernels
// include/flamegpu/runtime/AgentFunction.cuh:60-120: one thread per agent of the
// state list; the count is read from device memory so launches can be replayed
// from a captured graph while populations change.
namespace detail {
template <typename F>
__global__ void __launch_bounds__(256) agent_function_kernel(const FnArgs a) {
    const uint32_t n = a.d_n ? load_count(a.d_n) : a.n_cap;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (a.msg_out_flags) a.msg_out_flags[i] = a.msg_out_optional ? 0u : 1u;
    if (a.birth_flags) a.birth_flags[i] = 0u;
    DeviceAPI<typename F::MsgIn, typename F::MsgOut> api(a, i);
    const AGENT_STATUS s = F()(&api);
    if (a.death_flags)
        a.death_flags[i] = (uint32_t)s;
    else if (s == DEAD)
        report_error(a.d_error, DERR_DEATH_DISABLED, 0, i);
}

// Stages, for the block's 256 bin-sorted agents, every message of the bins their
// search neighbourhoods can touch: agents are grouped by their own cell row
// (cy, cz); group g needs the 9 (3 in 2-D) rows around it over the virtual x range
// [xlo-1, xhi+1] (wrapped onto the grid, which also covers the clamped strips of
// the non-wrapping iterator).  LDS holds, per virtual bin, its first staged slot
// and its first global message index, then the staged variables as SoA arrays.
// Returns false (block-uniform) when the block does not fit; its agents then read HBM.
template <int DIMS>
__device__ __forceinline__ bool stage_block(const FnArgs& a, uint32_t* smem, StageHeader& sh, uint32_t i, bool active) {
    __shared__ int s_row[256];
    __shared__ uint32_t s_w[4];
    const SpatialDev& m = a.in_meta;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t maxe = a.stage_max_entries;
    uint32_t* lstart = smem;
    uint32_t* gstart = smem + maxe + 1;
    uint32_t* vars = smem + 2 * (maxe + 1);
    const int gx = (int)m.grid[0], gy = (int)m.grid[1], gz = (int)m.grid[2];
    int cx = 0, cy = 0, cz = 0, row = -1;
    if (active) {
        cx = grid_pos(m, 0, a.stage_pos[0][i]);
        cy = grid_pos(m, 1, a.stage_pos[1][i]);
        cz = DIMS == 3 ? grid_pos(m, 2, a.stage_pos[2][i]) : 0;
        row = cz * gy + cy;
    }
    s_row[tid] = row;
    __syncthreads();
    const bool head = active && (tid == 0 || s_row[tid - 1] != row);
    const uint64_t hm = __ballot(head);
    if (lane == 0) s_w[w] = (uint32_t)__builtin_popcountll(hm);
    __syncthreads();
    uint32_t before = 0, groups = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k < w) before += s_w[k];
        groups += s_w[k];
    }
    if (groups > STAGE_MAX_GROUPS) return false;
    const int gidx = (int)(before + (uint32_t)__builtin_popcountll(hm & ((2ull << lane) - 1ull))) - 1;
    if (head) {
        sh.g[gidx].cy = cy;
        sh.g[gidx].cz = cz;
        sh.g[gidx].xlo = 0x7FFFFFFF;
        sh.g[gidx].xhi = -1;
    }
    __syncthreads();
    if (active) {
        atomicMin(&sh.g[gidx].xlo, cx);
        atomicMax(&sh.g[gidx].xhi, cx);
    }
    __syncthreads();
    constexpr uint32_t kRows = DIMS == 3 ? 9u : 3u;
    if (tid == 0) {
        uint32_t e = 0;
        for (uint32_t k = 0; k < groups; ++k) {
            StageGroup& g = sh.g[k];
            g.xbase = g.xlo - 1;
            g.width = g.xhi - g.xlo + 3;
            g.ebase = e;
            e += kRows * (uint32_t)g.width;
        }
        sh.ngroups = groups;
        sh.nentries = e;
        sh.ok = e <= maxe ? 1u : 0u;
    }
    __syncthreads();
    if (!sh.ok) return false;
    const uint32_t E = sh.nentries;
    // bin table: entry e = (group, neighbour row, virtual x); all loads issued before use
    uint32_t local = 0;
    uint32_t cnt[3], gs[3];
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        const uint32_t e = tid + q * 256u;
        cnt[q] = 0;
        gs[q] = 0;
        if (e < E) {
            uint32_t k = 0;
            while (k + 1 < groups && e >= sh.g[k + 1].ebase) ++k;
            const StageGroup& g = sh.g[k];
            const uint32_t rel = e - g.ebase;
            const int j = (int)(rel / (uint32_t)g.width);
            const int v = (int)(rel % (uint32_t)g.width);
            const int dy = DIMS == 3 ? j / 3 - 1 : j - 1;
            const int dz = DIMS == 3 ? j % 3 - 1 : 0;
            const int ay = wrap_cell(g.cy, dy, gy);
            const int az = DIMS == 3 ? wrap_cell(g.cz, dz, gz) : 0;
            int x = g.xbase + v;
            x = x < 0 ? x + gx : (x >= gx ? x - gx : x);
            const uint32_t h = (uint32_t)az * m.grid[0] * m.grid[1] + (uint32_t)ay * m.grid[0] + (uint32_t)x;
            gs[q] = m.pbm[h];
            cnt[q] = m.pbm[h + 1];
        }
    }
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        const uint32_t e = tid + q * 256u;
        if (e < E) {
            cnt[q] -= gs[q];
            gstart[e] = gs[q];
            lstart[e] = cnt[q];
        }
    }
    __syncthreads();
    // exclusive scan of the bin counts: thread t owns entries [3t, 3t+3)
    uint32_t mine[3];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        const uint32_t e = tid * 3u + q;
        mine[q] = e < E ? lstart[e] : 0u;
        sum += mine[q];
    }
    const uint32_t inc = wave_inclusive(sum);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t off = 0, total = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        if (k < w) off += s_w[k];
        total += s_w[k];
    }
    uint32_t run = off + inc - sum;
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        const uint32_t e = tid * 3u + q;
        if (e < E) lstart[e] = run;
        run += mine[q];
    }
    if (tid == 0) lstart[E] = total;
    if (total > a.stage_cap) return false;
    __syncthreads();
    // copy: staged slot l -> its bin (binary search over the bin starts) -> global index;
    // consecutive slots map to consecutive messages, so the loads coalesce.
    const uint32_t cap = a.stage_cap, nv = a.stage_nvars;
    for (uint32_t l = tid; l < total; l += 256u) {
        uint32_t lo = 0, hi = E;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (lstart[mid] <= l) lo = mid;
            else hi = mid;
        }
        const uint32_t g = gstart[lo] + (l - lstart[lo]);
        uint32_t val[STAGE_MAX_VARS];
#pragma unroll
        for (uint32_t v = 0; v < STAGE_MAX_VARS; ++v) val[v] = v < nv ? a.stage_src[v][g] : 0u;
#pragma unroll
        for (uint32_t v = 0; v < STAGE_MAX_VARS; ++v)
            if (v < nv) vars[v * cap + l] = val[v];
    }
    __syncthreads();
    return true;
}

// Spatial-input agent functions: neighbourhood staged in LDS, then the user function.
template <typename F>
__global__ void __launch_bounds__(256) agent_function_kernel_staged(const FnArgs a) {
    extern __shared__ uint32_t fgx_stage_smem[];
    __shared__ StageHeader sh;
    const uint32_t n = a.d_n ? load_count(a.d_n) : a.n_cap;
    const uint32_t b0 = blockIdx.x * blockDim.x;
    if (b0 >= n) return;
    const uint32_t i = b0 + threadIdx.x;
    const bool active = i < n;
    bool staged = false;
    if (a.stage_ok) staged = stage_block<F::MsgIn::In::dims>(a, fgx_stage_smem, sh, i, active);
    if (!active) return;
    if (a.msg_out_flags) a.msg_out_flags[i] = a.msg_out_optional ? 0u : 1u;
    if (a.birth_flags) a.birth_flags[i] = 0u;
    DeviceAPI<typename F::MsgIn, typename F::MsgOut> api(a, i);
    if (staged) api.message_in.enable_stage(fgx_stage_smem, &sh);
    const AGENT_STATUS s = F()(&api);
    if (a.death_flags)
        a.death_flags[i] = (uint32_t)s;
    else if (s == DEAD)
        report_error(a.d_error, DERR_DEATH_DISABLED, 0, i);
}

template <typename F>
void launch_agent_function(const FnArgs& a, uint32_t grid, uint32_t block, uint32_t smem, void* stream) {
    constexpr MsgKind k = F::MsgIn::kind;
    if constexpr (k == MsgKind::Spatial2D || k == MsgKind::Spatial3D) {
        hipLaunchKernelGGL(agent_function_kernel_staged<F>, dim3(grid), dim3(block), smem, (hipStream_t)stream, a);
    } else {
        (void)smem;
        hipLaunchKernelGGL(agent_function_kernel<F>, dim3(grid), dim3(block), 0, (hipStream_t)stream, a);
    }
}
}  // namespace detail

}  // namespace flamegpu

// ------------------------------------------------------------------ macros
// include/flamegpu/runtime/AgentFunction_shim.cuh:32-89
#define FLAMEGPU_AGENT_FUNCTION(funcName, message_in, message_out)                                                  \
    struct funcName##_impl {                                                                                        \
        typedef message_in MsgIn;                                                                                   \
        typedef message_out MsgOut;                                                                                 \
        __device__ __forceinline__ flamegpu::AGENT_STATUS operator()(                                               \
            flamegpu::DeviceAPI<message_in, message_out>* FLAMEGPU) const;                                          \
    };                                                                                                              \
    [[maybe_unused]] constexpr funcName##_impl funcName{};                                                           \
    __device__ __forceinline__ flamegpu::AGENT_STATUS funcName##_impl::operator()(                                  \
        flamegpu::DeviceAPI<message_in, message_out>* FLAMEGPU) const

#define FLAMEGPU_DEVICE_FUNCTION __device__ __forceinline__
#define FLAMEGPU_HOST_DEVICE_FUNCTION __host__ __device__ __forceinline__
