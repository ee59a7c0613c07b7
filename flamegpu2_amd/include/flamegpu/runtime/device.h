// Device runtime: DeviceAPI, message input iterators / output writers, AgentRandom,
// DeviceEnvironment and the agent-function kernels.
//
// Same surface as include/flamegpu/runtime/DeviceAPI.cuh:163-557,
// include/flamegpu/runtime/messaging/*/Message*Device.cuh and
// include/flamegpu/runtime/random/AgentRandom.cuh; the variable tables are
// compile-time hashed names resolved against kernel-argument bindings
// (flamegpu/detail/core.h) instead of the curve shared-memory hash table.
//
// A binding is resolved to (base pointer, index mask); an unknown name or a wrong
// type reports a device error and selects the zero page (reads) or the sink page
// (writes) with a zero mask, so execution continues without a data-dependent
// branch.  The checks depend only on kernel arguments and compile-time name
// hashes: inside message loops they are loop-invariant scalar work.
#pragma once

#include <cmath>
#include <limits>
#include <type_traits>
#include <utility>

#include "flamegpu/model.h"

namespace flamegpu {

namespace detail {

// Per-block LDS copy of the spatial neighbourhood (agent_function_kernel_staged).
struct StageView {
    const uint32_t* lstart;   // per virtual bin entry: first staged slot (exclusive scan)
    const uint32_t* gstart;   // per virtual bin entry: first global message index
    const uint32_t* vars;     // staged variables, SoA of `cap` slots: x, y (, z) (, id)
    uint32_t cap;
    uint32_t has_id;
    const StageHeader* sh;
};

// Everything one agent-function call needs besides its DeviceAPI: the launch
// arguments, the agent's list index, the values that differ between the
// simulations of a batched launch, and the call's error state.
struct ThreadCtx {
    const FnArgs& a;
    ErrState err;               // where this call reports device errors
    const char* zero_page;      // copies of a.zero_page / a.sink_page held in registers
    char* sink_page;
    uint32_t idx;
    const SpatialDev& meta;     // message-input spatial metadata
    const char* env;            // environment property buffer
    const uint32_t* d_step;     // step counter
    uint64_t seed;
    const StageView* stage;     // null outside the staged kernel
    uint32_t msg_off;           // brute-force input range [msg_off, msg_off + msg_n)
    uint32_t msg_n;
    uint32_t lidx;              // index among the agents this launch runs (output slots, flags)
    uint32_t stream_id;         // RNG stream of the call
    __attribute__((address_space(3))) uint32_t* wtab;   // this wave's LDS region for spatial unit tables (null: none)
};

// Per-wave LDS region of the spatial iterators (kWaveTabWords words):
//   [0]            claim word: 0 until the first spatial filter of the call takes the table
//   [4 + 4g ..]    group g (g < kTabGroups): first x cell, last x cell, table word offset, row key
//   [kTabHead ..]  the unit table: per group, per neighbour row, the bin-table entries of the
//                  group's x range (wrapping iterators: a start/end pair per cell; clamped
//                  ones: the bin boundaries of the row).
constexpr uint32_t kWaveTabWords = 1024;
constexpr uint32_t kTabGroups = 4;
constexpr uint32_t kTabHead = 32;

// Global-memory pointers (address space 1), so loads through pointers rebuilt from
// integers stay global loads instead of flat loads.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

struct Bound {
    uint64_t addr;   // base address: the bound array, or the zero/sink page
    uint32_t mask;   // all ones when resolved, zero when addr is the zero/sink page
    template <typename T>
    __device__ __forceinline__ gptr<T> as() const {
        return (gptr<T>)addr;
    }
};

template <typename T>
__device__ __forceinline__ Bound resolve(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, uint32_t h, uint32_t len,
                                         const ThreadCtx& c, uint32_t err_unknown, uint32_t err_type, bool write) {
    // one 16-byte load of the whole binding, then selects on the values (a select
    // of addresses would chain two dependent scalar loads)
    const uint4 raw = *reinterpret_cast<const uint4*>(&tab[table_slot(h, mul, VAR_SLOT_BITS)]);
    // name hash and element bytes (whose reserved upper half-word is zero) checked by
    // one 64-bit compare against a compile-time constant; one branch to the report
    const uint64_t got = ((uint64_t)raw.y << 32) | raw.x;
    const uint64_t want = ((uint64_t)((uint32_t)sizeof(T) * len) << 32) | h;
    const bool ok = got == want;
    if (__builtin_expect(!ok, 0)) report_error(c.err.d_error, raw.x != h ? err_unknown : err_type, h, c.err.index);
    const uint64_t bound = ((uint64_t)raw.w << 32) | raw.z;
    const uint64_t fallback = write ? (uint64_t)c.sink_page : (uint64_t)c.zero_page;
    return Bound{ok ? bound : fallback, ok ? 0xFFFFFFFFu : 0u};
}

template <typename T, unsigned N>
__device__ __forceinline__ T load_var(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, const char (&name)[N], uint32_t idx,
                                      const ThreadCtx& c, uint32_t err_unknown, uint32_t err_type) {
    const uint32_t h = name_hash(name, N);
    const Bound b = resolve<T>(tab, mul, h, 1, c, err_unknown, err_type, false);
    return b.as<const T>()[idx & b.mask];
}

template <typename T, unsigned N>
__device__ __forceinline__ T load_elem(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, const char (&name)[N], uint32_t idx,
                                       uint32_t len, uint32_t elem, const ThreadCtx& c, uint32_t err_unknown,
                                       uint32_t err_type) {
    const uint32_t h = name_hash(name, N);
    const Bound b = resolve<T>(tab, mul, h, len, c, err_unknown, err_type, false);
    const bool in_range = elem < len;
    c.err.note(!in_range, DERR_ARRAY_INDEX, h);
    const uint64_t m = (b.mask && in_range) ? ~0ull : 0ull;
    return b.as<const T>()[((uint64_t)idx * len + elem) & m];
}

template <typename T, unsigned N>
__device__ __forceinline__ void store_var(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, const char (&name)[N], uint32_t idx,
                                          T v, const ThreadCtx& c, uint32_t err_unknown, uint32_t err_type) {
    const uint32_t h = name_hash(name, N);
    const Bound b = resolve<T>(tab, mul, h, 1, c, err_unknown, err_type, true);
    b.as<T>()[idx & b.mask] = v;
}

template <typename T, unsigned N>
__device__ __forceinline__ void store_elem(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, const char (&name)[N], uint32_t idx,
                                           uint32_t len, uint32_t elem, T v, const ThreadCtx& c, uint32_t err_unknown,
                                           uint32_t err_type) {
    const uint32_t h = name_hash(name, N);
    const Bound b = resolve<T>(tab, mul, h, len, c, err_unknown, err_type, true);
    const bool in_range = elem < len;
    c.err.note(!in_range, DERR_ARRAY_INDEX, h);
    const uint64_t m = (b.mask && in_range) ? ~0ull : 0ull;
    b.as<T>()[((uint64_t)idx * len + elem) & m] = v;
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
    __device__ __forceinline__ explicit AgentRandom(const detail::ThreadCtx& c) : c_(c) {}
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
            const detail::FnArgs& a = c_.a;
            uint32_t c[4] = {blk, c_.stream_id, c_.d_step ? *c_.d_step : 0u, a.ids ? a.ids[c_.idx] : c_.idx};
            detail::philox4x32_10(c, (uint32_t)c_.seed, (uint32_t)(c_.seed >> 32));
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
    const detail::ThreadCtx& c_;
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
    c_.err.note(min > max, detail::DERR_RANDOM_RANGE, 0);
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
    __device__ __forceinline__ explicit DeviceEnvironment(const detail::ThreadCtx& c) : c_(c) {}
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
        return c_.a.env_b[detail::table_slot(h, c_.a.env_mul, detail::ENV_SLOT_BITS)].hash == h;
    }

 private:
    template <typename T, unsigned N>
    __device__ __forceinline__ T get(const char (&name)[N], uint32_t len, uint32_t i) const {
        const uint32_t h = detail::name_hash(name, N);
        const uint2 raw = *reinterpret_cast<const uint2*>(&c_.a.env_b[detail::table_slot(h, c_.a.env_mul, detail::ENV_SLOT_BITS)]);
        const bool known = raw.x == h;
        const bool typed = (raw.y >> 16) == (uint32_t)sizeof(T) * len && i < len;
        c_.err.note(!known, detail::DERR_UNKNOWN_ENV, h);
        c_.err.note(known && !typed, detail::DERR_ENV_TYPE, h);
        const bool ok = known && typed;
        const T* p = reinterpret_cast<const T*>(ok ? c_.env + (raw.y & 0xFFFFu) : c_.zero_page);
        return p[ok ? i : 0u];
    }
    const detail::ThreadCtx& c_;
};

// ------------------------------------------------------------------ message inputs
class MessageNone::In {
 public:
    __device__ __forceinline__ explicit In(const detail::ThreadCtx&) {}
};
class MessageNone::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx&) {}
};

// include/flamegpu/runtime/messaging/MessageBruteForce/MessageBruteForceDevice.cuh:19-172
class MessageBruteForce::In {
 public:
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c), off_(c.msg_off), len_(c.msg_n) {}
    class iterator;
    class Message {
     public:
        __device__ __forceinline__ Message(const detail::ThreadCtx& c, uint32_t i) : c_(&c), i_(i) {}
        template <typename T, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
            return detail::load_var<T>(c_->a.msg_in, c_->a.msg_in_mul, name, i_, *c_, detail::DERR_UNKNOWN_MSG_VAR,
                                       detail::DERR_MSG_VAR_TYPE);
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
            return detail::load_elem<T>(c_->a.msg_in, c_->a.msg_in_mul, name, i_, L, e, *c_, detail::DERR_UNKNOWN_MSG_VAR,
                                        detail::DERR_MSG_VAR_TYPE);
        }
        __device__ __forceinline__ size_type getIndex() const { return i_ - c_->msg_off; }

     private:
        friend class iterator;
        const detail::ThreadCtx* c_;
        uint32_t i_;
    };
    // operator* gives a Message& (MessageBruteForceDevice.cuh iterator), so loops may bind `auto&`
    class iterator {
     public:
        __device__ __forceinline__ iterator(const detail::ThreadCtx& c, uint32_t i) : m_(c, i) {}
        __device__ __forceinline__ iterator& operator++() {
            ++m_.i_;
            return *this;
        }
        __device__ __forceinline__ bool operator!=(const iterator& o) const { return m_.i_ != o.m_.i_; }
        __device__ __forceinline__ bool operator==(const iterator& o) const { return m_.i_ == o.m_.i_; }
        __device__ __forceinline__ Message& operator*() { return m_; }
        __device__ __forceinline__ Message* operator->() { return &m_; }

     private:
        Message m_;
    };
    __device__ __forceinline__ iterator begin() const { return iterator(c_, off_); }
    __device__ __forceinline__ iterator end() const { return iterator(c_, off_ + len_); }
    __device__ __forceinline__ size_type size() const { return len_; }
    __device__ __forceinline__ Message at(size_type i) const { return Message(c_, off_ + i); }

 private:
    const detail::ThreadCtx& c_;
    uint32_t off_, len_;
};

// include/flamegpu/runtime/messaging/MessageBruteForce/MessageBruteForceDevice.cuh:259-294
class MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : c_(c) {}
    template <typename T, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], T value) const {
        const bool on = c_.a.msg_out_flags != nullptr;
        c_.err.note(!on, detail::DERR_MSG_OUTPUT_DISABLED, 0);
        detail::store_var<T>(c_.a.msg_out, c_.a.msg_out_mul, name, c_.lidx, value, c_, detail::DERR_UNKNOWN_MSG_VAR,
                             detail::DERR_MSG_VAR_TYPE);
        if (on && c_.a.msg_out_optional) c_.a.msg_out_flags[c_.lidx] = 1u;
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], unsigned int e, T value) const {
        const bool on = c_.a.msg_out_flags != nullptr;
        c_.err.note(!on, detail::DERR_MSG_OUTPUT_DISABLED, 0);
        detail::store_elem<T>(c_.a.msg_out, c_.a.msg_out_mul, name, c_.lidx, L, e, value, c_, detail::DERR_UNKNOWN_MSG_VAR,
                              detail::DERR_MSG_VAR_TYPE);
        if (on && c_.a.msg_out_optional) c_.a.msg_out_flags[c_.lidx] = 1u;
    }

 protected:
    const detail::ThreadCtx& c_;
};

namespace detail {
// getGridPosition / getHash (MessageSpatial3DDevice.cuh:646-672).
__device__ __forceinline__ int grid_pos(const SpatialDev& m, int a, float p) {
    const int g = (int)floorf((p - m.min[a]) / m.cell);   // the bin width: the radius, or half of it
    const int hi = (int)m.grid[a] - 1;
    return g < 0 ? 0 : (g > hi ? hi : g);
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
// (c + d) mod g for any sign and size of d (a search reach may exceed a tiny grid)
__device__ __forceinline__ int wrap_mod(int c, int d, int g) {
    const int a = (c + d) % g;
    return a < 0 ? a + g : a;
}
constexpr uint32_t kHashX = name_hash("x", 2);
constexpr uint32_t kHashY = name_hash("y", 2);
constexpr uint32_t kHashZ = name_hash("z", 2);
constexpr uint32_t kHashId = name_hash("id", 3);
}  // namespace detail

// Spatial message iteration shared by the 2-D and 3-D filters.
//
// Visit order of the reference iterators:
//  Filter 3-D (MessageSpatial3DDevice.cuh:52-59,693-719): 9 x-strips, dy outer, dz inner;
//  WrapFilter 3-D (:264-279,727-749): 27 cells, dx outer, dy, dz inner, modulo grid;
//  Filter 2-D (MessageSpatial2DDevice.cuh:646-672): 3 strips; WrapFilter 2-D (:679-698): 9 cells.
// Units (strips or cells) are visited with up to three counters (c0 fastest) and the
// matching coordinate parts of the unit's first bin (pre-multiplied by their stride
// and wrapped in place for the wrapping iterators, raw row coordinates for the
// clamped ones), advanced with selects only: no divisions, no per-lane branches.
//
// The iterator keeps one candidate of look-ahead: while the body runs on message k,
// the load of message k+1's position and id (one 16-byte record, written with the
// list's index) is in flight, and on the per-lane walk the bin-table range of the unit
// after the current one is loaded when the current one starts.  Messages read through
// their position / id use those registers; other variables are loaded at access time.
// On the wave's unit table a lane starts with a bit mask of its non-empty units and
// takes the lowest remaining one whenever its current unit ends.
//
// Source: the message list in HBM (bin table pbm, packed records), or, in the staged
// kernel when every lane of the wave is covered by the block's LDS copy, the LDS tables
// (units then name virtual bin entries and the look-ahead walks LDS slots).
template <int DIMS, bool WRAP>
class SpatialFilter {
 public:
    // counter levels: wrap 3-D (dz, dy, dx), wrap 2-D (dy, dx), clamped 3-D (dz, dy), clamped 2-D (dy);
    // each counter runs over 2R+1 values for a reach of R bins either side (R = 1 in the
    // reference; 2 with half-width bins), so there are (2R+1)^levels units
    static constexpr int kLevels = WRAP ? (DIMS == 3 ? 3 : 2) : (DIMS == 3 ? 2 : 1);
    struct Unit {
        int c0, c1, c2;   // counters, 0..2
        int p0, p1, p2;   // coordinate parts of the unit's first bin
    };

    __device__ __forceinline__ SpatialFilter(const detail::ThreadCtx& c, float x, float y, float z, bool allow_stage)
        : c_(c) {
        const detail::SpatialDev& m = c.meta;
        loc_[0] = x;
        loc_[1] = y;
        loc_[2] = z;
        width_[0] = m.width[0];
        width_[1] = m.width[1];
        width_[2] = m.width[2];
        const int cx = detail::grid_pos(m, 0, x);
        const int cy = detail::grid_pos(m, 1, y);
        const int cz = DIMS == 3 ? detail::grid_pos(m, 2, z) : 0;
        // the hot variables come from the list's packed {x, y, z, id} records (an absent or
        // non-4-byte id reads as 0, and reading it reports the error below)
        const detail::VarBinding& bi = c.a.msg_in[detail::table_slot(detail::kHashId, c.a.msg_in_mul, detail::VAR_SLOT_BITS)];
        const bool id_known = bi.hash == detail::kHashId;
        const bool id_ok = id_known && bi.bytes == 4;
        id_err_ = id_ok ? 0u : (id_known ? (uint32_t)detail::DERR_MSG_VAR_TYPE : (uint32_t)detail::DERR_UNKNOWN_MSG_VAR);
        pbm_ = m.pbm;
        rec_ = reinterpret_cast<const uint4*>(c.a.in_meta.packed);   // the launch's (a batch: all simulations')
        sv_ = nullptr;
        tab_ = nullptr;
        tm_ = 0u;
        // LDS source when the whole wave is covered by the block's staged neighbourhood
        if (allow_stage && c.stage) {
            const detail::StageHeader* sh = c.stage->sh;
            int covered = 0, ebase = 0, xbase = 0, width = 0;
            for (uint32_t k = 0; k < sh->ngroups; ++k) {
                const detail::StageGroup& g = sh->g[k];
                if (g.cy == cy && g.cz == cz && cx >= g.xlo && cx <= g.xhi) {
                    covered = 1;
                    ebase = (int)g.ebase;
                    xbase = g.xbase;
                    width = g.width;
                }
            }
            if (__builtin_amdgcn_readfirstlane(__all(covered) ? 1 : 0)) {
                sv_ = c.stage;
                plan(m, cx, cy, cz, ebase, xbase, width);
                return;
            }
        }
        plan(m, cx, cy, cz, 0, 0, 0);
        if (c.wtab) build_table(cx, cy, cz);
    }

    class iterator;
    class Message {
     public:
        __device__ __forceinline__ Message(const SpatialFilter& f, uint32_t g, float x, float y, float z, uint32_t id)
            : f_(f), g_(g), x_(x), y_(y), z_(z), id_(id) {}
        template <typename T, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
            const uint32_t h = detail::name_hash(name, N);
            if (h == detail::kHashX || h == detail::kHashY || (DIMS == 3 && h == detail::kHashZ)) {
                if constexpr (std::is_same<T, float>::value) {
                    return h == detail::kHashX ? x_ : (h == detail::kHashY ? y_ : z_);
                } else {
                    f_.c_.err.note(true, detail::DERR_MSG_VAR_TYPE, h);
                    return T();
                }
            }
            if constexpr (sizeof(T) == 4) {
                if (h == detail::kHashId) {
                    f_.c_.err.note(f_.id_err_ != 0u, f_.id_err_, h);
                    T v;
                    __builtin_memcpy(&v, &id_, 4);
                    return v;
                }
            }
            return detail::load_var<T>(f_.c_.a.msg_in, f_.c_.a.msg_in_mul, name, g_, f_.c_, detail::DERR_UNKNOWN_MSG_VAR,
                                       detail::DERR_MSG_VAR_TYPE);
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
            return detail::load_elem<T>(f_.c_.a.msg_in, f_.c_.a.msg_in_mul, name, g_, L, e, f_.c_,
                                        detail::DERR_UNKNOWN_MSG_VAR, detail::DERR_MSG_VAR_TYPE);
        }
        __device__ __forceinline__ float getVirtualX() const { return getVirtualX(f_.loc_[0]); }
        __device__ __forceinline__ float getVirtualY() const { return getVirtualY(f_.loc_[1]); }
        __device__ __forceinline__ float getVirtualZ() const { return getVirtualZ(f_.loc_[2]); }
        __device__ __forceinline__ float getVirtualX(float x1) const { return detail::virtual_coord(x_, x1, f_.width_[0]); }
        __device__ __forceinline__ float getVirtualY(float y1) const { return detail::virtual_coord(y_, y1, f_.width_[1]); }
        __device__ __forceinline__ float getVirtualZ(float z1) const { return detail::virtual_coord(z_, z1, f_.width_[2]); }
        __device__ __forceinline__ size_type getIndex() const { return g_; }

     private:
        friend class iterator;
        const SpatialFilter& f_;
        uint32_t g_;
        float x_, y_, z_;
        uint32_t id_;
    };

    // operator* gives a Message& (MessageSpatial3DDevice.cuh iterator), so loops may bind `auto&`
    class iterator {
     public:
        // end sentinel
        __device__ __forceinline__ explicit iterator(const SpatialFilter& f) : f_(f), m_(f, 0, 0.f, 0.f, 0.f, 0), valid_(false) {}
        // first message
        __device__ __forceinline__ iterator(const SpatialFilter& f, int) : f_(f), m_(f, 0, 0.f, 0.f, 0.f, 0) {
            pu_ = 0;
            if (f_.tm_) {
                // the lane's non-empty units, one bit each (at most 32 units on the table path)
                uint32_t mask = 0;
                for (int u = 0; u < f_.units_; ++u) {
                    uint32_t s, e;
                    read_unit((uint32_t)u, s, e);
                    mask |= (e > s ? 1u : 0u) << u;
                }
                nm_ = mask;
                l_ = e_ = 0;
                take_unit();
            } else {
                f_.first_unit(k_);
                f_.load_unit(k_, ps_, pe_, po_);
                next_unit();
            }
            fetch();
            shift();
            step();
        }
        __device__ __forceinline__ iterator& operator++() {
            shift();
            step();
            return *this;
        }
        __device__ __forceinline__ bool operator!=(const iterator&) const { return valid_; }
        __device__ __forceinline__ Message& operator*() { return m_; }
        __device__ __forceinline__ Message* operator->() { return &m_; }

     private:
        // Move the look-ahead to the first message of the next non-empty unit.
        __device__ __forceinline__ void next_unit() {
            more_ = false;
#pragma clang loop unroll(disable)
            while (pu_ < f_.units_) {
                l_ = ps_;
                e_ = pe_;
                o_ = po_;
                ++pu_;
                if (pu_ < f_.units_) {
                    f_.next_unit(k_);
                    f_.load_unit(k_, ps_, pe_, po_);
                }
                if (l_ < e_) {
                    more_ = true;
                    return;
                }
            }
        }
        // Unit table: the range [s, e) of the lane's unit u.
        __device__ __forceinline__ void read_unit(uint32_t u, uint32_t& s, uint32_t& e) const {
            const __attribute__((address_space(3))) uint32_t* t = f_.tab_;
            const uint32_t at = f_.tlane_ + u * (WRAP ? 2u : 1u);
            s = t[at];
            e = t[WRAP ? at + 1u : at + f_.tstride_];
        }
        // Unit table: a lane whose look-ahead left its unit takes its next non-empty unit
        // (the lowest bit of nm_), with selects, not per-lane branches.  Every unit taken
        // holds a message, so one take per step is enough; with none left, l_ >= e_ stays.
        __device__ __forceinline__ void take_unit() {
            const bool need = l_ >= e_;
            const uint32_t u = nm_ ? (uint32_t)__builtin_ctz(nm_) : 0u;
            uint32_t s, e;
            read_unit(u, s, e);
            const bool take = need && nm_ != 0u;
            l_ = take ? s : l_;
            e_ = take ? e : e_;
            nm_ = need ? (nm_ & (nm_ - 1u)) : nm_;
            more_ = l_ < e_;
        }
        __device__ __forceinline__ void step() {
            ++l_;
            if (f_.tm_) take_unit();
            else if (l_ >= e_) next_unit();
            fetch();
        }
        // Unconditional (index 0 past the last message) so the loads issue without a branch.
        __device__ __forceinline__ void fetch() { f_.fetch(more_ ? l_ : 0u, nx_, ny_, nz_, nid_); }
        __device__ __forceinline__ void shift() {
            valid_ = more_;
            m_.g_ = l_ + o_;
            m_.x_ = nx_;
            m_.y_ = ny_;
            m_.z_ = nz_;
            m_.id_ = nid_;
        }

        const SpatialFilter& f_;
        Message m_;   // current message
        bool valid_ = false, more_ = false;
        // look-ahead: source index l_ in unit [.., e_), global index l_ + o_, its values
        uint32_t l_ = 0, e_ = 0, o_ = 0;
        float nx_ = 0.f, ny_ = 0.f, nz_ = 0.f;
        uint32_t nid_ = 0;
        // the unit after the look-ahead's: its number, counters and loaded range
        int pu_ = 0;
        Unit k_;
        uint32_t ps_ = 0, pe_ = 0, po_ = 0;
        // unit table: the lane's non-empty units not yet taken, one bit each
        uint32_t nm_ = 0;
    };
    __device__ __forceinline__ iterator begin() const { return iterator(*this, 0); }
    __device__ __forceinline__ iterator end() const { return iterator(*this); }

 private:
    // Per-lane search origin.  HBM: bins of the grid.  LDS (ebase/xbase/width of the
    // covering group): virtual bin entries, rows of `width` entries for the 9 (3)
    // neighbour rows, x counted from xbase.
    __device__ __forceinline__ void plan(const detail::SpatialDev& m, int cx, int cy, int cz, int ebase, int xbase,
                                         int width) {
        gx_ = (int)m.grid[0];
        gy_ = (int)m.grid[1];
        gz_ = (int)m.grid[2];
        const int R = m.span;
        last_ = 2 * R;
        units_ = kLevels == 3 ? (last_ + 1) * (last_ + 1) * (last_ + 1)
                              : (kLevels == 2 ? (last_ + 1) * (last_ + 1) : last_ + 1);
        cx_ = cx;
        ebase_ = ebase;
        xbase_ = xbase;
        ewidth_ = width;
        const int sy = gx_, sz = gx_ * gy_;
        if (WRAP) {
            xs_ = 0;
            span_ = 0;
            if (DIMS == 3) {   // c0: z, c1: y, c2: x
                st0_ = detail::wrap_mod(cz, -R, gz_) * sz;
                st1_ = detail::wrap_mod(cy, -R, gy_) * sy;
                st2_ = detail::wrap_mod(cx, -R, gx_);
                m0_ = (gz_ - 1) * sz;
                m1_ = (gy_ - 1) * sy;
                m2_ = gx_ - 1;
                s0_ = sz;
                s1_ = sy;
            } else {           // c0: y, c1: x
                st0_ = detail::wrap_mod(cy, -R, gy_) * sy;
                st1_ = detail::wrap_mod(cx, -R, gx_);
                st2_ = 0;
                m0_ = (gy_ - 1) * sy;
                m1_ = gx_ - 1;
                m2_ = 0;
                s0_ = sy;
                s1_ = 1;
            }
        } else {
            xs_ = cx - R < 0 ? 0 : cx - R;
            const int xe = cx + R > gx_ - 1 ? gx_ - 1 : cx + R;
            span_ = xe - xs_;
            // raw row coordinates; c0: z (3-D) or y (2-D), c1: y (3-D)
            st0_ = DIMS == 3 ? cz - R : cy - R;
            st1_ = cy - R;
            st2_ = 0;
            m0_ = m1_ = m2_ = 0;
            s0_ = s1_ = 1;
        }
    }
    // The wave's unit table (see the class comment).  Called by every lane that builds the
    // filter, after plan().  Leaves tab_ null (units from HBM) when the table is taken or
    // the rows do not fit.  Lanes exchange the group bounds through registers (shuffles and
    // readlane), and the table through volatile LDS accesses, which the compiler keeps in
    // program order (LDS executes one wave's accesses in order).
    __device__ __forceinline__ void build_table(int cx, int cy, int cz) {
        volatile __attribute__((address_space(3))) uint32_t* w = c_.wtab;
        if (units_ > 32) return;   // the walk keeps one bit per unit
        const uint32_t claimed = __builtin_amdgcn_readfirstlane(w[0]);
        if (claimed) return;
        w[0] = 1u;
        const int R = last_ / 2;
        const int reach = last_ + 1;                        // cells per axis in a neighbourhood
        const int nrows = DIMS == 3 ? reach * reach : reach;
        const int rk = DIMS == 3 ? cz * gy_ + cy : cy;      // the cell row
        const uint32_t lane = __lane_id();
        const uint64_t me = 1ull << lane;
        const uint64_t act = __ballot(1);
        // groups: runs of adjacent active lanes with the same row
        const int rk_prev = __shfl(rk, lane ? (int)lane - 1 : 0, 64);
        const bool lead = lane == 0 || !((act >> (lane - 1)) & 1ull) || rk_prev != rk;
        const uint64_t leads = __ballot(lead);
        const uint32_t ng = (uint32_t)__builtin_popcountll(leads);
        if (ng > detail::kTabGroups) return;
        const uint32_t g = (uint32_t)__builtin_popcountll(leads & (me | (me - 1ull))) - 1u;
        const int first = 63 - __builtin_clzll(leads & (me | (me - 1ull)));   // lane of this run's leader
        // x range of each run: an inclusive max/min scan within the run, read at its last lane
        int lo = cx, hi = cx;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int lo2 = __shfl_up(lo, o, 64), hi2 = __shfl_up(hi, o, 64);
            const bool in = (int)lane - o >= first;
            lo = in && lo2 < lo ? lo2 : lo;
            hi = in && hi2 > hi ? hi2 : hi;
        }
        const uint64_t lasts = __ballot(lane == 63 || !((act >> (lane + 1)) & 1ull) || ((leads >> (lane + 1)) & 1ull));
        constexpr uint32_t wpc = WRAP ? 2u : 1u;           // words per entry
        int gx0[detail::kTabGroups], growk[detail::kTabGroups];
        uint32_t gper[detail::kTabGroups], gcells[detail::kTabGroups], gbase[detail::kTabGroups];
        uint32_t off = detail::kTabHead, total = 0;
        uint64_t rest = lasts;
#pragma unroll
        for (uint32_t k = 0; k < detail::kTabGroups; ++k) {
            const bool on = k < ng;
            const int L = on ? __builtin_ctzll(rest) : 0;
            rest &= rest - 1ull;
            const int glo = __builtin_amdgcn_readlane(lo, L), ghi = __builtin_amdgcn_readlane(hi, L);
            gx0[k] = glo - R;
            growk[k] = on ? __builtin_amdgcn_readlane(rk, L) : 0;
            // wrapping: the cells [lo - R, hi + R]; clamped: the boundaries [lo - R, hi + R + 1]
            gper[k] = on ? (uint32_t)(ghi - glo + 2 * R + 1) + (WRAP ? 0u : 1u) : 0u;
            gcells[k] = gper[k] * (uint32_t)nrows;
            gbase[k] = off;
            off += gcells[k] * wpc;
            total += gcells[k];
        }
        if (off > detail::kWaveTabWords) return;
        // fill: the active lanes take the entries in turn, four loads in flight each
        const uint32_t rank = (uint32_t)__builtin_popcountll(act & (me - 1ull));
        const uint32_t nact = (uint32_t)__builtin_popcountll(act);
        for (uint32_t q0 = rank; q0 < total; q0 += 4u * nact) {
            uint32_t dst[4], src[4], v0[4], v1[4];
            bool ok[4];
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                uint32_t q = q0 + j * nact;
                ok[j] = q < total;
                q = ok[j] ? q : 0u;
                uint32_t k = 0;
#pragma unroll
                for (uint32_t t = 0; t + 1 < detail::kTabGroups; ++t) {
                    const bool past = k == t && q >= gcells[t];
                    q -= past ? gcells[t] : 0u;
                    k += past ? 1u : 0u;
                }
                int x0 = gx0[0], rkk = growk[0];
                uint32_t per = gper[0], base = gbase[0];
#pragma unroll
                for (uint32_t t = 1; t < detail::kTabGroups; ++t) {
                    x0 = k == t ? gx0[t] : x0;
                    rkk = k == t ? growk[t] : rkk;
                    per = k == t ? gper[t] : per;
                    base = k == t ? gbase[t] : base;
                }
                const uint32_t row = q / per, v = q - row * per;
                const int ry = DIMS == 3 ? rkk % gy_ : rkk, rz = DIMS == 3 ? rkk / gy_ : 0;
                const int dy = (DIMS == 3 ? (int)row / reach : (int)row) - R;
                const int dz = DIMS == 3 ? (int)row % reach - R : 0;
                const int vx = x0 + (int)v;
                uint32_t bin;
                bool in_grid = true;
                if (WRAP) {
                    const int y = detail::wrap_mod(ry, dy, gy_), z = DIMS == 3 ? detail::wrap_mod(rz, dz, gz_) : 0;
                    bin = (uint32_t)((z * gy_ + y) * gx_ + detail::wrap_mod(vx, 0, gx_));
                } else {
                    const int y = ry + dy, z = rz + dz;
                    in_grid = y >= 0 && y < gy_ && (DIMS == 2 || (z >= 0 && z < gz_));
                    const int xc = vx < 0 ? 0 : (vx > gx_ ? gx_ : vx);   // a boundary: [0, gx]
                    bin = in_grid ? (uint32_t)((z * gy_ + y) * gx_ + xc) : 0u;
                }
                dst[j] = base + (v * (uint32_t)nrows + row) * wpc;
                src[j] = bin;
                ok[j] = ok[j] && in_grid;
            }
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                v0[j] = ok[j] ? pbm_[src[j]] : 0u;
                v1[j] = (WRAP && ok[j]) ? pbm_[src[j] + 1u] : 0u;
            }
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                if (q0 + j * nact < total) {
                    w[dst[j]] = v0[j];
                    if (WRAP) w[dst[j] + 1u] = v1[j];
                }
            }
        }
        // this lane's row 0 entry for its first unit, and the words per row of its group
        uint32_t base = gbase[0], per = gper[0];
        int x0 = gx0[0];
#pragma unroll
        for (uint32_t t = 1; t < detail::kTabGroups; ++t) {
            base = g == t ? gbase[t] : base;
            per = g == t ? gper[t] : per;
            x0 = g == t ? gx0[t] : x0;
        }
        // Entries are stored column-major (x outer, neighbour row inner), so this lane's units
        // are consecutive entries: wrapping, unit u = (x cell, row) is entry tlane_ + u of cell
        // cx - R's column; clamped, unit = row, with the strip's end boundary tstride_ entries on.
        tlane_ = base + (uint32_t)((WRAP ? cx - R : xs_) - x0) * (uint32_t)nrows * wpc;
        tstride_ = WRAP ? 0u : (uint32_t)(span_ + 1) * (uint32_t)nrows;
        tab_ = c_.wtab;
        tm_ = 1u;
    }
    // Unit table: c0 counts neighbour rows, c1 cells along x (wrapping) and p0 is the entry.
    __device__ __forceinline__ void first_unit(Unit& k) const {
        k.c0 = k.c1 = k.c2 = 0;
        k.p0 = st0_;
        k.p1 = st1_;
        k.p2 = st2_;
    }
    __device__ __forceinline__ void next_unit(Unit& k) const {
        const bool w0 = k.c0 == last_;
        k.c0 = w0 ? 0 : k.c0 + 1;
        if (WRAP) {
            k.p0 = w0 ? st0_ : (k.p0 == m0_ ? 0 : k.p0 + s0_);
            const bool w1 = w0 && k.c1 == last_;
            k.c1 = w0 ? (w1 ? 0 : k.c1 + 1) : k.c1;
            k.p1 = w0 ? (w1 ? st1_ : (k.p1 == m1_ ? 0 : k.p1 + s1_)) : k.p1;
            if (kLevels == 3) {
                k.c2 = w1 ? k.c2 + 1 : k.c2;
                k.p2 = w1 ? (k.p2 == m2_ ? 0 : k.p2 + 1) : k.p2;
            }
        } else {
            k.p0 = w0 ? st0_ : k.p0 + 1;
            if (kLevels == 2) {
                k.c1 = w0 ? k.c1 + 1 : k.c1;
                k.p1 = w0 ? k.p1 + 1 : k.p1;
            }
        }
    }
    // Range [s, e) of unit k in the source's index space; o = global - source index.
    __device__ __forceinline__ void load_unit(const Unit& k, uint32_t& s, uint32_t& e, uint32_t& o) const {
        bool ok = true;
        if (!WRAP) {
            if (DIMS == 3) ok = k.p0 >= 0 && k.p0 < gz_ && k.p1 >= 0 && k.p1 < gy_;
            else ok = k.p0 >= 0 && k.p0 < gy_;
        }
        uint32_t first;
        if (sv_) {
            // rows: dy * 3 + dz (3-D) or dy (2-D); x: virtual, from xbase
            int row, vx;
            if (WRAP) {
                row = DIMS == 3 ? 3 * k.c1 + k.c0 : k.c0;
                vx = cx_ - 1 + (DIMS == 3 ? k.c2 : k.c1);
            } else {
                row = DIMS == 3 ? 3 * k.c1 + k.c0 : k.c0;
                vx = xs_;
            }
            first = ok ? (uint32_t)(ebase_ + row * ewidth_ + (vx - xbase_)) : 0u;
            s = sv_->lstart[first];
            const uint32_t t = sv_->lstart[first + (uint32_t)span_ + 1u];
            o = sv_->gstart[first] - s;
            e = ok ? t : s;
        } else {
            if (WRAP) {
                first = (uint32_t)(k.p0 + k.p1 + (DIMS == 3 ? k.p2 : 0));
            } else if (DIMS == 3) {
                first = ok ? (uint32_t)((k.p0 * gy_ + k.p1) * gx_ + xs_) : 0u;
            } else {
                first = ok ? (uint32_t)(k.p0 * gx_ + xs_) : 0u;
            }
            uint32_t t;
            if (WRAP) {
                uint32_t pair[2];
                __builtin_memcpy(pair, pbm_ + first, 8);
                s = pair[0];
                t = pair[1];
            } else {
                s = pbm_[first];
                t = pbm_[first + (uint32_t)span_ + 1u];
            }
            // use the loaded range here (the empty asm needs its registers), so that no global
            // load is outstanding for the range where the walk uses it: a wait there would also
            // wait for the message loads in flight, on every unit change
            asm volatile("" ::"v"(s), "v"(t));
            o = 0;
            e = ok ? t : s;
        }
    }
    __device__ __forceinline__ void fetch(uint32_t l, float& x, float& y, float& z, uint32_t& id) const {
        if (sv_) {
            const uint32_t cap = sv_->cap;
            const float* v = reinterpret_cast<const float*>(sv_->vars);
            x = v[l];
            y = v[cap + l];
            z = DIMS == 3 ? v[2 * cap + l] : 0.0f;
            id = sv_->has_id ? sv_->vars[(uint32_t)DIMS * cap + l] : 0u;
        } else {
            // one 16-byte load per candidate: the gathers of a wave touch several bins' lines
            // at once, and the texture-address unit pays per line and per load instruction
            const uint4 r = rec_[l];
            x = __builtin_bit_cast(float, r.x);
            y = __builtin_bit_cast(float, r.y);
            z = DIMS == 3 ? __builtin_bit_cast(float, r.z) : 0.0f;
            id = r.w;
        }
    }

    const detail::ThreadCtx& c_;
    const detail::StageView* sv_;
    const __attribute__((address_space(3))) uint32_t* tab_;   // the wave's unit table, or null
    uint32_t tm_;                   // 1 when units come from the table (wave-uniform)
    uint32_t tlane_, tstride_;      // table: this lane's first unit word; clamped: words to its strip end
    float loc_[3];
    float width_[3];
    int gx_, gy_, gz_;              // grid
    int cx_;                        // search origin's x cell (LDS source)
    int xs_, span_;                 // non-wrapping strips: first x, last x - first x
    int last_, units_;              // last counter value (2R) and unit count ((2R+1)^levels)
    int st0_, st1_, st2_;           // coordinate parts of the first unit
    int m0_, m1_, m2_, s0_, s1_;    // wrapping: last part value and stride per level
    int ebase_, xbase_, ewidth_;    // LDS source: covering group
    const uint32_t* pbm_;
    const uint4* rec_;              // packed {x, y, z, id} record per message (HBM source)
    uint32_t id_err_;
};

// include/flamegpu/runtime/messaging/MessageSpatial3D/MessageSpatial3DDevice.cuh:503-558
class MessageSpatial3D::In {
 public:
    static constexpr int dims = 3;
    typedef SpatialFilter<3, false> Filter;
    typedef SpatialFilter<3, true> WrapFilter;
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c) {}
    __device__ __forceinline__ Filter operator()(float x, float y, float z) const { return Filter(c_, x, y, z, true); }
    __device__ __forceinline__ WrapFilter wrap(float x, float y, float z) const {
        const detail::SpatialDev& m = c_.meta;
        const bool oob = x > m.max[0] || y > m.max[1] || z > m.max[2] || x < m.min[0] || y < m.min[1] || z < m.min[2];
        c_.err.note(oob, detail::DERR_WRAP_OUT_OF_BOUNDS, 0);
        c_.err.note(!m.wrap_ok, detail::DERR_WRAP_NOT_FACTOR, 0);
        return WrapFilter(c_, oob ? m.min[0] : x, oob ? m.min[1] : y, oob ? m.min[2] : z, !oob);
    }
    __device__ __forceinline__ float radius() const { return c_.meta.radius; }

 private:
    const detail::ThreadCtx& c_;
};

// include/flamegpu/runtime/messaging/MessageSpatial2D/MessageSpatial2DDevice.cuh
class MessageSpatial2D::In {
 public:
    static constexpr int dims = 2;
    typedef SpatialFilter<2, false> Filter;
    typedef SpatialFilter<2, true> WrapFilter;
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c) {}
    __device__ __forceinline__ Filter operator()(float x, float y) const { return Filter(c_, x, y, 0.0f, true); }
    __device__ __forceinline__ WrapFilter wrap(float x, float y) const {
        const detail::SpatialDev& m = c_.meta;
        const bool oob = x > m.max[0] || y > m.max[1] || x < m.min[0] || y < m.min[1];
        c_.err.note(oob, detail::DERR_WRAP_OUT_OF_BOUNDS, 0);
        c_.err.note(!m.wrap_ok, detail::DERR_WRAP_NOT_FACTOR, 0);
        return WrapFilter(c_, oob ? m.min[0] : x, oob ? m.min[1] : y, 0.0f, !oob);
    }
    __device__ __forceinline__ float radius() const { return c_.meta.radius; }

 private:
    const detail::ThreadCtx& c_;
};

class MessageSpatial2D::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : MessageBruteForce::Out(c) {}
    __device__ __forceinline__ void setLocation(float x, float y) const {
        setVariable<float>("x", x);
        setVariable<float>("y", y);
    }
};

// MessageSpatial3DDevice.cuh:674-684
class MessageSpatial3D::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : MessageBruteForce::Out(c) {}
    __device__ __forceinline__ void setLocation(float x, float y, float z) const {
        setVariable<float>("x", x);
        setVariable<float>("y", y);
        setVariable<float>("z", z);
    }
};

// ------------------------------------------------------------------ bucket messages
// include/flamegpu/runtime/messaging/MessageBucket/MessageBucketDevice.cuh:190-330.
// Keys [lower, upper] both inclusive; the bucket table holds keys + 1 starts.  The
// reference's Filter drops the last bucket (its bound test is endKey < max); here every
// bucket in the declared range is readable.
class MessageBucket::In {
 public:
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c) {}
    class Filter {
     public:
        class iterator;
        class Message {
         public:
            __device__ __forceinline__ Message(const detail::ThreadCtx& c, uint32_t i) : c_(&c), i_(i) {}
            template <typename T, unsigned N>
            __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
                return detail::load_var<T>(c_->a.msg_in, c_->a.msg_in_mul, name, i_, *c_, detail::DERR_UNKNOWN_MSG_VAR,
                                           detail::DERR_MSG_VAR_TYPE);
            }
            template <typename T, size_type L, unsigned N>
            __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
                return detail::load_elem<T>(c_->a.msg_in, c_->a.msg_in_mul, name, i_, L, e, *c_, detail::DERR_UNKNOWN_MSG_VAR,
                                            detail::DERR_MSG_VAR_TYPE);
            }

         private:
            friend class iterator;
            const detail::ThreadCtx* c_;
            uint32_t i_;
        };
        class iterator {
         public:
            __device__ __forceinline__ iterator(const detail::ThreadCtx& c, uint32_t i) : m_(c, i) {}
            __device__ __forceinline__ iterator& operator++() {
                ++m_.i_;
                return *this;
            }
            __device__ __forceinline__ bool operator!=(const iterator& o) const { return m_.i_ != o.m_.i_; }
            __device__ __forceinline__ bool operator==(const iterator& o) const { return m_.i_ == o.m_.i_; }
            __device__ __forceinline__ Message& operator*() { return m_; }
            __device__ __forceinline__ Message* operator->() { return &m_; }

         private:
            Message m_;
        };
        __device__ __forceinline__ Filter(const detail::ThreadCtx& c, uint32_t b, uint32_t e) : c_(c), b_(b), e_(e) {}
        __device__ __forceinline__ iterator begin() const { return iterator(c_, b_); }
        __device__ __forceinline__ iterator end() const { return iterator(c_, e_); }
        __device__ __forceinline__ size_type size() const { return e_ - b_; }

     private:
        const detail::ThreadCtx& c_;
        uint32_t b_, e_;
    };
    // message_in(key): the bucket of `key`
    __device__ __forceinline__ Filter operator()(int32_t key) const {
        const int64_t k = (int64_t)key - c_.meta.key_min;
        const bool bad = k < 0 || k >= (int64_t)c_.meta.keys;
        c_.err.note(bad, detail::DERR_BUCKET_KEY, (uint32_t)key);
        if (bad) return Filter(c_, 0, 0);
        return Filter(c_, c_.meta.pbm[k], c_.meta.pbm[k + 1]);
    }
    // message_in(begin, end): buckets [begin, end)
    __device__ __forceinline__ Filter operator()(int32_t begin_key, int32_t end_key) const {
        const int64_t b = (int64_t)begin_key - c_.meta.key_min, e = (int64_t)end_key - c_.meta.key_min;
        const bool bad = b < 0 || e > (int64_t)c_.meta.keys || e <= b;
        c_.err.note(bad, detail::DERR_BUCKET_KEY, (uint32_t)(bad && b >= 0 ? end_key : begin_key));
        if (bad) return Filter(c_, 0, 0);
        return Filter(c_, c_.meta.pbm[b], c_.meta.pbm[e]);
    }

 private:
    const detail::ThreadCtx& c_;
};

class MessageBucket::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : MessageBruteForce::Out(c) {}
    // MessageBucketDevice.cuh:303-318
    __device__ __forceinline__ void setKey(int32_t key) const {
        const int64_t k = (int64_t)key - c_.a.out_key_min;
        const bool bad = k < 0 || k >= (int64_t)c_.a.out_keys;
        c_.err.note(bad, detail::DERR_BUCKET_KEY, (uint32_t)key);
        if (bad) return;
        setVariable<int32_t>("_key", key);
    }
};

// ------------------------------------------------------------------ array messages
// include/flamegpu/runtime/messaging/MessageArray{,2D,3D}/MessageArray*Device.cuh.
namespace detail {
// A message of an array list at element i (i >= length: reads give zero).
class ArrayMessage {
 public:
    __device__ __forceinline__ ArrayMessage(const ThreadCtx& c, uint32_t i) : c_(c), i_(i) {}
    template <typename T, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
        const T v = load_var<T>(c_.a.msg_in, c_.a.msg_in_mul, name, i_ < len() ? i_ : 0u, c_, DERR_UNKNOWN_MSG_VAR,
                                DERR_MSG_VAR_TYPE);
        return i_ < len() ? v : T(0);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
        const T v = load_elem<T>(c_.a.msg_in, c_.a.msg_in_mul, name, i_ < len() ? i_ : 0u, L, e, c_, DERR_UNKNOWN_MSG_VAR,
                                 DERR_MSG_VAR_TYPE);
        return i_ < len() ? v : T(0);
    }
    __device__ __forceinline__ size_type getIndex() const { return i_; }
    __device__ __forceinline__ size_type getX() const { return i_ % c_.meta.grid[0]; }
    __device__ __forceinline__ size_type getY() const { return (i_ / c_.meta.grid[0]) % c_.meta.grid[1]; }
    __device__ __forceinline__ size_type getZ() const { return i_ / (c_.meta.grid[0] * c_.meta.grid[1]); }

 private:
    __device__ __forceinline__ uint32_t len() const { return c_.meta.grid[0] * c_.meta.grid[1] * c_.meta.grid[2]; }
    const ThreadCtx& c_;
    uint32_t i_;
};

// The Moore (or von Neumann) neighbourhood of an array element, in the reference's
// order: x offset outermost, then y, then z innermost, the origin skipped
// (MessageArray2DDevice.cuh:1380-1448, MessageArray3DDevice.cuh:1536-1571).  Clamped
// filters cover the offsets that stay inside the array; wrapping ones cover [-r, r]
// on every axis with the element taken modulo the dimension.
template <int D, bool WRAP, bool VN>
class ArrayFilter {
 public:
    class Message {
     public:
        __device__ __forceinline__ Message(const ArrayFilter& f, int x, int y, int z) : f_(f) {
            rel_[0] = x;
            rel_[1] = y;
            rel_[2] = z;
        }
        __device__ __forceinline__ bool operator==(const Message& o) const {
            return rel_[0] == o.rel_[0] && rel_[1] == o.rel_[1] && rel_[2] == o.rel_[2];
        }
        __device__ __forceinline__ bool operator!=(const Message& o) const { return !(*this == o); }
        __device__ __forceinline__ size_type getX() const { return coord(0); }
        __device__ __forceinline__ size_type getY() const { return coord(1); }
        __device__ __forceinline__ size_type getZ() const { return coord(2); }
        __device__ __forceinline__ int getOffsetX() const { return rel_[0]; }
        __device__ __forceinline__ int getOffsetY() const { return rel_[1]; }
        __device__ __forceinline__ int getOffsetZ() const { return rel_[2]; }
        template <typename T, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
            return load_var<T>(f_.c_.a.msg_in, f_.c_.a.msg_in_mul, name, index(), f_.c_, DERR_UNKNOWN_MSG_VAR,
                               DERR_MSG_VAR_TYPE);
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
            return load_elem<T>(f_.c_.a.msg_in, f_.c_.a.msg_in_mul, name, index(), L, e, f_.c_, DERR_UNKNOWN_MSG_VAR,
                                DERR_MSG_VAR_TYPE);
        }
        __device__ __forceinline__ Message& operator++() {
            do {
                if (D == 1) {
                    ++rel_[0];
                } else if (D == 2) {
                    if (rel_[1] >= f_.hi_[1]) {
                        rel_[1] = f_.lo_[1];
                        ++rel_[0];
                    } else {
                        ++rel_[1];
                    }
                } else {
                    if (rel_[2] >= f_.hi_[2]) {
                        rel_[2] = f_.lo_[2];
                        if (rel_[1] >= f_.hi_[1]) {
                            rel_[1] = f_.lo_[1];
                            ++rel_[0];
                        } else {
                            ++rel_[1];
                        }
                    } else {
                        ++rel_[2];
                    }
                }
            } while (rel_[0] <= f_.hi_[0] && skip());
            return *this;
        }
        __device__ __forceinline__ bool skip() const {
            const bool origin = rel_[0] == 0 && rel_[1] == 0 && rel_[2] == 0;
            const int l1 = abs(rel_[0]) + abs(rel_[1]) + abs(rel_[2]);
            return origin || (VN && l1 > f_.r_);
        }

     private:
        __device__ __forceinline__ uint32_t coord(int a) const {
            int v = (int)f_.loc_[a] + rel_[a];
            if (WRAP) {
                const int g = (int)f_.c_.meta.grid[a];
                v = v < 0 ? v + g : (v >= g ? v - g : v);
            }
            return (uint32_t)v;
        }
        __device__ __forceinline__ uint32_t index() const {
            uint32_t i = coord(0);
            if (D >= 2) i += coord(1) * f_.c_.meta.grid[0];
            if (D == 3) i += coord(2) * f_.c_.meta.grid[0] * f_.c_.meta.grid[1];
            return i;
        }
        const ArrayFilter& f_;
        int rel_[3];
    };
    class iterator {
     public:
        __device__ __forceinline__ explicit iterator(const Message& m) : m_(m) {}
        __device__ __forceinline__ iterator& operator++() {
            ++m_;
            return *this;
        }
        __device__ __forceinline__ bool operator==(const iterator& o) const { return m_ == o.m_; }
        __device__ __forceinline__ bool operator!=(const iterator& o) const { return m_ != o.m_; }
        __device__ __forceinline__ Message& operator*() { return m_; }
        __device__ __forceinline__ Message* operator->() { return &m_; }

     private:
        Message m_;
    };
    // valid == false: an empty filter (the call's arguments were reported as an error)
    __device__ __forceinline__ ArrayFilter(const ThreadCtx& c, const uint32_t* loc, uint32_t r, bool valid) : c_(c), r_((int)r) {
        for (int a = 0; a < 3; ++a) {
            loc_[a] = a < D ? loc[a] : 0u;
            if (a >= D) {
                lo_[a] = hi_[a] = 0;
            } else if (WRAP) {
                lo_[a] = -(int)r;
                hi_[a] = (int)r;
            } else {
                const int x = (int)loc_[a], g = (int)c.meta.grid[a];
                lo_[a] = x - (int)r < 0 ? -x : -(int)r;
                hi_[a] = x + (int)r >= g ? g - 1 - x : (int)r;
            }
        }
        if (!valid) {
            lo_[0] = 1;
            hi_[0] = 0;
        }
    }
    __device__ __forceinline__ iterator begin() const {
        Message m(*this, lo_[0], lo_[1], lo_[2]);
        if (lo_[0] <= hi_[0] && m.skip()) ++m;
        return iterator(m);
    }
    __device__ __forceinline__ iterator end() const { return iterator(Message(*this, hi_[0] + 1, lo_[1], lo_[2])); }

 private:
    const ThreadCtx& c_;
    uint32_t loc_[3];
    int lo_[3], hi_[3];
    int r_;
};

// Argument checks of the iteration entry points (MessageArray2DDevice.cuh:977-1092):
// radius 0, an origin outside the array, or (wrapping) a diameter above a dimension.
template <int D>
__device__ __forceinline__ bool array_args_ok(const ThreadCtx& c, const uint32_t* loc, uint32_t r, bool wrap) {
    bool in = true, fits = true;
    for (int a = 0; a < D; ++a) {
        in = in && loc[a] < c.meta.grid[a];
        fits = fits && 2u * r + 1u <= c.meta.grid[a];
    }
    c.err.note(!in, DERR_ARRAY_MSG_INDEX, loc[0]);
    c.err.note(r == 0u || (wrap && !fits), DERR_ARRAY_RADIUS, r);
    return in && r != 0u && (!wrap || fits);
}
}  // namespace detail

class MessageArray::In {
 public:
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c) {}
    typedef detail::ArrayMessage Message;
    typedef detail::ArrayFilter<1, false, false> Filter;
    typedef detail::ArrayFilter<1, true, false> WrapFilter;
    __device__ __forceinline__ size_type size() const { return c_.meta.grid[0]; }
    __device__ __forceinline__ Message at(size_type i) const {
        c_.err.note(i >= c_.meta.grid[0], detail::DERR_ARRAY_MSG_INDEX, i);
        return Message(c_, i >= c_.meta.grid[0] ? 0xFFFFFFFFu : i);
    }
    __device__ __forceinline__ Filter operator()(size_type x, size_type radius = 1) const {
        const uint32_t loc[1] = {x};
        const bool ok = detail::array_args_ok<1>(c_, loc, radius, false);
        return Filter(c_, loc, radius, ok);
    }
    __device__ __forceinline__ WrapFilter wrap(size_type x, size_type radius = 1) const {
        const uint32_t loc[1] = {x};
        const bool ok = detail::array_args_ok<1>(c_, loc, radius, true);
        return WrapFilter(c_, loc, radius, ok);
    }

 private:
    const detail::ThreadCtx& c_;
};

class MessageArray2D::In {
 public:
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c) {}
    typedef detail::ArrayMessage Message;
    typedef detail::ArrayFilter<2, false, false> Filter;
    typedef detail::ArrayFilter<2, true, false> WrapFilter;
    typedef detail::ArrayFilter<2, false, true> VonNeumannFilter;
    typedef detail::ArrayFilter<2, true, true> VonNeumannWrapFilter;
    __device__ __forceinline__ size_type getDimX() const { return c_.meta.grid[0]; }
    __device__ __forceinline__ size_type getDimY() const { return c_.meta.grid[1]; }
    __device__ __forceinline__ size_type size() const { return c_.meta.grid[0] * c_.meta.grid[1]; }
    __device__ __forceinline__ Message at(size_type x, size_type y) const {
        const bool bad = x >= c_.meta.grid[0] || y >= c_.meta.grid[1];
        c_.err.note(bad, detail::DERR_ARRAY_MSG_INDEX, x);
        return Message(c_, bad ? 0xFFFFFFFFu : y * c_.meta.grid[0] + x);
    }
    __device__ __forceinline__ Filter operator()(size_type x, size_type y, size_type radius = 1) const { return make<Filter>(x, y, radius, false); }
    __device__ __forceinline__ WrapFilter wrap(size_type x, size_type y, size_type radius = 1) const { return make<WrapFilter>(x, y, radius, true); }
    __device__ __forceinline__ VonNeumannFilter vn(size_type x, size_type y, size_type radius = 1) const { return make<VonNeumannFilter>(x, y, radius, false); }
    __device__ __forceinline__ VonNeumannWrapFilter vn_wrap(size_type x, size_type y, size_type radius = 1) const {
        return make<VonNeumannWrapFilter>(x, y, radius, true);
    }

 private:
    template <typename F>
    __device__ __forceinline__ F make(size_type x, size_type y, size_type r, bool wrap) const {
        const uint32_t loc[2] = {x, y};
        const bool ok = detail::array_args_ok<2>(c_, loc, r, wrap);
        return F(c_, loc, r, ok);
    }
    const detail::ThreadCtx& c_;
};

class MessageArray3D::In {
 public:
    __device__ __forceinline__ explicit In(const detail::ThreadCtx& c) : c_(c) {}
    typedef detail::ArrayMessage Message;
    typedef detail::ArrayFilter<3, false, false> Filter;
    typedef detail::ArrayFilter<3, true, false> WrapFilter;
    typedef detail::ArrayFilter<3, false, true> VonNeumannFilter;
    typedef detail::ArrayFilter<3, true, true> VonNeumannWrapFilter;
    __device__ __forceinline__ size_type getDimX() const { return c_.meta.grid[0]; }
    __device__ __forceinline__ size_type getDimY() const { return c_.meta.grid[1]; }
    __device__ __forceinline__ size_type getDimZ() const { return c_.meta.grid[2]; }
    __device__ __forceinline__ size_type size() const { return c_.meta.grid[0] * c_.meta.grid[1] * c_.meta.grid[2]; }
    __device__ __forceinline__ Message at(size_type x, size_type y, size_type z) const {
        const bool bad = x >= c_.meta.grid[0] || y >= c_.meta.grid[1] || z >= c_.meta.grid[2];
        c_.err.note(bad, detail::DERR_ARRAY_MSG_INDEX, x);
        return Message(c_, bad ? 0xFFFFFFFFu : (z * c_.meta.grid[1] + y) * c_.meta.grid[0] + x);
    }
    __device__ __forceinline__ Filter operator()(size_type x, size_type y, size_type z, size_type radius = 1) const {
        return make<Filter>(x, y, z, radius, false);
    }
    __device__ __forceinline__ WrapFilter wrap(size_type x, size_type y, size_type z, size_type radius = 1) const {
        return make<WrapFilter>(x, y, z, radius, true);
    }
    __device__ __forceinline__ VonNeumannFilter vn(size_type x, size_type y, size_type z, size_type radius = 1) const {
        return make<VonNeumannFilter>(x, y, z, radius, false);
    }
    __device__ __forceinline__ VonNeumannWrapFilter vn_wrap(size_type x, size_type y, size_type z, size_type radius = 1) const {
        return make<VonNeumannWrapFilter>(x, y, z, radius, true);
    }

 private:
    template <typename F>
    __device__ __forceinline__ F make(size_type x, size_type y, size_type z, size_type r, bool wrap) const {
        const uint32_t loc[3] = {x, y, z};
        const bool ok = detail::array_args_ok<3>(c_, loc, r, wrap);
        return F(c_, loc, r, ok);
    }
    const detail::ThreadCtx& c_;
};

// MessageArray*Device.cuh setIndex: the element this agent's message goes to; an index
// outside the array is reported and the message is dropped.
class MessageArray::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : MessageBruteForce::Out(c) {}
    __device__ __forceinline__ void setIndex(size_type i) const {
        const bool bad = i >= c_.a.out_dims[0];
        c_.err.note(bad, detail::DERR_ARRAY_MSG_INDEX, i);
        setVariable<uint32_t>("___INDEX", bad ? 0xFFFFFFFFu : i);
    }
};
class MessageArray2D::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : MessageBruteForce::Out(c) {}
    __device__ __forceinline__ void setIndex(size_type x, size_type y) const {
        const bool bad = x >= c_.a.out_dims[0] || y >= c_.a.out_dims[1];
        c_.err.note(bad, detail::DERR_ARRAY_MSG_INDEX, x);
        setVariable<uint32_t>("___INDEX", bad ? 0xFFFFFFFFu : y * c_.a.out_dims[0] + x);
    }
};
class MessageArray3D::Out : public MessageBruteForce::Out {
 public:
    __device__ __forceinline__ explicit Out(const detail::ThreadCtx& c) : MessageBruteForce::Out(c) {}
    __device__ __forceinline__ void setIndex(size_type x, size_type y, size_type z) const {
        const bool bad = x >= c_.a.out_dims[0] || y >= c_.a.out_dims[1] || z >= c_.a.out_dims[2];
        c_.err.note(bad, detail::DERR_ARRAY_MSG_INDEX, x);
        setVariable<uint32_t>("___INDEX", bad ? 0xFFFFFFFFu : (z * c_.a.out_dims[1] + y) * c_.a.out_dims[0] + x);
    }
};

// ------------------------------------------------------------------ DeviceAPI
template <typename MsgIn, typename MsgOut>
class DeviceAPI {
 public:
    // include/flamegpu/runtime/DeviceAPI.cuh:187-260 (AgentOut)
    class AgentOut {
     public:
        __device__ __forceinline__ explicit AgentOut(const detail::ThreadCtx& c) : c_(c) {}
        // (name[0] is a constant once the literal is inlined: the reserved-name test costs nothing)
        template <typename T, unsigned N>
        __device__ __forceinline__ void setVariable(const char (&name)[N], T value) const {
            if (name[0] == '_') {
                c_.err.note(true, detail::DERR_RESERVED_NAME, 0);
                return;
            }
            const bool on = c_.a.birth_flags != nullptr;
            c_.err.note(!on, detail::DERR_AGENT_OUTPUT_DISABLED, 0);
            detail::store_var<T>(c_.a.new_agent, c_.a.new_mul, name, c_.lidx, value, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                                 detail::DERR_AGENT_VAR_TYPE);
            if (on) c_.a.birth_flags[c_.lidx] = 1u;
        }
        template <typename T, size_type L, unsigned N>
        __device__ __forceinline__ void setVariable(const char (&name)[N], unsigned int e, T value) const {
            if (name[0] == '_') {
                c_.err.note(true, detail::DERR_RESERVED_NAME, 0);
                return;
            }
            const bool on = c_.a.birth_flags != nullptr;
            c_.err.note(!on, detail::DERR_AGENT_OUTPUT_DISABLED, 0);
            detail::store_elem<T>(c_.a.new_agent, c_.a.new_mul, name, c_.lidx, L, e, value, c_,
                                  detail::DERR_UNKNOWN_AGENT_VAR, detail::DERR_AGENT_VAR_TYPE);
            if (on) c_.a.birth_flags[c_.lidx] = 1u;
        }
        // Marks the birth (as setVariable does) and returns the child's ID.  Children of a
        // launch normally get nextID + their parent's rank among the parents that gave
        // birth, known only after the launch; once a parent reads the ID here, the launch
        // uses nextID + the parent's index among the launch's agents for every child
        // instead (the append after the launch reserves one ID per parent), so the value
        // returned here is the child's ID whatever the scheduling.
        __device__ __forceinline__ id_t getID() const {
            const bool on = c_.a.birth_flags != nullptr;
            c_.err.note(!on, detail::DERR_AGENT_OUTPUT_DISABLED, 0);
            if (!on) return ID_NOT_SET;
            // batched launches number each simulation's children after the launch
            c_.err.note(c_.a.d_id_slots == nullptr, detail::DERR_AGENT_OUT_ID_BATCH, 0);
            if (!c_.a.d_id_slots) return ID_NOT_SET;
            c_.a.birth_flags[c_.lidx] = 1u;
            *c_.a.d_id_slots = 1u;
            return *c_.a.d_new_next_id + c_.lidx;
        }

     private:
        const detail::ThreadCtx& c_;
    };

    __device__ __forceinline__ explicit DeviceAPI(const detail::ThreadCtx& c)
        : message_in(c), message_out(c), agent_out(c), random(c), environment(c), c_(c) {}

    template <typename T, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
        return detail::load_var<T>(c_.a.agent, c_.a.agent_mul, name, c_.idx, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                                   detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
        return detail::load_elem<T>(c_.a.agent, c_.a.agent_mul, name, c_.idx, L, e, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                                    detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], T value) const {
        if (name[0] == '_') {
            c_.err.note(true, detail::DERR_RESERVED_NAME, 0);
            return;
        }
        detail::store_var<T>(c_.a.agent, c_.a.agent_mul, name, c_.idx, value, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                             detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ void setVariable(const char (&name)[N], unsigned int e, T value) const {
        if (name[0] == '_') {
            c_.err.note(true, detail::DERR_RESERVED_NAME, 0);
            return;
        }
        detail::store_elem<T>(c_.a.agent, c_.a.agent_mul, name, c_.idx, L, e, value, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                              detail::DERR_AGENT_VAR_TYPE);
    }
    __device__ __forceinline__ id_t getID() const { return c_.a.ids ? c_.a.ids[c_.idx] : ID_NOT_SET; }
    __device__ __forceinline__ unsigned int getStepCounter() const { return c_.d_step ? *c_.d_step : 0u; }
    __device__ __forceinline__ unsigned int getThreadIndex() const { return c_.idx; }

    typename MsgIn::In message_in;
    typename MsgOut::Out message_out;
    AgentOut agent_out;
    AgentRandom random;
    DeviceEnvironment environment;

 private:
    const detail::ThreadCtx& c_;
};

// ------------------------------------------------------------------ ReadOnlyDeviceAPI
// include/flamegpu/runtime/DeviceAPI.cuh:40-160: the API of agent function conditions
// (variables read-only, no messages, no agent output).  Conditions draw from their own
// RNG stream (the function's stream with the top bit set).
class ReadOnlyDeviceAPI {
 public:
    __device__ __forceinline__ explicit ReadOnlyDeviceAPI(const detail::ThreadCtx& c) : random(c), environment(c), c_(c) {}
    template <typename T, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N]) const {
        return detail::load_var<T>(c_.a.agent, c_.a.agent_mul, name, c_.idx, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                                   detail::DERR_AGENT_VAR_TYPE);
    }
    template <typename T, size_type L, unsigned N>
    __device__ __forceinline__ T getVariable(const char (&name)[N], unsigned int e) const {
        return detail::load_elem<T>(c_.a.agent, c_.a.agent_mul, name, c_.idx, L, e, c_, detail::DERR_UNKNOWN_AGENT_VAR,
                                    detail::DERR_AGENT_VAR_TYPE);
    }
    __device__ __forceinline__ id_t getID() const { return c_.a.ids ? c_.a.ids[c_.idx] : ID_NOT_SET; }
    __device__ __forceinline__ unsigned int getStepCounter() const { return c_.d_step ? *c_.d_step : 0u; }
    __device__ __forceinline__ unsigned int getThreadIndex() const { return c_.idx; }

    AgentRandom random;
    DeviceEnvironment environment;

 private:
    const detail::ThreadCtx& c_;
};

// ------------------------------------------------------------------ the agent-function kernels
// include/flamegpu/runtime/AgentFunction.cuh:60-120: one thread per agent of the
// state list; the count is read from device memory so launches can be replayed
// from a captured graph while populations change.
namespace detail {

template <typename F>
__device__ __forceinline__ void run_agent(const ThreadCtx& c) {
    const FnArgs& a = c.a;
    const uint32_t o = c.lidx;
    if (a.msg_out_flags) a.msg_out_flags[o] = a.msg_out_optional ? 0u : 1u;
    if (a.birth_flags) a.birth_flags[o] = 0u;
    if (c.wtab) *static_cast<volatile __attribute__((address_space(3))) uint32_t*>(c.wtab) = 0u;   // free for this call's first spatial filter
    DeviceAPI<typename F::MsgIn, typename F::MsgOut> api(c);
    const AGENT_STATUS s = F()(&api);
    if (a.death_flags)
        a.death_flags[o] = (uint32_t)s;
    else
        c.err.note(s == DEAD, DERR_DEATH_DISABLED, 0);
}

// Workgroups go to the 8 XCDs round-robin (workgroup w runs on XCD w % 8), and each
// XCD has its own L2.  This renumbering gives XCD x one contiguous range of logical
// blocks, started in order, so the bin-sorted agents that share neighbourhoods (and
// so messages) also share an L2: XCD x gets q + (x < r) blocks, q = G / 8, r = G % 8,
// starting at x*q + min(x, r).
__device__ __forceinline__ uint32_t xcd_block() {
    constexpr uint32_t kXcds = 8;
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t x = b % kXcds, k = b / kXcds, q = g / kXcds, r = g % kXcds;
    return x * q + (x < r ? x : r) + k;
}

// Agents [first, n) of the list run, first = the agents a function condition disabled.
template <typename F>
constexpr bool reads_spatial() {
    return F::MsgIn::kind == MsgKind::Spatial2D || F::MsgIn::kind == MsgKind::Spatial3D;
}

// Dynamic LDS of the non-staged kernels: one unit-table region per wave for spatial input.
template <typename F>
constexpr uint32_t fn_lds_bytes(uint32_t block) {
    return reads_spatial<F>() ? (block / 64u) * kWaveTabWords * 4u : 0u;
}

// This wave's unit-table region, or null (no spatial input).
template <typename F>
__device__ __forceinline__ __attribute__((address_space(3))) uint32_t* wave_table() {
    if constexpr (reads_spatial<F>()) {
        extern __shared__ uint32_t fgx_stage_smem[];
        return (__attribute__((address_space(3))) uint32_t*)fgx_stage_smem +
               __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * kWaveTabWords;
    }
    return nullptr;
}

// One block of an agent-function launch: block `blk` of the function runs agents
// [first + blk * blockDim, ...) of its list.  Shared by the function's own kernel (its
// launch record in the kernel-argument segment) and the layer kernel (its record in
// device memory).
template <typename F>
__device__ __forceinline__ void run_function_block(const FnArgs& a, uint32_t blk) {
    const uint32_t first = a.d_first ? load_count(a.d_first) : 0u;
    const uint32_t n = a.d_n ? load_count(a.d_n) : a.n_cap;
    const uint32_t o = blk * blockDim.x + threadIdx.x;
    if (a.d_msg_count_out && o == 0) *a.d_msg_count_out = n;   // one message per agent (no condition here)
    if (first + o >= n) return;
    const uint32_t i = first + o;
    uint32_t mn = 0u;
    if constexpr (F::MsgIn::kind == MsgKind::BruteForce) mn = a.d_msg_in_count ? load_count(a.d_msg_in_count) : 0u;
    const ThreadCtx c{a,  ErrState{a.d_error, i}, a.zero_page, a.sink_page, i, a.in_meta, a.env, a.d_step, a.seed, nullptr,
                      0u, mn,                     o,           a.stream_id, a.unit_tables ? wave_table<F>() : nullptr};
    run_agent<F>(c);
}

template <typename F>
__global__ void __launch_bounds__(256) agent_function_kernel(const FnArgs a) {
    run_function_block<F>(a, xcd_block());
}

// The same kernel built for 8 resident waves per SIMD (at most 64 VGPRs), for neighbour
// walks: their loops wait on gathers, and a kernel a few registers above 64 keeps only 7.
template <typename F>
__global__ void __launch_bounds__(256, 8) agent_function_kernel_w8(const FnArgs a) {
    run_function_block<F>(a, xcd_block());
}

// Whether a function's launches take agent_function_kernel_w8: when the plain build needs
// more than 64 and at most 72 VGPRs (so 7 waves per SIMD) and the 8-wave build spills at
// most 64 bytes per lane.  Read once from the compiled kernels' attributes.
template <typename F>
bool use_w8_kernel() {
    static const bool use = [] {
        hipFuncAttributes plain{}, w8{};
        if (hipFuncGetAttributes(&plain, reinterpret_cast<const void*>(&agent_function_kernel<F>)) != hipSuccess ||
            hipFuncGetAttributes(&w8, reinterpret_cast<const void*>(&agent_function_kernel_w8<F>)) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return plain.numRegs > 64 && plain.numRegs <= 72 && w8.localSizeBytes <= 64;
    }();
    return use;
}

// ---------------------------------------------------------------- layer kernel
// Compile-time registry of a translation unit's agent functions.  FLAMEGPU_AGENT_FUNCTION
// takes a __COUNTER__ value N and instantiates fn_put<N, F>, whose friend definitions give
// fn_key<N> a definition; fn_collect<0, End> then lists, in definition order, every F
// registered below End.  The layer kernel is instantiated over that list where the
// model library closes its translation unit (flamegpu/model_abi.h).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wundefined-inline"
#pragma clang diagnostic ignored "-Wundefined-internal"
template <int N>
struct fn_key {
    friend constexpr int fn_defined(fn_key<N>);
    friend auto fn_type(fn_key<N>);
};
template <int N, typename F>
struct fn_put {
    friend constexpr int fn_defined(fn_key<N>) { return 1; }
    friend auto fn_type(fn_key<N>) { return static_cast<F*>(nullptr); }
    static constexpr bool ok = true;
};
template <int N, int = fn_defined(fn_key<N>{})>
constexpr bool fn_registered(int) { return true; }
template <int N>
constexpr bool fn_registered(...) { return false; }
#pragma clang diagnostic pop

template <typename... Fs>
struct fn_list {};
template <typename A, typename B>
struct fn_concat;
template <typename... A, typename... B>
struct fn_concat<fn_list<A...>, fn_list<B...>> {
    using type = fn_list<A..., B...>;
};
template <int N, bool On = fn_registered<N>(0)>
struct fn_at {
    using type = fn_list<>;
};
template <int N>
struct fn_at<N, true> {
    using type = fn_list<std::remove_pointer_t<decltype(fn_type(fn_key<N>{}))>>;
};
// the slots [Lo, Hi), split in halves so that instantiation depth stays logarithmic
template <int Lo, int Hi, bool One = (Hi - Lo <= 1)>
struct fn_collect {
    using type = typename fn_concat<typename fn_collect<Lo, (Lo + Hi) / 2>::type,
                                    typename fn_collect<(Lo + Hi) / 2, Hi>::type>::type;
};
template <int Lo, int Hi>
struct fn_collect<Lo, Hi, true> {
    using type = std::conditional_t<(Hi > Lo), typename fn_at<Lo>::type, fn_list<>>;
};

template <typename... Fs, size_t... I>
__device__ __forceinline__ void run_layer_fn(uint32_t fn, const FnArgs& a, uint32_t blk, std::index_sequence<I...>) {
    ((fn == I ? run_function_block<Fs>(a, blk) : (void)0), ...);
}

// One launch for the functions of a layer (fgx_layer_launch builds the image; function k
// has begin[k + 1] - begin[k] blocks).  Each function's blocks are split evenly over the 8
// XCDs, in contiguous chunks as xcd_block() gives a launch of its own (neighbouring agents
// share an L2): workgroup w runs on XCD x = w % 8 as that XCD's block j = w / 8, and XCD x
// runs function 0's chunk, then function 1's, and so on.  A whole-grid contiguous split
// would put a small function's (mostly idle) blocks on some XCDs and the busy ones on others.
template <typename... Fs>
__global__ void __launch_bounds__(256) layer_kernel(const char* __restrict__ image) {
    constexpr uint32_t kXcds = 8;
    const fgx_layer_header& h = *reinterpret_cast<const fgx_layer_header*>(image);
    const uint32_t x = blockIdx.x % kXcds;
    uint32_t j = blockIdx.x / kXcds;
    const uint32_t nf = h.nfns;
    uint32_t k = FGX_LAYER_MAX_FNS, local = 0;
#pragma unroll
    for (uint32_t f = 0; f < FGX_LAYER_MAX_FNS; ++f) {
        if (f >= nf || k != FGX_LAYER_MAX_FNS) continue;
        const uint32_t nb = h.begin[f + 1] - h.begin[f];
        const uint32_t lo = (uint32_t)(((uint64_t)x * nb) / kXcds), hi = (uint32_t)(((uint64_t)(x + 1) * nb) / kXcds);
        if (j < hi - lo) {
            k = f;
            local = lo + j;
        } else {
            j -= hi - lo;
        }
    }
    if (k == FGX_LAYER_MAX_FNS) return;   // this XCD's share is done
    const FnArgs& a = *reinterpret_cast<const FnArgs*>(image + FGX_LAYER_ARGS_OFFSET + (size_t)k * h.stride);
    run_layer_fn<Fs...>(h.fn[k], a, local, std::index_sequence_for<Fs...>{});
}

// Batched launch: the simulations of a batch share one launch.  Tile t (blockDim x
// apt agents) belongs to the simulation s with segs[s].tile_off <= t <
// segs[s+1].tile_off and runs with that simulation's metadata, environment, step,
// seed and error words; each thread runs apt agents, blockDim apart.  A grid
// smaller than the tile count makes the blocks persistent: logical block b (in
// XCD-contiguous order) runs tiles [b*T/G, (b+1)*T/G).
template <typename F>
__global__ void __launch_bounds__(1024) agent_function_kernel_seg(const FnArgs a, const SegDev* __restrict__ segs,
                                                                  uint32_t nsegs, uint32_t apt) {
    // segs[nsegs] is the sentinel record: its tile_off is the batch's tile count, kept on
    // the device when births and deaths change the populations
    const uint32_t ntiles = segs[nsegs].tile_off;
    const uint32_t lb = xcd_block();
    const uint32_t t0 = (uint32_t)((uint64_t)lb * ntiles / gridDim.x);
    const uint32_t t1 = (uint32_t)((uint64_t)(lb + 1) * ntiles / gridDim.x);
    for (uint32_t t = t0; t < t1; ++t) {
        uint32_t lo = 0, hi = nsegs;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (segs[mid].tile_off <= t) lo = mid;
            else hi = mid;
        }
        const SegDev& s = segs[lo];
        const uint32_t base = (t - s.tile_off) * blockDim.x * apt + threadIdx.x;
        for (uint32_t k = 0; k < apt; ++k) {
            const uint32_t local = base + k * blockDim.x;
            if (local < s.n) {
                const uint32_t i = s.agent_off + local;
                const ThreadCtx c{a,        ErrState{s.d_error, i}, a.zero_page, a.sink_page, i, s.meta, s.env, s.d_step,
                                  s.seed,   nullptr,                s.msg_off,   s.msg_n,     i,     a.stream_id,
                                  a.unit_tables ? wave_table<F>() : nullptr};
                run_agent<F>(c);
            }
        }
    }
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

// Spatial-input agent functions with the neighbourhood staged in LDS, then the user
// function (waves whose lanes are all covered read LDS, the others HBM).
template <typename F>
__global__ void __launch_bounds__(256) agent_function_kernel_staged(const FnArgs a) {
    extern __shared__ uint32_t fgx_stage_smem[];
    __shared__ StageHeader sh;
    const uint32_t n = a.d_n ? load_count(a.d_n) : a.n_cap;
    const uint32_t b0 = xcd_block() * blockDim.x;
    if (a.d_msg_count_out && b0 == 0 && threadIdx.x == 0) *a.d_msg_count_out = n;
    if (b0 >= n) return;
    const uint32_t i = b0 + threadIdx.x;
    const bool active = i < n;
    const bool staged = stage_block<F::MsgIn::In::dims>(a, fgx_stage_smem, sh, i, active);
    if (!active) return;
    const uint32_t maxe = a.stage_max_entries;
    const StageView sv{fgx_stage_smem, fgx_stage_smem + maxe + 1, fgx_stage_smem + 2 * (maxe + 1), a.stage_cap,
                       a.stage_nvars > (uint32_t)F::MsgIn::In::dims ? 1u : 0u, &sh};
    const ThreadCtx c{a,         ErrState{a.d_error, i}, a.zero_page, a.sink_page, i, a.in_meta, a.env, a.d_step,
                      a.seed,    staged ? &sv : nullptr, 0u, 0u, i, a.stream_id, nullptr};
    run_agent<F>(c);
}

// include/flamegpu/runtime/AgentFunctionCondition.cuh:44-91: flags[i] = !condition, so the
// stable partition that follows puts the disabled agents first.
template <typename C>
__global__ void __launch_bounds__(256) agent_condition_kernel(const FnArgs a, uint32_t* __restrict__ flags) {
    const uint32_t n = a.d_n ? load_count(a.d_n) : a.n_cap;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ThreadCtx c{a, ErrState{a.d_error, i}, a.zero_page, a.sink_page, i, a.in_meta, a.env, a.d_step, a.seed, nullptr,
                      0u, 0u, i, a.stream_id | 0x80000000u, nullptr};
    ReadOnlyDeviceAPI api(c);
    flags[i] = C()(&api) ? 0u : 1u;
}

template <typename C>
void launch_condition(const FnArgs& a, uint32_t* flags, uint32_t grid, void* stream) {
    hipLaunchKernelGGL(agent_condition_kernel<C>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, flags);
}

template <typename F>
void launch_agent_function(const FnArgs& a, uint32_t grid, uint32_t block, uint32_t smem, void* stream) {
    constexpr MsgKind k = F::MsgIn::kind;
    if constexpr (k == MsgKind::Spatial2D || k == MsgKind::Spatial3D) {
        if (a.stage_ok) {
            hipLaunchKernelGGL(agent_function_kernel_staged<F>, dim3(grid), dim3(block), smem, (hipStream_t)stream, a);
            return;
        }
    }
    (void)smem;
    if constexpr (k == MsgKind::Spatial2D || k == MsgKind::Spatial3D) {
        if (block <= 256 && use_w8_kernel<F>()) {
            hipLaunchKernelGGL(agent_function_kernel_w8<F>, dim3(grid), dim3(block), fn_lds_bytes<F>(block),
                               (hipStream_t)stream, a);
            return;
        }
    }
    hipLaunchKernelGGL(agent_function_kernel<F>, dim3(grid), dim3(block), fn_lds_bytes<F>(block), (hipStream_t)stream,
                       a);
}

template <typename F>
void launch_agent_function_seg(const FnArgs& a, const SegDev* segs, uint32_t nsegs, uint32_t apt, uint32_t grid,
                               uint32_t block, void* stream) {
    hipLaunchKernelGGL(agent_function_kernel_seg<F>, dim3(grid), dim3(block), fn_lds_bytes<F>(block), (hipStream_t)stream,
                       a, segs, nsegs, apt);
}

// Host view of the layer kernel over the functions Fs (see LayerKernel).
template <typename L>
struct LayerKernelOf;
template <typename... Fs>
struct LayerKernelOf<fn_list<Fs...>> {
    static int index_of(LaunchFn launch) {
        const LaunchFn fns[] = {&launch_agent_function<Fs>..., nullptr};
        for (int k = 0; k < (int)sizeof...(Fs); ++k)
            if (fns[k] == launch) return k;
        return -1;
    }
    static uint32_t lds_bytes(int index, uint32_t block) {
        const uint32_t b[] = {fn_lds_bytes<Fs>(block)..., 0u};
        return index >= 0 && index < (int)sizeof...(Fs) ? b[index] : 0u;
    }
    static LayerKernel get() {
        LayerKernel k;
        if constexpr (sizeof...(Fs) > 0) k.kernel = reinterpret_cast<const void*>(&layer_kernel<Fs...>);
        k.index_of = &index_of;
        k.lds_bytes = &lds_bytes;
        return k;
    }
};
}  // namespace detail

}  // namespace flamegpu

// ------------------------------------------------------------------ macros
// include/flamegpu/runtime/AgentFunction_shim.cuh:32-89
#define FLAMEGPU_AGENT_FUNCTION(funcName, message_in, message_out)                                                  \
    struct funcName##_impl {                                                                                        \
        typedef message_in MsgIn;                                                                                   \
        typedef message_out MsgOut;                                                                                 \
        static constexpr int fgx_slot = __COUNTER__;                                                                \
        __device__ __forceinline__ flamegpu::AGENT_STATUS operator()(                                               \
            flamegpu::DeviceAPI<message_in, message_out>* FLAMEGPU) const;                                          \
    };                                                                                                              \
    static_assert(flamegpu::detail::fn_put<funcName##_impl::fgx_slot, funcName##_impl>::ok, "registry");            \
    [[maybe_unused]] constexpr funcName##_impl funcName{};                                                           \
    __device__ __forceinline__ flamegpu::AGENT_STATUS funcName##_impl::operator()(                                  \
        flamegpu::DeviceAPI<message_in, message_out>* FLAMEGPU) const

// include/flamegpu/runtime/AgentFunctionCondition_shim.cuh
#define FLAMEGPU_AGENT_FUNCTION_CONDITION(funcName)                                                                 \
    struct funcName##_cdn_impl {                                                                                    \
        __device__ __forceinline__ bool operator()(flamegpu::ReadOnlyDeviceAPI* FLAMEGPU) const;                   \
    };                                                                                                              \
    [[maybe_unused]] constexpr funcName##_cdn_impl funcName{};                                                       \
    __device__ __forceinline__ bool funcName##_cdn_impl::operator()(flamegpu::ReadOnlyDeviceAPI* FLAMEGPU) const

#define FLAMEGPU_DEVICE_FUNCTION __device__ __forceinline__
#define FLAMEGPU_HOST_DEVICE_FUNCTION __host__ __device__ __forceinline__
