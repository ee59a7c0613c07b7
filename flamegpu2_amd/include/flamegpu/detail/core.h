// Core definitions shared by the host runtime and the generated agent kernels.
//
// Replaces include/flamegpu/defines.h, include/flamegpu/exception/FLAMEGPUException.h
// and the curve hash table (include/flamegpu/runtime/detail/curve/DeviceCurve.cuh):
// a variable name is hashed at compile time (FNV-1a over the string literal) and
// resolved against a small binding table that travels in the kernel argument
// segment, so every lookup is a handful of scalar compares that the compiler
// hoists out of message loops.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <stdexcept>
#include <string>
#include <typeindex>

#include "fgx.h"

namespace flamegpu {

// include/flamegpu/defines.h
typedef unsigned int id_t;
typedef unsigned int size_type;
constexpr id_t ID_NOT_SET = 0;

// include/flamegpu/runtime/AgentFunction.cuh:19
enum AGENT_STATUS { ALIVE = 1, DEAD = 0 };
typedef AGENT_STATUS FLAMEGPU_AGENT_STATUS;
// include/flamegpu/runtime/HostAPI_macros.h
enum CONDITION_RESULT { CONTINUE = 0, EXIT = 1 };
typedef CONDITION_RESULT FLAMEGPU_CONDITION_RESULT;

namespace exception {
// include/flamegpu/exception/FLAMEGPUException.h:60-300 (the subset raised on this path)
class FLAMEGPUException : public std::exception {
 public:
    explicit FLAMEGPUException(std::string msg = "") : msg_(std::move(msg)) {}
    const char* what() const noexcept override { return msg_.c_str(); }
    virtual const char* exception_type() const { return "FLAMEGPUException"; }

 protected:
    std::string msg_;
};
#define FGX_DERIVED_EXCEPTION(name)                                                    \
    class name : public FLAMEGPUException {                                            \
     public:                                                                           \
        explicit name(const std::string& m = "") : FLAMEGPUException(m) {}             \
        const char* exception_type() const override { return #name; }                 \
    };
FGX_DERIVED_EXCEPTION(CUDAError)
FGX_DERIVED_EXCEPTION(InvalidArgument)
FGX_DERIVED_EXCEPTION(InvalidAgentVar)
FGX_DERIVED_EXCEPTION(InvalidAgentName)
FGX_DERIVED_EXCEPTION(InvalidAgentState)
FGX_DERIVED_EXCEPTION(InvalidAgentFunc)
FGX_DERIVED_EXCEPTION(InvalidFuncLayerIndx)
FGX_DERIVED_EXCEPTION(InvalidLayerMember)
FGX_DERIVED_EXCEPTION(InvalidMessageName)
FGX_DERIVED_EXCEPTION(InvalidMessageType)
FGX_DERIVED_EXCEPTION(InvalidMessageVar)
FGX_DERIVED_EXCEPTION(InvalidEnvProperty)
FGX_DERIVED_EXCEPTION(InvalidEnvPropertyType)
FGX_DERIVED_EXCEPTION(InvalidDependencyGraph)
FGX_DERIVED_EXCEPTION(InvalidOperation)
FGX_DERIVED_EXCEPTION(InvalidVarType)
FGX_DERIVED_EXCEPTION(InvalidCUDAdevice)
FGX_DERIVED_EXCEPTION(DeviceError)
FGX_DERIVED_EXCEPTION(OutOfBoundsException)
FGX_DERIVED_EXCEPTION(UnsupportedVarType)
FGX_DERIVED_EXCEPTION(EnsembleError)
FGX_DERIVED_EXCEPTION(InvalidStateName)
FGX_DERIVED_EXCEPTION(ReservedName)
FGX_DERIVED_EXCEPTION(DuplicateEnvProperty)
FGX_DERIVED_EXCEPTION(AgentIDCollision)
FGX_DERIVED_EXCEPTION(ArrayMessageWriteConflict)
FGX_DERIVED_EXCEPTION(InvalidMessage)
FGX_DERIVED_EXCEPTION(InvalidInputFile)
FGX_DERIVED_EXCEPTION(UnknownInternalError)
FGX_DERIVED_EXCEPTION(InvalidHostFunc)
FGX_DERIVED_EXCEPTION(InvalidPopulationData)
FGX_DERIVED_EXCEPTION(InvalidConfig)
#undef FGX_DERIVED_EXCEPTION
}  // namespace exception

namespace detail {

inline std::string format(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return std::string(buf);
}

// Raise a C-ABI failure as flamegpu::exception::CUDAError (or the mapped type).
inline void fgx_check(int rc, const char* what) {
    if (rc == FGX_OK) return;
    const std::string m = std::string(what) + ": " + fgx_last_error();
    if (rc == FGX_ERR_ARG) throw exception::InvalidArgument(m);
    if (rc == FGX_ERR_DEVICE) throw exception::DeviceError(m);
    throw exception::CUDAError(m);
}

// FNV-1a over the characters of a name (compile-time for string literals).
__host__ __device__ constexpr uint32_t name_hash(const char* s, unsigned n) {
    uint32_t h = 2166136261u;
    for (unsigned i = 0; i < n && s[i]; ++i) {
        h ^= (uint8_t)s[i];
        h *= 16777619u;
    }
    return h;
}
inline uint32_t name_hash(const std::string& s) { return name_hash(s.c_str(), (unsigned)s.size()); }

// ------------------------------------------------------------------ kernel ABI
// Binding tables are perfect hashes built on the host: a name with hash h lives
// in slot (h * mul) >> (32 - log2(SLOTS)).  A device lookup is one multiply, one
// shift and one scalar load from the kernel argument segment, free of loops, so
// the compiler hoists it out of message loops.
constexpr unsigned MAX_AGENT_VARS = 16;
constexpr unsigned MAX_MSG_VARS = 16;
constexpr unsigned MAX_ENV_PROPS = 32;
constexpr unsigned VAR_SLOTS = 32;
constexpr unsigned VAR_SLOT_BITS = 5;
constexpr unsigned ENV_SLOTS = 64;
constexpr unsigned ENV_SLOT_BITS = 6;

struct alignas(16) VarBinding {
    uint32_t hash;
    uint16_t bytes;   // element bytes (type size x array length)
    uint16_t reserved;
    char* ptr;
};

// LDS staging of spatial message neighbourhoods (see agent_function_kernel_staged).
// Staged variables, in order: x, y (, z), then "id" when the message has a 4-byte id.
constexpr unsigned STAGE_MAX_GROUPS = 4;
constexpr unsigned STAGE_MAX_VARS = 4;
constexpr unsigned STAGE_MAX_ENTRIES = 768;   // staged virtual bins per block (3 per thread)
struct StageGroup {
    int cy, cz;        // the group's own row
    int xlo, xhi;      // x extent of its agents' cells
    int xbase;         // first staged (virtual) x
    int width;         // staged virtual bins per neighbour row
    uint32_t ebase;    // first entry of the group
};
struct StageHeader {
    uint32_t ngroups;
    uint32_t nentries;
    uint32_t ok;
    uint32_t total;
    StageGroup g[STAGE_MAX_GROUPS];
};

struct alignas(8) EnvBinding {
    uint32_t hash;
    uint16_t offset;
    uint16_t bytes;
};

// Device mirror of the spatial message metadata (MessageSpatial3D.h:38-68).
// Message-input metadata.  Spatial: min/max/radius/width/grid/wrap_ok and the bin
// table pbm.  Bucket: key_min, keys and pbm (keys + 1 bucket starts).  Array: grid =
// the array dimensions (unused ones 1).
struct SpatialDev {
    float min[3];
    float max[3];
    float radius;
    float width[3];
    uint32_t grid[3];
    int32_t wrap_ok;
    const uint32_t* pbm;
    int32_t key_min;
    uint32_t keys;
    float cell;      // bin width (the radius, or half of it for half-width bins)
    int32_t span;    // bins searched either side of the query's bin (1, or 2 for half-width bins)
    const void* packed;   // 16-byte {x, y, z, id} record per message of the built list (launch arguments only)
};

// Device error codes (the device half of exception::DeviceError, replacing DTHROW).
enum DeviceErrorCode : uint32_t {
    DERR_NONE = 0,
    DERR_UNKNOWN_AGENT_VAR = 1,
    DERR_AGENT_VAR_TYPE = 2,
    DERR_UNKNOWN_MSG_VAR = 3,
    DERR_MSG_VAR_TYPE = 4,
    DERR_WRAP_OUT_OF_BOUNDS = 5,
    DERR_WRAP_NOT_FACTOR = 6,
    DERR_DEATH_DISABLED = 7,
    DERR_AGENT_OUTPUT_DISABLED = 8,
    DERR_RESERVED_NAME = 9,
    DERR_UNKNOWN_ENV = 10,
    DERR_ENV_TYPE = 11,
    DERR_ARRAY_INDEX = 13,
    DERR_RANDOM_RANGE = 14,
    DERR_MSG_OUTPUT_DISABLED = 15,
    DERR_BUCKET_KEY = 16,
    DERR_ARRAY_MSG_INDEX = 17,
    DERR_ARRAY_CONFLICT = 18,
    DERR_ARRAY_RADIUS = 19,
    DERR_AGENT_OUT_ID_BATCH = 20,
};

inline const char* device_error_text(uint32_t code) {
    switch (code) {
        case DERR_UNKNOWN_AGENT_VAR: return "unknown agent variable";
        case DERR_AGENT_VAR_TYPE: return "agent variable accessed with the wrong type or length";
        case DERR_UNKNOWN_MSG_VAR: return "unknown message variable";
        case DERR_MSG_VAR_TYPE: return "message variable accessed with the wrong type or length";
        case DERR_BUCKET_KEY: return "bucket message key out of range (MessageBucket)";
        case DERR_ARRAY_MSG_INDEX: return "array message index out of bounds (MessageArray*::setIndex / iteration origin)";
        case DERR_ARRAY_CONFLICT: return "multiple messages output to the same array message index, or more messages than array elements";
        case DERR_ARRAY_RADIUS: return "invalid radius for array message iteration";
        case DERR_AGENT_OUT_ID_BATCH: return "AgentOut::getID() is not available in batched simulations";
        case DERR_WRAP_OUT_OF_BOUNDS: return "location exceeds environment bounds, unsupported by the wrapped iterator (MessageSpatial::In::wrap())";
        case DERR_WRAP_NOT_FACTOR: return "spatial messaging radius is not a factor of the environment dimensions, unsupported by the wrapped iterator";
        case DERR_DEATH_DISABLED: return "agent death must be enabled per agent function when defining the model";
        case DERR_AGENT_OUTPUT_DISABLED: return "agent output must be enabled per agent function when defining the model";
        case DERR_RESERVED_NAME: return "variable names starting with '_' are reserved for internal use";
        case DERR_UNKNOWN_ENV: return "unknown environment property";
        case DERR_ENV_TYPE: return "environment property accessed with the wrong type or length";
        case DERR_ARRAY_INDEX: return "array variable index out of bounds";
        case DERR_RANDOM_RANGE: return "invalid arguments passed to AgentRandom::uniform()";
        case DERR_MSG_OUTPUT_DISABLED: return "message output used by a function without message output";
        default: return "unknown device error";
    }
}

struct FnArgs {
    uint32_t n_cap;              // launch bound (state list capacity)
    const char* zero_page;       // 64 zero bytes: reads through an unresolved binding land here
    char* sink_page;             // 64 bytes: writes through an unresolved binding land here
    uint32_t stream_id;          // RNG stream of this agent function
    const uint32_t* d_n;         // device-resident agent count
    const uint32_t* d_step;      // device-resident step counter
    uint64_t seed;
    uint32_t* d_error;           // [code, detail hash, agent index]
    uint32_t* death_flags;       // non-null when agent death is enabled
    uint32_t* msg_out_flags;     // non-null with message output
    uint32_t* birth_flags;       // non-null with agent output
    const uint32_t* d_new_next_id;  // agent output: the output agent's next ID (constant during the launch)
    uint32_t* d_id_slots;        // agent output: set to 1 when a parent reads its child's ID
    const char* env;             // environment property buffer (device)
    const uint32_t* d_msg_in_count;
    const uint32_t* ids;         // the agent list's "_id" array
    const uint32_t* d_first;     // function condition: agents [*d_first, n) run (NULL: all)
    int32_t msg_out_optional;
    int32_t unit_tables;         // spatial input: build the per-wave unit tables in LDS (0: bin table from HBM)
    uint32_t* d_msg_count_out;   // mandatory output: the launch writes the message count (agent count) here
    uint32_t agent_mul, msg_in_mul, msg_out_mul, new_mul, env_mul;
    // neighbourhood staging (spatial message input)
    int32_t stage_ok;                 // agent x/y(/z) present and staged vars chosen
    uint32_t stage_nvars;             // staged 4-byte message variables (dims, + 1 with "id")
    uint32_t stage_max_entries;       // virtual bins per block
    uint32_t stage_cap;               // staged messages per block
    const float* stage_pos[3];        // agent x, y, z (the assumed search origin)
    const uint32_t* stage_src[STAGE_MAX_VARS];
    SpatialDev in_meta;
    // message output checks: array dimensions / bucket key range of the output list
    uint32_t out_dims[3];
    int32_t out_key_min;
    uint32_t out_keys;
    VarBinding agent[VAR_SLOTS];
    VarBinding msg_in[VAR_SLOTS];
    VarBinding msg_out[VAR_SLOTS];
    VarBinding new_agent[VAR_SLOTS];
    EnvBinding env_b[ENV_SLOTS];
};

// One simulation of a batch in a segmented launch.  The batch's agent and message
// lists are the concatenation of its simulations' lists; a simulation owns the
// agents [agent_off, agent_off + n) and, for its message input, the bins of
// meta.pbm (which points at the simulation's part of the batch's bin table and
// holds message indices into the concatenated message list).
struct SegDev {
    uint32_t agent_off;
    uint32_t n;
    uint32_t tile_off;           // first tile (blockDim x agents-per-thread agents) of the simulation
    uint32_t msg_off;            // brute-force input: the simulation's messages [msg_off, msg_off + msg_n)
    uint32_t msg_n;
    uint32_t pad0;
    uint64_t seed;
    const uint32_t* d_step;      // the simulation's step counter
    const char* env;             // the simulation's environment properties
    uint32_t* d_error;           // the simulation's error words [code, hash, index]
    SpatialDev meta;
};

__device__ __forceinline__ void report_error(uint32_t* d_error, uint32_t code, uint32_t detail, uint32_t index);

// Device error reporting for one agent function call.  Accessors test their
// (wave-uniform) binding checks with a scalar branch around the report, whose
// global atomics keep the compiler from turning it into selects: the checks cost
// a compare and a branch, and no per-lane state is carried through message loops.
struct ErrState {
    uint32_t* d_error;
    uint32_t index;
    __device__ __forceinline__ void note(bool bad, uint32_t c, uint32_t h) const {
        if (bad) report_error(d_error, c, h, index);
    }
};

struct Found {
    char* ptr;
    uint32_t bytes;
};

__host__ __device__ constexpr uint32_t table_slot(uint32_t h, uint32_t mul, unsigned bits) {
    return (h * mul) >> (32u - bits);
}

__device__ __forceinline__ Found find_binding(const VarBinding (&tab)[VAR_SLOTS], uint32_t mul, uint32_t h) {
    const VarBinding& b = tab[table_slot(h, mul, VAR_SLOT_BITS)];
    if (b.hash != h) return Found{nullptr, 0};
    return Found{b.ptr, (uint32_t)b.bytes};
}

__device__ __forceinline__ void report_error(uint32_t* d_error, uint32_t code, uint32_t detail, uint32_t index) {
    if (!d_error) return;
    if (atomicCAS(&d_error[0], 0u, code) == 0u) {
        d_error[1] = detail;
        d_error[2] = index;
    }
}

// Host: pick a multiplier that places every key in its own slot.
inline uint32_t perfect_hash_mul(const uint32_t* hashes, unsigned n, unsigned bits) {
    const unsigned slots = 1u << bits;
    for (uint32_t k = 0; k < 1000000u; ++k) {
        const uint32_t mul = 2654435761u + 2u * k * 40503u;
        uint64_t used[2] = {0, 0};
        bool ok = true;
        for (unsigned i = 0; i < n && ok; ++i) {
            const uint32_t s = table_slot(hashes[i], mul, bits);
            if (used[s >> 6] & (1ull << (s & 63))) ok = false;
            used[s >> 6] |= 1ull << (s & 63);
        }
        if (ok) return mul;
    }
    (void)slots;
    throw exception::InvalidOperation("could not build a collision-free variable table");
}

// Philox4x32-10 (Salmon et al., SC'11), used for the counter-based agent RNG.
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
#else
        const uint32_t hi0 = (uint32_t)(((uint64_t)0xD2511F53u * c[0]) >> 32);
        const uint32_t hi1 = (uint32_t)(((uint64_t)0xCD9E8D57u * c[2]) >> 32);
#endif
        const uint32_t lo0 = 0xD2511F53u * c[0];
        const uint32_t lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

}  // namespace detail
}  // namespace flamegpu
