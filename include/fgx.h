/*
 * fgx.h - C ABI of the MI355X (gfx950) hot-path primitives (libfgx.so).
 *
 * Plain pointers and sizes only: device pointers are `void*`/typed pointers
 * into HBM, streams are `void*` (hipStream_t).  Every call is stream-ordered,
 * allocates nothing on the hot path (scratch is sized by fgx_scratch_bytes and
 * owned by the caller) and performs no hidden host synchronisation, so a whole
 * simulation step can be captured into a hipGraph.  Counts that change on the
 * device (compaction, births) stay in device memory.
 *
 * Return value: 0 on success, otherwise an FGX_ERR_* code; the message is
 * available from fgx_last_error() (thread-local).  This replaces the
 * reference's C++ exceptions (include/flamegpu/exception/FLAMEGPUException.h:16,
 * include/flamegpu/simulation/detail/CUDAErrorChecking.cuh:18-48); the C++
 * facade rethrows them as flamegpu::exception::* types.
 *
 * Calls on one context may run on several streams at once as long as each
 * concurrent call has its own scratch buffer.  Scratch contents need no
 * preparation: every call initialises what it reads (fgx_scratch_clear is
 * kept for callers that want defined contents).
 */
#ifndef FGX_H_
#define FGX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FGX_OK = 0,
    FGX_ERR_ARG = 1,       /* invalid argument (reference: exception::InvalidArgument) */
    FGX_ERR_HIP = 2,       /* HIP runtime error (reference: exception::CUDAError) */
    FGX_ERR_CAPACITY = 3,  /* a device buffer is too small for the requested work */
    FGX_ERR_DEVICE = 4,    /* device-side check failed (reference: exception::DeviceError) */
    FGX_ERR_STATE = 5      /* call made in the wrong state (e.g. graph not begun) */
};

#define FGX_MAX_VARS 32

typedef struct fgx_ctx fgx_ctx;
typedef struct fgx_graph fgx_graph;

/* ---------------------------------------------------------------- context
 * Replaces the per-simulation CUDA context / stream set up by
 * CUDASimulation::initialiseSingletons (src/flamegpu/simulation/CUDASimulation.cu:1596-1701). */
int fgx_ctx_create(int device, fgx_ctx** out);
int fgx_ctx_destroy(fgx_ctx* ctx);
const char* fgx_last_error(void);
/* Device properties used for launch shaping. */
int fgx_device_info(fgx_ctx* ctx, int* cu_count, int* xcd_count, size_t* hbm_bytes);
/* Context options (no counterpart in the reference).
 * FGX_OPT_KEEP_SCRATCH_ZERO (default 0): the caller promises that every scratch buffer
 * it passes to fgx_spatial_build, fgx_bucket_build and fgx_agent_sort on this context
 * was zeroed when allocated (fgx_scratch_clear) and is written by nothing but those
 * calls.  The bin sorts then leave their histogram zeroed behind them, and a sort on a
 * buffer that fgx_scratch_clear zeroed, or that the last sort used with the same bin
 * count, skips the clearing launch.  Under stream capture a sort skips it only when
 * every earlier sort of the buffer used the same bin count; a buffer whose graphs are
 * replayed must then keep that one bin count in later sorts too (call fgx_scratch_clear
 * and capture again after changing it). */
enum { FGX_OPT_KEEP_SCRATCH_ZERO = 1 };
int fgx_ctx_set_option(fgx_ctx* ctx, int option, int value);

/* ---------------------------------------------------------------- memory, streams, events, graphs
 * Thin wrappers so the C++ runtime never includes HIP headers. */
int fgx_malloc(fgx_ctx* ctx, size_t bytes, void** out);
int fgx_free(fgx_ctx* ctx, void* p);
int fgx_host_alloc(size_t bytes, void** out);   /* pinned host memory */
int fgx_host_free(void* p);
int fgx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int fgx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int fgx_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* Fill n 32-bit words with `value` using a kernel (graph-capture safe). */
int fgx_fill_u32(uint32_t* dst, uint32_t value, size_t n, void* stream);
/* *p += delta on the device (step counters inside captured graphs). */
int fgx_add_u32(uint32_t* p, uint32_t delta, void* stream);
/* Step bookkeeping in one launch (the end of CUDASimulation::step,
 * src/flamegpu/simulation/CUDASimulation.cu:619-630: message counts reset, step counter
 * incremented): *dst[k] = values[k] for k < n, then *inc += 1 when inc != NULL.
 * n <= FGX_SET_COUNTS_MAX; the pointers are device memory. */
#define FGX_SET_COUNTS_MAX 32
int fgx_set_counts(uint32_t n, uint32_t* const* dst, const uint32_t* values, uint32_t* inc, void* stream);
/* Fill `nvars` variable arrays of n elements with per-variable default values
 * (element bytes var_bytes[v], default value bytes at defaults + def_off[v]);
 * replaces CUDAScatter::broadcastInit (src/flamegpu/simulation/detail/CUDAScatter.cu:404-422). */
int fgx_fill_vars(uint32_t nvars, void* const* vars, const uint32_t* var_bytes, const void* defaults,
                  const uint32_t* def_off, uint32_t n, void* stream);
int fgx_stream_create(void** out);
int fgx_stream_destroy(void* stream);
int fgx_stream_synchronize(void* stream);
int fgx_device_synchronize(void);
int fgx_event_create(void** out);
int fgx_event_destroy(void* ev);
int fgx_event_record(void* ev, void* stream);
int fgx_stream_wait_event(void* stream, void* ev);
int fgx_event_elapsed_ms(void* start, void* stop, float* ms);
/* Stream capture of whole steps (replaces the reference's per-layer host
 * orchestration, src/flamegpu/simulation/CUDASimulation.cu:644-1109). */
int fgx_graph_begin(void* stream);
int fgx_graph_end(void* stream, fgx_graph** out);
int fgx_graph_launch(fgx_graph* g, void* stream);
int fgx_graph_destroy(fgx_graph* g);

/* ---------------------------------------------------------------- fused layer launch
 * One launch for the independent agent functions of a layer (replaces the
 * reference's one launch per function, each on its own stream:
 * src/flamegpu/simulation/CUDASimulation.cu:931-1038, streams at :2036-2048).
 *
 * `kernel` is a layer kernel of the model library (the host stub of a __global__
 * function taking one argument, `const void* image`): it finds its function from its
 * block index and the image header below, and runs it with that function's launch
 * record.  The grid is 8 x sum over functions of ceil(blocks / 8): workgroup w runs on XCD
 * w % 8, and each XCD runs its share of every function (blocks [x*B/8, (x+1)*B/8) of a
 * function of B blocks), in contiguous chunks.  Each fgx_fn_launch gives the function's index among the kernel's
 * functions, its blocks (of `block` threads), its dynamic LDS per block, and its
 * launch record in host memory (`args_bytes` bytes, the same size for every function).
 *
 * Images are content-addressed: the first call with a given image copies it (with
 * hipMemcpyAsync on `stream`, from a pinned copy that is never rewritten) into device
 * space that fgx_ctx_create allocates and the context keeps until it is destroyed; later calls with the same bytes reuse it, and an image is never changed
 * once written, so the call may be captured into a graph and replayed.  When that space
 * (FGX_LAYER_SPACE_BYTES) is full the call returns FGX_ERR_CAPACITY and launches
 * nothing: the caller launches the functions one by one instead. */
#define FGX_LAYER_MAX_FNS 8
#define FGX_LAYER_MAX_ARGS_BYTES 8192u
#define FGX_LAYER_ARGS_OFFSET 128u      /* launch record k at ARGS_OFFSET + k * stride */
#define FGX_LAYER_SPACE_BYTES (4u << 20)   /* per context, device and pinned host */
typedef struct {
    uint32_t nfns;
    uint32_t stride;                          /* bytes between launch records (16-aligned) */
    uint32_t begin[FGX_LAYER_MAX_FNS + 1];    /* first block of function k; begin[nfns] = grid */
    uint32_t fn[FGX_LAYER_MAX_FNS];           /* the functions' indices */
} fgx_layer_header;
typedef struct {
    const void* args;      /* host: the function's launch record */
    uint32_t fn;           /* index among the layer kernel's functions */
    uint32_t blocks;       /* blocks of `block` threads (at least 1) */
    uint32_t lds_bytes;    /* dynamic LDS per block the function needs */
} fgx_fn_launch;
int fgx_layer_launch(fgx_ctx* ctx, const void* kernel, uint32_t block, const fgx_fn_launch* fns, uint32_t nfns,
                     uint32_t args_bytes, void* stream);
/* Layer images held by the context: count and bytes used (tests, diagnostics). */
int fgx_layer_space_info(fgx_ctx* ctx, uint32_t* images, size_t* bytes);

/* ---------------------------------------------------------------- scratch sizing
 * CUB-style two-phase sizing (replaces CubTemporaryMemory,
 * src/flamegpu/simulation/detail/CubTemporaryMemory.cu:23-32 and the
 * scan/compaction buffers of CUDAScanCompaction, include/flamegpu/simulation/detail/CUDAScanCompaction.h:72-138). */
enum {
    FGX_OP_SPATIAL_BUILD = 0,  /* n = max messages, bins = grid bins */
    FGX_OP_COMPACT = 1,        /* n = max items */
    FGX_OP_BIRTHS = 2,         /* n = max parents */
    FGX_OP_AGENT_SORT = 3,     /* n = max agents, bins = 2^key_bits */
    FGX_OP_SCAN = 4,           /* n = max items */
    FGX_OP_STATS = 5           /* fgx_reduce_stats; n and bins unused */
};
int fgx_scratch_bytes(int op, uint32_t n, uint32_t bins, size_t* bytes);
int fgx_scratch_clear(void* scratch, size_t bytes, void* stream);

/* ---------------------------------------------------------------- spatial grid metadata
 * MessageSpatial3D::CUDAModelHandler ctor (src/flamegpu/runtime/messaging/MessageSpatial3D.cu:31-52),
 * 2-D twin (src/flamegpu/runtime/messaging/MessageSpatial2D.cu:36-54),
 * MetaData (include/flamegpu/runtime/messaging/MessageSpatial3D.h:38-68). */
typedef struct {
    int32_t dims;          /* 2 or 3 */
    float min[3], max[3];
    float radius;
    float width[3];
    uint32_t grid[3];      /* grid[2] == 1 in 2-D */
    uint32_t bins;
    int32_t wrap_ok;       /* approxExactlyDivisible(width, cell) on every axis */
    float cell;            /* bin width: radius x bin-width multiplier */
    int32_t span;          /* bins searched either side of a query's bin: 1, or 2 for half-width bins */
} fgx_spatial_meta;
int fgx_spatial_meta_init(int dims, const float* min, const float* max, float radius, fgx_spatial_meta* out);
/* Half-width bins (mult = 0.5; mult = 1 is fgx_spatial_meta_init).  No counterpart in the
 * reference, whose bin width is always the radius: bins of width radius/2 (exact in FP32),
 * grid ceil(width / cell), and the iterators search 2 bins either side in the reference's
 * order, so a query visits 5^3 (5^2) bins of 1/8 (1/4) the volume instead of 3^3 (3^2).
 * The in-radius neighbours are the same; their visit order, so FP summation order, is not. */
int fgx_spatial_meta_init_mult(int dims, const float* min, const float* max, float radius, float mult,
                               fgx_spatial_meta* out);

/* ---------------------------------------------------------------- message-list build (PBM)
 * Replaces MessageSpatial3D/2D::CUDAModelHandler::buildIndex
 * (src/flamegpu/runtime/messaging/MessageSpatial3D.cu:113-146,
 *  src/flamegpu/runtime/messaging/MessageSpatial2D.cu:113-146):
 * atomicHistogram3D (:54-72) + cub::DeviceScan::ExclusiveSum (:134) +
 * CUDAScatter::pbm_reorder (src/flamegpu/simulation/detail/CUDAScatter.cu:276-337).
 * Bins each message (x,y,z; z ignored in 2-D), writes pbm[bins+1] (bin start
 * table, pbm[bins] = n) and scatters every variable into vars_out in bin
 * order, STABLY (input order within a bin).  n = *d_n if d_n != NULL
 * (device-resident count, must be <= n_cap), else n_cap. */
int fgx_spatial_build(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                      const float* x, const float* y, const float* z,
                      uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                      uint32_t* pbm, void* scratch, size_t scratch_bytes, void* stream);
/* As fgx_spatial_build, and also writes the list's 16-byte records in bin order while it
 * scatters: word k of record i is 4-byte variable rec_vars[k] of the sorted element i, or
 * 0 when rec_vars[k] is -1 (the records fgx_pack_records would write from vars_out,
 * without its extra pass).  records holds 16 * n_cap bytes, 16-byte aligned.  No
 * counterpart in the reference (see fgx_pack_records). */
int fgx_spatial_build_records(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                              const float* x, const float* y, const float* z, uint32_t nvars,
                              const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                              const int32_t* rec_vars, void* records, uint32_t* pbm, void* scratch,
                              size_t scratch_bytes, void* stream);
/* Packs four 4-byte variables of a list into 16-byte records: out[i] = {w0[i], w1[i],
 * w2[i], w3[i]} for i < n, a NULL source giving 0 words.  Used for the position and id
 * of each message of a built spatial list, which the neighbour walk then reads with one
 * load per candidate instead of four.  No counterpart in the reference (its iterators
 * read each variable separately, MessageSpatial3DDevice.cuh:592-632).  n = *d_n if d_n
 * != NULL (<= n_cap), else n_cap; out holds 16 * n_cap bytes, 16-byte aligned. */
int fgx_pack_records(const uint32_t* d_n, uint32_t n_cap, const void* w0, const void* w1, const void* w2,
                     const void* w3, void* out, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * fgx_spatial_build with its working space taken from a device array that the library
 * holds on each GPU for the life of the process (FGX_NOALLOC_BUILD_BYTES) instead of a
 * caller buffer.  The caller passes the inputs and outputs only.  n_cap is limited to
 * FGX_NOALLOC_MAX_MESSAGES and the bin count to FGX_NOALLOC_MAX_BINS (FGX_ERR_CAPACITY
 * beyond).  Allocates nothing.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other host threads) race and must be ordered by the
 * caller.  Results are those of fgx_spatial_build. */
#define FGX_NOALLOC_MAX_MESSAGES 2097152u
#define FGX_NOALLOC_MAX_BINS 2097152u
#define FGX_NOALLOC_BUILD_BYTES (34u << 20)
int fgx_spatial_build_noalloc(fgx_ctx* ctx, const fgx_spatial_meta* meta, const uint32_t* d_n, uint32_t n_cap,
                              const float* x, const float* y, const float* z,
                              uint32_t nvars, const void* const* vars_in, void* const* vars_out,
                              const uint32_t* var_bytes, uint32_t* pbm, void* stream);

/* Bucket message index (src/flamegpu/runtime/messaging/MessageBucket.cu:58-150,
 * MessageBucket::CUDAModelHandler::buildIndex): a stable sort of the messages by integer
 * key into `keys` buckets [key_min, key_min + keys), with pbm[0..keys] the bucket start
 * table (pbm[keys] = n).  Keys outside the range are clamped (the device API reports
 * them when they are set).  Within a bucket messages keep their output order, where
 * the reference's order is that of atomics.  Scratch: fgx_scratch_bytes(FGX_OP_SPATIAL_BUILD,
 * n_cap, keys). */
int fgx_bucket_build(fgx_ctx* ctx, int32_t key_min, uint32_t keys, const uint32_t* d_n, uint32_t n_cap, const int32_t* key,
                     uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                     uint32_t* pbm, void* scratch, size_t scratch_bytes, void* stream);

/* Array message list (src/flamegpu/runtime/messaging/MessageArray.cu buildIndex,
 * src/flamegpu/simulation/detail/CUDAScatter.cu:540-640 arrayMessageReorder): message i
 * is copied to element index[i] of the output arrays (length elements, every element
 * nobody wrote is zero).  Two messages to one element, or more messages than elements,
 * set the error word d_error[0..2] to {18, element or count, 0xFFFFFFFF}
 * (exception::ArrayMessageWriteConflict at the next host check).  counts: `length`
 * words of caller scratch. */
int fgx_array_build(fgx_ctx* ctx, uint32_t length, const uint32_t* d_n, uint32_t n_cap, const uint32_t* index,
                    uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                    uint32_t* counts, uint32_t* d_error, void* stream);

/* ---------------------------------------------------------------- auto spatial agent sort
 * Replaces CUDASimulation::spatialSortAgent_async (src/flamegpu/simulation/CUDASimulation.cu:463-573):
 * key = floor(((p-min)/width)*grid) unclamped, masked to floor(log2(bins))+1 bits
 * (calculateSpatialHash, :372-408), stable sort by key (cub::DeviceRadixSort::SortPairs,
 * include/flamegpu/runtime/agent/HostAgentAPI.cuh:858-924), gather of every
 * agent variable (scatterSort_async, src/flamegpu/simulation/detail/CUDAFatAgentStateList.cu:228-243). */
int fgx_agent_sort_key_bits(int dims, const float* min, const float* max, float radius, uint32_t* key_bits);
int fgx_agent_sort(fgx_ctx* ctx, int dims, const float* min, const float* max, float radius,
                   const uint32_t* d_n, uint32_t n_cap, const float* x, const float* y, const float* z,
                   uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                   void* scratch, size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------- compaction
 * Replaces CUDAScatter::scatter + the death scan (scatter_generic,
 * src/flamegpu/simulation/detail/CUDAScatter.cu:68-88,140-179;
 * CUDAFatAgent::processDeath, src/flamegpu/simulation/detail/CUDAFatAgent.cu:105-137):
 * keeps the items whose flag != 0 equals keep != 0, stably, over every variable,
 * in passes ordered by kernel boundaries (per-tile counts, their scan, then ranks
 * and copies; no waits between blocks).  The kept count is written to
 * *d_count_out (device memory), which may be d_n itself. */
int fgx_compact(fgx_ctx* ctx, const uint32_t* flags, uint32_t keep, const uint32_t* d_n, uint32_t n_cap,
                uint32_t nvars, const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes,
                uint32_t* d_count_out, void* scratch, size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------- birth append
 * Replaces CUDAAgentStateList::scatterNew (src/flamegpu/simulation/detail/CUDAAgentStateList.cu:189-250)
 * and the device ID counter (CUDAFatAgent::getDeviceNextID/notifyDeviceBirths,
 * src/flamegpu/simulation/detail/CUDAFatAgent.cu:320-346; DeviceAPI::AgentOut::genID,
 * include/flamegpu/runtime/DeviceAPI.cuh:546-557).  flags == NULL appends every
 * parent (state transition append, CUDAFatAgent::transitionState,
 * src/flamegpu/simulation/detail/CUDAFatAgent.cu:139-184).  Parents with flags[i] != 0
 * append newbuf[*][i] after *d_dst_count entries of dst, in parent order;
 * variable id_var (-1: none) receives *d_next_id + rank.  *d_dst_count and
 * *d_next_id advance by the number of births.  Births beyond dst_capacity
 * are dropped and reported through *d_overflow (set to 1) when non-NULL.
 * d_id_slots (may be NULL) is a device word that an agent function sets to 1 when
 * a parent reads its child's ID in-kernel (AgentOut::getID): the launch then
 * reserved one ID per parent, so every child of the append receives
 * *d_next_id + its parent's index, *d_next_id advances by the parent count, and
 * the append clears the word.  Either way IDs are unique and do not depend on
 * scheduling. */
int fgx_append_births(fgx_ctx* ctx, const uint32_t* flags, const uint32_t* d_n_parents, uint32_t n_parents_cap,
                      uint32_t nvars, const void* const* newbuf, void* const* dst, const uint32_t* var_bytes,
                      int32_t id_var, uint32_t* d_dst_count, uint32_t dst_capacity, uint32_t* d_next_id,
                      uint32_t* d_id_slots, uint32_t* d_overflow, void* scratch, size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------- agent function conditions
 * CUDAFatAgent::processFunctionCondition (src/flamegpu/simulation/detail/CUDAFatAgent.cu:186-236):
 * stable partition of every variable, items with flags[i] != 0 (condition false) first, then
 * the others, both in input order.  *d_first_count receives the size of the first part (the
 * disabled agents); d_work is one device word of working space.  Scratch as fgx_compact. */
int fgx_partition(fgx_ctx* ctx, const uint32_t* flags, const uint32_t* d_n, uint32_t n_cap, uint32_t nvars,
                  const void* const* vars_in, void* const* vars_out, const uint32_t* var_bytes, uint32_t* d_first_count,
                  uint32_t* d_work, void* scratch, size_t scratch_bytes, void* stream);
/* Death compaction under a function condition (CUDAFatAgentStateList::scatterDeath with
 * disabled agents, src/flamegpu/simulation/detail/CUDAFatAgentStateList.cu:155-178): the first
 * *d_prefix items (disabled) are kept in front, item i >= *d_prefix is kept when
 * (flags[i - *d_prefix] != 0) == (keep != 0). */
int fgx_compact_prefix(fgx_ctx* ctx, const uint32_t* flags, uint32_t keep, const uint32_t* d_prefix, const uint32_t* d_n,
                       uint32_t n_cap, uint32_t nvars, const void* const* vars_in, void* const* vars_out,
                       const uint32_t* var_bytes, uint32_t* d_count_out, void* scratch, size_t scratch_bytes, void* stream);
/* State transition under a function condition (CUDAFatAgent::transitionState,
 * src/flamegpu/simulation/detail/CUDAFatAgent.cu:139-184, scatterAll from data_condition):
 * appends items [*d_src_off, *d_n) of src after *d_dst_count items of dst. */
int fgx_append_range(fgx_ctx* ctx, const uint32_t* d_src_off, const uint32_t* d_n, uint32_t n_cap, uint32_t nvars,
                     const void* const* src, void* const* dst, const uint32_t* var_bytes, uint32_t* d_dst_count,
                     uint32_t dst_capacity, uint32_t* d_overflow, void* scratch, size_t scratch_bytes, void* stream);
/* *dst = max(*a - *b, 0) on the device (enabled-agent counts). */
int fgx_sub_u32(uint32_t* dst, const uint32_t* a, const uint32_t* b, void* stream);

/* ---------------------------------------------------------------- scan
 * Exclusive prefix sum of n u32 (cub::DeviceScan::ExclusiveSum call sites,
 * e.g. src/flamegpu/runtime/messaging/MessageSpatial3D.cu:134). */
int fgx_exclusive_scan_u32(fgx_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                           void* scratch, size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------- ID assignment
 * allocateIDs (src/flamegpu/simulation/detail/CUDAFatAgent.cu:311-319): ids[i] = next_id + i
 * where ids[i] == 0 (ID_NOT_SET). */
int fgx_assign_ids(uint32_t* ids, uint32_t n, uint32_t next_id, void* stream);

/* ---------------------------------------------------------------- reductions (host agent API)
 * Deterministic on-device reductions for HostAgentAPI::count/sum
 * (include/flamegpu/runtime/agent/HostAgentAPI.cuh:174,689): fixed-shape
 * tree so the result does not depend on scheduling.  type: 0 f32, 1 f64, 2 i32, 3 u32, 4 i64, 5 u64.
 * Writes a double (f32/f64) or 64-bit integer to *d_out. */
int fgx_reduce_sum(fgx_ctx* ctx, int type, const void* in, const uint32_t* d_n, uint32_t n_cap,
                   void* d_out, void* scratch, size_t scratch_bytes, void* stream);

/* Statistics of one agent variable for the step and exit logs
 * (include/flamegpu/simulation/AgentLoggingConfig.h: getAgentVariableSum/Min/Max/Mean/
 * StandardDevFunc; src/flamegpu/simulation/CUDASimulation.cu:1949-2021 processStepLog).
 * Writes four 8-byte words to d_out (8-byte aligned, device memory):
 *   [0] sum, [1] min, [2] max  as double (f32/f64), int64 (i32/i64) or uint64 (u32/u64);
 *   [3] sum over items of ((double)x - sum / n)^2 as double (0 when n == 0).
 * min/max of an empty list are the type's largest / lowest value (the reduction identities).
 * Deterministic: a fixed-shape tree, so the bits do not depend on scheduling.  Four
 * launches, no host synchronisation.  Scratch: fgx_scratch_bytes(FGX_OP_STATS, 0, 0). */
int fgx_reduce_stats(fgx_ctx* ctx, int type, const void* in, const uint32_t* d_n, uint32_t n_cap,
                     void* d_out, void* scratch, size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------- batched (segmented) primitives
 * A batch runs many independent simulations in the same launches.  Their lists
 * are packed one after another (simulation s owns elements [elem_off, elem_off + n)),
 * and their bin tables likewise (keys [key_off, key_off + keys)).  The reference has
 * no batched path: CUDAEnsemble runs each simulation with its own launches
 * (src/flamegpu/simulation/CUDAEnsemble.cu, src/flamegpu/simulation/detail/SimRunner.cu:51-90);
 * these entries replace that per-simulation launch sequence for the message-list
 * build and the agent sort.  Each simulation's results equal a separate run of
 * fgx_spatial_build / fgx_agent_sort on its own list. */
typedef struct {
    uint32_t elem_off;     /* first element of the simulation (filled by fgx_segments_pack) */
    uint32_t n;            /* elements */
    uint32_t key_off;      /* first key / bin of the simulation (filled by fgx_segments_pack) */
    uint32_t keys;         /* keys (sort) or bins (build) of the simulation (set by fgx_seg_key_init) */
    uint32_t tile_off;     /* first 2048-element tile (filled by fgx_segments_pack) */
    int32_t dims;
    float min[3];
    float radius;
    float width[3];
    uint32_t grid[3];
    uint32_t key_mask;     /* agent sort: keys - 1 */
    uint32_t pad[3];
} fgx_seg_key;             /* 80 bytes */

enum { FGX_SEG_MESSAGE_BINS = 0, FGX_SEG_AGENT_SORT = 1 };

/* Geometry of one simulation's keys: message bins (as fgx_spatial_meta_init) or the
 * agent sort key (as fgx_agent_sort).  Leaves the offsets and n untouched. */
int fgx_seg_key_init(int kind, int dims, const float* min, const float* max, float radius, fgx_seg_key* out);
/* Packs the simulations in order: fills elem_off, key_off and tile_off from n and keys,
 * and returns the totals (elements, keys, 2048-element tiles). */
int fgx_segments_pack(fgx_seg_key* segs, uint32_t nsegs, uint32_t* n_total, uint32_t* keys_total, uint32_t* tiles_total);
/* Segmented message-list build: d_segs (device copy of the packed table) with kind
 * FGX_SEG_MESSAGE_BINS; pbm holds keys_total + 1 entries of absolute message indices.
 * n_total bounds the packed elements (and sizes the scratch); d_n_total, when given, is
 * the device-resident count (populations that change on the device, whose tables
 * fgx_seg_layout keeps; tiles_total then bounds the tiles).
 * Scratch: fgx_scratch_bytes(FGX_OP_SPATIAL_BUILD, n_total, keys_total). */
int fgx_spatial_build_seg(fgx_ctx* ctx, const fgx_seg_key* d_segs, uint32_t nsegs, uint32_t tiles_total,
                          uint32_t n_total, const uint32_t* d_n_total, uint32_t keys_total, const float* x,
                          const float* y, const float* z, uint32_t nvars, const void* const* vars_in,
                          void* const* vars_out, const uint32_t* var_bytes, uint32_t* pbm, void* scratch,
                          size_t scratch_bytes, void* stream);
/* Segmented agent sort (kind FGX_SEG_AGENT_SORT), counts as above.  Scratch:
 * fgx_scratch_bytes(FGX_OP_AGENT_SORT, n_total, keys_total). */
int fgx_agent_sort_seg(fgx_ctx* ctx, const fgx_seg_key* d_segs, uint32_t nsegs, uint32_t tiles_total,
                       uint32_t n_total, const uint32_t* d_n_total, uint32_t keys_total, const float* x,
                       const float* y, const float* z, uint32_t nvars, const void* const* vars_in,
                       void* const* vars_out, const uint32_t* var_bytes, void* scratch, size_t scratch_bytes,
                       void* stream);
/* p[i] += delta for i < n (per-simulation step counters of a batch). */
int fgx_add_u32_n(uint32_t* p, uint32_t n, uint32_t delta, void* stream);
/* Copies element ranges of several variables in one launch: range r moves
 * d_ranges[3r+2] elements from src offset d_ranges[3r] to dst offset d_ranges[3r+1].
 * Destination ranges must be sorted by offset and not overlap; n_total is the
 * sum of the range lengths (repacking a batch when simulations join or leave). */
int fgx_copy_ranges(uint32_t nvars, const void* const* src, void* const* dst, const uint32_t* var_bytes,
                    const uint32_t* d_ranges, uint32_t nranges, uint32_t n_total, void* stream);

/* ---------------------------------------------------------------- batches whose populations change
 * A batch packs its simulations' lists one after another; when deaths and births change
 * the populations, each simulation's count stays on the device (reference: each
 * simulation of an ensemble compacts and appends on its own, CUDAFatAgent.cu:105-137 and
 * CUDAAgentStateList.cu:189-250).  A stable compaction of the whole packed list keeps
 * every simulation's survivors together and in order; these entries recount the
 * simulations, append each one's new agents after its own, and patch every table that
 * records the layout (per-simulation offset, count and first tile). */
typedef struct {
    void* table;          /* device records, one per simulation (+ 1 sentinel record when sentinel != 0) */
    uint32_t stride;      /* bytes per record */
    int32_t off_field;    /* byte offset of the uint32 element offset in a record, or -1 */
    int32_t n_field;      /* byte offset of the uint32 element count, or -1 */
    int32_t tile_field;   /* byte offset of the uint32 first tile, or -1 */
    uint32_t tile_elems;  /* elements per tile when tile_field >= 0 */
    int32_t sentinel;     /* record nsegs receives the totals (offset = elements, tile = tiles, count 0) */
} fgx_seg_target;
#define FGX_SEG_MAX_TARGETS 8
/* d_cnt[s] (+ d_add[s] when d_add is given, stored back into d_cnt) -> d_off[s] (exclusive
 * offsets), *d_total, and the fields of every target record.  One block of 1024 threads. */
int fgx_seg_layout(uint32_t nsegs, uint32_t* d_cnt, const uint32_t* d_add, uint32_t* d_off, uint32_t* d_total,
                   const fgx_seg_target* targets, uint32_t ntargets, void* stream);
/* d_out[s] = flags equal to `keep` in [d_off[s], d_off[s] + d_cnt[s]); d_out may be d_cnt. */
int fgx_seg_count(uint32_t nsegs, const uint32_t* d_off, const uint32_t* d_cnt, const uint32_t* flags, uint32_t keep,
                  uint32_t* d_out, void* stream);
/* Writes the list with each simulation's d_born[s] new agents after its own d_cnt[s]
 * agents into list_out (capacity n_cap_out elements).  born_in holds the new agents of all
 * simulations compacted in order (a stable fgx_compact of the new-agent buffers by birth
 * flag), so simulation s's are at the exclusive prefix of d_born.  With id_var >= 0, new
 * agent r of simulation s gets ID d_next_id[s] + r and d_next_id[s] advances by d_born[s]
 * (CUDAAgentStateList::scatterNew numbering, per simulation).  d_work: 2 * nsegs + 2 words.
 * The caller then runs fgx_seg_layout with d_add = d_born. */
int fgx_seg_append(uint32_t nsegs, const uint32_t* d_off, const uint32_t* d_cnt, const uint32_t* d_born, uint32_t n_cap_out,
                   uint32_t nvars, const void* const* list_in, const void* const* born_in, void* const* list_out,
                   const uint32_t* var_bytes, int32_t id_var, uint32_t* d_next_id, uint32_t* d_work, void* stream);

#ifdef __cplusplus
}
#endif
#endif  /* FGX_H_ */
