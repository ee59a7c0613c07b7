/*
 * Serial CPU oracle for the per-step spatial-messaging hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * Every function restates the reference's semantics (FLAME GPU 2, paths
 * relative to the reference root) in plain serial C++:
 *   - grid metadata / bin hash / PBM / neighbour visit order:
 *       src/flamegpu/runtime/messaging/MessageSpatial3D.cu:31-146,
 *       src/flamegpu/runtime/messaging/MessageSpatial2D.cu:36-146,
 *       include/flamegpu/runtime/messaging/MessageSpatial3D/MessageSpatial3DDevice.cuh:52-100,646-749,
 *       include/flamegpu/runtime/messaging/MessageSpatial2D/MessageSpatial2DDevice.cuh:605-697,
 *       include/flamegpu/detail/numeric.h:12-20
 *   - compaction / birth append: src/flamegpu/simulation/detail/CUDAScatter.cu:68-117,
 *       src/flamegpu/simulation/detail/CUDAFatAgent.cu:105-137,311-391,
 *       src/flamegpu/simulation/detail/CUDAAgentStateList.cu:189-250
 *   - agent spatial sort key: src/flamegpu/simulation/CUDASimulation.cu:335-573
 *   - RNG float mapping: include/flamegpu/runtime/random/AgentRandom.cuh:76-84
 *   - step / layer order: src/flamegpu/simulation/CUDASimulation.cu:575-1109
 *
 * Two conventions replace non-deterministic parts of the reference (whose
 * within-bin message order and birth IDs come from atomicInc):
 *   - within a bin, messages keep their output order (stable counting sort);
 *   - new agents get ID = nextID + (exclusive rank of the parent among the
 *     parents that gave birth), appended after the survivors in parent order.
 * The RNG is Philox4x32-10 (Salmon et al., SC'11) keyed by
 * (seed, agent ID, step, function stream, draw block).
 */
#ifndef FGX_ORACLE_H_
#define FGX_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 and the keyed agent stream ------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* k-th 32-bit draw of agent `id` in function stream `stream` at `step`. */
uint32_t orc_rng_u32(uint64_t seed, uint32_t id, uint32_t step, uint32_t stream, uint32_t k);
float orc_rng_uniform(uint64_t seed, uint32_t id, uint32_t step, uint32_t stream, uint32_t k);

/* ---- Spatial grid metadata -------------------------------------------- */
typedef struct {
    int32_t dims;          /* 2 or 3 */
    float min[3], max[3];
    float radius;
    float width[3];
    uint32_t grid[3];      /* grid[2] == 1 for 2-D */
    uint32_t bins;
    int32_t wrap_ok;
    float cell;            /* bin width: radius x bin-width multiplier (radius when the multiplier is 1) */
    int32_t span;          /* bins searched either side of the query's bin: ceil(1 / multiplier) */
} orc_spatial_meta;

int orc_spatial_meta_init(int dims, const float* mn, const float* mx, float radius, orc_spatial_meta* out);
/* Bins of width radius x mult (mult = 1 or 1/2): the grid is ceil(width / (radius x mult)) and
 * the iterators search span = ceil(1/mult) bins either side, in the reference's order
 * (strips of 2*span+1 bins along x, dy outer, dz inner; wrapped cells dx, dy, dz).
 * There is no multiplier in the reference; mult = 1 is its grid exactly. */
int orc_spatial_meta_init_mult(int dims, const float* mn, const float* mx, float radius, float mult,
                               orc_spatial_meta* out);
uint32_t orc_spatial_hash(const orc_spatial_meta* m, float x, float y, float z);

/* Stable counting sort of a message list by bin.  pbm has bins+1 entries.
 * perm (optional, n entries) receives, for each output slot, its input index. */
int orc_spatial_build(const orc_spatial_meta* m, uint32_t n,
                      const float* x, const float* y, const float* z,
                      uint32_t nvars, const void* const* vars_in, const uint32_t* var_bytes,
                      void* const* vars_out, uint32_t* pbm, uint32_t* perm);

/* Message indices (into the sorted list) visited by the reference iterator
 * for a query at (x,y,z): mode 0 = Filter (clamped Moore strips),
 * mode 1 = WrapFilter (27/9 cells modulo grid).  Returns the count; writes
 * at most max_out indices. */
uint32_t orc_spatial_candidates(const orc_spatial_meta* m, const uint32_t* pbm, int mode,
                                float x, float y, float z, uint32_t* out, uint32_t max_out);

/* Exact mean candidate count per message for the Filter iterator when the
 * query points are the messages themselves: sum_b n_b * sum_{b' in nbhd(b)} n_b' / N. */
double orc_spatial_mean_candidates(const orc_spatial_meta* m, const uint32_t* pbm, int mode);

/* ---- Compaction, births, agent sort ----------------------------------- */
/* Stable partition keeping items whose flag == keep (1 = ALIVE). Returns count. */
uint32_t orc_compact(uint32_t n, const uint32_t* flags, uint32_t keep,
                     uint32_t nvars, const void* const* in, const uint32_t* var_bytes, void* const* out);
/* Append the flagged entries of a per-parent new-agent buffer after
 * dst_count entries of dst, writing ID = next_id + rank into id_var.
 * Returns the new dst count; *next_id advances by the number of births. */
uint32_t orc_births(uint32_t n_parents, const uint32_t* flags,
                    uint32_t nvars, const void* const* newbuf, const uint32_t* var_bytes,
                    void* const* dst, uint32_t dst_count, uint32_t id_var, uint32_t* next_id);
/* The same, with by_parent_index != 0: the launch's parents reserved one ID each, so a
 * child's ID is next_id + its parent's index and *next_id advances by n_parents (the
 * library's rule once a parent reads AgentOut::getID() in-kernel). */
uint32_t orc_births_ex(uint32_t n_parents, const uint32_t* flags,
                       uint32_t nvars, const void* const* newbuf, const uint32_t* var_bytes,
                       void* const* dst, uint32_t dst_count, uint32_t id_var, uint32_t* next_id,
                       int32_t by_parent_index);
/* Auto spatial sort key (unclamped) masked to the sort's bit range. */
void orc_agent_sort_keys(int dims, const float* mn, const float* mx, float radius, uint32_t n,
                         const float* x, const float* y, const float* z, uint32_t* keys, uint32_t* key_bits);
/* Stable sort permutation of keys (perm[i] = source index of output i). */
void orc_stable_sort_perm(uint32_t n, const uint32_t* keys, uint32_t* perm);

/* ---- Whole-model simulations ------------------------------------------ */
/* model: 0 = circles spatial-3D, 1 = boids spatial-3D (wrapped),
 *        2 = predator-prey spatial-2D, 3 = boids brute-force (wrapped distances off),
 *        4 = circles brute-force.
 * params: model specific floats (see oracle/models.cpp for the order). */
typedef struct orc_sim orc_sim;
orc_sim* orc_sim_create(int model, const float* params, uint32_t nparams, uint64_t seed);
void orc_sim_destroy(orc_sim* s);
/* Agent type t (0 = first agent; predator-prey: 0 prey, 1 predator).
 * vars: one pointer per agent variable in the model's declared order
 * (orc_sim_var_count / orc_sim_var_bytes); ids may be NULL (assigned 1..n). */
uint32_t orc_sim_var_count(const orc_sim* s, int t);
uint32_t orc_sim_var_bytes(const orc_sim* s, int t, uint32_t v);
int orc_sim_set_agents(orc_sim* s, int t, uint32_t n, const void* const* vars, const uint32_t* ids);
uint32_t orc_sim_count(const orc_sim* s, int t);
int orc_sim_get_agents(const orc_sim* s, int t, void* const* vars, uint32_t* ids);
int orc_sim_step(orc_sim* s, uint32_t nsteps);
/* PBM of message list `msg` as last built (bins+1 entries), and its bin count. */
uint32_t orc_sim_bins(const orc_sim* s, int msg);
/* Rebuild every spatial message grid with bin width radius x mult (before the first step). */
int orc_sim_set_bin_mult(orc_sim* s, float mult);
int orc_sim_get_pbm(const orc_sim* s, int msg, uint32_t* pbm);
/* Mean candidates per agent of the last neighbour pass of message list msg. */
double orc_sim_last_candidates(const orc_sim* s, int msg);

#ifdef __cplusplus
}
#endif
#endif  /* FGX_ORACLE_H_ */
