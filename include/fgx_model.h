/*
 * fgx_model.h - C ABI exported by every model library (libfgxm_<model>.so).
 *
 * A model library is a translation unit written against the FLAME GPU 2 C++
 * API (flamegpu/flamegpu.h) that defines
 *     flamegpu::ModelDescription fgxm_build_model(const char* variant,
 *                                                 const float* params, uint32_t nparams);
 * and includes flamegpu/model_abi.h.  These entry points drive the model's
 * CUDASimulation from any host language (ctypes, cgo, JNI ...); they mirror
 * CUDASimulation::setPopulationData/getPopulationData/step
 * (include/flamegpu/simulation/CUDASimulation.h:153-369) with host buffers
 * holding one array per agent variable.
 *
 * Return value: 0 on success, else an FGX_ERR_* code with the message in
 * fgxm_last_error().
 */
#ifndef FGX_MODEL_H_
#define FGX_MODEL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fgxm_sim fgxm_sim;

const char* fgxm_last_error(void);
/* Build the model (variant selects one of several models in a library, may be
 * NULL) and its CUDASimulation on `device` with RNG seed `seed`. */
int fgxm_create(const char* variant, const float* params, uint32_t nparams, uint64_t seed, int device, fgxm_sim** out);
int fgxm_destroy(fgxm_sim* sim);
/* Agent variables in declaration order (without the internal _id).
 * type: 0 f32, 1 f64, 2 i32, 3 u32, 4 i64, 5 u64, 6 i8, 7 u8, 8 i16, 9 u16, -1 other;
 * elements: array length (1 for scalars); bytes = element size x elements. */
int fgxm_var_count(fgxm_sim* sim, const char* agent, uint32_t* nvars);
int fgxm_var_info(fgxm_sim* sim, const char* agent, uint32_t v, const char** name, int32_t* type, uint32_t* elements,
                  uint32_t* bytes);
/* state may be NULL or "" for the agent's initial state; ids may be NULL (assigned). */
int fgxm_set_agents(fgxm_sim* sim, const char* agent, const char* state, uint32_t n, const void* const* vars,
                    const uint32_t* ids);
int fgxm_count(fgxm_sim* sim, const char* agent, const char* state, uint32_t* out);
int fgxm_get_agents(fgxm_sim* sim, const char* agent, const char* state, void* const* vars, uint32_t* ids);
int fgxm_step(fgxm_sim* sim, uint32_t nsteps);
/* options: "graphs" (capture steps as hipGraphs, default 1), "concurrency"
 * (in-layer streams, default 1), "kernel_timing" (HIP events around every
 * launch, eager steps only, default 0), "staging" (LDS staging of spatial
 * message neighbourhoods, default 0). */
int fgxm_set_option(fgxm_sim* sim, const char* key, int value);
int fgxm_sync(fgxm_sim* sim);
/* Bin start table of a spatial message list as last built. */
int fgxm_get_pbm(fgxm_sim* sim, const char* msg, uint32_t* pbm, uint32_t capacity, uint32_t* bins);
/* Per-label kernel timings since the last reset as text lines "label total_ms count". */
int fgxm_kernel_times(fgxm_sim* sim, char* buf, uint32_t buflen, int reset);
/* Runs `steps` steps between two HIP events on the simulation stream and
 * returns the elapsed milliseconds (inputs already resident on the device). */
int fgxm_timed_steps(fgxm_sim* sim, uint32_t steps, double* ms);
int fgxm_step_counter(fgxm_sim* sim, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* FGX_MODEL_H_ */
