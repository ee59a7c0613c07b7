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
 * message neighbourhoods, default 0), "bins_per_radius" (1: spatial bins as wide
 * as the radius, the reference's grid; 2: half-width bins, searched 2 bins either
 * side; no counterpart in the reference; before the first step). */
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
/* Runtime counters (no counterpart in the reference): "fused_layer_launches" (layers
 * enqueued as one launch, graph replays not counted), "layer_images" (fused-layer images
 * the simulation's context holds). */
int fgxm_stat(fgxm_sim* sim, const char* key, uint64_t* out);
/* CUDAEnsemble::simulate (src/flamegpu/simulation/CUDAEnsemble.cu) over `nplans` runs of
 * the model `variant`: run i with seed seeds[i] for steps[i] steps.  options: "key=value"
 * pairs separated by ';': concurrent (runs at once on the per-run path), error_level
 * (0 Off, 1 Slow, 2 Fast), batched (0: every run as its own simulation).  When sum_agent
 * and sum_var are given, every run's exit log sums that variable into out_sums[i] (NaN for
 * a run that failed or did not run).  out_status[i]: 0 done, 1 failed, 2 not started;
 * out_counts: final population per run and agent (agents in name order).  Returns the
 * EnsembleError code when runs failed and the error level is not Off, after filling the
 * outputs; *out_batched = 1 when the runs ran as one batch. */
int fgxm_ensemble_simulate(const char* variant, const float* params, uint32_t nparams, uint32_t nplans,
                           const uint64_t* seeds, const uint32_t* steps, const char* options, const char* sum_agent,
                           const char* sum_var, int32_t* out_status, uint32_t* out_counts, double* out_sums,
                           int32_t* out_batched);
/* CUDASimulation::simulate() for `steps` steps: init functions, step log frame 0,
 * the steps, exit functions, exit log (src/flamegpu/simulation/CUDASimulation.cu:1220-1344). */
int fgxm_simulate(fgxm_sim* sim, uint32_t steps);
/* Step / exit logging (include/flamegpu/simulation/LoggingConfig.h, AgentLoggingConfig.h):
 * exit_log 0 configures the step log, 1 the exit log.  what:
 *   "count" (agent, state), "mean" | "standard_deviation" | "min" | "max" | "sum" (agent,
 *   state, name = variable), "env" (name = property), "timing" (value 0/1),
 *   "frequency" (step log only, value = steps).  state NULL or "" = "default". */
int fgxm_log_config(fgxm_sim* sim, int exit_log, const char* what, const char* agent, const char* state,
                    const char* name, uint32_t value);
/* State files (.xml / .json by extension; src/flamegpu/io/{XML,JSON}State{Reader,Writer}.cu):
 * export = CUDASimulation::exportData; load = SimulationConfig().input_file + applyConfig
 * (configuration, environment, every agent population; the step counter resumes from
 * the file's step count).  fgxm_state_summary parses a file against the model on the
 * host only and writes a JSON description (counts, per-variable sums) to buf. */
int fgxm_export_state(fgxm_sim* sim, const char* path, int pretty);
int fgxm_load_state(fgxm_sim* sim, const char* path);
int fgxm_state_summary(const char* variant, const float* params, uint32_t nparams, const char* path, char* buf,
                       uint32_t buflen);
/* Writes the run log as JSON (src/flamegpu/io/JSONLogger.cu layout). */
int fgxm_export_log(fgxm_sim* sim, const char* path, int steps, int exit_log, int step_time, int exit_time);

/* ---------------------------------------------------------------- batches
 * Many simulations of this model co-scheduled in shared segmented launches
 * (flamegpu::BatchSimulation).  The reference has no batched path (CUDAEnsemble
 * launches each simulation separately, src/flamegpu/simulation/CUDAEnsemble.cu);
 * this is the plan-level equivalent of CUDAEnsemble::simulate with RunPlan seeds,
 * step counts and properties, where each simulation's `params` (as for fgxm_create)
 * give its environment and spatial message geometry.  Simulations are staged by
 * fgxm_batch_add + fgxm_batch_set_agents and join at the next fgxm_batch_step;
 * finished ones stay readable until fgxm_batch_remove. */
typedef struct fgxm_batch fgxm_batch;
int fgxm_batch_create(const char* variant, const float* params, uint32_t nparams, int device, fgxm_batch** out);
int fgxm_batch_destroy(fgxm_batch* b);
/* As fgxm_var_count / fgxm_var_info, for the batch's model. */
int fgxm_batch_var_count(fgxm_batch* b, const char* agent, uint32_t* nvars);
int fgxm_batch_var_info(fgxm_batch* b, const char* agent, uint32_t v, const char** name, int32_t* type,
                        uint32_t* elements, uint32_t* bytes);
int fgxm_batch_add(fgxm_batch* b, const float* params, uint32_t nparams, uint64_t seed, uint32_t steps, int32_t* handle);
int fgxm_batch_set_agents(fgxm_batch* b, int32_t handle, const char* agent, uint32_t n, const void* const* vars,
                          const uint32_t* ids);
/* HBM bytes the staged simulation will occupy once resident. */
int fgxm_batch_estimate_bytes(fgxm_batch* b, int32_t handle, uint64_t* bytes);
int fgxm_batch_step(fgxm_batch* b, uint32_t nsteps);
/* Handles whose steps are done. */
int fgxm_batch_finished(fgxm_batch* b, int32_t* handles, uint32_t capacity, uint32_t* count);
int fgxm_batch_steps_done(fgxm_batch* b, int32_t handle, uint32_t* steps);
int fgxm_batch_count(fgxm_batch* b, int32_t handle, const char* agent, uint32_t* n);
int fgxm_batch_get_agents(fgxm_batch* b, int32_t handle, const char* agent, void* const* vars, uint32_t* ids);
int fgxm_batch_remove(fgxm_batch* b, int32_t handle);
/* Launch shape of the batched agent functions: block size (multiple of 64, <= 1024),
 * agents per thread, persistent blocks per CU (0: one block per tile). */
int fgxm_batch_set_tuning(fgxm_batch* b, uint32_t block, uint32_t agents_per_thread, uint32_t blocks_per_cu);
/* options: "graphs" (capture batched steps as hipGraphs, default 1), "bins_per_radius"
 * (1 or 2, as fgxm_set_option, for the simulations added afterwards). */
int fgxm_batch_set_option(fgxm_batch* b, const char* key, int value);
int fgxm_batch_info(fgxm_batch* b, uint32_t* resident, uint32_t* staged, uint64_t* resident_agents,
                    uint64_t* resident_bytes);
int fgxm_batch_sync(fgxm_batch* b);
/* Runs two untimed steps (repack, graph capture), then `steps` batched steps
 * between two HIP events on the batch stream. */
int fgxm_batch_timed_steps(fgxm_batch* b, uint32_t steps, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* FGX_MODEL_H_ */
