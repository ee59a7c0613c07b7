// Implementation of include/fgx_model.h for one model translation unit.
// Include once, after defining fgxm_build_model().
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <typeindex>

#include "fgx_model.h"
#include "flamegpu/flamegpu.h"

flamegpu::ModelDescription fgxm_build_model(const char* variant, const float* params, uint32_t nparams);

struct fgxm_sim {
    std::unique_ptr<flamegpu::ModelDescription> model;
    std::unique_ptr<flamegpu::CUDASimulation> sim;
    void* ev0 = nullptr;
    void* ev1 = nullptr;
    std::unique_ptr<flamegpu::StepLoggingConfig> step_log;
    std::unique_ptr<flamegpu::LoggingConfig> exit_log;
};

namespace fgxm_detail {
inline thread_local std::string last_error;

// The layer kernel of this translation unit: every agent function defined above.
inline flamegpu::detail::LayerKernel layer_kernel() {
    return flamegpu::detail::LayerKernelOf<flamegpu::detail::fn_collect<0, __COUNTER__>::type>::get();
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return FGX_OK;
    } catch (const flamegpu::exception::InvalidArgument& e) {
        last_error = std::string(e.exception_type()) + ": " + e.what();
        return FGX_ERR_ARG;
    } catch (const flamegpu::exception::DeviceError& e) {
        last_error = std::string(e.exception_type()) + ": " + e.what();
        return FGX_ERR_DEVICE;
    } catch (const flamegpu::exception::CUDAError& e) {
        last_error = std::string(e.exception_type()) + ": " + e.what();
        return FGX_ERR_HIP;
    } catch (const flamegpu::exception::FLAMEGPUException& e) {
        last_error = std::string(e.exception_type()) + ": " + e.what();
        return FGX_ERR_STATE;
    } catch (const std::exception& e) {
        last_error = e.what();
        return FGX_ERR_STATE;
    }
}

inline std::shared_ptr<flamegpu::AgentData> agent(fgxm_sim* s, const char* name) {
    auto& agents = s->model->data()->agents;
    auto it = agents.find(name ? name : "");
    if (it == agents.end()) throw flamegpu::exception::InvalidAgentName(std::string("unknown agent ") + (name ? name : "(null)"));
    return it->second;
}
}  // namespace fgxm_detail

namespace fgxm_detail {
inline void var_info(const std::shared_ptr<flamegpu::AgentData>& a, uint32_t v, const char** name, int32_t* type,
                     uint32_t* elements, uint32_t* bytes) {
    {
        if (v >= a->declared.size()) throw flamegpu::exception::OutOfBoundsException("variable index out of range");
        const auto& var = a->variables.at(a->declared[v]);
        *name = var.name.c_str();
        *bytes = (uint32_t)var.bytes();
        *elements = var.elements;
        const std::type_index t = var.type;
        int32_t code = -1;
        if (t == typeid(float)) code = 0;
        else if (t == typeid(double)) code = 1;
        else if (t == typeid(int32_t)) code = 2;
        else if (t == typeid(uint32_t)) code = 3;
        else if (t == typeid(int64_t) || t == typeid(long long)) code = 4;
        else if (t == typeid(uint64_t) || t == typeid(unsigned long long)) code = 5;
        else if (t == typeid(int8_t)) code = 6;
        else if (t == typeid(uint8_t)) code = 7;
        else if (t == typeid(int16_t)) code = 8;
        else if (t == typeid(uint16_t)) code = 9;
        *type = code;
    }
}
}  // namespace fgxm_detail

#define FGXM_EXPORT __attribute__((visibility("default")))

extern "C" {

FGXM_EXPORT const char* fgxm_last_error(void) { return fgxm_detail::last_error.c_str(); }

FGXM_EXPORT int fgxm_create(const char* variant, const float* params, uint32_t nparams, uint64_t seed, int device, fgxm_sim** out) {
    return fgxm_detail::guard([&] {
        auto* s = new fgxm_sim();
        try {
            s->model.reset(new flamegpu::ModelDescription(fgxm_build_model(variant, params, nparams)));
            s->sim.reset(new flamegpu::CUDASimulation(*s->model));
            s->sim->SimulationConfig().random_seed = seed;
            s->sim->CUDAConfig().device_id = device;
            s->sim->applyConfig();
            s->sim->setLayerKernel(fgxm_detail::layer_kernel());
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

FGXM_EXPORT int fgxm_destroy(fgxm_sim* s) {
    return fgxm_detail::guard([&] {
        if (s->ev0) fgx_event_destroy(s->ev0);
        if (s->ev1) fgx_event_destroy(s->ev1);
        delete s;
    });
}

FGXM_EXPORT int fgxm_var_count(fgxm_sim* s, const char* agent, uint32_t* n) {
    return fgxm_detail::guard([&] { *n = (uint32_t)fgxm_detail::agent(s, agent)->declared.size(); });
}

FGXM_EXPORT int fgxm_var_info(fgxm_sim* s, const char* agent, uint32_t v, const char** name, int32_t* type, uint32_t* elements,
                  uint32_t* bytes) {
    return fgxm_detail::guard([&] { fgxm_detail::var_info(fgxm_detail::agent(s, agent), v, name, type, elements, bytes); });
}


FGXM_EXPORT int fgxm_set_agents(fgxm_sim* s, const char* agent, const char* state, uint32_t n, const void* const* vars,
                    const uint32_t* ids) {
    return fgxm_detail::guard([&] {
        auto a = fgxm_detail::agent(s, agent);
        flamegpu::AgentVector pop(flamegpu::AgentDescription(a), n);
        for (size_t v = 0; v < a->declared.size(); ++v) {
            auto& raw = pop.raw(a->declared[v]);
            std::memcpy(raw.data(), vars[v], raw.size());
        }
        if (ids) std::memcpy(pop.raw("_id").data(), ids, (size_t)n * 4);
        s->sim->setPopulationData(pop, state ? state : "");
    });
}

FGXM_EXPORT int fgxm_count(fgxm_sim* s, const char* agent, const char* state, uint32_t* out) {
    return fgxm_detail::guard([&] {
        flamegpu::HostAPI api(*s->sim, 0);
        *out = api.agent(agent, state ? state : "").count();
    });
}

FGXM_EXPORT int fgxm_get_agents(fgxm_sim* s, const char* agent, const char* state, void* const* vars, uint32_t* ids) {
    return fgxm_detail::guard([&] {
        auto a = fgxm_detail::agent(s, agent);
        flamegpu::AgentVector pop{flamegpu::AgentDescription(a)};
        s->sim->getPopulationData(pop, state ? state : "");
        for (size_t v = 0; v < a->declared.size(); ++v) {
            auto& raw = pop.raw(a->declared[v]);
            if (vars[v]) std::memcpy(vars[v], raw.data(), raw.size());
        }
        if (ids) std::memcpy(ids, pop.raw("_id").data(), pop.size() * 4);
    });
}

FGXM_EXPORT int fgxm_step(fgxm_sim* s, uint32_t nsteps) {
    return fgxm_detail::guard([&] {
        for (uint32_t i = 0; i < nsteps; ++i) s->sim->step();
        s->sim->check_device_error();
    });
}

// option "bins_per_radius": 1 (bins as wide as the radius) or 2 (half-width bins)
inline float bins_per_radius_mult(int value) {
    if (value != 1 && value != 2) throw flamegpu::exception::InvalidArgument("bins_per_radius must be 1 or 2");
    return value == 1 ? 1.0f : 0.5f;
}

FGXM_EXPORT int fgxm_set_option(fgxm_sim* s, const char* key, int value) {
    return fgxm_detail::guard([&] {
        const std::string k = key ? key : "";
        if (k == "graphs") s->sim->CUDAConfig().graphs = value != 0;
        else if (k == "concurrency") s->sim->CUDAConfig().inLayerConcurrency = value != 0;
        else if (k == "kernel_timing") s->sim->setKernelTiming(value != 0);
        else if (k == "staging") s->sim->setNeighbourhoodStaging(value != 0);
        else if (k == "unit_tables") s->sim->setUnitTables(value != 0);
        else if (k == "layer_fusion") s->sim->setLayerFusion(value != 0);
        else if (k == "bins_per_radius") s->sim->setBinWidthMultiplier(bins_per_radius_mult(value));
        else throw flamegpu::exception::InvalidArgument("unknown option " + k);
    });
}

FGXM_EXPORT int fgxm_sync(fgxm_sim* s) {
    return fgxm_detail::guard([&] {
        s->sim->synchronize();
        s->sim->check_device_error();
    });
}

FGXM_EXPORT int fgxm_get_pbm(fgxm_sim* s, const char* msg, uint32_t* pbm, uint32_t capacity, uint32_t* bins) {
    return fgxm_detail::guard([&] {
        const uint32_t b = s->sim->messageBins(msg ? msg : "");
        *bins = b;
        if (capacity < b + 1) throw flamegpu::exception::InvalidArgument("PBM buffer too small");
        s->sim->getPBM(msg, pbm);
    });
}

FGXM_EXPORT int fgxm_kernel_times(fgxm_sim* s, char* buf, uint32_t buflen, int reset) {
    return fgxm_detail::guard([&] {
        s->sim->synchronize();
        std::ostringstream os;
        for (auto& kv : s->sim->getKernelTimes()) os << kv.first << " " << kv.second.ms << " " << kv.second.count << "\n";
        const std::string t = os.str();
        if (buf && buflen) {
            std::strncpy(buf, t.c_str(), buflen - 1);
            buf[buflen - 1] = 0;
        }
        if (reset) s->sim->resetKernelTimes();
    });
}

FGXM_EXPORT int fgxm_timed_steps(fgxm_sim* s, uint32_t steps, double* ms) {
    return fgxm_detail::guard([&] {
        if (!s->ev0) flamegpu::detail::fgx_check(fgx_event_create(&s->ev0), "event");
        if (!s->ev1) flamegpu::detail::fgx_check(fgx_event_create(&s->ev1), "event");
        void* st = s->sim->stream();
        s->sim->synchronize();
        flamegpu::detail::fgx_check(fgx_event_record(s->ev0, st), "event record");
        for (uint32_t i = 0; i < steps; ++i) s->sim->step();
        flamegpu::detail::fgx_check(fgx_event_record(s->ev1, st), "event record");
        float f = 0.0f;
        flamegpu::detail::fgx_check(fgx_event_elapsed_ms(s->ev0, s->ev1, &f), "event elapsed");
        *ms = f;
        s->sim->check_device_error();
    });
}

FGXM_EXPORT int fgxm_step_counter(fgxm_sim* s, uint32_t* out) {
    return fgxm_detail::guard([&] { *out = s->sim->getStepCounter(); });
}

FGXM_EXPORT int fgxm_ensemble_simulate(const char* variant, const float* params, uint32_t nparams, uint32_t nplans,
                                       const uint64_t* seeds, const uint32_t* steps, const char* options,
                                       const char* sum_agent, const char* sum_var, int32_t* out_status,
                                       uint32_t* out_counts, double* out_sums, int32_t* out_batched) {
    return fgxm_detail::guard([&] {
        flamegpu::ModelDescription model = fgxm_build_model(variant, params, nparams);
        flamegpu::RunPlanVector plans(model, nplans);
        for (uint32_t i = 0; i < nplans; ++i) {
            plans[i].setRandomSimulationSeed(seeds[i]);
            plans[i].setSteps(steps[i]);
        }
        flamegpu::CUDAEnsemble ens(model);
        std::istringstream opts(options ? options : "");
        std::string kv;
        while (std::getline(opts, kv, ';')) {
            const auto eq = kv.find('=');
            if (kv.empty() || eq == std::string::npos) continue;
            const std::string k = kv.substr(0, eq);
            const long v = std::stol(kv.substr(eq + 1));
            if (k == "concurrent") ens.Config().concurrent_runs = (unsigned)v;
            else if (k == "error_level") ens.Config().error_level = (flamegpu::CUDAEnsemble::EnsembleConfig::ErrorLevel)v;
            else if (k == "batched") ens.Config().batched = v != 0;
            else throw flamegpu::exception::InvalidArgument("unknown ensemble option " + k);
        }
        // sum_agent empty: sum_var names an environment property, logged at exit
        const bool sums = sum_agent && *sum_agent && sum_var && *sum_var;
        const bool env_value = !(sum_agent && *sum_agent) && sum_var && *sum_var;
        std::type_index sum_type = typeid(void);
        if (env_value) {
            auto pit = model.data()->env.props.find(sum_var);
            if (pit == model.data()->env.props.end())
                throw flamegpu::exception::InvalidEnvProperty(std::string("unknown environment property ") + sum_var);
            sum_type = pit->second.type;
            if (sum_type != typeid(uint32_t) && sum_type != typeid(int32_t) && sum_type != typeid(float))
                throw flamegpu::exception::UnsupportedVarType("logged environment values of 4-byte types only");
            flamegpu::LoggingConfig cfg(model);
            cfg.logEnvironment(sum_var);
            ens.setExitLog(cfg);
        }
        if (sums) {
            flamegpu::LoggingConfig cfg(model);
            auto ait = model.data()->agents.find(sum_agent);
            if (ait == model.data()->agents.end()) throw flamegpu::exception::InvalidAgentName(std::string("unknown agent ") + sum_agent);
            auto vit = ait->second->variables.find(sum_var);
            if (vit == ait->second->variables.end()) throw flamegpu::exception::InvalidAgentVar(std::string("unknown variable ") + sum_var);
            sum_type = vit->second.type;
            if (sum_type == typeid(uint32_t)) cfg.agent(sum_agent).logSum<uint32_t>(sum_var);
            else if (sum_type == typeid(int32_t)) cfg.agent(sum_agent).logSum<int32_t>(sum_var);
            else if (sum_type == typeid(float)) cfg.agent(sum_agent).logSum<float>(sum_var);
            else throw flamegpu::exception::UnsupportedVarType("exit sums of 4-byte variables only");
            ens.setExitLog(cfg);
        }
        std::string err;
        try {
            ens.simulate(plans);
        } catch (const flamegpu::exception::EnsembleError& e) {
            err = e.what();
        }
        const auto& recs = ens.getRunRecords();
        const auto& logs = ens.getLogs();
        for (uint32_t i = 0; i < nplans; ++i) {
            const auto& r = recs[i];
            out_status[i] = !r.started ? 2 : (r.error.empty() ? 0 : 1);
            uint32_t k = 0;
            for (auto& a : model.data()->agents) {
                auto it = r.final_counts.find(a.first);
                if (out_counts) out_counts[(size_t)i * model.data()->agents.size() + k] = it == r.final_counts.end() ? 0u : it->second;
                ++k;
            }
            if (out_sums) {
                out_sums[i] = std::nan("");
                auto lt = logs.find(i);
                if (sums && out_status[i] == 0 && lt != logs.end()) {
                    const auto f = lt->second.getExitLog().getAgent(sum_agent);
                    if (sum_type == typeid(uint32_t)) out_sums[i] = (double)f.getSum<uint32_t>(sum_var);
                    else if (sum_type == typeid(int32_t)) out_sums[i] = (double)f.getSum<int32_t>(sum_var);
                    else out_sums[i] = (double)f.getSum<float>(sum_var);
                } else if (env_value && out_status[i] == 0 && lt != logs.end()) {
                    const auto& f = lt->second.getExitLog();
                    if (sum_type == typeid(uint32_t)) out_sums[i] = (double)f.getEnvironmentProperty<uint32_t>(sum_var);
                    else if (sum_type == typeid(int32_t)) out_sums[i] = (double)f.getEnvironmentProperty<int32_t>(sum_var);
                    else out_sums[i] = (double)f.getEnvironmentProperty<float>(sum_var);
                }
            }
        }
        if (out_batched) *out_batched = ens.lastRunBatched() ? 1 : 0;
        if (!err.empty()) throw flamegpu::exception::EnsembleError(err);
    });
}

FGXM_EXPORT int fgxm_stat(fgxm_sim* s, const char* key, uint64_t* out) {
    return fgxm_detail::guard([&] {
        const std::string k = key ? key : "";
        if (k == "fused_layer_launches") *out = s->sim->fusedLayerLaunches();
        else if (k == "layer_images") *out = s->sim->layerImages();
        else throw flamegpu::exception::InvalidArgument("unknown counter " + k);
    });
}

FGXM_EXPORT int fgxm_simulate(fgxm_sim* s, uint32_t steps) {
    return fgxm_detail::guard([&] {
        s->sim->SimulationConfig().steps = steps;
        s->sim->simulate();
    });
}

FGXM_EXPORT int fgxm_log_config(fgxm_sim* s, int exit_log, const char* what, const char* agent, const char* state,
                                const char* name, uint32_t value) {
    return fgxm_detail::guard([&] {
        if (!s->step_log) s->step_log.reset(new flamegpu::StepLoggingConfig(*s->model));
        if (!s->exit_log) s->exit_log.reset(new flamegpu::LoggingConfig(*s->model));
        flamegpu::LoggingConfig& cfg = exit_log ? *s->exit_log : static_cast<flamegpu::LoggingConfig&>(*s->step_log);
        const std::string w = what ? what : "";
        const std::string st = state && *state ? state : flamegpu::ModelData::DEFAULT_STATE();
        if (w == "frequency") {
            if (exit_log) throw flamegpu::exception::InvalidArgument("an exit log has no frequency");
            s->step_log->setFrequency(value);
        } else if (w == "timing") {
            cfg.logTiming(value != 0);
        } else if (w == "env") {
            cfg.logEnvironment(name ? name : "");
        } else if (w == "count") {
            cfg.agent(agent ? agent : "", st).logCount();
        } else {
            auto a = fgxm_detail::agent(s, agent);
            auto vit = a->variables.find(name ? name : "");
            if (vit == a->variables.end()) throw flamegpu::exception::InvalidAgentVar(std::string("unknown variable ") + (name ? name : ""));
            auto alc = cfg.agent(agent, st);
            auto dispatch = [&](auto tag) {
                using T = decltype(tag);
                if (w == "mean") alc.logMean<T>(name);
                else if (w == "standard_deviation") alc.logStandardDev<T>(name);
                else if (w == "min") alc.logMin<T>(name);
                else if (w == "max") alc.logMax<T>(name);
                else if (w == "sum") alc.logSum<T>(name);
                else throw flamegpu::exception::InvalidArgument("unknown log item " + w);
            };
            const std::type_index t = vit->second.type;
            if (t == typeid(float)) dispatch(float());
            else if (t == typeid(double)) dispatch(double());
            else if (t == typeid(int32_t)) dispatch(int32_t());
            else if (t == typeid(uint32_t)) dispatch(uint32_t());
            else if (t == typeid(int64_t)) dispatch(int64_t());
            else if (t == typeid(uint64_t)) dispatch(uint64_t());
            else throw flamegpu::exception::UnsupportedVarType("variable type has no logged reductions");
        }
        if (exit_log) s->sim->setExitLog(*s->exit_log);
        else s->sim->setStepLog(*s->step_log);
    });
}

FGXM_EXPORT int fgxm_export_state(fgxm_sim* s, const char* path, int pretty) {
    return fgxm_detail::guard([&] { s->sim->exportData(path ? path : "", pretty != 0); });
}

FGXM_EXPORT int fgxm_load_state(fgxm_sim* s, const char* path) {
    return fgxm_detail::guard([&] {
        s->sim->SimulationConfig().input_file = path ? path : "";
        s->sim->applyConfig();
    });
}

// Parses a state file against the model without a simulation (host only) and
// describes it as JSON: {"step": N|null, "env": {name: [values]}, "agents":
// {agent: {state: {"n": count, "vars": {name: [sum of element k over agents]}}}}}.
FGXM_EXPORT int fgxm_state_summary(const char* variant, const float* params, uint32_t nparams, const char* path, char* buf,
                                   uint32_t buflen) {
    return fgxm_detail::guard([&] {
        flamegpu::ModelDescription model = fgxm_build_model(variant, params, nparams);
        const auto md = model.data();
        flamegpu::io::StateData d = flamegpu::io::read_state(path ? path : "");
        std::string o = "{\"step\":" + (d.has_step ? std::to_string(d.step_count) : std::string("null")) + ",\"env\":{";
        bool first = true;
        for (auto& kv : d.environment) {
            auto it = md->env.props.find(kv.first);
            if (it == md->env.props.end()) throw flamegpu::exception::InvalidEnvProperty("unknown property " + kv.first);
            std::vector<char> b(it->second.type_size * it->second.elements);
            o += std::string(first ? "" : ",") + "\"" + kv.first + "\":[";
            first = false;
            for (size_t e = 0; e < kv.second.size(); ++e) {
                flamegpu::io::detail::parse_token(it->second.type, kv.second[e], b.data(), e);
                o += (e ? "," : "") + flamegpu::io::detail::format_value(it->second.type, b.data(), e);
            }
            o += "]";
        }
        o += "},\"agents\":{";
        std::string cur;
        first = true;
        for (auto& kv : d.agents) {
            auto a = md->agents.find(kv.first.first);
            if (a == md->agents.end()) throw flamegpu::exception::InvalidAgentName("unknown agent " + kv.first.first);
            if (first || kv.first.first != cur) {
                if (!first) o += "},";
                o += "\"" + kv.first.first + "\":{";
                cur = kv.first.first;
            } else {
                o += ",";
            }
            first = false;
            const std::string st = kv.first.second.empty() ? a->second->initial_state : kv.first.second;
            o += "\"" + st + "\":{\"n\":" + std::to_string(kv.second.size()) + ",\"vars\":{";
            std::map<std::string, std::vector<double>> sums;
            for (auto& rec : kv.second)
                for (auto& v : rec.vars) {
                    auto vit = a->second->variables.find(v.first);
                    if (vit == a->second->variables.end()) throw flamegpu::exception::InvalidAgentVar("unknown variable " + v.first);
                    std::vector<char> b(vit->second.bytes());
                    auto& acc = sums[v.first];
                    acc.resize(vit->second.elements, 0.0);
                    for (size_t e = 0; e < v.second.size(); ++e) {
                        flamegpu::io::detail::parse_token(vit->second.type, v.second[e], b.data(), e);
                        acc[e] += flamegpu::io::detail::as_double(vit->second.type, b.data(), e);
                    }
                }
            bool fv = true;
            for (auto& sv : sums) {
                o += std::string(fv ? "" : ",") + "\"" + sv.first + "\":[";
                fv = false;
                for (size_t e = 0; e < sv.second.size(); ++e) {
                    char t[64];
                    snprintf(t, sizeof t, "%.17g", sv.second[e]);
                    o += (e ? "," : "") + std::string(t);
                }
                o += "]";
            }
            o += "}}";
        }
        if (!first) o += "}";
        o += "}}";
        if (!buf || buflen < o.size() + 1) throw flamegpu::exception::OutOfBoundsException("summary buffer too small");
        std::memcpy(buf, o.c_str(), o.size() + 1);
    });
}

FGXM_EXPORT int fgxm_export_log(fgxm_sim* s, const char* path, int steps, int exit_log, int step_time, int exit_time) {
    return fgxm_detail::guard([&] { s->sim->exportLog(path ? path : "", steps != 0, exit_log != 0, step_time != 0, exit_time != 0); });
}

}  // extern "C"

// ------------------------------------------------------------------ batches
struct fgxm_batch {
    std::string variant;
    std::unique_ptr<flamegpu::ModelDescription> model;
    std::unique_ptr<flamegpu::BatchSimulation> batch;
    std::map<int32_t, flamegpu::BatchSimulation::Plan> staged;   // added, not yet committed
    std::map<int32_t, int> live;                                 // ABI handle -> batch handle
    int32_t next = 1;
    void* ev0 = nullptr;
    void* ev1 = nullptr;
};

namespace fgxm_detail {
// A simulation's plan from its params: the environment buffer and spatial message
// geometry of the model built from them; the structure must match the batch's model.
inline flamegpu::BatchSimulation::Plan plan_from_params(fgxm_batch* b, const float* params, uint32_t nparams) {
    flamegpu::ModelDescription m = fgxm_build_model(b->variant.empty() ? nullptr : b->variant.c_str(), params, nparams);
    const auto& d = *m.data();
    const auto& base = *b->model->data();
    if (d.env.buffer.size() != base.env.buffer.size() || d.agents.size() != base.agents.size() ||
        d.messages.size() != base.messages.size())
        throw flamegpu::exception::InvalidArgument("batch simulation params build a model of a different structure");
    flamegpu::BatchSimulation::Plan p;
    p.env = d.env.buffer;
    for (auto& kv : d.messages) {
        flamegpu::SpatialBounds sb;
        for (int a = 0; a < 3; ++a) {
            sb.min[a] = kv.second->min[a];
            sb.max[a] = kv.second->max[a];
        }
        sb.radius = kv.second->radius;
        p.messages[kv.first] = sb;
    }
    return p;
}
inline void commit(fgxm_batch* b) {
    for (auto& kv : b->staged) b->live[kv.first] = b->batch->add(kv.second);
    b->staged.clear();
}
inline int handle(fgxm_batch* b, int32_t h) {
    commit(b);
    auto it = b->live.find(h);
    if (it == b->live.end()) throw flamegpu::exception::InvalidArgument("unknown batch simulation handle");
    return it->second;
}
inline std::shared_ptr<flamegpu::AgentData> batch_agent(fgxm_batch* b, const char* name) {
    auto& agents = b->model->data()->agents;
    auto it = agents.find(name ? name : "");
    if (it == agents.end()) throw flamegpu::exception::InvalidAgentName(std::string("unknown agent ") + (name ? name : "(null)"));
    return it->second;
}
}  // namespace fgxm_detail

extern "C" {

FGXM_EXPORT int fgxm_batch_create(const char* variant, const float* params, uint32_t nparams, int device, fgxm_batch** out) {
    return fgxm_detail::guard([&] {
        auto* b = new fgxm_batch();
        try {
            b->variant = variant ? variant : "";
            b->model.reset(new flamegpu::ModelDescription(fgxm_build_model(variant, params, nparams)));
            b->batch.reset(new flamegpu::BatchSimulation(*b->model, device));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

FGXM_EXPORT int fgxm_batch_destroy(fgxm_batch* b) {
    return fgxm_detail::guard([&] {
        if (b->ev0) fgx_event_destroy(b->ev0);
        if (b->ev1) fgx_event_destroy(b->ev1);
        delete b;
    });
}

FGXM_EXPORT int fgxm_batch_var_count(fgxm_batch* b, const char* agent, uint32_t* n) {
    return fgxm_detail::guard([&] { *n = (uint32_t)fgxm_detail::batch_agent(b, agent)->declared.size(); });
}

FGXM_EXPORT int fgxm_batch_var_info(fgxm_batch* b, const char* agent, uint32_t v, const char** name, int32_t* type,
                                    uint32_t* elements, uint32_t* bytes) {
    return fgxm_detail::guard([&] { fgxm_detail::var_info(fgxm_detail::batch_agent(b, agent), v, name, type, elements, bytes); });
}

FGXM_EXPORT int fgxm_batch_add(fgxm_batch* b, const float* params, uint32_t nparams, uint64_t seed, uint32_t steps,
                               int32_t* handle) {
    return fgxm_detail::guard([&] {
        auto p = fgxm_detail::plan_from_params(b, params, nparams);
        p.seed = seed;
        p.steps = steps;
        const int32_t h = b->next++;
        b->staged[h] = std::move(p);
        *handle = h;
    });
}

FGXM_EXPORT int fgxm_batch_set_agents(fgxm_batch* b, int32_t h, const char* agent, uint32_t n, const void* const* vars,
                                      const uint32_t* ids) {
    return fgxm_detail::guard([&] {
        auto it = b->staged.find(h);
        if (it == b->staged.end())
            throw flamegpu::exception::InvalidOperation("populations can only be set before the simulation joins the batch");
        auto a = fgxm_detail::batch_agent(b, agent);
        auto pop = std::make_shared<flamegpu::AgentVector>(flamegpu::AgentDescription(a), n);
        for (size_t v = 0; v < a->declared.size(); ++v) {
            auto& raw = pop->raw(a->declared[v]);
            std::memcpy(raw.data(), vars[v], raw.size());
        }
        if (ids) std::memcpy(pop->raw("_id").data(), ids, (size_t)n * 4);
        it->second.populations[a->name] = pop;
    });
}

FGXM_EXPORT int fgxm_batch_estimate_bytes(fgxm_batch* b, int32_t h, uint64_t* bytes) {
    return fgxm_detail::guard([&] {
        auto it = b->staged.find(h);
        if (it == b->staged.end()) throw flamegpu::exception::InvalidArgument("estimate needs a staged simulation");
        *bytes = b->batch->estimateBytes(it->second);
    });
}

FGXM_EXPORT int fgxm_batch_step(fgxm_batch* b, uint32_t nsteps) {
    return fgxm_detail::guard([&] {
        fgxm_detail::commit(b);
        for (uint32_t i = 0; i < nsteps; ++i) b->batch->step();
    });
}

FGXM_EXPORT int fgxm_batch_finished(fgxm_batch* b, int32_t* handles, uint32_t capacity, uint32_t* count) {
    return fgxm_detail::guard([&] {
        const auto f = b->batch->finished();
        uint32_t k = 0;
        for (auto& kv : b->live)
            if (std::find(f.begin(), f.end(), kv.second) != f.end()) {
                if (k < capacity && handles) handles[k] = kv.first;
                ++k;
            }
        *count = k;
    });
}

FGXM_EXPORT int fgxm_batch_steps_done(fgxm_batch* b, int32_t h, uint32_t* steps) {
    return fgxm_detail::guard([&] {
        if (b->staged.count(h)) {
            *steps = 0;
            return;
        }
        *steps = b->batch->stepsDone(fgxm_detail::handle(b, h));
    });
}

FGXM_EXPORT int fgxm_batch_count(fgxm_batch* b, int32_t h, const char* agent, uint32_t* n) {
    return fgxm_detail::guard([&] {
        auto a = fgxm_detail::batch_agent(b, agent);
        flamegpu::AgentVector pop{flamegpu::AgentDescription(a)};
        b->batch->getPopulationData(fgxm_detail::handle(b, h), pop);
        *n = (uint32_t)pop.size();
    });
}

FGXM_EXPORT int fgxm_batch_get_agents(fgxm_batch* b, int32_t h, const char* agent, void* const* vars, uint32_t* ids) {
    return fgxm_detail::guard([&] {
        auto a = fgxm_detail::batch_agent(b, agent);
        flamegpu::AgentVector pop{flamegpu::AgentDescription(a)};
        b->batch->getPopulationData(fgxm_detail::handle(b, h), pop);
        for (size_t v = 0; v < a->declared.size(); ++v) {
            auto& raw = pop.raw(a->declared[v]);
            if (vars[v]) std::memcpy(vars[v], raw.data(), raw.size());
        }
        if (ids) std::memcpy(ids, pop.raw("_id").data(), pop.size() * 4);
    });
}

FGXM_EXPORT int fgxm_batch_remove(fgxm_batch* b, int32_t h) {
    return fgxm_detail::guard([&] {
        if (b->staged.erase(h)) return;
        b->batch->remove(fgxm_detail::handle(b, h));
        b->live.erase(h);
    });
}

FGXM_EXPORT int fgxm_batch_set_tuning(fgxm_batch* b, uint32_t block, uint32_t agents_per_thread, uint32_t blocks_per_cu) {
    return fgxm_detail::guard([&] {
        flamegpu::BatchSimulation::Tuning t;
        t.block = block;
        t.agents_per_thread = agents_per_thread;
        t.blocks_per_cu = blocks_per_cu;
        b->batch->setTuning(t);
    });
}

FGXM_EXPORT int fgxm_batch_set_option(fgxm_batch* b, const char* key, int value) {
    return fgxm_detail::guard([&] {
        const std::string k = key ? key : "";
        if (k == "graphs") b->batch->setGraphs(value != 0);
        else if (k == "unit_tables") b->batch->setUnitTables(value != 0);
        else if (k == "bins_per_radius") b->batch->setBinWidthMultiplier(bins_per_radius_mult(value));
        else throw flamegpu::exception::InvalidArgument("unknown batch option " + k);
    });
}

FGXM_EXPORT int fgxm_batch_info(fgxm_batch* b, uint32_t* resident, uint32_t* staged, uint64_t* resident_agents,
                                uint64_t* resident_bytes) {
    return fgxm_detail::guard([&] {
        if (resident) *resident = b->batch->residentCount();
        if (staged) *staged = (uint32_t)b->staged.size() + b->batch->pendingCount();
        if (resident_agents) *resident_agents = b->batch->residentAgents();
        if (resident_bytes) *resident_bytes = b->batch->residentBytes();
    });
}

FGXM_EXPORT int fgxm_batch_sync(fgxm_batch* b) {
    return fgxm_detail::guard([&] { b->batch->synchronize(); });
}

FGXM_EXPORT int fgxm_batch_timed_steps(fgxm_batch* b, uint32_t steps, double* ms) {
    return fgxm_detail::guard([&] {
        fgxm_detail::commit(b);
        if (!b->ev0) flamegpu::detail::fgx_check(fgx_event_create(&b->ev0), "event");
        if (!b->ev1) flamegpu::detail::fgx_check(fgx_event_create(&b->ev1), "event");
        // two untimed steps first: joiners and leavers repack, and each of the two
        // buffer arrangements a step alternates between is captured as a graph
        b->batch->step();
        b->batch->step();
        void* st = b->batch->stream();
        b->batch->synchronize();
        flamegpu::detail::fgx_check(fgx_event_record(b->ev0, st), "event record");
        for (uint32_t i = 0; i < steps; ++i) b->batch->step();
        flamegpu::detail::fgx_check(fgx_event_record(b->ev1, st), "event record");
        float f = 0.0f;
        flamegpu::detail::fgx_check(fgx_event_elapsed_ms(b->ev0, b->ev1, &f), "event elapsed");
        *ms = f;
    });
}

}  // extern "C"
