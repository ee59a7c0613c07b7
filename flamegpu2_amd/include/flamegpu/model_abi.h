// Implementation of include/fgx_model.h for one model translation unit.
// Include once, after defining fgxm_build_model().
#pragma once

#include <cstring>
#include <memory>
#include <sstream>
#include <string>

#include "fgx_model.h"
#include "flamegpu/flamegpu.h"

flamegpu::ModelDescription fgxm_build_model(const char* variant, const float* params, uint32_t nparams);

struct fgxm_sim {
    std::unique_ptr<flamegpu::ModelDescription> model;
    std::unique_ptr<flamegpu::CUDASimulation> sim;
    void* ev0 = nullptr;
    void* ev1 = nullptr;
};

namespace fgxm_detail {
inline thread_local std::string last_error;

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
    return fgxm_detail::guard([&] {
        auto a = fgxm_detail::agent(s, agent);
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
    });
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

FGXM_EXPORT int fgxm_set_option(fgxm_sim* s, const char* key, int value) {
    return fgxm_detail::guard([&] {
        const std::string k = key ? key : "";
        if (k == "graphs") s->sim->CUDAConfig().graphs = value != 0;
        else if (k == "concurrency") s->sim->CUDAConfig().inLayerConcurrency = value != 0;
        else if (k == "kernel_timing") s->sim->setKernelTiming(value != 0);
        else if (k == "staging") s->sim->setNeighbourhoodStaging(value != 0);
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

}  // extern "C"
