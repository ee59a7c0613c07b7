// Host agent reductions (HostAgentAPI::count / sum / min / max / meanStandardDeviation,
// include/flamegpu/runtime/agent/HostAgentAPI.cuh; tests/test_cases/runtime/agent/host_reduction/)
// called from a step function, which stores the results in environment properties
// (read back through the exit log).  Variant "reductions".
#include <string>

#include "flamegpu/flamegpu.h"

using flamegpu::ALIVE;
using flamegpu::MessageNone;

FLAMEGPU_AGENT_FUNCTION(host_noop, MessageNone, MessageNone) { return ALIVE; }

template <typename T>
static void reduce_into(flamegpu::HostAPI* FLAMEGPU, const char* var, const std::string& tag) {
    auto a = FLAMEGPU->agent("agent");
    FLAMEGPU->environment.setProperty<double>("min_" + tag, static_cast<double>(a.min<T>(var)));
    FLAMEGPU->environment.setProperty<double>("max_" + tag, static_cast<double>(a.max<T>(var)));
    FLAMEGPU->environment.setProperty<double>("sum_" + tag, static_cast<double>(a.sum<T, typename flamegpu::sum_input_t<T>::result_t>(var)));
    const auto ms = a.meanStandardDeviation<T>(var);
    FLAMEGPU->environment.setProperty<double>("mean_" + tag, ms.first);
    FLAMEGPU->environment.setProperty<double>("sd_" + tag, ms.second);
}

FLAMEGPU_STEP_FUNCTION(reductions_step) {
    reduce_into<float>(FLAMEGPU, "f", "f");
    reduce_into<double>(FLAMEGPU, "d", "d");
    reduce_into<int32_t>(FLAMEGPU, "i", "i");
    reduce_into<uint32_t>(FLAMEGPU, "u", "u");
    reduce_into<int64_t>(FLAMEGPU, "l", "l");
    reduce_into<uint64_t>(FLAMEGPU, "ul", "ul");
    FLAMEGPU->environment.setProperty<unsigned int>("count", FLAMEGPU->agent("agent").count());
}

flamegpu::ModelDescription fgxm_build_model(const char* variant_c, const float*, uint32_t) {
    const std::string v = variant_c ? variant_c : "";
    flamegpu::ModelDescription model("kat_host_" + v);
    if (v != "reductions") throw flamegpu::exception::InvalidArgument("unknown test model variant '" + v + "'");
    auto a = model.newAgent("agent");
    a.newVariable<float>("f");
    a.newVariable<double>("d");
    a.newVariable<int32_t>("i");
    a.newVariable<uint32_t>("u");
    a.newVariable<int64_t>("l");
    a.newVariable<uint64_t>("ul");
    model.newLayer().addAgentFunction(a.newFunction("noop", host_noop));
    model.addStepFunction(reductions_step);
    auto env = model.Environment();
    for (const char* t : {"f", "d", "i", "u", "l", "ul"})
        for (const char* r : {"min_", "max_", "sum_", "mean_", "sd_"}) env.newProperty<double>(std::string(r) + t, 0.0);
    env.newProperty<unsigned int>("count", 0u);
    return model;
}

#include "flamegpu/model_abi.h"
