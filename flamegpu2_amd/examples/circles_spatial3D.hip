// Circles, spatial-3D messaging: examples/cpp/circles_spatial3D/src/main.cu:5-54,
// with the agent clamped to the [0, ENV_MAX]^3 box after moving.
#include <cfloat>
#include <cstdio>
#include <random>

#include "flamegpu/flamegpu.h"

FLAMEGPU_AGENT_FUNCTION(output_message, flamegpu::MessageNone, flamegpu::MessageSpatial3D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setLocation(
        FLAMEGPU->getVariable<float>("x"),
        FLAMEGPU->getVariable<float>("y"),
        FLAMEGPU->getVariable<float>("z"));
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(move, flamegpu::MessageSpatial3D, flamegpu::MessageNone) {
    const flamegpu::id_t ID = FLAMEGPU->getID();
    const float REPULSE_FACTOR = FLAMEGPU->environment.getProperty<float>("repulse");
    const float ENV_MAX = FLAMEGPU->environment.getProperty<float>("ENV_MAX");
    const float RADIUS = FLAMEGPU->message_in.radius();
    float fx = 0.0;
    float fy = 0.0;
    float fz = 0.0;
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    const float z1 = FLAMEGPU->getVariable<float>("z");
    int count = 0;
    for (const auto &message : FLAMEGPU->message_in(x1, y1, z1)) {
        if (message.getVariable<flamegpu::id_t>("id") != ID) {
            const float x2 = message.getVariable<float>("x");
            const float y2 = message.getVariable<float>("y");
            const float z2 = message.getVariable<float>("z");
            float x21 = x2 - x1;
            float y21 = y2 - y1;
            float z21 = z2 - z1;
            const float separation = sqrtf(x21 * x21 + y21 * y21 + z21 * z21);
            if (separation < RADIUS && separation > 0.0f) {
                float k = sinf((separation / RADIUS) * 3.141f * -2) * REPULSE_FACTOR;
                x21 /= separation;
                y21 /= separation;
                z21 /= separation;
                fx += k * x21;
                fy += k * y21;
                fz += k * z21;
                count++;
            }
        }
    }
    fx /= count > 0 ? count : 1;
    fy /= count > 0 ? count : 1;
    fz /= count > 0 ? count : 1;
    float nx = x1 + fx, ny = y1 + fy, nz = z1 + fz;
    nx = nx < 0.0f ? 0.0f : (nx > ENV_MAX ? ENV_MAX : nx);
    ny = ny < 0.0f ? 0.0f : (ny > ENV_MAX ? ENV_MAX : ny);
    nz = nz < 0.0f ? 0.0f : (nz > ENV_MAX ? ENV_MAX : nz);
    FLAMEGPU->setVariable<float>("x", nx);
    FLAMEGPU->setVariable<float>("y", ny);
    FLAMEGPU->setVariable<float>("z", nz);
    FLAMEGPU->setVariable<float>("drift", sqrtf(fx * fx + fy * fy + fz * fz));
    return flamegpu::ALIVE;
}

// params: {ENV_MAX, radius, repulse}
flamegpu::ModelDescription fgxm_build_model(const char*, const float* p, uint32_t n) {
    const float ENV_MAX = n > 0 ? p[0] : 25.0f;
    const float RADIUS = n > 1 ? p[1] : 2.0f;
    const float REPULSE = n > 2 ? p[2] : 0.05f;
    flamegpu::ModelDescription model("Circles Spatial3D");
    {
        flamegpu::MessageSpatial3D::Description message = model.newMessage<flamegpu::MessageSpatial3D>("location");
        message.newVariable<flamegpu::id_t>("id");
        message.setRadius(RADIUS);
        message.setMin(0, 0, 0);
        message.setMax(ENV_MAX, ENV_MAX, ENV_MAX);
    }
    {
        flamegpu::AgentDescription agent = model.newAgent("Circle");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("z");
        agent.newVariable<float>("drift");
        agent.newFunction("output_message", output_message).setMessageOutput("location");
        agent.newFunction("move", move).setMessageInput("location");
    }
    {
        flamegpu::EnvironmentDescription env = model.Environment();
        env.newProperty("repulse", REPULSE);
        env.newProperty("ENV_MAX", ENV_MAX);
    }
    {
        flamegpu::LayerDescription layer = model.newLayer();
        layer.addAgentFunction(output_message);
    }
    {
        flamegpu::LayerDescription layer = model.newLayer();
        layer.addAgentFunction(move);
    }
    return model;
}

#ifdef FGX_MODEL_ABI
#include "flamegpu/model_abi.h"
#else
FLAMEGPU_STEP_FUNCTION(Validation) {
    static float prevTotalDrift = FLT_MAX;
    static unsigned int driftDropped = 0;
    static unsigned int driftIncreased = 0;
    float totalDrift = FLAMEGPU->agent("Circle").sum<float>("drift");
    if (totalDrift <= prevTotalDrift)
        driftDropped++;
    else
        driftIncreased++;
    prevTotalDrift = totalDrift;
    printf("%.2f%% Drift correct\n", 100 * driftDropped / static_cast<float>(driftDropped + driftIncreased));
}

int main(int argc, const char** argv) {
    const unsigned int AGENT_COUNT = 16384;
    const float ENV_MAX = static_cast<float>(floor(cbrt(AGENT_COUNT)));
    const float params[3] = {ENV_MAX, 2.0f, 0.05f};
    flamegpu::ModelDescription model = fgxm_build_model(nullptr, params, 3);
    model.addStepFunction(Validation);
    flamegpu::CUDASimulation cudaSimulation(model, argc, argv);
    if (cudaSimulation.getSimulationConfig().input_file.empty()) {
        std::mt19937_64 rng;
        std::uniform_real_distribution<float> dist(0.0f, ENV_MAX);
        flamegpu::AgentVector population(model.Agent("Circle"), AGENT_COUNT);
        for (unsigned int i = 0; i < AGENT_COUNT; i++) {
            flamegpu::AgentVector::Agent instance = population[i];
            instance.setVariable<float>("x", dist(rng));
            instance.setVariable<float>("y", dist(rng));
            instance.setVariable<float>("z", dist(rng));
        }
        cudaSimulation.setPopulationData(population);
    }
    cudaSimulation.simulate();
    flamegpu::util::cleanup();
    return 0;
}
#endif
