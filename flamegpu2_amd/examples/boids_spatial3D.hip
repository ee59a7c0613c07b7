// Boids, spatial-3D messaging with wrap-around boundaries:
// examples/python_native/boids_spatial3D_wrapped/boids_spatial3D.py:79-240 written against
// the C++ API (wrap() iterator, virtual coordinates, position wrap).  The speed
// bounds the reference hard-codes (1.0 / 0.5) are the MAX_SPEED / MIN_SPEED
// environment properties, and the visualisation-only wing variables are dropped.
#include <cstdio>
#include <random>

#include "flamegpu/flamegpu.h"

FLAMEGPU_HOST_DEVICE_FUNCTION float vec3Length(const float x, const float y, const float z) {
    return sqrtf(x * x + y * y + z * z);
}

FLAMEGPU_AGENT_FUNCTION(outputdata, flamegpu::MessageNone, flamegpu::MessageSpatial3D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setVariable<float>("x", FLAMEGPU->getVariable<float>("x"));
    FLAMEGPU->message_out.setVariable<float>("y", FLAMEGPU->getVariable<float>("y"));
    FLAMEGPU->message_out.setVariable<float>("z", FLAMEGPU->getVariable<float>("z"));
    FLAMEGPU->message_out.setVariable<float>("fx", FLAMEGPU->getVariable<float>("fx"));
    FLAMEGPU->message_out.setVariable<float>("fy", FLAMEGPU->getVariable<float>("fy"));
    FLAMEGPU->message_out.setVariable<float>("fz", FLAMEGPU->getVariable<float>("fz"));
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(inputdata, flamegpu::MessageSpatial3D, flamegpu::MessageNone) {
    const flamegpu::id_t id = FLAMEGPU->getID();
    float agent_x = FLAMEGPU->getVariable<float>("x");
    float agent_y = FLAMEGPU->getVariable<float>("y");
    float agent_z = FLAMEGPU->getVariable<float>("z");
    float agent_fx = FLAMEGPU->getVariable<float>("fx");
    float agent_fy = FLAMEGPU->getVariable<float>("fy");
    float agent_fz = FLAMEGPU->getVariable<float>("fz");

    float perceived_centre_x = 0.0f;
    float perceived_centre_y = 0.0f;
    float perceived_centre_z = 0.0f;
    int perceived_count = 0;
    float global_velocity_x = 0.0f;
    float global_velocity_y = 0.0f;
    float global_velocity_z = 0.0f;
    float velocity_change_x = 0.f;
    float velocity_change_y = 0.f;
    float velocity_change_z = 0.f;

    const float INTERACTION_RADIUS = FLAMEGPU->environment.getProperty<float>("INTERACTION_RADIUS");
    const float SEPARATION_RADIUS = FLAMEGPU->environment.getProperty<float>("SEPARATION_RADIUS");
    const float COLLISION_SCALE = FLAMEGPU->environment.getProperty<float>("COLLISION_SCALE");
    for (const auto &message : FLAMEGPU->message_in.wrap(agent_x, agent_y, agent_z)) {
        if (message.getVariable<flamegpu::id_t>("id") != id) {
            const float message_x = message.getVirtualX();
            const float message_y = message.getVirtualY();
            const float message_z = message.getVirtualZ();
            const float separation = vec3Length(agent_x - message_x, agent_y - message_y, agent_z - message_z);
            if (separation < INTERACTION_RADIUS) {
                perceived_centre_x += message_x;
                perceived_centre_y += message_y;
                perceived_centre_z += message_z;
                perceived_count++;
                global_velocity_x += message.getVariable<float>("fx");
                global_velocity_y += message.getVariable<float>("fy");
                global_velocity_z += message.getVariable<float>("fz");
                if (separation < SEPARATION_RADIUS) {
                    const float normalizedSeparation = (separation / SEPARATION_RADIUS);
                    const float invNormSep = (1.0f - normalizedSeparation);
                    const float invSqSep = invNormSep * invNormSep;
                    velocity_change_x += COLLISION_SCALE * (agent_x - message_x) * invSqSep;
                    velocity_change_y += COLLISION_SCALE * (agent_y - message_y) * invSqSep;
                    velocity_change_z += COLLISION_SCALE * (agent_z - message_z) * invSqSep;
                }
            }
        }
    }
    if (perceived_count) {
        const float c = static_cast<float>(perceived_count);
        perceived_centre_x /= c;
        perceived_centre_y /= c;
        perceived_centre_z /= c;
        global_velocity_x /= c;
        global_velocity_y /= c;
        global_velocity_z /= c;
        const float STEER_SCALE = FLAMEGPU->environment.getProperty<float>("STEER_SCALE");
        velocity_change_x += (perceived_centre_x - agent_x) * STEER_SCALE;
        velocity_change_y += (perceived_centre_y - agent_y) * STEER_SCALE;
        velocity_change_z += (perceived_centre_z - agent_z) * STEER_SCALE;
        const float MATCH_SCALE = FLAMEGPU->environment.getProperty<float>("MATCH_SCALE");
        velocity_change_x += global_velocity_x * MATCH_SCALE - agent_fx;
        velocity_change_y += global_velocity_y * MATCH_SCALE - agent_fy;
        velocity_change_z += global_velocity_z * MATCH_SCALE - agent_fz;
    }
    const float GLOBAL_SCALE = FLAMEGPU->environment.getProperty<float>("GLOBAL_SCALE");
    velocity_change_x *= GLOBAL_SCALE;
    velocity_change_y *= GLOBAL_SCALE;
    velocity_change_z *= GLOBAL_SCALE;
    agent_fx += velocity_change_x;
    agent_fy += velocity_change_y;
    agent_fz += velocity_change_z;
    const float agent_fscale = vec3Length(agent_fx, agent_fy, agent_fz);
    const float MAX_SPEED = FLAMEGPU->environment.getProperty<float>("MAX_SPEED");
    const float MIN_SPEED = FLAMEGPU->environment.getProperty<float>("MIN_SPEED");
    if (agent_fscale > MAX_SPEED) {
        agent_fx /= agent_fscale;
        agent_fy /= agent_fscale;
        agent_fz /= agent_fscale;
        agent_fx *= MAX_SPEED;
        agent_fy *= MAX_SPEED;
        agent_fz *= MAX_SPEED;
    }
    if (agent_fscale < MIN_SPEED) {
        agent_fx /= agent_fscale;
        agent_fy /= agent_fscale;
        agent_fz /= agent_fscale;
        agent_fx *= MIN_SPEED;
        agent_fy *= MIN_SPEED;
        agent_fz *= MIN_SPEED;
    }
    const float TIME_SCALE = FLAMEGPU->environment.getProperty<float>("TIME_SCALE");
    agent_x += agent_fx * TIME_SCALE;
    agent_y += agent_fy * TIME_SCALE;
    agent_z += agent_fz * TIME_SCALE;
    const float minPosition = FLAMEGPU->environment.getProperty<float>("MIN_POSITION");
    const float maxPosition = FLAMEGPU->environment.getProperty<float>("MAX_POSITION");
    const float width = maxPosition - minPosition;
    if (agent_x < minPosition) agent_x += width;
    if (agent_y < minPosition) agent_y += width;
    if (agent_z < minPosition) agent_z += width;
    if (agent_x > maxPosition) agent_x -= width;
    if (agent_y > maxPosition) agent_y -= width;
    if (agent_z > maxPosition) agent_z -= width;
    FLAMEGPU->setVariable<float>("x", agent_x);
    FLAMEGPU->setVariable<float>("y", agent_y);
    FLAMEGPU->setVariable<float>("z", agent_z);
    FLAMEGPU->setVariable<float>("fx", agent_fx);
    FLAMEGPU->setVariable<float>("fy", agent_fy);
    FLAMEGPU->setVariable<float>("fz", agent_fz);
    return flamegpu::ALIVE;
}

static const char* const kBoidsParams[] = {"MIN_POSITION", "MAX_POSITION", "INTERACTION_RADIUS", "SEPARATION_RADIUS",
                                           "TIME_SCALE", "GLOBAL_SCALE", "STEER_SCALE", "COLLISION_SCALE",
                                           "MATCH_SCALE", "MIN_SPEED", "MAX_SPEED"};

// params: the 11 environment properties above, in that order.
flamegpu::ModelDescription fgxm_build_model(const char*, const float* p, uint32_t n) {
    // Defaults: the wrapped reference example (boids_spatial3D.py:256-282).
    float v[11] = {-0.5f, 0.5f, 0.05f, 0.01f, 0.001f, 0.25f, 0.055f, 10.0f, 0.015f, 0.5f, 1.0f};
    for (uint32_t i = 0; i < n && i < 11; ++i) v[i] = p[i];
    flamegpu::ModelDescription model("Boids Spatial3D (wrapped)");
    flamegpu::EnvironmentDescription env = model.Environment();
    for (int i = 0; i < 11; ++i) env.newProperty(kBoidsParams[i], v[i]);
    {
        flamegpu::MessageSpatial3D::Description message = model.newMessage<flamegpu::MessageSpatial3D>("location");
        message.setRadius(env.getProperty<float>("INTERACTION_RADIUS"));
        message.setMin(v[0], v[0], v[0]);
        message.setMax(v[1], v[1], v[1]);
        message.newVariable<flamegpu::id_t>("id");
        message.newVariable<float>("fx");
        message.newVariable<float>("fy");
        message.newVariable<float>("fz");
    }
    flamegpu::AgentDescription agent = model.newAgent("Boid");
    agent.newVariable<float>("x");
    agent.newVariable<float>("y");
    agent.newVariable<float>("z");
    agent.newVariable<float>("fx");
    agent.newVariable<float>("fy");
    agent.newVariable<float>("fz");
    flamegpu::AgentFunctionDescription outputdataDescription = agent.newFunction("outputdata", outputdata);
    outputdataDescription.setMessageOutput("location");
    flamegpu::AgentFunctionDescription inputdataDescription = agent.newFunction("inputdata", inputdata);
    inputdataDescription.setMessageInput("location");
    inputdataDescription.dependsOn(outputdataDescription);
    model.addExecutionRoot(outputdataDescription);
    model.generateLayers();
    return model;
}

#ifdef FGX_MODEL_ABI
#include "flamegpu/model_abi.h"
#else
int main(int argc, const char** argv) {
    flamegpu::ModelDescription model = fgxm_build_model(nullptr, nullptr, 0);
    flamegpu::CUDASimulation cudaSimulation(model, argc, argv);
    flamegpu::EnvironmentDescription env = model.Environment();
    std::mt19937_64 rngEngine(cudaSimulation.getSimulationConfig().random_seed);
    std::uniform_real_distribution<float> position_distribution(env.getProperty<float>("MIN_POSITION"),
                                                                env.getProperty<float>("MAX_POSITION"));
    std::uniform_real_distribution<float> velocity_distribution(-1, 1);
    std::uniform_real_distribution<float> velocity_magnitude_distribution(0.1f, 1.0f);
    const unsigned int populationSize = 40000;
    flamegpu::AgentVector population(model.Agent("Boid"), populationSize);
    for (unsigned int i = 0; i < populationSize; i++) {
        flamegpu::AgentVector::Agent instance = population[i];
        instance.setVariable<float>("x", position_distribution(rngEngine));
        instance.setVariable<float>("y", position_distribution(rngEngine));
        instance.setVariable<float>("z", position_distribution(rngEngine));
        float fx = velocity_distribution(rngEngine);
        float fy = velocity_distribution(rngEngine);
        float fz = velocity_distribution(rngEngine);
        const float len = vec3Length(fx, fy, fz);
        const float mag = velocity_magnitude_distribution(rngEngine);
        instance.setVariable<float>("fx", fx / len * mag);
        instance.setVariable<float>("fy", fy / len * mag);
        instance.setVariable<float>("fz", fz / len * mag);
    }
    cudaSimulation.setPopulationData(population);
    cudaSimulation.simulate();
    flamegpu::util::cleanup();
    return 0;
}
#endif
