// Predator-prey, spatial-2D messaging with agent birth and death.
// Five layers per step:
//   [prey_output, pred_output] [prey_flee, pred_hunt] [prey_move, pred_move]
//   [prey_reproduce, pred_reproduce] [pred_starve]
// Prey die when a predator is within EAT_RADIUS or with probability P_PREY_DEATH
// and reproduce with P_PREY_REPRO (child at the parent +-JITTER/2); predators
// that ate this step reproduce with P_PRED_REPRO and starve after STARVE_STEPS
// steps without food.  Positions wrap in [0, WIDTH)^2.
#include <cstdio>
#include <random>

#include "flamegpu/flamegpu.h"

FLAMEGPU_DEVICE_FUNCTION float wrap_coord(float v, float w) {
    if (v < 0.0f) v += w;
    if (v >= w) v -= w;
    return v;
}

FLAMEGPU_AGENT_FUNCTION(prey_output, flamegpu::MessageNone, flamegpu::MessageSpatial2D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setLocation(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"));
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(prey_flee, flamegpu::MessageSpatial2D, flamegpu::MessageNone) {
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    const float R = FLAMEGPU->message_in.radius();
    const float EAT_R = FLAMEGPU->environment.getProperty<float>("EAT_RADIUS");
    float fx = 0.0f, fy = 0.0f;
    int eaten = 0;
    for (const auto& m : FLAMEGPU->message_in(x1, y1)) {
        const float dx = x1 - m.getVariable<float>("x");
        const float dy = y1 - m.getVariable<float>("y");
        const float d = sqrtf(dx * dx + dy * dy);
        if (d < R && d > 0.0f) {
            fx += dx / d;
            fy += dy / d;
        }
        if (d < EAT_R) eaten = 1;
    }
    const float FLEE = FLAMEGPU->environment.getProperty<float>("FLEE_SCALE");
    float vx = FLAMEGPU->getVariable<float>("vx") + fx * FLEE;
    float vy = FLAMEGPU->getVariable<float>("vy") + fy * FLEE;
    const float s = sqrtf(vx * vx + vy * vy);
    if (s > 1.0f) {
        vx /= s;
        vy /= s;
    }
    FLAMEGPU->setVariable<float>("vx", vx);
    FLAMEGPU->setVariable<float>("vy", vy);
    if (eaten) return flamegpu::DEAD;
    if (FLAMEGPU->random.uniform<float>() < FLAMEGPU->environment.getProperty<float>("P_PREY_DEATH")) return flamegpu::DEAD;
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(prey_move, flamegpu::MessageNone, flamegpu::MessageNone) {
    const float W = FLAMEGPU->environment.getProperty<float>("WIDTH");
    FLAMEGPU->setVariable<float>("x", wrap_coord(FLAMEGPU->getVariable<float>("x") + FLAMEGPU->getVariable<float>("vx"), W));
    FLAMEGPU->setVariable<float>("y", wrap_coord(FLAMEGPU->getVariable<float>("y") + FLAMEGPU->getVariable<float>("vy"), W));
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(prey_reproduce, flamegpu::MessageNone, flamegpu::MessageNone) {
    if (FLAMEGPU->random.uniform<float>() < FLAMEGPU->environment.getProperty<float>("P_PREY_REPRO")) {
        const float W = FLAMEGPU->environment.getProperty<float>("WIDTH");
        const float J = FLAMEGPU->environment.getProperty<float>("JITTER");
        const float jx = FLAMEGPU->random.uniform<float>();
        const float jy = FLAMEGPU->random.uniform<float>();
        const float rvx = FLAMEGPU->random.uniform<float>();
        const float rvy = FLAMEGPU->random.uniform<float>();
        FLAMEGPU->agent_out.setVariable<float>("x", wrap_coord(FLAMEGPU->getVariable<float>("x") + (jx - 0.5f) * J, W));
        FLAMEGPU->agent_out.setVariable<float>("y", wrap_coord(FLAMEGPU->getVariable<float>("y") + (jy - 0.5f) * J, W));
        FLAMEGPU->agent_out.setVariable<float>("vx", rvx * 2.0f - 1.0f);
        FLAMEGPU->agent_out.setVariable<float>("vy", rvy * 2.0f - 1.0f);
    }
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(pred_output, flamegpu::MessageNone, flamegpu::MessageSpatial2D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setLocation(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"));
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(pred_hunt, flamegpu::MessageSpatial2D, flamegpu::MessageNone) {
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    float best = FLAMEGPU->message_in.radius();
    float bx = 0.0f, by = 0.0f;
    int found = 0;
    for (const auto& m : FLAMEGPU->message_in(x1, y1)) {
        const float dx = m.getVariable<float>("x") - x1;
        const float dy = m.getVariable<float>("y") - y1;
        const float d = sqrtf(dx * dx + dy * dy);
        if (d < best) {
            best = d;
            bx = dx;
            by = dy;
            found = 1;
        }
    }
    const float HUNT = FLAMEGPU->environment.getProperty<float>("HUNT_SCALE");
    float vx = FLAMEGPU->getVariable<float>("vx");
    float vy = FLAMEGPU->getVariable<float>("vy");
    if (found && best > 0.0f) {
        vx += bx / best * HUNT;
        vy += by / best * HUNT;
    }
    const float s = sqrtf(vx * vx + vy * vy);
    if (s > 1.0f) {
        vx /= s;
        vy /= s;
    }
    FLAMEGPU->setVariable<float>("vx", vx);
    FLAMEGPU->setVariable<float>("vy", vy);
    const float EAT_R = FLAMEGPU->environment.getProperty<float>("EAT_RADIUS");
    const unsigned int life = FLAMEGPU->getVariable<unsigned int>("life");
    FLAMEGPU->setVariable<unsigned int>("life", (found && best < EAT_R) ? 0u : life + 1u);
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(pred_move, flamegpu::MessageNone, flamegpu::MessageNone) {
    const float W = FLAMEGPU->environment.getProperty<float>("WIDTH");
    FLAMEGPU->setVariable<float>("x", wrap_coord(FLAMEGPU->getVariable<float>("x") + FLAMEGPU->getVariable<float>("vx"), W));
    FLAMEGPU->setVariable<float>("y", wrap_coord(FLAMEGPU->getVariable<float>("y") + FLAMEGPU->getVariable<float>("vy"), W));
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(pred_reproduce, flamegpu::MessageNone, flamegpu::MessageNone) {
    const float r = FLAMEGPU->random.uniform<float>();
    if (FLAMEGPU->getVariable<unsigned int>("life") == 0u && r < FLAMEGPU->environment.getProperty<float>("P_PRED_REPRO")) {
        const float W = FLAMEGPU->environment.getProperty<float>("WIDTH");
        const float J = FLAMEGPU->environment.getProperty<float>("JITTER");
        const float jx = FLAMEGPU->random.uniform<float>();
        const float jy = FLAMEGPU->random.uniform<float>();
        const float rvx = FLAMEGPU->random.uniform<float>();
        const float rvy = FLAMEGPU->random.uniform<float>();
        FLAMEGPU->agent_out.setVariable<float>("x", wrap_coord(FLAMEGPU->getVariable<float>("x") + (jx - 0.5f) * J, W));
        FLAMEGPU->agent_out.setVariable<float>("y", wrap_coord(FLAMEGPU->getVariable<float>("y") + (jy - 0.5f) * J, W));
        FLAMEGPU->agent_out.setVariable<float>("vx", rvx * 2.0f - 1.0f);
        FLAMEGPU->agent_out.setVariable<float>("vy", rvy * 2.0f - 1.0f);
        FLAMEGPU->agent_out.setVariable<unsigned int>("life", 0u);
    }
    return flamegpu::ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(pred_starve, flamegpu::MessageNone, flamegpu::MessageNone) {
    return FLAMEGPU->getVariable<unsigned int>("life") >= FLAMEGPU->environment.getProperty<unsigned int>("STARVE_STEPS")
               ? flamegpu::DEAD
               : flamegpu::ALIVE;
}

// params: {WIDTH, RADIUS, EAT_RADIUS, P_PREY_DEATH, P_PREY_REPRO, P_PRED_REPRO, STARVE_STEPS, JITTER,
//          FLEE_SCALE, HUNT_SCALE}
flamegpu::ModelDescription fgxm_build_model(const char*, const float* p, uint32_t n) {
    float v[10] = {1000.0f, 5.0f, 1.0f, 0.005f, 0.02f, 0.01f, 50.0f, 1.0f, 0.5f, 0.5f};
    for (uint32_t i = 0; i < n && i < 10; ++i) v[i] = p[i];
    flamegpu::ModelDescription model("Predator Prey Spatial2D");
    flamegpu::EnvironmentDescription env = model.Environment();
    env.newProperty("WIDTH", v[0]);
    env.newProperty("EAT_RADIUS", v[2]);
    env.newProperty("P_PREY_DEATH", v[3]);
    env.newProperty("P_PREY_REPRO", v[4]);
    env.newProperty("P_PRED_REPRO", v[5]);
    env.newProperty("STARVE_STEPS", static_cast<unsigned int>(v[6]));
    env.newProperty("JITTER", v[7]);
    env.newProperty("FLEE_SCALE", v[8]);
    env.newProperty("HUNT_SCALE", v[9]);
    for (const char* name : {"prey_location", "predator_location"}) {
        auto m = model.newMessage<flamegpu::MessageSpatial2D>(name);
        m.newVariable<flamegpu::id_t>("id");
        m.setRadius(v[1]);
        m.setMin(0.0f, 0.0f);
        m.setMax(v[0], v[0]);
    }
    flamegpu::AgentDescription prey = model.newAgent("Prey");
    prey.newVariable<float>("x");
    prey.newVariable<float>("y");
    prey.newVariable<float>("vx");
    prey.newVariable<float>("vy");
    flamegpu::AgentDescription pred = model.newAgent("Predator");
    pred.newVariable<float>("x");
    pred.newVariable<float>("y");
    pred.newVariable<float>("vx");
    pred.newVariable<float>("vy");
    pred.newVariable<unsigned int>("life", 0u);
    // Definition order fixes each function's RNG stream.
    auto f_prey_output = prey.newFunction("prey_output", prey_output);
    f_prey_output.setMessageOutput("prey_location");
    auto f_prey_flee = prey.newFunction("prey_flee", prey_flee);
    f_prey_flee.setMessageInput("predator_location");
    f_prey_flee.setAllowAgentDeath(true);
    auto f_prey_move = prey.newFunction("prey_move", prey_move);
    auto f_prey_repro = prey.newFunction("prey_reproduce", prey_reproduce);
    f_prey_repro.setAgentOutput(prey);
    auto f_pred_output = pred.newFunction("pred_output", pred_output);
    f_pred_output.setMessageOutput("predator_location");
    auto f_pred_hunt = pred.newFunction("pred_hunt", pred_hunt);
    f_pred_hunt.setMessageInput("prey_location");
    auto f_pred_move = pred.newFunction("pred_move", pred_move);
    auto f_pred_repro = pred.newFunction("pred_reproduce", pred_reproduce);
    f_pred_repro.setAgentOutput(pred);
    auto f_pred_starve = pred.newFunction("pred_starve", pred_starve);
    f_pred_starve.setAllowAgentDeath(true);
    {
        auto l = model.newLayer("output");
        l.addAgentFunction(f_prey_output);
        l.addAgentFunction(f_pred_output);
    }
    {
        auto l = model.newLayer("interact");
        l.addAgentFunction(f_prey_flee);
        l.addAgentFunction(f_pred_hunt);
    }
    {
        auto l = model.newLayer("move");
        l.addAgentFunction(f_prey_move);
        l.addAgentFunction(f_pred_move);
    }
    {
        auto l = model.newLayer("reproduce");
        l.addAgentFunction(f_prey_repro);
        l.addAgentFunction(f_pred_repro);
    }
    model.newLayer("starve").addAgentFunction(f_pred_starve);
    return model;
}

#ifdef FGX_MODEL_ABI
#include "flamegpu/model_abi.h"
#else
int main(int argc, const char** argv) {
    flamegpu::ModelDescription model = fgxm_build_model(nullptr, nullptr, 0);
    flamegpu::CUDASimulation sim(model, argc, argv);
    std::mt19937_64 rng(sim.getSimulationConfig().random_seed);
    std::uniform_real_distribution<float> pos(0.0f, 1000.0f), vel(-1.0f, 1.0f);
    for (auto& spec : {std::make_pair("Prey", 100000u), std::make_pair("Predator", 25000u)}) {
        flamegpu::AgentVector pop(model.Agent(spec.first), spec.second);
        for (auto a : pop) {
            a.setVariable<float>("x", pos(rng));
            a.setVariable<float>("y", pos(rng));
            a.setVariable<float>("vx", vel(rng));
            a.setVariable<float>("vy", vel(rng));
        }
        sim.setPopulationData(pop);
    }
    sim.simulate();
    flamegpu::HostAPI api(sim, 0);
    printf("prey %u predators %u\n", api.agent("Prey").count(), api.agent("Predator").count());
    return 0;
}
#endif
