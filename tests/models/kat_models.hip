// Known-answer test models, restated from the reference's own GoogleTest suite and
// written against the drop-in C++ API.  One library, one model per `variant`:
//   mandatory3d / optional3d / optionalnone3d / mandatory2d
//       tests/test_cases/runtime/messaging/test_spatial_3d.cu:72-447, test_spatial_2d.cu:66-320
//   wrapped3d / wrapped2d (params: x, y, z offset)
//       test_spatial_3d.cu:855-975, test_spatial_2d.cu:884-955
//   bounds3d            test_spatial_3d.cu:1100-1189
//   sort3d (params: sort period)
//       tests/test_cases/runtime/agent/detail/test_spatial_agent_sort.cu:23-115
//   birth_mandatory / birth_optional / birth_mandatory_death / birth_optional_death
//       tests/test_cases/runtime/agent/test_device_agent_creation.cu:23-300
//   transition          tests/test_cases/runtime/agent/detail/test_agent_state_transition.cu:67-174
//   transition_src0_dest10 / transition_src10_dest0 / transition_src10_dest10
//                       test_agent_state_transition.cu:67-250
//   concurrency         tests/test_cases/simulation/test_cuda_simulation_concurrency.cu:322-390
//   two_radii           one list sorted for two spatial inputs of different radii
//   ensemble_simulate / ensemble_errors
//                       tests/test_cases/simulation/test_cuda_ensemble.cu:251-322, 630-748
//   bruteforce          tests/test_cases/runtime/messaging/test_brute_force.cu:73-200
//   random              tests/test_cases/runtime/random/test_agent_random.cu:28-185
//   death_disabled / wrap_out_of_bounds   device errors (test_spatial_3d.cu:976-1039)
//   agent_id_states / agent_id_two_types  in-kernel child IDs (AgentOut::getID)
//       tests/test_cases/runtime/agent/test_device_agent_creation.cu:1195-1340
#include <string>

#include "flamegpu/flamegpu.h"

using flamegpu::ALIVE;
using flamegpu::DEAD;
using flamegpu::MessageBruteForce;
using flamegpu::MessageNone;
using flamegpu::MessageSpatial2D;
using flamegpu::MessageSpatial3D;

// ---------------------------------------------------------------- spatial 3-D
FLAMEGPU_AGENT_FUNCTION(out_mandatory3D, MessageNone, MessageSpatial3D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setLocation(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"),
                                      FLAMEGPU->getVariable<float>("z"));
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(out_optional3D, MessageNone, MessageSpatial3D) {
    if (FLAMEGPU->getVariable<int>("do_output")) {
        FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
        FLAMEGPU->message_out.setLocation(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"),
                                          FLAMEGPU->getVariable<float>("z"));
    }
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(out_optional3DNone, MessageNone, MessageSpatial3D) { return ALIVE; }
FLAMEGPU_AGENT_FUNCTION(in3D, MessageSpatial3D, MessageNone) {
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    const float z1 = FLAMEGPU->getVariable<float>("z");
    unsigned int count = 0;
    unsigned int badCount = 0;
    int myBin[3] = {static_cast<int>(x1), static_cast<int>(y1), static_cast<int>(z1)};
    for (const auto& message : FLAMEGPU->message_in(x1, y1, z1)) {
        int messageBin[3] = {static_cast<int>(message.getVariable<float>("x")),
                             static_cast<int>(message.getVariable<float>("y")),
                             static_cast<int>(message.getVariable<float>("z"))};
        bool isBad = false;
        for (unsigned int i = 0; i < 3; ++i) {
            int binDiff = myBin[i] - messageBin[i];
            if (binDiff > 1 || binDiff < -1) isBad = true;
        }
        count++;
        badCount = isBad ? badCount + 1 : badCount;
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    FLAMEGPU->setVariable<unsigned int>("badCount", badCount);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(inWrapped3D, MessageSpatial3D, MessageNone) {
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    const float z1 = FLAMEGPU->getVariable<float>("z");
    const flamegpu::id_t ID = FLAMEGPU->getID();
    unsigned int count = 0;
    unsigned int badCount = 0;
    float xSum = 0, ySum = 0, zSum = 0;
    for (const auto& message : FLAMEGPU->message_in.wrap(x1, y1, z1)) {
        const float x2 = message.getVirtualX(x1);
        const float y2 = message.getVirtualY(y1);
        const float z2 = message.getVirtualZ(z1);
        float x21 = x2 - x1;
        float y21 = y2 - y1;
        float z21 = z2 - z1;
        const float distance = sqrtf(x21 * x21 + y21 * y21 + z21 * z21);
        if (distance > FLAMEGPU->message_in.radius() || (fabsf(x21) != 2.0f && x2 != x1) ||
            (fabsf(y21) != 2.0f && y2 != y1) || (fabsf(z21) != 2.0f && z2 != z1)) {
            badCount++;
        } else {
            count++;
            if (message.getVariable<flamegpu::id_t>("id") != ID) {
                xSum += x21;
                ySum += y21;
                zSum += z21;
            }
        }
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    FLAMEGPU->setVariable<unsigned int>("badCount", badCount);
    FLAMEGPU->setVariable<float>("result_x", xSum);
    FLAMEGPU->setVariable<float>("result_y", ySum);
    FLAMEGPU->setVariable<float>("result_z", zSum);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(in_count3D, MessageSpatial3D, MessageNone) {
    unsigned int count = 0;
    for (const auto& message : FLAMEGPU->message_in(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"),
                                                    FLAMEGPU->getVariable<float>("z"))) {
        (void)message;
        ++count;
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(in_wrap_any3D, MessageSpatial3D, MessageNone) {
    unsigned int count = 0;
    for (const auto& message : FLAMEGPU->message_in.wrap(FLAMEGPU->getVariable<float>("x"),
                                                         FLAMEGPU->getVariable<float>("y"),
                                                         FLAMEGPU->getVariable<float>("z"))) {
        (void)message;
        ++count;
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    return ALIVE;
}

// ---------------------------------------------------------------- spatial 2-D
FLAMEGPU_AGENT_FUNCTION(out_mandatory2D, MessageNone, MessageSpatial2D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setLocation(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"));
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(in2D, MessageSpatial2D, MessageNone) {
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    unsigned int count = 0;
    unsigned int badCount = 0;
    int myBin[2] = {static_cast<int>(x1), static_cast<int>(y1)};
    for (const auto& message : FLAMEGPU->message_in(x1, y1)) {
        int messageBin[2] = {static_cast<int>(message.getVariable<float>("x")),
                             static_cast<int>(message.getVariable<float>("y"))};
        bool isBad = false;
        for (unsigned int i = 0; i < 2; ++i) {
            int binDiff = myBin[i] - messageBin[i];
            if (binDiff > 1 || binDiff < -1) isBad = true;
        }
        count++;
        badCount = isBad ? badCount + 1 : badCount;
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    FLAMEGPU->setVariable<unsigned int>("badCount", badCount);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(inWrapped2D, MessageSpatial2D, MessageNone) {
    const float x1 = FLAMEGPU->getVariable<float>("x");
    const float y1 = FLAMEGPU->getVariable<float>("y");
    const flamegpu::id_t ID = FLAMEGPU->getID();
    unsigned int count = 0;
    unsigned int badCount = 0;
    float xSum = 0, ySum = 0;
    for (const auto& message : FLAMEGPU->message_in.wrap(x1, y1)) {
        const float x2 = message.getVirtualX(x1);
        const float y2 = message.getVirtualY(y1);
        float x21 = x2 - x1;
        float y21 = y2 - y1;
        const float distance = sqrtf(x21 * x21 + y21 * y21);
        if (distance > FLAMEGPU->message_in.radius() || (fabsf(x21) != 2.0f && x2 != x1) ||
            (fabsf(y21) != 2.0f && y2 != y1)) {
            badCount++;
        } else {
            count++;
            if (message.getVariable<flamegpu::id_t>("id") != ID) {
                xSum += x21;
                ySum += y21;
            }
        }
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    FLAMEGPU->setVariable<unsigned int>("badCount", badCount);
    FLAMEGPU->setVariable<float>("result_x", xSum);
    FLAMEGPU->setVariable<float>("result_y", ySum);
    return ALIVE;
}

// ---------------------------------------------------------------- sort, births, transitions, brute force
FLAMEGPU_AGENT_FUNCTION(dummySpatialFunc_3D, MessageSpatial3D, MessageNone) { return ALIVE; }

FLAMEGPU_AGENT_FUNCTION(MandatoryOutput, MessageNone, MessageNone) {
    unsigned int id = FLAMEGPU->getVariable<unsigned int>("id") + 1;
    FLAMEGPU->agent_out.setVariable<float>("x", id + 12.0f);
    FLAMEGPU->agent_out.setVariable<unsigned int>("id", id);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(OptionalOutput, MessageNone, MessageNone) {
    unsigned int id = FLAMEGPU->getVariable<unsigned int>("id") + 1;
    if (threadIdx.x % 2 == 0) {
        FLAMEGPU->agent_out.setVariable<float>("x", id + 12.0f);
        FLAMEGPU->agent_out.setVariable<unsigned int>("id", id);
    }
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(MandatoryOutputWithDeath, MessageNone, MessageNone) {
    unsigned int id = FLAMEGPU->getVariable<unsigned int>("id") + 1;
    FLAMEGPU->agent_out.setVariable<float>("x", id + 12.0f);
    FLAMEGPU->agent_out.setVariable<unsigned int>("id", id);
    return DEAD;
}
FLAMEGPU_AGENT_FUNCTION(OptionalOutputWithDeath, MessageNone, MessageNone) {
    unsigned int id = FLAMEGPU->getVariable<unsigned int>("id") + 1;
    if (threadIdx.x % 2 == 0) {
        FLAMEGPU->agent_out.setVariable<float>("x", id + 12.0f);
        FLAMEGPU->agent_out.setVariable<unsigned int>("id", id);
    } else {
        return DEAD;
    }
    return ALIVE;
}
// test_agent_state_transition.cu:29-62
FLAMEGPU_AGENT_FUNCTION(AgentGood, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<int>("x", 11);
    for (int k = 0; k < 4; ++k) FLAMEGPU->setVariable<int, 4>("y", k, 23 + k);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(AgentBad, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<int>("x", 13);
    for (int k = 0; k < 4; ++k) FLAMEGPU->setVariable<int, 4>("y", k, 3 + k);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(AgentDecrement, MessageNone, MessageNone) {
    const unsigned int x = FLAMEGPU->getVariable<unsigned int>("x");
    FLAMEGPU->setVariable<unsigned int>("x", x == 0 ? 0 : x - 1);
    for (int k = 0; k < 4; ++k) FLAMEGPU->setVariable<int, 4>("z", k, 23 + k);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(AgentNull, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<unsigned int>("x", 0xFFFFFFFFu);
    for (int k = 0; k < 4; ++k) FLAMEGPU->setVariable<int, 4>("z", k, 3 + k);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION_CONDITION(Zero_X) { return FLAMEGPU->getVariable<unsigned int>("x") == 0; }

// tests/test_cases/simulation/test_cuda_simulation_concurrency.cu:322-334: slow, uninteresting work
FLAMEGPU_AGENT_FUNCTION(SlowAgentFunction, MessageNone, MessageNone) {
    for (int i = 0; i < 65536; i++) {
        const float v = FLAMEGPU->getVariable<float>("v");
        FLAMEGPU->setVariable<float>("v", v + v);
    }
    return ALIVE;
}

// tests/test_cases/simulation/test_cuda_ensemble.cu:251-278, 630-652
FLAMEGPU_AGENT_FUNCTION(simulateAgentFn, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<uint32_t>("counter", FLAMEGPU->getVariable<uint32_t>("counter") + 1);
    return ALIVE;
}
FLAMEGPU_INIT_FUNCTION(simulateInit) {
    const uint32_t n = FLAMEGPU->environment.getProperty<uint32_t>("POPULATION_TO_GENERATE");
    auto agent = FLAMEGPU->agent("Agent");
    for (uint32_t i = 0; i < n; ++i) agent.newAgent().setVariable<uint32_t>("counter", 0);
}
// the reference adds the sum to a file-scope atomic; here it goes to an environment
// property, which the exit log reports per run
FLAMEGPU_EXIT_FUNCTION(simulateExit) {
    const uint32_t total = FLAMEGPU->agent("Agent").sum<uint32_t, uint32_t>("counter");
    FLAMEGPU->environment.setProperty<uint32_t>("EXIT_SUM", total);
}
// every second run fails (the reference counts calls in a static; here the run's seed decides)
FLAMEGPU_STEP_FUNCTION(throwException) {
    if (FLAMEGPU->random.getSeed() % 2 == 1) throw flamegpu::exception::UnknownInternalError("Dummy Exception");
}

// one agent list sorted by two functions whose spatial inputs have different radii (so
// different bin counts share the list's sort scratch)
FLAMEGPU_AGENT_FUNCTION(out_loc, MessageNone, MessageSpatial3D) {
    FLAMEGPU->message_out.setVariable<flamegpu::id_t>("id", FLAMEGPU->getID());
    FLAMEGPU->message_out.setLocation(FLAMEGPU->getVariable<float>("x"), FLAMEGPU->getVariable<float>("y"),
                                      FLAMEGPU->getVariable<float>("z"));
    return ALIVE;
}
template <int R10>
__device__ __forceinline__ unsigned int count_within(flamegpu::DeviceAPI<MessageSpatial3D, MessageNone>* FLAMEGPU) {
    const float x = FLAMEGPU->getVariable<float>("x"), y = FLAMEGPU->getVariable<float>("y"),
                z = FLAMEGPU->getVariable<float>("z");
    const float r = R10 / 10.0f;
    unsigned int c = 0;
    for (const auto& m : FLAMEGPU->message_in(x, y, z)) {
        const float dx = m.getVariable<float>("x") - x, dy = m.getVariable<float>("y") - y,
                    dz = m.getVariable<float>("z") - z;
        c += dx * dx + dy * dy + dz * dz <= r * r ? 1u : 0u;
    }
    return c;
}
FLAMEGPU_AGENT_FUNCTION(in_near, MessageSpatial3D, MessageNone) {
    FLAMEGPU->setVariable<unsigned int>("near", count_within<10>(FLAMEGPU));
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(in_far, MessageSpatial3D, MessageNone) {
    FLAMEGPU->setVariable<unsigned int>("far", count_within<25>(FLAMEGPU));
    return ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(AddOne, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<unsigned int>("x", FLAMEGPU->getVariable<unsigned int>("x") + 1);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(out_bf, MessageNone, MessageBruteForce) {
    FLAMEGPU->message_out.setVariable<int>("x", FLAMEGPU->getVariable<int>("x"));
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(in_bf, MessageBruteForce, MessageNone) {
    int sum = 0;
    int product = 1;
    for (const auto& m : FLAMEGPU->message_in) {
        const int x = m.getVariable<int>("x");
        sum += x;
        product *= x;
        product = product > 1000000 ? 1 : product;
    }
    FLAMEGPU->setVariable<int>("sum", sum);
    FLAMEGPU->setVariable<int>("product", product);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(draw_random, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<float>("a", FLAMEGPU->random.uniform<float>());
    FLAMEGPU->setVariable<float>("b", FLAMEGPU->random.uniform<float>());
    FLAMEGPU->setVariable<int>("c", FLAMEGPU->random.uniform<int>(1, 6));
    FLAMEGPU->setVariable<float>("d", FLAMEGPU->random.normal<float>());
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(die_without_permission, MessageNone, MessageNone) { return DEAD; }

// ---------------------------------------------------------------- function conditions
// tests/test_cases/runtime/test_agent_function_conditions.cu:26-47
FLAMEGPU_AGENT_FUNCTION(NullFn1, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<int>("x", FLAMEGPU->getVariable<int>("x") + 1);
    FLAMEGPU->setVariable<int, 4>("y", 0, 3);
    FLAMEGPU->setVariable<int, 4>("y", 1, 4);
    FLAMEGPU->setVariable<int, 4>("y", 2, 5);
    FLAMEGPU->setVariable<int, 4>("y", 3, 6);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(NullFn2, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<int>("x", FLAMEGPU->getVariable<int>("x") - 1);
    FLAMEGPU->setVariable<int, 4>("y", 0, 23);
    FLAMEGPU->setVariable<int, 4>("y", 1, 24);
    FLAMEGPU->setVariable<int, 4>("y", 2, 25);
    FLAMEGPU->setVariable<int, 4>("y", 3, 26);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION_CONDITION(Condition1) { return FLAMEGPU->getVariable<int>("x") == 1; }
FLAMEGPU_AGENT_FUNCTION_CONDITION(Condition2) { return FLAMEGPU->getVariable<int>("x") != 1; }
FLAMEGPU_AGENT_FUNCTION_CONDITION(AllFail) { return false; }
// a condition with death, births and message output under it
FLAMEGPU_AGENT_FUNCTION_CONDITION(NotMultipleOf3) { return FLAMEGPU->getVariable<int>("x") % 3 != 0; }
FLAMEGPU_AGENT_FUNCTION(act_under_condition, MessageNone, MessageBruteForce) {
    const int x = FLAMEGPU->getVariable<int>("x");
    FLAMEGPU->message_out.setVariable<int>("x", x);
    if (x % 5 == 0) FLAMEGPU->agent_out.setVariable<int>("x", x + 1000);
    return x % 2 == 0 ? DEAD : ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(sum_messages, MessageBruteForce, MessageNone) {
    int sum = 0;
    for (const auto& m : FLAMEGPU->message_in) sum += m.getVariable<int>("x");
    FLAMEGPU->setVariable<int>("sum", sum);
    return ALIVE;
}

FLAMEGPU_AGENT_FUNCTION(AddFloatOne, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<float>("float_var", FLAMEGPU->getVariable<float>("float_var") + 1.0f);
    return ALIVE;
}
// tests/test_cases/io/test_logging.cu:18-36
FLAMEGPU_AGENT_FUNCTION(agent_fn1, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<float>("float_var", FLAMEGPU->getVariable<float>("float_var") + 1.0f);
    FLAMEGPU->setVariable<int>("int_var", FLAMEGPU->getVariable<int>("int_var") + 1);
    FLAMEGPU->setVariable<unsigned int>("uint_var", FLAMEGPU->getVariable<unsigned int>("uint_var") + 1);
    return ALIVE;
}
FLAMEGPU_STEP_FUNCTION(step_fn1) {
    FLAMEGPU->environment.setProperty<float>("float_prop", FLAMEGPU->environment.getProperty<float>("float_prop") + 1.0f);
    FLAMEGPU->environment.setProperty<int>("int_prop", FLAMEGPU->environment.getProperty<int>("int_prop") + 1);
    FLAMEGPU->environment.setProperty<unsigned int>("uint_prop", FLAMEGPU->environment.getProperty<unsigned int>("uint_prop") + 1);
    auto a = FLAMEGPU->environment.getProperty<float, 2>("float_prop_array");
    auto b = FLAMEGPU->environment.getProperty<int, 3>("int_prop_array");
    auto c = FLAMEGPU->environment.getProperty<unsigned int, 4>("uint_prop_array");
    FLAMEGPU->environment.setProperty<float, 2>("float_prop_array", {a[0] + 1.0f, a[1] + 1.0f});
    FLAMEGPU->environment.setProperty<int, 3>("int_prop_array", {b[0] + 1, b[1] + 1, b[2] + 1});
    FLAMEGPU->environment.setProperty<unsigned int, 4>("uint_prop_array", {c[0] + 1, c[1] + 1, c[2] + 1, c[3] + 1});
}

// test_device_agent_creation.cu:1195-1203
FLAMEGPU_AGENT_FUNCTION(CopyID, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<flamegpu::id_t>("id_copy", FLAMEGPU->getID());
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(DeviceBirth, MessageNone, MessageNone) {
    FLAMEGPU->setVariable<flamegpu::id_t>("id_other", FLAMEGPU->agent_out.getID());
    FLAMEGPU->agent_out.setVariable<flamegpu::id_t>("id_other", FLAMEGPU->getID());
    return ALIVE;
}

flamegpu::ModelDescription fgxm_build_model(const char* variant_c, const float* p, uint32_t np) {
    const std::string v = variant_c ? variant_c : "";
    flamegpu::ModelDescription model("kat_" + v);
    auto param = [&](uint32_t i, float d) { return i < np ? p[i] : d; };
    if (v == "mandatory3d" || v == "optional3d" || v == "optionalnone3d") {
        auto message = model.newMessage<MessageSpatial3D>("location");
        message.setMin(0, 0, 0);
        message.setMax(5, 5, 5);
        message.setRadius(1);
        message.newVariable<flamegpu::id_t>("id");
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("z");
        agent.newVariable<int>("do_output");
        agent.newVariable<unsigned int>("count");
        agent.newVariable<unsigned int>("badCount");
        flamegpu::AgentFunctionDescription out = v == "mandatory3d" ? agent.newFunction("out", out_mandatory3D)
                                               : v == "optional3d" ? agent.newFunction("out", out_optional3D)
                                                                   : agent.newFunction("out", out_optional3DNone);
        out.setMessageOutput("location");
        out.setMessageOutputOptional(v != "mandatory3d");
        auto in = agent.newFunction("in", in3D);
        in.setMessageInput("location");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "mandatory2d") {
        auto message = model.newMessage<MessageSpatial2D>("location");
        message.setMin(0, 0);
        message.setMax(5, 5);
        message.setRadius(1);
        message.newVariable<flamegpu::id_t>("id");
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<unsigned int>("count");
        agent.newVariable<unsigned int>("badCount");
        auto out = agent.newFunction("out", out_mandatory2D);
        out.setMessageOutput("location");
        auto in = agent.newFunction("in", in2D);
        in.setMessageInput("location");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "wrapped3d") {
        const float ox = param(0, 0), oy = param(1, 0), oz = param(2, 0);
        auto message = model.newMessage<MessageSpatial3D>("location");
        message.setMin(0 + ox, 0 + oy, 0 + oz);
        message.setMax(70 + ox, 70 + oy, 70 + oz);
        message.setRadius(3.5);
        message.newVariable<flamegpu::id_t>("id");
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("z");
        agent.newVariable<float>("result_x");
        agent.newVariable<float>("result_y");
        agent.newVariable<float>("result_z");
        agent.newVariable<unsigned int>("count");
        agent.newVariable<unsigned int>("badCount");
        auto out = agent.newFunction("out", out_mandatory3D);
        out.setMessageOutput("location");
        auto in = agent.newFunction("in", inWrapped3D);
        in.setMessageInput("location");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "wrapped2d") {
        const float ox = param(0, 0), oy = param(1, 0);
        auto message = model.newMessage<MessageSpatial2D>("location");
        message.setMin(0 + ox, 0 + oy);
        message.setMax(70 + ox, 70 + oy);
        message.setRadius(3.5);
        message.newVariable<flamegpu::id_t>("id");
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("result_x");
        agent.newVariable<float>("result_y");
        agent.newVariable<unsigned int>("count");
        agent.newVariable<unsigned int>("badCount");
        auto out = agent.newFunction("out", out_mandatory2D);
        out.setMessageOutput("location");
        auto in = agent.newFunction("in", inWrapped2D);
        in.setMessageInput("location");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "bounds3d" || v == "wrap_out_of_bounds") {
        auto message = model.newMessage<MessageSpatial3D>("location");
        if (v == "bounds3d") {
            message.setMin(0, 0, 0);
            message.setMax(50.1f, 50.1f, 50.1f);
            message.setRadius(10);
        } else {
            message.setMin(0, 0, 0);
            message.setMax(10, 10, 10);
            message.setRadius(1);
        }
        message.newVariable<flamegpu::id_t>("id");
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("z");
        agent.newVariable<unsigned int>("count", 0);
        auto out = agent.newFunction("out", out_mandatory3D);
        out.setMessageOutput(message);
        auto in = v == "bounds3d" ? agent.newFunction("in", in_count3D) : agent.newFunction("in", in_wrap_any3D);
        in.setMessageInput(message);
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "sort3d") {
        auto agent = model.newAgent("agent");
        agent.newVariable<int>("initial_order");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("z");
        agent.setSortPeriod((unsigned int)param(0, 1));
        auto loc = model.newMessage<MessageSpatial3D>("location");
        loc.setMin(-5, -5, -5);
        loc.setMax(5, 5, 5);
        loc.setRadius(0.2f);
        auto f = agent.newFunction("dummySpatialFunc", dummySpatialFunc_3D);
        f.setMessageInput("location");
        model.newLayer().addAgentFunction(f);
    } else if (v.rfind("birth_", 0) == 0) {
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<unsigned int>("id");
        flamegpu::AgentFunctionDescription f =
            v == "birth_mandatory"         ? agent.newFunction("output", MandatoryOutput)
            : v == "birth_optional"        ? agent.newFunction("output", OptionalOutput)
            : v == "birth_mandatory_death" ? agent.newFunction("output", MandatoryOutputWithDeath)
                                           : agent.newFunction("output", OptionalOutputWithDeath);
        f.setAgentOutput(agent);
        f.setAllowAgentDeath(v == "birth_mandatory_death" || v == "birth_optional_death");
        model.newLayer().addAgentFunction(f);
    } else if (v == "transition") {
        auto agent = model.newAgent("agent");
        agent.newState("a");
        agent.newState("b");
        agent.newVariable<unsigned int>("x");
        auto f = agent.newFunction("add_one", AddOne);
        f.setInitialState("a");
        f.setEndState("b");
        model.newLayer().addAgentFunction(f);
    } else if (v == "transition_src0_dest10" || v == "transition_src10_dest0") {
        // test_agent_state_transition.cu:67-171 (Src_0_Dest_10, Src_10_Dest_0)
        auto agent = model.newAgent("agent");
        agent.newState("Start");
        agent.newState("End");
        agent.newState("End2");
        agent.setInitialState("Start");
        agent.newVariable<int>("x");
        agent.newVariable<int, 4>("y");
        auto f1 = agent.newFunction("Function1", AgentGood);
        f1.setInitialState("Start");
        f1.setEndState("End");
        if (v == "transition_src0_dest10") {
            model.newLayer().addAgentFunction(f1);
        } else {
            auto f2 = agent.newFunction("Function2", AgentBad);
            f2.setInitialState("End");
            f2.setEndState("End2");
            model.newLayer().addAgentFunction(f2);
            model.newLayer().addAgentFunction(f1);
        }
    } else if (v == "transition_src10_dest10") {
        // test_agent_state_transition.cu:172-250: a conditional transition, so lists never swap
        auto agent = model.newAgent("agent");
        agent.newState("Start");
        agent.newState("End");
        agent.setInitialState("Start");
        agent.newVariable<unsigned int>("x");
        agent.newVariable<unsigned int>("y");
        agent.newVariable<int, 4>("z");
        auto f1 = agent.newFunction("Function1", AgentDecrement);
        f1.setInitialState("Start");
        f1.setEndState("Start");
        auto f2 = agent.newFunction("Function2", AgentNull);
        f2.setInitialState("Start");
        f2.setEndState("End");
        f2.setFunctionCondition(Zero_X);
        model.newLayer().addAgentFunction(f1);
        model.newLayer().addAgentFunction(f2);
    } else if (v == "concurrency") {
        // test_cuda_simulation_concurrency.cu:353-390 (LayerConcurrency): one layer, one slow
        // function for each of 4 agent types
        auto layer = model.newLayer();
        for (int i = 0; i < 4; ++i) {
            const std::string name = "agent_" + std::to_string(i);
            auto a = model.newAgent(name);
            a.newVariable<float>("v");
            layer.addAgentFunction(a.newFunction(name + "_slowAgentFunction", SlowAgentFunction));
        }
    } else if (v == "ensemble_simulate" || v == "ensemble_errors") {
        // test_cuda_ensemble.cu:279-322 (simulate), 653-748 (ErrorOff / ErrorSlow / ErrorFast)
        model.Environment().newProperty<uint32_t>("POPULATION_TO_GENERATE", v == "ensemble_simulate" ? 32u : 1u, true);
        model.Environment().newProperty<uint32_t>("EXIT_SUM", 0u);
        auto agent = model.newAgent("Agent");
        agent.newVariable<uint32_t>("counter", 0);
        model.addInitFunction(simulateInit);
        if (v == "ensemble_simulate") {
            auto f = agent.newFunction("simulateAgentFn", simulateAgentFn);
            model.newLayer().addAgentFunction(f);
            model.addExitFunction(simulateExit);
        } else {
            model.addStepFunction(throwException);
        }
    } else if (v == "two_radii") {
        for (auto spec : {std::make_pair("near", 1.0f), std::make_pair("far", 2.5f)}) {
            auto m = model.newMessage<MessageSpatial3D>(spec.first);
            m.setMin(0, 0, 0);
            m.setMax(10, 10, 10);
            m.setRadius(spec.second);
            m.newVariable<flamegpu::id_t>("id");
        }
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        agent.newVariable<float>("y");
        agent.newVariable<float>("z");
        agent.newVariable<unsigned int>("near");
        agent.newVariable<unsigned int>("far");
        auto o1 = agent.newFunction("out_near", out_loc);
        o1.setMessageOutput("near");
        auto o2 = agent.newFunction("out_far", out_loc);
        o2.setMessageOutput("far");
        auto i1 = agent.newFunction("in_near", in_near);
        i1.setMessageInput("near");
        auto i2 = agent.newFunction("in_far", in_far);
        i2.setMessageInput("far");
        model.newLayer().addAgentFunction(o1);
        model.newLayer().addAgentFunction(o2);
        model.newLayer().addAgentFunction(i1);
        model.newLayer().addAgentFunction(i2);
    } else if (v == "bruteforce") {
        auto message = model.newMessage("msg");
        message.newVariable<int>("x");
        auto agent = model.newAgent("agent");
        agent.newVariable<int>("x");
        agent.newVariable<int>("sum");
        agent.newVariable<int>("product");
        auto out = agent.newFunction("out", out_bf);
        out.setMessageOutput("msg");
        auto in = agent.newFunction("in", in_bf);
        in.setMessageInput("msg");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "random") {
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("a");
        agent.newVariable<float>("b");
        agent.newVariable<int>("c");
        agent.newVariable<float>("d");
        model.newLayer().addAgentFunction(agent.newFunction("draw", draw_random));
    } else if (v == "agent_id_states") {
        // AgentID_MultipleStatesUniqueIDs (test_device_agent_creation.cu:1204-1244)
        auto agent = model.newAgent("agent");
        agent.newVariable<flamegpu::id_t>("id_copy", flamegpu::ID_NOT_SET);
        agent.newVariable<flamegpu::id_t>("id_other", flamegpu::ID_NOT_SET);
        agent.newState("a");
        agent.newState("b");
        auto af1_a = agent.newFunction("birth", DeviceBirth);
        af1_a.setAgentOutput(agent, "a");
        af1_a.setInitialState("a");
        af1_a.setEndState("a");
        auto af1_b = agent.newFunction("birth2", DeviceBirth);
        af1_b.setAgentOutput(agent, "b");
        af1_b.setInitialState("b");
        af1_b.setEndState("b");
        auto af_a = agent.newFunction("copy_id", CopyID);
        af_a.setInitialState("a");
        af_a.setEndState("a");
        auto af_b = agent.newFunction("copy_id2", CopyID);
        af_b.setInitialState("b");
        af_b.setEndState("b");
        model.newLayer().addAgentFunction(af1_a);
        model.newLayer().addAgentFunction(af1_b);
        auto layer2 = model.newLayer();
        layer2.addAgentFunction(af_a);
        layer2.addAgentFunction(af_b);
    } else if (v == "agent_id_two_types") {
        // AgentID_DeviceBirth_MultipleAgents (test_device_agent_creation.cu:1297-1339)
        auto agentA = model.newAgent("agentA");
        agentA.newVariable<flamegpu::id_t>("id_other", flamegpu::ID_NOT_SET);
        auto agentB = model.newAgent("agentB");
        agentB.newVariable<flamegpu::id_t>("id_other", flamegpu::ID_NOT_SET);
        auto af1_a = agentA.newFunction("birth", DeviceBirth);
        af1_a.setAgentOutput(agentB);
        model.newLayer().addAgentFunction(af1_a);
    } else if (v == "death_disabled") {
        auto agent = model.newAgent("agent");
        agent.newVariable<float>("x");
        model.newLayer().addAgentFunction(agent.newFunction("die", die_without_permission));
    } else if (v == "cond_split") {
        // TestAgentFunctionConditions.SplitAgents (test_agent_function_conditions.cu:48-102)
        auto a = model.newAgent("agent");
        a.newVariable<int>("x");
        a.newVariable<int, 4>("y");
        a.newState("Start");
        a.newState("End");
        a.newState("End2");
        auto af1 = a.newFunction("Function1", NullFn1);
        af1.setInitialState("Start");
        af1.setEndState("End");
        af1.setFunctionCondition(Condition1);
        auto af2 = a.newFunction("Function2", NullFn2);
        af2.setInitialState("Start");
        af2.setEndState("End2");
        af2.setFunctionCondition(Condition2);
        model.newLayer().addAgentFunction(af1);
        model.newLayer().addAgentFunction(af2);
    } else if (v == "cond_all_disabled") {
        // TestAgentFunctionConditions.AllDisabled (test_agent_function_conditions.cu:107-129)
        auto a = model.newAgent("agent");
        a.newVariable<int>("x");
        a.newVariable<int, 4>("y");
        auto af1 = a.newFunction("Function1", NullFn1);
        af1.setFunctionCondition(AllFail);
        auto af2 = a.newFunction("Function2", NullFn2);
        model.newLayer().addAgentFunction(af1);
        model.newLayer().addAgentFunction(af2);
    } else if (v == "cond_mixed") {
        auto message = model.newMessage("msg");
        message.newVariable<int>("x");
        auto a = model.newAgent("agent");
        a.newVariable<int>("x");
        a.newVariable<int>("sum");
        auto act = a.newFunction("act", act_under_condition);
        act.setFunctionCondition(NotMultipleOf3);
        act.setMessageOutput("msg");
        act.setAgentOutput(a);
        act.setAllowAgentDeath(true);
        auto sum = a.newFunction("sum", sum_messages);
        sum.setMessageInput("msg");
        model.newLayer().addAgentFunction(act);
        model.newLayer().addAgentFunction(sum);
    } else if (v == "logging" || v == "logging_empty") {
        // tests/test_cases/io/test_logging.cu:49-78 (CUDASimulationStep), :336-356 (EmptyMean)
        auto a = model.newAgent("Agent1");
        a.newVariable<float>("float_var");
        if (v == "logging") {
            a.newVariable<int>("int_var");
            a.newVariable<unsigned int>("uint_var");
            auto f1 = a.newFunction("Func1", agent_fn1);
            model.newLayer().addAgentFunction(f1);
            model.addStepFunction(step_fn1);
            model.Environment().newProperty<float>("float_prop", 1.0f);
            model.Environment().newProperty<int>("int_prop", 1);
            model.Environment().newProperty<unsigned int>("uint_prop", 1);
            model.Environment().newProperty<float, 2>("float_prop_array", {1.0f, 2.0f});
            model.Environment().newProperty<int, 3>("int_prop_array", {2, 3, 4});
            model.Environment().newProperty<unsigned int, 4>("uint_prop_array", {3, 4, 5, 6});
        } else {
            auto f1 = a.newFunction("Func1", AddFloatOne);
            model.newLayer().addAgentFunction(f1);
        }
    } else {
        throw flamegpu::exception::InvalidArgument("unknown test model variant '" + v + "'");
    }
    return model;
}

#include "flamegpu/model_abi.h"
