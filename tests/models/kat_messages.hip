// Bucket and array message models for the known-answer tests, written against the
// drop-in C++ API.  One library, one model per `variant`:
//   bucket_mandatory / bucket_optional / bucket_optional_none / bucket_range
//       tests/test_cases/runtime/messaging/test_bucket.cu:60-330,612-690
//   array (params: dims, X, Y, Z, radius, mode, optional)
//       tests/test_cases/runtime/messaging/test_array.cu, test_array_2d.cu, test_array_3d.cu:
//       at(), Moore / wrapped Moore / von Neumann / wrapped von Neumann neighbourhoods.
//       mode 0 clamped Moore, 1 wrapped Moore, 2 von Neumann, 3 wrapped von Neumann.
//       Each agent records the value read through at() of its own element and an
//       order-sensitive hash and count of its neighbourhood.
//   truncate / append_keep / optional_truncate / optional_append / append_resize
//       tests/test_cases/runtime/messaging/test_append_truncate.cu:72-310
//   persistence_off / persistence_on
//       tests/test_cases/runtime/messaging/test_brute_force.cu:440-600
#include <string>

#include "flamegpu/flamegpu.h"

using flamegpu::ALIVE;
using flamegpu::MessageArray;
using flamegpu::MessageArray2D;
using flamegpu::MessageArray3D;
using flamegpu::MessageBruteForce;
using flamegpu::MessageBucket;
using flamegpu::MessageNone;

// ---------------------------------------------------------------- append / truncate
// (test_append_truncate.cu:22-67, 229-256)
FLAMEGPU_AGENT_FUNCTION(Out_AppendTruncate, MessageNone, MessageBruteForce) {
    FLAMEGPU->message_out.setVariable("x", 0);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(Out_AppendTruncate2, MessageNone, MessageBruteForce) {
    FLAMEGPU->message_out.setVariable("x", 1);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(In_AppendTruncate, MessageBruteForce, MessageNone) {
    unsigned int count = 0;
    for (auto& message : FLAMEGPU->message_in) {
        (void)message;
        count++;
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(In_AppendTruncate2, MessageBruteForce, MessageNone) {
    unsigned int count0 = 0, count1 = 0;
    for (auto& message : FLAMEGPU->message_in) {
        const int x = message.getVariable<int>("x");
        if (x == 0) count0++;
        else if (x == 1) count1++;
    }
    FLAMEGPU->setVariable<unsigned int>("count0", count0);
    FLAMEGPU->setVariable<unsigned int>("count1", count1);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(OptionalOut_AppendTruncate, MessageNone, MessageBruteForce) {
    if (FLAMEGPU->getVariable<unsigned int>("do_out") > 0) {
        FLAMEGPU->message_out.setVariable("x", 0);
        FLAMEGPU->setVariable<unsigned int>("do_out", 0);
    } else {
        FLAMEGPU->setVariable<unsigned int>("do_out", 1);
    }
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(OptionalOut_AppendTruncate2, MessageNone, MessageBruteForce) {
    if (FLAMEGPU->getVariable<unsigned int>("do_out") > 0) FLAMEGPU->message_out.setVariable("x", 1);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(Out_1, MessageNone, MessageBruteForce) {
    FLAMEGPU->message_out.setVariable("x", 1);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(Out_2, MessageNone, MessageBruteForce) {
    FLAMEGPU->message_out.setVariable("x", 2);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(In_012, MessageBruteForce, MessageNone) {
    unsigned int c0 = 0, c1 = 0, c2 = 0;
    for (auto& message : FLAMEGPU->message_in) {
        const int x = message.getVariable<int>("x");
        if (x == 1) c1++;
        else if (x == 2) c2++;
        else c0++;
    }
    FLAMEGPU->setVariable<unsigned int>("0", c0);
    FLAMEGPU->setVariable<unsigned int>("1", c1);
    FLAMEGPU->setVariable<unsigned int>("2", c2);
    return ALIVE;
}

// ---------------------------------------------------------------- persistence
// (test_brute_force.cu:447-470): output on even steps only, read every step
FLAMEGPU_AGENT_FUNCTION_CONDITION(EvenOnlyCondition) { return FLAMEGPU->getStepCounter() % 2 == 0; }
FLAMEGPU_AGENT_FUNCTION(out_simple, MessageNone, MessageBruteForce) {
    FLAMEGPU->message_out.setVariable<int>("id", FLAMEGPU->getVariable<int>("id"));
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(in_simple, MessageBruteForce, MessageNone) {
    unsigned int count = 0, sum = 0;
    for (auto& m : FLAMEGPU->message_in) {
        count++;
        sum += m.getVariable<int>("id");
    }
    FLAMEGPU->setVariable<unsigned int>("count", count);
    FLAMEGPU->setVariable<unsigned int>("sum", sum);
    return ALIVE;
}

// ---------------------------------------------------------------- bucket (test_bucket.cu:60-98)
FLAMEGPU_AGENT_FUNCTION(bucket_out_mandatory, MessageNone, MessageBucket) {
    const int id = FLAMEGPU->getVariable<int>("id");
    FLAMEGPU->message_out.setVariable<int>("id", id);
    FLAMEGPU->message_out.setKey(12 + (id / 2));
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(bucket_out_optional, MessageNone, MessageBucket) {
    if (FLAMEGPU->getVariable<int>("do_output")) {
        const int id = FLAMEGPU->getVariable<int>("id");
        FLAMEGPU->message_out.setVariable<int>("id", id);
        FLAMEGPU->message_out.setKey(12 + (id / 2));
    }
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(bucket_out_none, MessageNone, MessageBucket) { return ALIVE; }
FLAMEGPU_AGENT_FUNCTION(bucket_in, MessageBucket, MessageNone) {
    const int id = FLAMEGPU->getVariable<int>("id");
    const int id_m1 = id == 0 ? 0 : id - 1;
    unsigned int count = 0;
    unsigned int sum = 0;
    for (auto& m : FLAMEGPU->message_in(12 + (id_m1 / 2))) {
        count++;
        sum += m.getVariable<int>("id");
    }
    FLAMEGPU->setVariable<unsigned int>("count1", count);
    FLAMEGPU->setVariable<unsigned int>("count2", FLAMEGPU->message_in(12 + (id_m1 / 2)).size());
    FLAMEGPU->setVariable<unsigned int>("sum", sum);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(bucket_in_range, MessageBucket, MessageNone) {
    const int id = FLAMEGPU->getVariable<int>("id");
    const int id_m4 = 12 + ((id / 8) * 4);
    unsigned int count = 0;
    unsigned int sum = 0;
    for (auto& m : FLAMEGPU->message_in(id_m4, id_m4 + 4)) {
        count++;
        sum += m.getVariable<int>("id");
    }
    FLAMEGPU->setVariable<unsigned int>("count1", count);
    FLAMEGPU->setVariable<unsigned int>("count2", FLAMEGPU->message_in(12 + id / 2).size());
    FLAMEGPU->setVariable<unsigned int>("sum", sum);
    return ALIVE;
}

// ---------------------------------------------------------------- arrays
template <typename Out>
__device__ __forceinline__ bool array_should_write(const Out&, unsigned int write, unsigned int optional) {
    return !optional || (write % 2u) == 0u;
}
FLAMEGPU_AGENT_FUNCTION(array1_out, MessageNone, MessageArray) {
    const unsigned int w = FLAMEGPU->getVariable<unsigned int>("write");
    if (array_should_write(FLAMEGPU->message_out, w, FLAMEGPU->environment.getProperty<unsigned int>("optional"))) {
        FLAMEGPU->message_out.setVariable<unsigned int>("v", w * 3u + 1u);
        FLAMEGPU->message_out.setIndex(w);
    }
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(array2_out, MessageNone, MessageArray2D) {
    const unsigned int w = FLAMEGPU->getVariable<unsigned int>("write");
    if (array_should_write(FLAMEGPU->message_out, w, FLAMEGPU->environment.getProperty<unsigned int>("optional"))) {
        const unsigned int X = FLAMEGPU->environment.getProperty<unsigned int>("X");
        FLAMEGPU->message_out.setVariable<unsigned int>("v", w * 3u + 1u);
        FLAMEGPU->message_out.setIndex(w % X, w / X);
    }
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(array3_out, MessageNone, MessageArray3D) {
    const unsigned int w = FLAMEGPU->getVariable<unsigned int>("write");
    if (array_should_write(FLAMEGPU->message_out, w, FLAMEGPU->environment.getProperty<unsigned int>("optional"))) {
        const unsigned int X = FLAMEGPU->environment.getProperty<unsigned int>("X");
        const unsigned int Y = FLAMEGPU->environment.getProperty<unsigned int>("Y");
        FLAMEGPU->message_out.setVariable<unsigned int>("v", w * 3u + 1u);
        FLAMEGPU->message_out.setIndex(w % X, (w / X) % Y, w / (X * Y));
    }
    return ALIVE;
}

// order-sensitive hash of the visited elements' values and their offsets
template <typename F>
__device__ __forceinline__ void array_walk(F&& filter, unsigned int& h, unsigned int& n) {
    for (auto& m : filter) {
        h = h * 31u + m.template getVariable<unsigned int>("v");
        ++n;
    }
}

FLAMEGPU_AGENT_FUNCTION(array1_in, MessageArray, MessageNone) {
    const unsigned int i = FLAMEGPU->getVariable<unsigned int>("index");
    const unsigned int r = FLAMEGPU->environment.getProperty<unsigned int>("radius");
    const unsigned int mode = FLAMEGPU->environment.getProperty<unsigned int>("mode");
    FLAMEGPU->setVariable<unsigned int>("read", FLAMEGPU->message_in.at(i).getVariable<unsigned int>("v"));
    unsigned int h = 0, n = 0;
    if (mode == 1)
        array_walk(FLAMEGPU->message_in.wrap(i, r), h, n);
    else
        array_walk(FLAMEGPU->message_in(i, r), h, n);
    FLAMEGPU->setVariable<unsigned int>("hash", h);
    FLAMEGPU->setVariable<unsigned int>("count", n);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(array2_in, MessageArray2D, MessageNone) {
    const unsigned int i = FLAMEGPU->getVariable<unsigned int>("index");
    const unsigned int r = FLAMEGPU->environment.getProperty<unsigned int>("radius");
    const unsigned int mode = FLAMEGPU->environment.getProperty<unsigned int>("mode");
    const unsigned int X = FLAMEGPU->message_in.getDimX();
    const unsigned int x = i % X, y = i / X;
    FLAMEGPU->setVariable<unsigned int>("read", FLAMEGPU->message_in.at(x, y).getVariable<unsigned int>("v"));
    unsigned int h = 0, n = 0;
    if (mode == 0) array_walk(FLAMEGPU->message_in(x, y, r), h, n);
    else if (mode == 1) array_walk(FLAMEGPU->message_in.wrap(x, y, r), h, n);
    else if (mode == 2) array_walk(FLAMEGPU->message_in.vn(x, y, r), h, n);
    else array_walk(FLAMEGPU->message_in.vn_wrap(x, y, r), h, n);
    FLAMEGPU->setVariable<unsigned int>("hash", h);
    FLAMEGPU->setVariable<unsigned int>("count", n);
    return ALIVE;
}
FLAMEGPU_AGENT_FUNCTION(array3_in, MessageArray3D, MessageNone) {
    const unsigned int i = FLAMEGPU->getVariable<unsigned int>("index");
    const unsigned int r = FLAMEGPU->environment.getProperty<unsigned int>("radius");
    const unsigned int mode = FLAMEGPU->environment.getProperty<unsigned int>("mode");
    const unsigned int X = FLAMEGPU->message_in.getDimX(), Y = FLAMEGPU->message_in.getDimY();
    const unsigned int x = i % X, y = (i / X) % Y, z = i / (X * Y);
    FLAMEGPU->setVariable<unsigned int>("read", FLAMEGPU->message_in.at(x, y, z).getVariable<unsigned int>("v"));
    unsigned int h = 0, n = 0;
    if (mode == 0) array_walk(FLAMEGPU->message_in(x, y, z, r), h, n);
    else if (mode == 1) array_walk(FLAMEGPU->message_in.wrap(x, y, z, r), h, n);
    else if (mode == 2) array_walk(FLAMEGPU->message_in.vn(x, y, z, r), h, n);
    else array_walk(FLAMEGPU->message_in.vn_wrap(x, y, z, r), h, n);
    FLAMEGPU->setVariable<unsigned int>("hash", h);
    FLAMEGPU->setVariable<unsigned int>("count", n);
    return ALIVE;
}

flamegpu::ModelDescription fgxm_build_model(const char* variant_c, const float* p, uint32_t np) {
    const std::string v = variant_c ? variant_c : "";
    flamegpu::ModelDescription model("kat_messages_" + v);
    auto param = [&](uint32_t i, float d) { return i < np ? p[i] : d; };
    if (v == "bucket_mandatory" || v == "bucket_optional" || v == "bucket_optional_none" || v == "bucket_range") {
        const int agent_count = (int)param(0, 1024.0f);
        auto message = model.newMessage<MessageBucket>("bucket");
        message.setBounds(12, 12 + agent_count / 2);
        message.newVariable<int>("id");
        auto agent = model.newAgent("agent");
        agent.newVariable<int>("id");
        agent.newVariable<int>("do_output");
        agent.newVariable<unsigned int>("count1", 0);
        agent.newVariable<unsigned int>("count2", 0);
        agent.newVariable<unsigned int>("sum", 0);
        flamegpu::AgentFunctionDescription out = v == "bucket_optional" ? agent.newFunction("out", bucket_out_optional)
                                               : v == "bucket_optional_none" ? agent.newFunction("out", bucket_out_none)
                                                                             : agent.newFunction("out", bucket_out_mandatory);
        out.setMessageOutput("bucket");
        if (v == "bucket_optional" || v == "bucket_optional_none") out.setMessageOutputOptional(true);
        flamegpu::AgentFunctionDescription in = v == "bucket_range" ? agent.newFunction("in", bucket_in_range)
                                                                    : agent.newFunction("in", bucket_in);
        in.setMessageInput("bucket");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "array") {
        const unsigned dims = (unsigned)param(0, 2.0f);
        const unsigned X = (unsigned)param(1, 12.0f), Y = (unsigned)param(2, 13.0f), Z = (unsigned)param(3, 1.0f);
        auto env = model.Environment();
        env.newProperty<unsigned int>("X", X);
        env.newProperty<unsigned int>("Y", Y);
        env.newProperty<unsigned int>("radius", (unsigned)param(4, 1.0f));
        env.newProperty<unsigned int>("mode", (unsigned)param(5, 0.0f));
        env.newProperty<unsigned int>("optional", (unsigned)param(6, 0.0f));
        auto agent = model.newAgent("agent");
        agent.newVariable<unsigned int>("index");
        agent.newVariable<unsigned int>("write");
        agent.newVariable<unsigned int>("read", 0xFFFFFFFFu);
        agent.newVariable<unsigned int>("hash", 0);
        agent.newVariable<unsigned int>("count", 0);
        if (dims == 1) {
            auto m = model.newMessage<MessageArray>("array");
            m.setLength(X);
            m.newVariable<unsigned int>("v");
        } else if (dims == 2) {
            auto m = model.newMessage<MessageArray2D>("array");
            m.setDimensions(X, Y);
            m.newVariable<unsigned int>("v");
        } else {
            auto m = model.newMessage<MessageArray3D>("array");
            m.setDimensions(X, Y, Z);
            m.newVariable<unsigned int>("v");
        }
        flamegpu::AgentFunctionDescription out = dims == 1 ? agent.newFunction("out", array1_out)
                                               : dims == 2 ? agent.newFunction("out", array2_out)
                                                           : agent.newFunction("out", array3_out);
        flamegpu::AgentFunctionDescription in = dims == 1 ? agent.newFunction("in", array1_in)
                                              : dims == 2 ? agent.newFunction("in", array2_in)
                                                          : agent.newFunction("in", array3_in);
        out.setMessageOutput("array");
        if (param(6, 0.0f) != 0.0f) out.setMessageOutputOptional(true);
        in.setMessageInput("array");
        model.newLayer().addAgentFunction(out);
        model.newLayer().addAgentFunction(in);
    } else if (v == "truncate" || v == "append_keep" || v == "optional_truncate" || v == "optional_append") {
        const bool optional = v == "optional_truncate" || v == "optional_append";
        const bool two = v == "append_keep" || v == "optional_append";
        auto message = model.newMessage("Message");
        message.newVariable<int>("x");
        auto a = model.newAgent("Agent");
        if (two) {
            a.newVariable<unsigned int>("count0");
            a.newVariable<unsigned int>("count1");
        } else {
            a.newVariable<unsigned int>("count");
        }
        if (optional) a.newVariable<unsigned int>("do_out");
        flamegpu::AgentFunctionDescription fo = optional ? a.newFunction("Out_AppendTruncate", OptionalOut_AppendTruncate)
                                                         : a.newFunction("Out_AppendTruncate", Out_AppendTruncate);
        fo.setMessageOutputOptional(optional);
        fo.setMessageOutput(message);
        model.newLayer("OutLayer").addAgentFunction(fo);
        if (two) {
            flamegpu::AgentFunctionDescription fo2 = optional
                                                         ? a.newFunction("Out_AppendTruncate2", OptionalOut_AppendTruncate2)
                                                         : a.newFunction("Out_AppendTruncate2", Out_AppendTruncate2);
            fo2.setMessageOutputOptional(optional);
            fo2.setMessageOutput(message);
            model.newLayer("OutLayer2").addAgentFunction(fo2);
        }
        flamegpu::AgentFunctionDescription fi = two ? a.newFunction("In_AppendTruncate", In_AppendTruncate2)
                                                    : a.newFunction("In_AppendTruncate", In_AppendTruncate);
        fi.setMessageInput(message);
        model.newLayer("InLayer").addAgentFunction(fi);
    } else if (v == "append_resize") {
        auto message = model.newMessage("Message");
        message.newVariable<int>("x");
        auto a = model.newAgent("a");
        auto fa = a.newFunction("Out_1", Out_1);
        fa.setMessageOutput(message);
        auto b = model.newAgent("b");
        auto fb = b.newFunction("Out_2", Out_2);
        fb.setMessageOutput(message);
        auto c = model.newAgent("c");
        auto fc = c.newFunction("In", In_012);
        fc.setMessageInput(message);
        c.newVariable<unsigned int>("0", 0);
        c.newVariable<unsigned int>("1", 0);
        c.newVariable<unsigned int>("2", 0);
        model.newLayer("OutLayer").addAgentFunction(fa);
        model.newLayer("OutLayer2").addAgentFunction(fb);
        model.newLayer("InLayer").addAgentFunction(fc);
    } else if (v == "persistence_off" || v == "persistence_on") {
        auto message = model.newMessage<MessageBruteForce>("msg");
        message.newVariable<int>("id");
        message.setPersistent(v == "persistence_on");
        auto agent = model.newAgent("agent");
        agent.newVariable<int>("id");
        agent.newVariable<unsigned int>("count", 0);
        agent.newVariable<unsigned int>("sum", 0);
        auto af = agent.newFunction("out", out_simple);
        af.setMessageOutput("msg");
        af.setMessageOutputOptional(true);
        af.setFunctionCondition(EvenOnlyCondition);
        auto in = agent.newFunction("in", in_simple);
        in.setMessageInput("msg");
        model.newLayer().addAgentFunction(af);
        model.newLayer().addAgentFunction(in);
    } else {
        throw flamegpu::exception::InvalidArgument("unknown test model variant '" + v + "'");
    }
    return model;
}

#include "flamegpu/model_abi.h"
