// Model description API: ModelDescription, AgentDescription, AgentFunctionDescription,
// LayerDescription, EnvironmentDescription and the message-type descriptions.
//
// Mirrors include/flamegpu/model/{ModelDescription,AgentDescription,
// AgentFunctionDescription,LayerDescription,EnvironmentDescription}.h and the
// message hosts (include/flamegpu/runtime/messaging/MessageSpatial3D/MessageSpatial3DHost.h:272-310 etc.)
// with the same names, argument meaning and exceptions.  The description
// objects are value handles onto shared model data, as in the reference.
#pragma once

#include <algorithm>
#include <array>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <type_traits>
#include <typeindex>
#include <vector>

#include "flamegpu/detail/core.h"

namespace flamegpu {

enum class MsgKind { None = 0, BruteForce = 1, Spatial2D = 2, Spatial3D = 3, Bucket = 4, Array = 5, Array2D = 6, Array3D = 7 };
inline bool is_array_kind(MsgKind k) { return k == MsgKind::Array || k == MsgKind::Array2D || k == MsgKind::Array3D; }

struct Variable {
    std::string name;
    std::type_index type = std::type_index(typeid(void));
    size_t type_size = 0;
    size_type elements = 1;
    std::vector<char> default_value;
    size_t bytes() const { return type_size * elements; }
};
// Variables are kept in name order, as the reference's std::map (include/flamegpu/model/Variable.h:128).
typedef std::map<std::string, Variable> VariableMap;

struct ModelData;
struct AgentData;
struct AgentFunctionData;
struct MessageData;
struct LayerData;
class HostAPI;
class AgentFunctionDescription;

typedef void (*FLAMEGPU_HOST_FUNCTION_POINTER)(HostAPI*);
typedef FLAMEGPU_HOST_FUNCTION_POINTER FLAMEGPU_INIT_FUNCTION_POINTER;
typedef FLAMEGPU_HOST_FUNCTION_POINTER FLAMEGPU_STEP_FUNCTION_POINTER;
typedef FLAMEGPU_HOST_FUNCTION_POINTER FLAMEGPU_EXIT_FUNCTION_POINTER;
typedef CONDITION_RESULT (*FLAMEGPU_EXIT_CONDITION_POINTER)(HostAPI*);

namespace detail {
typedef void (*LaunchFn)(const FnArgs& args, uint32_t grid, uint32_t block, uint32_t smem, void* stream);
typedef void (*CondLaunchFn)(const FnArgs& args, uint32_t* flags, uint32_t grid, void* stream);
// The layer kernel of a translation unit: one __global__ function that runs any of the
// unit's agent functions, chosen per block from a fused-layer image (fgx_layer_launch).
struct LayerKernel {
    const void* kernel = nullptr;                        // host stub of the layer kernel
    int (*index_of)(LaunchFn launch) = nullptr;          // a function's index in it, or -1
    uint32_t (*lds_bytes)(int index, uint32_t block) = nullptr;   // its dynamic LDS per block
};
typedef void (*LaunchSegFn)(const FnArgs& args, const SegDev* segs, uint32_t nsegs, uint32_t apt, uint32_t grid,
                            uint32_t block, void* stream);
template <typename C>
void launch_condition(const FnArgs& a, uint32_t* flags, uint32_t grid, void* stream);
template <typename T>
Variable make_variable(const std::string& name, size_type elements, const T* defaults) {
    static_assert(std::is_arithmetic<T>::value, "agent/message variables must be arithmetic types");
    Variable v;
    v.name = name;
    v.type = std::type_index(typeid(T));
    v.type_size = sizeof(T);
    v.elements = elements;
    v.default_value.assign(sizeof(T) * elements, 0);
    if (defaults) std::memcpy(v.default_value.data(), defaults, sizeof(T) * elements);
    return v;
}
}  // namespace detail

struct MessageData {
    std::string name;
    MsgKind kind = MsgKind::BruteForce;
    VariableMap variables;
    float radius = 0.0f;
    float min[3] = {0, 0, 0};
    float max[3] = {0, 0, 0};
    bool radius_set = false, min_set = false, max_set = false;
    bool persistent = false;
    // bucket: keys [lower, upper] (both inclusive; include/flamegpu/runtime/messaging/MessageBucket/MessageBucketHost.h)
    int32_t lower = 0;
    int32_t upper = 0x7FFFFFFF;
    bool upper_set = false;
    // array: dimensions (1-D uses dims[0]; unused ones 1)
    uint32_t dims[3] = {0, 1, 1};
    ModelData* model = nullptr;
};

struct AgentFunctionData {
    std::string name;
    AgentData* parent = nullptr;
    detail::LaunchFn launch = nullptr;
    detail::LaunchSegFn launch_seg = nullptr;   // batched (segmented) launch
    detail::CondLaunchFn condition = nullptr;   // agent function condition (NULL: none)
    MsgKind in_kind = MsgKind::None, out_kind = MsgKind::None;
    std::weak_ptr<MessageData> message_input, message_output;
    bool message_output_optional = false;
    std::weak_ptr<AgentData> agent_output;
    std::string agent_output_state;
    std::string initial_state, end_state;
    bool has_agent_death = false;
    std::vector<AgentFunctionData*> dependencies;   // this function depends on these
    std::vector<AgentFunctionData*> dependents;
    uint32_t stream_id = 0;
};

struct AgentData {
    std::string name;
    VariableMap variables;
    std::vector<std::string> declared;   // declaration order (without _id)
    std::set<std::string> states;
    std::string initial_state;
    bool explicit_states = false;
    std::map<std::string, std::shared_ptr<AgentFunctionData>> functions;
    unsigned int sortPeriod = 1;         // src/flamegpu/model/AgentData.cpp:17
    ModelData* model = nullptr;
};

struct LayerData {
    std::string name;
    size_t index = 0;
    std::vector<std::shared_ptr<AgentFunctionData>> agent_functions;
    std::vector<FLAMEGPU_HOST_FUNCTION_POINTER> host_functions;
};

struct EnvProp {
    std::type_index type = std::type_index(typeid(void));
    size_t type_size = 0;
    size_type elements = 1;
    size_t offset = 0;
    bool isConst = false;
};

struct EnvironmentData {
    std::map<std::string, EnvProp> props;
    std::vector<char> buffer;
};

struct ModelData {
    static const char* DEFAULT_STATE() { return "default"; }
    std::string name;
    std::map<std::string, std::shared_ptr<AgentData>> agents;
    std::map<std::string, std::shared_ptr<MessageData>> messages;
    std::vector<std::shared_ptr<LayerData>> layers;
    EnvironmentData env;
    std::vector<FLAMEGPU_INIT_FUNCTION_POINTER> init_functions;
    std::vector<FLAMEGPU_STEP_FUNCTION_POINTER> step_functions;
    std::vector<FLAMEGPU_EXIT_FUNCTION_POINTER> exit_functions;
    std::vector<FLAMEGPU_EXIT_CONDITION_POINTER> exit_conditions;
    std::vector<AgentFunctionData*> execution_roots;
    uint32_t next_stream_id = 0;
};

// ------------------------------------------------------------------ messages (host side)
struct MessageNone {
    static constexpr MsgKind kind = MsgKind::None;
    class In;
    class Out;
};

struct MessageBruteForce {
    static constexpr MsgKind kind = MsgKind::BruteForce;
    class In;
    class Out;
    // include/flamegpu/runtime/messaging/MessageBruteForce/MessageBruteForceHost.h
    class Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : d_(std::move(d)) {}
        template <typename T>
        void newVariable(const std::string& name) {
            add(detail::make_variable<T>(name, 1, nullptr));
        }
        template <typename T, size_type N>
        void newVariable(const std::string& name) {
            add(detail::make_variable<T>(name, N, nullptr));
        }
        std::string getName() const { return d_->name; }
        void setPersistent(bool p) { d_->persistent = p; }
        bool getPersistent() const { return d_->persistent; }
        bool hasVariable(const std::string& n) const { return d_->variables.count(n) != 0; }
        size_type getVariablesCount() const { return (size_type)d_->variables.size(); }
        std::shared_ptr<MessageData> data() const { return d_; }

     protected:
        void add(Variable v) {
            if (d_->variables.count(v.name))
                throw exception::InvalidMessageVar("Message '" + d_->name + "' already contains variable '" + v.name +
                                                   "', in MessageDescription::newVariable()");
            if (!v.name.empty() && v.name[0] == '_')
                throw exception::ReservedName("Message variable names starting with '_' are reserved");
            d_->variables.emplace(v.name, v);
        }
        std::shared_ptr<MessageData> d_;
    };
    typedef Description CDescription;
};

struct MessageSpatial2D {
    static constexpr MsgKind kind = MsgKind::Spatial2D;
    class In;
    class Out;
    // include/flamegpu/runtime/messaging/MessageSpatial2D/MessageSpatial2DHost.h
    class Description : public MessageBruteForce::Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : MessageBruteForce::Description(std::move(d)) {}
        void setRadius(float r) {
            if (!(r > 0.0f)) throw exception::InvalidArgument("Spatial messaging radius must be a positive value");
            d_->radius = r;
            d_->radius_set = true;
        }
        void setMin(float x, float y) {
            if (d_->max_set && (x >= d_->max[0] || y >= d_->max[1]))
                throw exception::InvalidArgument("Spatial messaging min bounds must be lower than max bounds");
            d_->min[0] = x; d_->min[1] = y;
            d_->min_set = true;
        }
        void setMax(float x, float y) {
            if (d_->min_set && (x <= d_->min[0] || y <= d_->min[1]))
                throw exception::InvalidArgument("Spatial messaging max bounds must be greater than min bounds");
            d_->max[0] = x; d_->max[1] = y;
            d_->max_set = true;
        }
        float getRadius() const { return d_->radius; }
        float getMinX() const { return d_->min[0]; }
        float getMinY() const { return d_->min[1]; }
        float getMaxX() const { return d_->max[0]; }
        float getMaxY() const { return d_->max[1]; }
    };
    typedef Description CDescription;
};

struct MessageSpatial3D {
    static constexpr MsgKind kind = MsgKind::Spatial3D;
    class In;
    class Out;
    // include/flamegpu/runtime/messaging/MessageSpatial3D/MessageSpatial3DHost.h:272-310
    class Description : public MessageBruteForce::Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : MessageBruteForce::Description(std::move(d)) {}
        void setRadius(float r) {
            if (!(r > 0.0f)) throw exception::InvalidArgument("Spatial messaging radius must be a positive value");
            d_->radius = r;
            d_->radius_set = true;
        }
        void setMin(float x, float y, float z) {
            if (d_->max_set && (x >= d_->max[0] || y >= d_->max[1] || z >= d_->max[2]))
                throw exception::InvalidArgument("Spatial messaging min bounds must be lower than max bounds");
            d_->min[0] = x; d_->min[1] = y; d_->min[2] = z;
            d_->min_set = true;
        }
        void setMax(float x, float y, float z) {
            if (d_->min_set && (x <= d_->min[0] || y <= d_->min[1] || z <= d_->min[2]))
                throw exception::InvalidArgument("Spatial messaging max bounds must be greater than min bounds");
            d_->max[0] = x; d_->max[1] = y; d_->max[2] = z;
            d_->max_set = true;
        }
        float getRadius() const { return d_->radius; }
        float getMinX() const { return d_->min[0]; }
        float getMinY() const { return d_->min[1]; }
        float getMinZ() const { return d_->min[2]; }
        float getMaxX() const { return d_->max[0]; }
        float getMaxY() const { return d_->max[1]; }
        float getMaxZ() const { return d_->max[2]; }
    };
    typedef Description CDescription;
};

// include/flamegpu/runtime/messaging/MessageBucket/MessageBucketHost.h;
// src/flamegpu/runtime/messaging/MessageBucket.cu:190-215 (bound validation)
struct MessageBucket {
    static constexpr MsgKind kind = MsgKind::Bucket;
    class In;
    class Out;
    class Description : public MessageBruteForce::Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : MessageBruteForce::Description(std::move(d)) {}
        void setLowerBound(int32_t key) {
            if (d_->upper_set && key >= d_->upper)
                throw exception::InvalidArgument("Bucket messaging minimum bound must be lower than upper bound");
            d_->lower = key;
        }
        void setUpperBound(int32_t key) {
            if (key <= d_->lower)
                throw exception::InvalidArgument("Bucket messaging upperBound bound must be greater than lower bound");
            d_->upper = key;
            d_->upper_set = true;
        }
        void setBounds(int32_t min, int32_t max) {
            if (min >= max) throw exception::InvalidArgument("Bucket messaging minimum bound must be lower than upper bound");
            d_->lower = min;
            d_->upper = max;
            d_->upper_set = true;
        }
        int32_t getLowerBound() const { return d_->lower; }
        int32_t getUpperBound() const { return d_->upper; }
    };
    typedef Description CDescription;
};

// include/flamegpu/runtime/messaging/MessageArray/MessageArrayHost.h; MessageArray.cu setLength
struct MessageArray {
    static constexpr MsgKind kind = MsgKind::Array;
    class In;
    class Out;
    class Description : public MessageBruteForce::Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : MessageBruteForce::Description(std::move(d)) {}
        void setLength(size_type len) {
            if (len == 0) throw exception::InvalidArgument("Array messaging length must not be zero");
            d_->dims[0] = len;
        }
        size_type getLength() const { return d_->dims[0]; }
    };
    typedef Description CDescription;
};

// include/flamegpu/runtime/messaging/MessageArray2D/MessageArray2DHost.h
struct MessageArray2D {
    static constexpr MsgKind kind = MsgKind::Array2D;
    class In;
    class Out;
    class Description : public MessageBruteForce::Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : MessageBruteForce::Description(std::move(d)) {}
        void setDimensions(size_type x, size_type y) {
            if (x == 0 || y == 0) throw exception::InvalidArgument("All dimensions must be above zero in array2D message");
            d_->dims[0] = x;
            d_->dims[1] = y;
        }
        void setDimensions(const std::array<size_type, 2>& d) { setDimensions(d[0], d[1]); }
        std::array<size_type, 2> getDimensions() const { return {d_->dims[0], d_->dims[1]}; }
        size_type getDimX() const { return d_->dims[0]; }
        size_type getDimY() const { return d_->dims[1]; }
    };
    typedef Description CDescription;
};

// include/flamegpu/runtime/messaging/MessageArray3D/MessageArray3DHost.h
struct MessageArray3D {
    static constexpr MsgKind kind = MsgKind::Array3D;
    class In;
    class Out;
    class Description : public MessageBruteForce::Description {
     public:
        explicit Description(std::shared_ptr<MessageData> d) : MessageBruteForce::Description(std::move(d)) {}
        void setDimensions(size_type x, size_type y, size_type z) {
            if (x == 0 || y == 0 || z == 0) throw exception::InvalidArgument("All dimensions must be above zero in array3D message");
            d_->dims[0] = x;
            d_->dims[1] = y;
            d_->dims[2] = z;
        }
        void setDimensions(const std::array<size_type, 3>& d) { setDimensions(d[0], d[1], d[2]); }
        std::array<size_type, 3> getDimensions() const { return {d_->dims[0], d_->dims[1], d_->dims[2]}; }
        size_type getDimX() const { return d_->dims[0]; }
        size_type getDimY() const { return d_->dims[1]; }
        size_type getDimZ() const { return d_->dims[2]; }
    };
    typedef Description CDescription;
};

// ------------------------------------------------------------------ agent functions
class AgentDescription;

class AgentFunctionDescription {
 public:
    explicit AgentFunctionDescription(std::shared_ptr<AgentFunctionData> d) : d_(std::move(d)) {}
    // include/flamegpu/model/AgentFunctionDescription.h:190-304
    void setInitialState(const std::string& s) {
        check_state(s, "setInitialState");
        d_->initial_state = s;
    }
    void setEndState(const std::string& s) {
        check_state(s, "setEndState");
        d_->end_state = s;
    }
    void setMessageInput(const std::string& name) { setMessageInput(find_message(name, "setMessageInput")); }
    void setMessageInput(const MessageBruteForce::Description& m) { setMessageInput(m.data()); }
    void setMessageOutput(const std::string& name) { setMessageOutput(find_message(name, "setMessageOutput")); }
    void setMessageOutput(const MessageBruteForce::Description& m) { setMessageOutput(m.data()); }
    void setMessageOutputOptional(bool o) { d_->message_output_optional = o; }
    void setAgentOutput(const std::string& agent_name, const std::string& state = ModelData::DEFAULT_STATE());
    void setAgentOutput(const AgentDescription& agent, const std::string& state = ModelData::DEFAULT_STATE());
    void setAllowAgentDeath(bool d) { d_->has_agent_death = d; }
    // include/flamegpu/model/AgentFunctionDescription.h: setFunctionCondition(FLAMEGPU_AGENT_FUNCTION_CONDITION)
    template <typename C>
    void setFunctionCondition(C) {
        d_->condition = &detail::launch_condition<C>;
    }
    bool hasFunctionCondition() const { return d_->condition != nullptr; }
    void dependsOn(const AgentFunctionDescription& other) {
        if (other.d_.get() == d_.get()) throw exception::InvalidDependencyGraph("A function cannot depend on itself");
        d_->dependencies.push_back(other.d_.get());
        other.d_->dependents.push_back(d_.get());
    }
    std::string getName() const { return d_->name; }
    std::string getInitialState() const { return d_->initial_state; }
    std::string getEndState() const { return d_->end_state; }
    bool getMessageOutputOptional() const { return d_->message_output_optional; }
    bool getAllowAgentDeath() const { return d_->has_agent_death; }
    bool hasMessageInput() const { return !d_->message_input.expired(); }
    bool hasMessageOutput() const { return !d_->message_output.expired(); }
    bool hasAgentOutput() const { return !d_->agent_output.expired(); }
    std::shared_ptr<AgentFunctionData> data() const { return d_; }

 private:
    void check_state(const std::string& s, const char* where) const {
        if (!d_->parent->states.count(s))
            throw exception::InvalidStateName("Agent ('" + d_->parent->name + "') does not contain state '" + s +
                                              "', in AgentFunctionDescription::" + where + "()");
    }
    std::shared_ptr<MessageData> find_message(const std::string& name, const char* where) const {
        auto& msgs = d_->parent->model->messages;
        auto it = msgs.find(name);
        if (it == msgs.end())
            throw exception::InvalidMessageName("Model ('" + d_->parent->model->name + "') does not contain message '" +
                                                name + "', in AgentFunctionDescription::" + where + "()");
        return it->second;
    }
    void setMessageInput(const std::shared_ptr<MessageData>& m) {
        if (d_->in_kind != m->kind)
            throw exception::InvalidMessageType("Message '" + m->name + "' is not of the input type of function '" +
                                                d_->name + "', in AgentFunctionDescription::setMessageInput()");
        if (auto o = d_->message_output.lock())
            if (o == m)
                throw exception::InvalidMessageName("Message '" + m->name + "' is already bound as message output in "
                                                    "agent function " + d_->name);
        d_->message_input = m;
    }
    void setMessageOutput(const std::shared_ptr<MessageData>& m) {
        if (d_->out_kind != m->kind)
            throw exception::InvalidMessageType("Message '" + m->name + "' is not of the output type of function '" +
                                                d_->name + "', in AgentFunctionDescription::setMessageOutput()");
        if (auto i = d_->message_input.lock())
            if (i == m)
                throw exception::InvalidMessageName("Message '" + m->name + "' is already bound as message input in "
                                                    "agent function " + d_->name);
        d_->message_output = m;
    }
    std::shared_ptr<AgentFunctionData> d_;
};
typedef AgentFunctionDescription CAgentFunctionDescription;

namespace detail {
template <typename F>
void launch_agent_function(const FnArgs& a, uint32_t grid, uint32_t block, uint32_t smem, void* stream);
template <typename F>
void launch_agent_function_seg(const FnArgs& a, const SegDev* segs, uint32_t nsegs, uint32_t apt, uint32_t grid,
                               uint32_t block, void* stream);
}

// ------------------------------------------------------------------ agents
class AgentDescription {
 public:
    AgentDescription(std::shared_ptr<AgentData> d) : d_(std::move(d)) {}  // NOLINT
    // include/flamegpu/model/AgentDescription.h:209-310
    void newState(const std::string& s) {
        if (!d_->explicit_states) {
            // The implicit "default" state is replaced by the first named state.
            d_->states.clear();
            d_->explicit_states = true;
            d_->initial_state = s;
            for (auto& f : d_->functions) {
                f.second->initial_state = s;
                f.second->end_state = s;
            }
        }
        if (d_->states.count(s))
            throw exception::InvalidStateName("Agent ('" + d_->name + "') already contains state '" + s + "'");
        d_->states.insert(s);
    }
    void setInitialState(const std::string& s) {
        if (!d_->states.count(s))
            throw exception::InvalidStateName("Agent ('" + d_->name + "') does not contain state '" + s + "'");
        d_->initial_state = s;
    }
    template <typename T>
    void newVariable(const std::string& name, T default_value = T()) {
        add(detail::make_variable<T>(name, 1, &default_value));
    }
    template <typename T, size_type N>
    void newVariable(const std::string& name, const std::array<T, N>& default_value = {}) {
        add(detail::make_variable<T>(name, N, default_value.data()));
    }
    template <typename F>
    AgentFunctionDescription newFunction(const std::string& name, F = F()) {
        if (d_->functions.count(name))
            throw exception::InvalidAgentFunc("Agent ('" + d_->name + "') already contains function '" + name +
                                              "', in AgentDescription::newFunction()");
        auto f = std::make_shared<AgentFunctionData>();
        f->name = name;
        f->parent = d_.get();
        f->launch = &detail::launch_agent_function<F>;
        f->launch_seg = &detail::launch_agent_function_seg<F>;
        f->in_kind = F::MsgIn::kind;
        f->out_kind = F::MsgOut::kind;
        f->initial_state = d_->initial_state;
        f->end_state = d_->initial_state;
        f->stream_id = d_->model->next_stream_id++;
        d_->functions.emplace(name, f);
        return AgentFunctionDescription(f);
    }
    AgentFunctionDescription Function(const std::string& name) const {
        auto it = d_->functions.find(name);
        if (it == d_->functions.end())
            throw exception::InvalidAgentFunc("Agent ('" + d_->name + "') does not contain function '" + name + "'");
        return AgentFunctionDescription(it->second);
    }
    void setSortPeriod(unsigned int p) { d_->sortPeriod = p; }
    unsigned int getSortPeriod() const { return d_->sortPeriod; }
    std::string getName() const { return d_->name; }
    std::string getInitialState() const { return d_->initial_state; }
    bool hasState(const std::string& s) const { return d_->states.count(s) != 0; }
    bool hasVariable(const std::string& n) const { return d_->variables.count(n) != 0; }
    bool hasFunction(const std::string& n) const { return d_->functions.count(n) != 0; }
    size_type getVariablesCount() const { return (size_type)d_->variables.size() - 1; }  // excludes _id
    size_type getFunctionsCount() const { return (size_type)d_->functions.size(); }
    std::type_index getVariableType(const std::string& n) const { return var(n).type; }
    size_t getVariableSize(const std::string& n) const { return var(n).type_size; }
    size_type getVariableLength(const std::string& n) const { return var(n).elements; }
    std::shared_ptr<AgentData> data() const { return d_; }

 private:
    const Variable& var(const std::string& n) const {
        auto it = d_->variables.find(n);
        if (it == d_->variables.end())
            throw exception::InvalidAgentVar("Agent ('" + d_->name + "') does not contain variable '" + n + "'");
        return it->second;
    }
    void add(Variable v) {
        if (!v.name.empty() && v.name[0] == '_')
            throw exception::ReservedName("Agent variable names cannot begin with '_', this is reserved for internal usage, "
                                          "in AgentDescription::newVariable()");
        if (d_->variables.count(v.name))
            throw exception::InvalidAgentVar("Agent ('" + d_->name + "') already contains variable '" + v.name +
                                             "', in AgentDescription::newVariable()");
        if (d_->variables.size() >= detail::MAX_AGENT_VARS)
            throw exception::InvalidAgentVar("Agent ('" + d_->name + "') has too many variables");
        d_->declared.push_back(v.name);
        d_->variables.emplace(v.name, v);
    }
    std::shared_ptr<AgentData> d_;
};
typedef AgentDescription CAgentDescription;

inline void AgentFunctionDescription::setAgentOutput(const std::string& agent_name, const std::string& state) {
    auto& agents = d_->parent->model->agents;
    auto it = agents.find(agent_name);
    if (it == agents.end())
        throw exception::InvalidAgentName("Model does not contain agent '" + agent_name +
                                          "', in AgentFunctionDescription::setAgentOutput()");
    if (!it->second->states.count(state))
        throw exception::InvalidStateName("Agent ('" + agent_name + "') does not contain state '" + state +
                                          "', in AgentFunctionDescription::setAgentOutput()");
    d_->agent_output = it->second;
    d_->agent_output_state = state;
}
inline void AgentFunctionDescription::setAgentOutput(const AgentDescription& agent, const std::string& state) {
    setAgentOutput(agent.getName(), state);
}

// ------------------------------------------------------------------ environment
class EnvironmentDescription {
 public:
    explicit EnvironmentDescription(ModelData* m) : m_(m) {}
    // include/flamegpu/model/EnvironmentDescription.h
    template <typename T>
    void newProperty(const std::string& name, T value, bool isConst = false) {
        add<T>(name, &value, 1, isConst);
    }
    template <typename T, size_type N>
    void newProperty(const std::string& name, const std::array<T, N>& value, bool isConst = false) {
        add<T>(name, value.data(), N, isConst);
    }
    template <typename T>
    T getProperty(const std::string& name) const {
        const EnvProp& p = prop<T>(name, 1);
        T v;
        std::memcpy(&v, m_->env.buffer.data() + p.offset, sizeof(T));
        return v;
    }
    template <typename T>
    T setProperty(const std::string& name, T value) {
        const EnvProp& p = prop<T>(name, 1);
        T old;
        std::memcpy(&old, m_->env.buffer.data() + p.offset, sizeof(T));
        std::memcpy(m_->env.buffer.data() + p.offset, &value, sizeof(T));
        return old;
    }
    bool hasProperty(const std::string& n) const { return m_->env.props.count(n) != 0; }

 private:
    template <typename T>
    const EnvProp& prop(const std::string& name, size_type n) const {
        auto it = m_->env.props.find(name);
        if (it == m_->env.props.end())
            throw exception::InvalidEnvProperty("Environment property '" + name + "' was not found");
        if (it->second.type != std::type_index(typeid(T)) || it->second.elements < n)
            throw exception::InvalidEnvPropertyType("Environment property '" + name + "' type mismatch");
        return it->second;
    }
    template <typename T>
    void add(const std::string& name, const T* v, size_type n, bool isConst) {
        static_assert(std::is_arithmetic<T>::value, "environment properties must be arithmetic");
        if (!name.empty() && name[0] == '_')
            throw exception::ReservedName("Environment property names cannot begin with '_'");
        if (m_->env.props.count(name))
            throw exception::DuplicateEnvProperty("Environment property '" + name + "' already exists");
        if (m_->env.props.size() >= detail::MAX_ENV_PROPS)
            throw exception::InvalidEnvProperty("Too many environment properties");
        auto& buf = m_->env.buffer;
        size_t off = (buf.size() + alignof(T) - 1) / alignof(T) * alignof(T);
        off = (off + 7) & ~size_t(7);
        buf.resize(off + sizeof(T) * n);
        std::memcpy(buf.data() + off, v, sizeof(T) * n);
        EnvProp p;
        p.type = std::type_index(typeid(T));
        p.type_size = sizeof(T);
        p.elements = n;
        p.offset = off;
        p.isConst = isConst;
        m_->env.props.emplace(name, p);
    }
    ModelData* m_;
};

// ------------------------------------------------------------------ layers
class LayerDescription {
 public:
    LayerDescription(ModelData* m, std::shared_ptr<LayerData> l) : m_(m), l_(std::move(l)) {}
    // src/flamegpu/model/LayerDescription.cpp:93-170 admission rules
    void addAgentFunction(const AgentFunctionDescription& f) { add(f.data()); }
    void addAgentFunction(const std::string& agent, const std::string& fn) {
        auto a = m_->agents.find(agent);
        if (a == m_->agents.end())
            throw exception::InvalidAgentFunc("Agent '" + agent + "' was not found, in LayerDescription::addAgentFunction()");
        auto f = a->second->functions.find(fn);
        if (f == a->second->functions.end())
            throw exception::InvalidAgentFunc("Agent function '" + fn + "' of agent '" + agent +
                                              "' was not found, in LayerDescription::addAgentFunction()");
        add(f->second);
    }
    template <typename F>
    void addAgentFunction(F = F()) {
        const detail::LaunchFn target = &detail::launch_agent_function<F>;
        std::shared_ptr<AgentFunctionData> match;
        for (auto& a : m_->agents)
            for (auto& f : a.second->functions)
                if (f.second->launch == target) {
                    if (match)
                        throw exception::InvalidAgentFunc("Agent function is registered by more than one agent; use "
                                                          "addAgentFunction(agent_name, function_name)");
                    match = f.second;
                }
        if (!match)
            throw exception::InvalidAgentFunc("Agent function was not found in the model, in LayerDescription::addAgentFunction()");
        add(match);
    }
    void addHostFunction(FLAMEGPU_HOST_FUNCTION_POINTER fn) {
        if (!l_->agent_functions.empty())
            throw exception::InvalidLayerMember("A layer containing agent functions may not also contain host functions");
        if (std::find(l_->host_functions.begin(), l_->host_functions.end(), fn) != l_->host_functions.end())
            throw exception::InvalidHostFunc("Host function already exists in the layer");
        l_->host_functions.push_back(fn);
    }
    std::string getName() const { return l_->name; }
    size_t getIndex() const { return l_->index; }
    size_type getAgentFunctionsCount() const { return (size_type)l_->agent_functions.size(); }

 private:
    void add(const std::shared_ptr<AgentFunctionData>& f) {
        if (!l_->host_functions.empty())
            throw exception::InvalidLayerMember("A layer containing host functions, may not also contain agent functions, "
                                                "in LayerDescription::addAgentFunction()");
        auto f_agent_out = f->agent_output.lock();
        auto f_msg_out = f->message_output.lock();
        auto f_msg_in = f->message_input.lock();
        for (const auto& b : l_->agent_functions) {
            if (b == f)
                throw exception::InvalidAgentFunc("Attempted to add agent function '" + f->name + "' owned by agent '" +
                                                  f->parent->name + "' to same layer twice");
            if (b->parent == f->parent &&
                (b->initial_state == f->initial_state || b->initial_state == f->end_state ||
                 b->end_state == f->initial_state || b->end_state == f->end_state))
                throw exception::InvalidAgentFunc("Agent function '" + f->name + "' owned by agent '" + f->parent->name +
                                                  "' cannot be added to this layer as agent function '" + b->name +
                                                  "' within the layer shares an input or output state");
            if (f_agent_out && f_agent_out.get() == b->parent && b->initial_state == f->agent_output_state)
                throw exception::InvalidLayerMember("Agent function '" + f->name + "' births into the state that '" +
                                                    b->name + "' in the same layer requires as input");
            if (auto b_out = b->agent_output.lock())
                if (b_out.get() == f->parent && f->initial_state == b->agent_output_state)
                    throw exception::InvalidLayerMember("Agent function '" + b->name + "' births into the state that '" +
                                                        f->name + "' requires as input");
            auto b_msg_out = b->message_output.lock();
            auto b_msg_in = b->message_input.lock();
            if (f_msg_out && (f_msg_out == b_msg_out || f_msg_out == b_msg_in))
                throw exception::InvalidLayerMember("Agent function '" + f->name + "' outputs message '" + f_msg_out->name +
                                                    "' used by '" + b->name + "' in the same layer");
            if (f_msg_in && f_msg_in == b_msg_out)
                throw exception::InvalidLayerMember("Agent function '" + f->name + "' reads message '" + f_msg_in->name +
                                                    "' output by '" + b->name + "' in the same layer");
        }
        l_->agent_functions.push_back(f);
    }
    ModelData* m_;
    std::shared_ptr<LayerData> l_;
};

// ------------------------------------------------------------------ model
class ModelDescription {
 public:
    explicit ModelDescription(const std::string& name) : m_(std::make_shared<ModelData>()) { m_->name = name; }
    // include/flamegpu/model/ModelDescription.h:47-273
    AgentDescription newAgent(const std::string& name) {
        if (m_->agents.count(name))
            throw exception::InvalidAgentName("Agent with name '" + name + "' already exists, in ModelDescription::newAgent()");
        auto a = std::make_shared<AgentData>();
        a->name = name;
        a->model = m_.get();
        a->initial_state = ModelData::DEFAULT_STATE();
        a->states.insert(a->initial_state);
        a->variables.emplace("_id", detail::make_variable<id_t>("_id", 1, nullptr));
        m_->agents.emplace(name, a);
        return AgentDescription(a);
    }
    AgentDescription Agent(const std::string& name) const {
        auto it = m_->agents.find(name);
        if (it == m_->agents.end())
            throw exception::InvalidAgentName("Agent ('" + name + "') was not found, in ModelDescription::Agent()");
        return AgentDescription(it->second);
    }
    bool hasAgent(const std::string& n) const { return m_->agents.count(n) != 0; }
    template <typename M>
    typename M::Description newMessage(const std::string& name) {
        static_assert(M::kind != MsgKind::None, "MessageNone cannot be instantiated");
        if (m_->messages.count(name))
            throw exception::InvalidMessageName("Message with name '" + name + "' already exists, in ModelDescription::newMessage()");
        auto d = std::make_shared<MessageData>();
        d->name = name;
        d->kind = M::kind;
        d->model = m_.get();
        if (M::kind == MsgKind::Spatial2D || M::kind == MsgKind::Spatial3D) {
            d->variables.emplace("x", detail::make_variable<float>("x", 1, nullptr));
            d->variables.emplace("y", detail::make_variable<float>("y", 1, nullptr));
            if (M::kind == MsgKind::Spatial3D) d->variables.emplace("z", detail::make_variable<float>("z", 1, nullptr));
        }
        // hidden variables: the bucket key (MessageBucket.cu Data()), the array element index
        // (MessageArray.cu Data(): "___INDEX")
        if (M::kind == MsgKind::Bucket) d->variables.emplace("_key", detail::make_variable<int32_t>("_key", 1, nullptr));
        if (is_array_kind(M::kind)) d->variables.emplace("___INDEX", detail::make_variable<uint32_t>("___INDEX", 1, nullptr));
        m_->messages.emplace(name, d);
        return typename M::Description(d);
    }
    MessageBruteForce::Description newMessage(const std::string& name) { return newMessage<MessageBruteForce>(name); }
    template <typename M = MessageBruteForce>
    typename M::Description Message(const std::string& name) const {
        auto it = m_->messages.find(name);
        if (it == m_->messages.end())
            throw exception::InvalidMessageName("Message ('" + name + "') was not found, in ModelDescription::Message()");
        if (M::kind != MsgKind::BruteForce && it->second->kind != M::kind)
            throw exception::InvalidMessageType("Message ('" + name + "') is not of the requested type");
        return typename M::Description(it->second);
    }
    bool hasMessage(const std::string& n) const { return m_->messages.count(n) != 0; }
    EnvironmentDescription Environment() const { return EnvironmentDescription(m_.get()); }
    LayerDescription newLayer(const std::string& name = "") {
        for (auto& l : m_->layers)
            if (!name.empty() && l->name == name)
                throw exception::InvalidFuncLayerIndx("Layer '" + name + "' already exists, in ModelDescription::newLayer()");
        auto l = std::make_shared<LayerData>();
        l->name = name;
        l->index = m_->layers.size();
        m_->layers.push_back(l);
        return LayerDescription(m_.get(), l);
    }
    LayerDescription Layer(size_type i) const {
        if (i >= m_->layers.size()) throw exception::OutOfBoundsException("Layer index out of bounds");
        return LayerDescription(m_.get(), m_->layers[i]);
    }
    LayerDescription Layer(const std::string& name) const {
        for (auto& l : m_->layers)
            if (l->name == name) return LayerDescription(m_.get(), l);
        throw exception::InvalidFuncLayerIndx("Layer '" + name + "' was not found");
    }
    size_type getLayersCount() const { return (size_type)m_->layers.size(); }
    size_type getAgentsCount() const { return (size_type)m_->agents.size(); }
    size_type getMessagesCount() const { return (size_type)m_->messages.size(); }
    void addInitFunction(FLAMEGPU_INIT_FUNCTION_POINTER f) { m_->init_functions.push_back(f); }
    void addStepFunction(FLAMEGPU_STEP_FUNCTION_POINTER f) { m_->step_functions.push_back(f); }
    void addExitFunction(FLAMEGPU_EXIT_FUNCTION_POINTER f) { m_->exit_functions.push_back(f); }
    void addExitCondition(FLAMEGPU_EXIT_CONDITION_POINTER f) { m_->exit_conditions.push_back(f); }
    void addExecutionRoot(const AgentFunctionDescription& f) { m_->execution_roots.push_back(f.data().get()); }
    // src/flamegpu/model/DependencyGraph.cpp:100-228: layer = longest dependency path from a
    // root; a function the layer rules refuse goes to a new layer appended at the end.
    void generateLayers() {
        if (m_->execution_roots.empty())
            throw exception::InvalidDependencyGraph("Dependency graph has no execution roots, in ModelDescription::generateLayers()");
        std::map<AgentFunctionData*, int> depth;
        std::set<AgentFunctionData*> on_stack;
        std::function<void(AgentFunctionData*, int)> visit = [&](AgentFunctionData* n, int d) {
            if (on_stack.count(n)) throw exception::InvalidDependencyGraph("Dependency graph contains a cycle");
            auto it = depth.find(n);
            if (it != depth.end() && it->second >= d) return;
            depth[n] = d;
            on_stack.insert(n);
            for (auto c : n->dependents) visit(c, d + 1);
            on_stack.erase(n);
        };
        for (auto r : m_->execution_roots) visit(r, 0);
        std::vector<std::vector<AgentFunctionData*>> ideal;
        std::function<void(AgentFunctionData*)> place = [&](AgentFunctionData* n) {
            const size_t d = (size_t)depth[n];
            while (ideal.size() <= d) ideal.emplace_back();
            if (std::find(ideal[d].begin(), ideal[d].end(), n) == ideal[d].end()) ideal[d].push_back(n);
            for (auto c : n->dependents) place(c);
        };
        for (auto r : m_->execution_roots) place(r);
        auto shared_of = [&](AgentFunctionData* n) {
            return n->parent->functions.at(n->name);
        };
        for (auto& lay : ideal) {
            LayerDescription layer = newLayer();
            for (auto n : lay) {
                try {
                    layer.addAgentFunction(AgentFunctionDescription(shared_of(n)));
                } catch (const exception::InvalidAgentFunc&) {
                    newLayer().addAgentFunction(AgentFunctionDescription(shared_of(n)));
                } catch (const exception::InvalidLayerMember&) {
                    newLayer().addAgentFunction(AgentFunctionDescription(shared_of(n)));
                }
            }
        }
    }
    std::string getName() const { return m_->name; }
    std::shared_ptr<ModelData> data() const { return m_; }

 private:
    std::shared_ptr<ModelData> m_;
};

}  // namespace flamegpu
