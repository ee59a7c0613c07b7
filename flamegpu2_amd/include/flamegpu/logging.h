// Step and exit logs: LoggingConfig, StepLoggingConfig, AgentLoggingConfig,
// LogFrame / StepLogFrame / ExitLogFrame / AgentLogFrame and RunLog, with the
// surface of include/flamegpu/simulation/{LoggingConfig,AgentLoggingConfig,LogFrame}.h.
//
// The reference evaluates every logged reduction on the host after each logged step
// (src/flamegpu/simulation/CUDASimulation.cu:1949-2021), each a CUB reduction followed
// by a synchronising copy.  Here the simulation enqueues one fgx_reduce_stats per
// logged variable on its stream (sum, min, max and squared deviation together), writing
// into a device-side log buffer; the host reads the buffer back only when the log is
// asked for or the buffer is full.  Logged steps therefore add no host synchronisation.
#pragma once

#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <typeindex>
#include <utility>
#include <vector>

#include "flamegpu/model.h"

namespace flamegpu {

// include/flamegpu/simulation/AgentLoggingConfig_SumReturn.h: the type a sum is returned in.
template <typename T>
struct sum_input_t {
    typedef T result_t;
};
template <> struct sum_input_t<float> { typedef double result_t; };
template <> struct sum_input_t<double> { typedef double result_t; };
template <> struct sum_input_t<char> { typedef uint64_t result_t; };
template <> struct sum_input_t<uint8_t> { typedef uint64_t result_t; };
template <> struct sum_input_t<uint16_t> { typedef uint64_t result_t; };
template <> struct sum_input_t<uint32_t> { typedef uint64_t result_t; };
template <> struct sum_input_t<uint64_t> { typedef uint64_t result_t; };
template <> struct sum_input_t<int8_t> { typedef int64_t result_t; };
template <> struct sum_input_t<int16_t> { typedef int64_t result_t; };
template <> struct sum_input_t<int32_t> { typedef int64_t result_t; };
template <> struct sum_input_t<int64_t> { typedef int64_t result_t; };

namespace detail {
// A logged value: type, element count and bytes (the reference's detail::Any).
struct LogValue {
    std::type_index type = std::type_index(typeid(void));
    unsigned int elements = 1;
    std::vector<char> bytes;
    template <typename T>
    static LogValue of(T v) {
        LogValue r;
        r.type = std::type_index(typeid(T));
        r.bytes.resize(sizeof(T));
        std::memcpy(r.bytes.data(), &v, sizeof(T));
        return r;
    }
    template <typename T>
    T get(unsigned int i = 0) const {
        T v;
        std::memcpy(&v, bytes.data() + (size_t)i * sizeof(T), sizeof(T));
        return v;
    }
};

// fgx_reduce_stats type code of a variable type, or -1 when it has none.
inline int stats_type_code(const std::type_index& t) {
    if (t == std::type_index(typeid(float))) return 0;
    if (t == std::type_index(typeid(double))) return 1;
    if (t == std::type_index(typeid(int32_t))) return 2;
    if (t == std::type_index(typeid(uint32_t))) return 3;
    if (t == std::type_index(typeid(int64_t)) || t == std::type_index(typeid(long long))) return 4;
    if (t == std::type_index(typeid(uint64_t)) || t == std::type_index(typeid(unsigned long long))) return 5;
    return -1;
}
}  // namespace detail

class AgentLoggingConfig;

// include/flamegpu/simulation/LoggingConfig.h:27-170
class LoggingConfig {
 public:
    enum Reduction { Mean, StandardDev, Min, Max, Sum };
    static constexpr const char* toString(const Reduction& r) {
        return r == Mean ? "mean"
             : r == StandardDev ? "standard_deviation"
             : r == Min ? "min"
             : r == Max ? "max"
             : r == Sum ? "sum" : "unknown";
    }
    struct NameReduction {
        std::string name;
        Reduction reduction;
        bool operator<(const NameReduction& o) const {
            return name == o.name ? reduction < o.reduction : name < o.name;
        }
    };
    struct AgentEntry {
        std::set<NameReduction> reductions;
        bool count = false;
    };

    explicit LoggingConfig(const ModelDescription& model) : model_(model.data()) {}
    LoggingConfig(const LoggingConfig& o)
        : model_(o.model_), environment_(o.environment_), log_timing_(o.log_timing_) {
        // deep copy, so the copy's later log*() calls do not change the original
        for (auto& kv : o.agents_) agents_[kv.first] = std::make_shared<AgentEntry>(*kv.second);
    }
    virtual ~LoggingConfig() = default;

    // src/flamegpu/simulation/LoggingConfig.cu: agent()
    AgentLoggingConfig agent(const std::string& agent_name, const std::string& agent_state = ModelData::DEFAULT_STATE());
    // src/flamegpu/simulation/LoggingConfig.cu: logEnvironment()
    void logEnvironment(const std::string& property_name) {
        if (!model_->env.props.count(property_name))
            throw exception::InvalidEnvProperty("Environment property '" + property_name +
                                                "' was not found in the model description, in LoggingConfig::logEnvironment()");
        if (!environment_.insert(property_name).second)
            throw exception::InvalidEnvProperty("Environment property '" + property_name +
                                                "' has already been marked for logging, in LoggingConfig::logEnvironment()");
    }
    void logTiming(bool doLogTiming) { log_timing_ = doLogTiming; }

    // internal
    const std::shared_ptr<ModelData>& model() const { return model_; }
    const std::set<std::string>& environment() const { return environment_; }
    const std::map<std::pair<std::string, std::string>, std::shared_ptr<AgentEntry>>& agents() const { return agents_; }
    bool timing() const { return log_timing_; }

 protected:
    std::shared_ptr<ModelData> model_;
    std::set<std::string> environment_;
    std::map<std::pair<std::string, std::string>, std::shared_ptr<AgentEntry>> agents_;
    bool log_timing_ = false;
};

// include/flamegpu/simulation/LoggingConfig.h:172-215
class StepLoggingConfig : public LoggingConfig {
 public:
    explicit StepLoggingConfig(const ModelDescription& model) : LoggingConfig(model) {}
    StepLoggingConfig(const StepLoggingConfig& o) : LoggingConfig(o), frequency_(o.frequency_) {}
    explicit StepLoggingConfig(const LoggingConfig& o) : LoggingConfig(o) {}
    void setFrequency(unsigned int steps) { frequency_ = steps; }
    unsigned int frequency() const { return frequency_; }

 private:
    unsigned int frequency_ = 1;
};

// include/flamegpu/simulation/AgentLoggingConfig.h:20-110;
// src/flamegpu/simulation/AgentLoggingConfig.cu: log() validation.
class AgentLoggingConfig {
 public:
    AgentLoggingConfig(std::shared_ptr<AgentData> agent, std::shared_ptr<LoggingConfig::AgentEntry> entry)
        : agent_(std::move(agent)), entry_(std::move(entry)) {}
    void logCount() { entry_->count = true; }
    template <typename T>
    void logMean(const std::string& v) { log<T>(v, LoggingConfig::Mean, "Mean"); }
    template <typename T>
    void logStandardDev(const std::string& v) { log<T>(v, LoggingConfig::StandardDev, "StandardDev"); }
    template <typename T>
    void logMin(const std::string& v) { log<T>(v, LoggingConfig::Min, "Min"); }
    template <typename T>
    void logMax(const std::string& v) { log<T>(v, LoggingConfig::Max, "Max"); }
    template <typename T>
    void logSum(const std::string& v) { log<T>(v, LoggingConfig::Sum, "Sum"); }

 private:
    template <typename T>
    void log(const std::string& name, LoggingConfig::Reduction r, const char* method) {
        auto it = agent_->variables.find(name);
        if (it == agent_->variables.end())
            throw exception::InvalidAgentVar("Agent ('" + agent_->name + "') variable '" + name +
                                             "' was not found in the model description, in AgentLoggingConfig::log" + method + "()");
        if (it->second.type != std::type_index(typeid(T)))
            throw exception::InvalidVarType("Agent ('" + agent_->name + "') variable '" + name +
                                            "' has a different type, in AgentLoggingConfig::log" + method + "()");
        if (it->second.elements != 1)
            throw exception::InvalidVarType("Agent ('" + agent_->name + "') variable '" + name +
                                            "' is an array variable, in AgentLoggingConfig::log" + method + "()");
        if (detail::stats_type_code(it->second.type) < 0)
            throw exception::UnsupportedVarType("Agent ('" + agent_->name + "') variable '" + name +
                                                "': logged reductions support float, double and 32/64-bit integers");
        if (!entry_->reductions.insert({name, r}).second)
            throw exception::InvalidArgument("Agent ('" + agent_->name + "') variable '" + name + "' " + method +
                                             " has already been marked for logging, in AgentLoggingConfig::log" + method + "()");
    }
    std::shared_ptr<AgentData> agent_;
    std::shared_ptr<LoggingConfig::AgentEntry> entry_;
};

inline AgentLoggingConfig LoggingConfig::agent(const std::string& agent_name, const std::string& agent_state) {
    auto a = model_->agents.find(agent_name);
    if (a == model_->agents.end())
        throw exception::InvalidAgentName("Agent '" + agent_name + "' was not found in the model description, in LoggingConfig::agent()");
    if (!a->second->states.count(agent_state))
        throw exception::InvalidAgentState("State '" + agent_state + "' was not found within agent '" + agent_name +
                                           "' in the model description, in LoggingConfig::agent()");
    auto& e = agents_[{agent_name, agent_state}];
    if (!e) e = std::make_shared<AgentEntry>();
    return AgentLoggingConfig(a->second, e);
}

// include/flamegpu/simulation/LogFrame.h:230-340
class AgentLogFrame {
 public:
    AgentLogFrame(const std::map<LoggingConfig::NameReduction, detail::LogValue>& data, unsigned int count)
        : data_(data), count_(count) {}
    unsigned int getCount() const {
        if (count_ != UINT_MAX) return count_;
        throw exception::InvalidOperation("Count of agents in state was not found in the log, in AgentLogFrame::getCount()");
    }
    template <typename T>
    T getMin(const std::string& v) const { return typed<T>(v, LoggingConfig::Min, "getMin"); }
    template <typename T>
    T getMax(const std::string& v) const { return typed<T>(v, LoggingConfig::Max, "getMax"); }
    template <typename T>
    typename sum_input_t<T>::result_t getSum(const std::string& v) const {
        return typed<typename sum_input_t<T>::result_t>(v, LoggingConfig::Sum, "getSum");
    }
    double getMean(const std::string& v) const { return find(v, LoggingConfig::Mean, "getMean").get<double>(); }
    double getStandardDev(const std::string& v) const {
        return find(v, LoggingConfig::StandardDev, "getStandardDev").get<double>();
    }

 private:
    const detail::LogValue& find(const std::string& v, LoggingConfig::Reduction r, const char* method) const {
        auto it = data_.find({v, r});
        if (it == data_.end())
            throw exception::InvalidAgentVar(std::string(LoggingConfig::toString(r)) + " of agent variable '" + v +
                                             "' was not found in the log, in AgentLogFrame::" + method + "()");
        return it->second;
    }
    template <typename T>
    T typed(const std::string& v, LoggingConfig::Reduction r, const char* method) const {
        const auto& x = find(v, r, method);
        if (x.type != std::type_index(typeid(T)))
            throw exception::InvalidVarType("Agent variable '" + v + "' was logged with a different type, in AgentLogFrame::" +
                                            method + "()");
        return x.get<T>();
    }
    const std::map<LoggingConfig::NameReduction, detail::LogValue>& data_;
    unsigned int count_;
};

// include/flamegpu/simulation/LogFrame.h:24-160
class LogFrame {
 public:
    typedef std::map<LoggingConfig::NameReduction, detail::LogValue> AgentData_;
    typedef std::map<std::pair<std::string, std::string>, std::pair<AgentData_, unsigned int>> Agents;

    unsigned int getStepCount() const { return step_count; }
    bool hasEnvironmentProperty(const std::string& name) const { return environment.count(name) != 0; }
    template <typename T>
    T getEnvironmentProperty(const std::string& name) const {
        const auto& v = env(name, typeid(T));
        if (v.elements != 1)
            throw exception::InvalidEnvPropertyType("Environment property '" + name +
                                                    "' is an array, use alternate function with array interface");
        return v.get<T>();
    }
    template <typename T, unsigned int N>
    std::array<T, N> getEnvironmentProperty(const std::string& name) const {
        const auto& v = env(name, typeid(T));
        if (v.elements != N)
            throw exception::InvalidEnvPropertyType("Environment property '" + name + "' has a different length");
        std::array<T, N> r;
        std::memcpy(r.data(), v.bytes.data(), sizeof(T) * N);
        return r;
    }
    AgentLogFrame getAgent(const std::string& agent_name, const std::string& state_name = ModelData::DEFAULT_STATE()) const {
        auto it = agents.find({agent_name, state_name});
        if (it == agents.end())
            throw exception::InvalidAgentState("Log data for agent '" + agent_name + "' state '" + state_name +
                                               "' was not found, in LogFrame::getAgent()");
        return AgentLogFrame(it->second.first, it->second.second);
    }
    const std::map<std::string, detail::LogValue>& getEnvironment() const { return environment; }
    const Agents& getAgents() const { return agents; }

    // internal
    std::map<std::string, detail::LogValue> environment;
    Agents agents;
    unsigned int step_count = 0;

 private:
    const detail::LogValue& env(const std::string& name, const std::type_info& t) const {
        auto it = environment.find(name);
        if (it == environment.end())
            throw exception::InvalidEnvProperty("Environment property '" + name + "' was not found in the log");
        if (it->second.type != std::type_index(t))
            throw exception::InvalidEnvPropertyType("Environment property '" + name + "' has a different type");
        return it->second;
    }
};

struct StepLogFrame : public LogFrame {
    double getStepTime() const { return step_time; }
    double step_time = 0.0;
};

struct ExitLogFrame : public LogFrame {
    double getRTCTime() const { return rtc_time; }
    double getInitTime() const { return init_time; }
    double getExitTime() const { return exit_time; }
    double getTotalTime() const { return total_time; }
    double rtc_time = 0.0, init_time = 0.0, exit_time = 0.0, total_time = 0.0;
};

// include/flamegpu/simulation/LogFrame.h:170-228
struct RunLog {
    struct PerformanceSpecs {
        std::string device_name;
        int device_cc_major = 0;
        int device_cc_minor = 0;
        int cuda_version = 0;
        bool seatbelts = true;
        std::string flamegpu_version = "fgx";
    };
    const ExitLogFrame& getExitLog() const { return exit; }
    const std::list<StepLogFrame>& getStepLog() const { return step; }
    uint64_t getRandomSeed() const { return random_seed; }
    unsigned int getStepLogFrequency() const { return step_log_frequency; }
    PerformanceSpecs getPerformanceSpecs() const { return performance_specs; }

    ExitLogFrame exit;
    std::list<StepLogFrame> step;
    uint64_t random_seed = 0;
    unsigned int step_log_frequency = 0;
    PerformanceSpecs performance_specs;
};

namespace detail {
template <typename T, typename Acc>
inline void decode_stats(const char* w, unsigned int n, LoggingConfig::Reduction r, LogValue& out) {
    Acc sum, mn, mx;
    double sq;
    std::memcpy(&sum, w, 8);
    std::memcpy(&mn, w + 8, 8);
    std::memcpy(&mx, w + 16, 8);
    std::memcpy(&sq, w + 24, 8);
    switch (r) {
        case LoggingConfig::Sum: out = LogValue::of<typename sum_input_t<T>::result_t>((typename sum_input_t<T>::result_t)sum); break;
        case LoggingConfig::Min: out = LogValue::of<T>((T)mn); break;
        case LoggingConfig::Max: out = LogValue::of<T>((T)mx); break;
        // include/flamegpu/simulation/AgentLoggingConfig.h: mean 0 and standard deviation 0 for an empty list
        case LoggingConfig::Mean: out = LogValue::of<double>(n ? (double)sum / (double)n : 0.0); break;
        case LoggingConfig::StandardDev: out = LogValue::of<double>(n ? std::sqrt(sq / (double)n) : 0.0); break;
    }
}
}  // namespace detail

namespace detail {
// JSON text of a run log, in the layout of src/flamegpu/io/JSONLogger.cu (config,
// performance_specs, steps[], exit; per frame step_index, env, agents[name][state][]
// with count and variables[name][reduction]).  Differs in one place: "count" is written
// as the number itself.
inline void json_value(std::string& o, const LogValue& v) {
    auto one = [&](unsigned int i) {
        char b[64];
        if (v.type == std::type_index(typeid(float))) snprintf(b, sizeof b, "%.9g", (double)v.get<float>(i));
        else if (v.type == std::type_index(typeid(double))) snprintf(b, sizeof b, "%.17g", v.get<double>(i));
        else if (v.type == std::type_index(typeid(int64_t))) snprintf(b, sizeof b, "%lld", (long long)v.get<int64_t>(i));
        else if (v.type == std::type_index(typeid(uint64_t))) snprintf(b, sizeof b, "%llu", (unsigned long long)v.get<uint64_t>(i));
        else if (v.type == std::type_index(typeid(int32_t))) snprintf(b, sizeof b, "%d", v.get<int32_t>(i));
        else if (v.type == std::type_index(typeid(uint32_t))) snprintf(b, sizeof b, "%u", v.get<uint32_t>(i));
        else if (v.type == std::type_index(typeid(int16_t))) snprintf(b, sizeof b, "%d", (int)v.get<int16_t>(i));
        else if (v.type == std::type_index(typeid(uint16_t))) snprintf(b, sizeof b, "%u", (unsigned)v.get<uint16_t>(i));
        else if (v.type == std::type_index(typeid(int8_t))) snprintf(b, sizeof b, "%d", (int)v.get<int8_t>(i));
        else if (v.type == std::type_index(typeid(uint8_t))) snprintf(b, sizeof b, "%u", (unsigned)v.get<uint8_t>(i));
        else if (v.type == std::type_index(typeid(char))) snprintf(b, sizeof b, "%d", (int)v.get<char>(i));
        else throw exception::InvalidArgument("log value of unsupported type");
        // non-finite values are not JSON numbers
        if (!std::strcmp(b, "nan") || !std::strcmp(b, "-nan") || !std::strcmp(b, "inf") || !std::strcmp(b, "-inf"))
            snprintf(b, sizeof b, "null");
        o += b;
    };
    if (v.elements == 1) {
        one(0);
        return;
    }
    o += "[";
    for (unsigned int i = 0; i < v.elements; ++i) {
        if (i) o += ",";
        one(i);
    }
    o += "]";
}

inline void json_frame_body(std::string& o, const LogFrame& f) {
    o += "\"step_index\":" + std::to_string(f.getStepCount());
    if (!f.getEnvironment().empty()) {
        o += ",\"env\":{";
        bool first = true;
        for (auto& kv : f.getEnvironment()) {
            if (!first) o += ",";
            first = false;
            o += "\"" + kv.first + "\":";
            json_value(o, kv.second);
        }
        o += "}";
    }
    if (!f.getAgents().empty()) {
        o += ",\"agents\":{";
        std::string cur_agent;
        bool first_agent = true;
        for (auto it = f.getAgents().begin(); it != f.getAgents().end(); ++it) {
            const auto& key = it->first;
            if (first_agent || key.first != cur_agent) {
                if (!first_agent) o += "},";
                o += "\"" + key.first + "\":{";
                cur_agent = key.first;
                first_agent = false;
            } else {
                o += ",";
            }
            o += "\"" + key.second + "\":[{";
            bool sep = false;
            if (it->second.second != UINT_MAX) {
                o += "\"count\":" + std::to_string(it->second.second);
                sep = true;
            }
            if (!it->second.first.empty()) {
                if (sep) o += ",";
                o += "\"variables\":{";
                std::string cur_var;
                bool first_var = true;
                for (auto& r : it->second.first) {
                    if (first_var || r.first.name != cur_var) {
                        if (!first_var) o += "},";
                        o += "\"" + r.first.name + "\":{";
                        cur_var = r.first.name;
                        first_var = false;
                    } else {
                        o += ",";
                    }
                    o += std::string("\"") + LoggingConfig::toString(r.first.reduction) + "\":";
                    json_value(o, r.second);
                }
                o += "}}";
            }
            o += "}]";
        }
        o += "}}";
    }
}

inline std::string run_log_json(const RunLog& log, bool steps, bool exit, bool step_time, bool exit_time) {
    std::string o = "{\"config\":{\"random_seed\":" + std::to_string(log.getRandomSeed()) + "}";
    char b[64];
    if (step_time || exit_time) {
        const auto& p = log.performance_specs;
        o += ",\"performance_specs\":{\"device_name\":\"" + p.device_name + "\",\"device_cc_major\":" +
             std::to_string(p.device_cc_major) + ",\"device_cc_minor\":" + std::to_string(p.device_cc_minor) +
             ",\"cuda_version\":" + std::to_string(p.cuda_version) + ",\"seatbelts\":" + (p.seatbelts ? "true" : "false") +
             ",\"flamegpu_version\":\"" + p.flamegpu_version + "\"}";
    }
    if (steps) {
        o += ",\"steps\":[";
        bool first = true;
        for (auto& s : log.getStepLog()) {
            if (!first) o += ",";
            first = false;
            o += "{";
            if (step_time) {
                snprintf(b, sizeof b, "\"step_time\":%.17g,", s.getStepTime());
                o += b;
            }
            json_frame_body(o, s);
            o += "}";
        }
        o += "]";
    }
    if (exit) {
        o += ",\"exit\":{";
        const auto& e = log.getExitLog();
        if (exit_time) {
            snprintf(b, sizeof b, "\"rtc_time\":%.17g,", e.getRTCTime());
            o += b;
            snprintf(b, sizeof b, "\"init_time\":%.17g,", e.getInitTime());
            o += b;
            snprintf(b, sizeof b, "\"exit_time\":%.17g,", e.getExitTime());
            o += b;
            snprintf(b, sizeof b, "\"total_time\":%.17g,", e.getTotalTime());
            o += b;
        }
        json_frame_body(o, e);
        o += "}";
    }
    o += "}";
    return o;
}
}  // namespace detail

}  // namespace flamegpu
