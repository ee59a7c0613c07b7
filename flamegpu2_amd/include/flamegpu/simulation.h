// Host runtime: AgentVector, HostAPI, CUDASimulation, RunPlan(Vector), CUDAEnsemble.
//
// Same surface as include/flamegpu/simulation/{CUDASimulation,AgentVector,RunPlan,
// RunPlanVector,CUDAEnsemble}.h and include/flamegpu/runtime/HostAPI.h.  The step
// (src/flamegpu/simulation/CUDASimulation.cu:575-1109) is re-designed:
//   * all device work goes through the libfgx C ABI, stream-ordered, with every
//     agent/message count kept in device memory (no per-compaction readback);
//   * the host only synchronises once per step, and only when a model can grow
//     (births, state transitions), to size buffers for the step's upper bound;
//   * functions of one layer run on their own streams (fork/join events);
//   * whole steps are captured into a hipGraph and replayed.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "flamegpu/runtime/device.h"
#include "flamegpu/logging.h"
#include "flamegpu/io.h"

namespace flamegpu {

class CUDASimulation;

// ------------------------------------------------------------------ AgentVector
// include/flamegpu/simulation/AgentVector.h: host-side population, stored as
// one contiguous array per variable.
class AgentVector {
 public:
    class Agent {
     public:
        Agent(AgentVector* v, size_t i) : v_(v), i_(i) {}
        template <typename T>
        void setVariable(const std::string& name, T value) {
            std::memcpy(v_->elem<T>(name, 1, 0, i_), &value, sizeof(T));
        }
        template <typename T, size_type N>
        void setVariable(const std::string& name, const std::array<T, N>& value) {
            std::memcpy(v_->elem<T>(name, N, 0, i_), value.data(), sizeof(T) * N);
        }
        template <typename T, size_type N = 0>
        void setVariable(const std::string& name, unsigned int index, T value) {
            const size_type len = v_->length(name);
            if (index >= len) throw exception::OutOfBoundsException("Index out of bounds for variable '" + name + "'");
            std::memcpy(v_->elem<T>(name, len, index, i_), &value, sizeof(T));
        }
        template <typename T>
        T getVariable(const std::string& name) const {
            T r;
            std::memcpy(&r, v_->elem<T>(name, 1, 0, i_), sizeof(T));
            return r;
        }
        template <typename T, size_type N>
        std::array<T, N> getVariable(const std::string& name) const {
            std::array<T, N> r;
            std::memcpy(r.data(), v_->elem<T>(name, N, 0, i_), sizeof(T) * N);
            return r;
        }
        template <typename T, size_type N = 0>
        T getVariable(const std::string& name, unsigned int index) const {
            const size_type len = v_->length(name);
            if (index >= len) throw exception::OutOfBoundsException("Index out of bounds for variable '" + name + "'");
            T r;
            std::memcpy(&r, v_->elem<T>(name, len, index, i_), sizeof(T));
            return r;
        }
        id_t getID() const { return getVariable<id_t>("_id"); }
        size_t getIndex() const { return i_; }

     private:
        AgentVector* v_;
        size_t i_;
    };
    typedef Agent CAgent;

    class iterator {
     public:
        iterator(AgentVector* v, size_t i) : v_(v), i_(i) {}
        Agent operator*() const { return Agent(v_, i_); }
        iterator& operator++() {
            ++i_;
            return *this;
        }
        bool operator!=(const iterator& o) const { return i_ != o.i_; }
        bool operator==(const iterator& o) const { return i_ == o.i_; }

     private:
        AgentVector* v_;
        size_t i_;
    };

    explicit AgentVector(const CAgentDescription& agent, size_t count = 0) : a_(agent.data()) {
        for (auto& kv : a_->variables) data_[kv.first];
        resize(count);
    }
    Agent operator[](size_t i) { return Agent(this, i); }
    Agent at(size_t i) {
        if (i >= n_) throw exception::OutOfBoundsException("AgentVector index out of bounds");
        return Agent(this, i);
    }
    Agent front() { return at(0); }
    Agent back() { return at(n_ - 1); }
    iterator begin() { return iterator(this, 0); }
    iterator end() { return iterator(this, n_); }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    void clear() { resize(0); }
    void push_back() { resize(n_ + 1); }
    void resize(size_t n) {
        for (auto& kv : a_->variables) {
            const Variable& v = kv.second;
            auto& buf = data_[kv.first];
            const size_t b = v.bytes();
            buf.resize(n * b);
            for (size_t i = n_; i < n; ++i) std::memcpy(buf.data() + i * b, v.default_value.data(), b);
        }
        n_ = n;
    }
    std::string getAgentName() const { return a_->name; }
    const std::shared_ptr<AgentData>& agentData() const { return a_; }
    // Raw access used by the runtime.
    std::vector<char>& raw(const std::string& name) { return data_.at(name); }

 private:
    size_type length(const std::string& name) const {
        auto it = a_->variables.find(name);
        if (it == a_->variables.end())
            throw exception::InvalidAgentVar("Agent '" + a_->name + "' does not contain variable '" + name + "'");
        return it->second.elements;
    }
    template <typename T>
    char* elem(const std::string& name, size_type n, size_type e, size_t i) {
        auto it = a_->variables.find(name);
        if (it == a_->variables.end())
            throw exception::InvalidAgentVar("Agent '" + a_->name + "' does not contain variable '" + name + "'");
        if (it->second.type != std::type_index(typeid(T)) || it->second.elements != n)
            throw exception::InvalidVarType("Variable '" + name + "' accessed with the wrong type or length");
        if (i >= n_) throw exception::OutOfBoundsException("AgentVector index out of bounds");
        return data_.at(name).data() + i * it->second.bytes() + (size_t)e * sizeof(T);
    }
    std::shared_ptr<AgentData> a_;
    std::map<std::string, std::vector<char>> data_;
    size_t n_ = 0;
};

namespace detail {

struct VarBuf {
    std::string name;
    uint32_t bytes = 0;
    void* cur = nullptr;
    void* alt = nullptr;
    std::vector<char> def;
};

struct StateList {
    std::string state;
    std::vector<VarBuf> vars;     // name order, includes _id
    int id_slot = -1;
    uint32_t cap = 0;
    uint32_t* d_count = nullptr;
    uint32_t h_count = 0;
    void* sort_scratch = nullptr;
    size_t sort_scratch_bytes = 0;
    uint32_t sort_scratch_keys = 0;   // the largest key range the scratch was sized for
    int var_index(const std::string& n) const {
        for (size_t i = 0; i < vars.size(); ++i)
            if (vars[i].name == n) return (int)i;
        return -1;
    }
};

struct AgentRT {
    std::shared_ptr<AgentData> d;
    std::map<std::string, StateList> states;
    uint32_t* d_next_id = nullptr;
    uint32_t h_next_id = 1;
};

struct MessageRT {
    std::shared_ptr<MessageData> d;
    std::vector<VarBuf> vars;
    uint32_t cap = 0;
    uint32_t* d_count = nullptr;
    fgx_spatial_meta meta{};
    bool spatial = false;
    MsgKind kind = MsgKind::BruteForce;
    int32_t key_min = 0;            // bucket: keys [key_min, key_min + keys)
    uint32_t keys = 0;
    uint32_t dims[3] = {1, 1, 1};   // array: dimensions; length = their product
    uint32_t length = 0;
    uint32_t* counts = nullptr;     // array: per-element write counts (build scratch)
    uint32_t* pbm = nullptr;        // spatial bin table / bucket start table
    void* build_scratch = nullptr;
    size_t build_scratch_bytes = 0;
    void* packed = nullptr;         // spatial: 16-byte {x, y, z, id} record per message, rebuilt with the index
    bool truncate = true;
    bool pbm_dirty = true;
    int var_index(const std::string& n) const {
        for (size_t i = 0; i < vars.size(); ++i)
            if (vars[i].name == n) return (int)i;
        return -1;
    }
};

struct FnRT {
    std::shared_ptr<AgentFunctionData> d;
    AgentRT* agent = nullptr;
    std::string in_state, end_state;
    MessageRT* min = nullptr;
    MessageRT* mout = nullptr;
    AgentRT* out_agent = nullptr;
    std::string out_state;
    int sort_dims = 0;              // 2/3 when the function triggers the auto spatial sort
    int fused_fn = -2;              // index in the layer kernel (-1: launched on its own, -2: not looked up)
    bool swap_transition = false;   // this step: the end state is known empty, so the lists swap
    uint32_t cap = 0;               // per-launch buffers sized for this many agents
    uint32_t* death_flags = nullptr;
    uint32_t* birth_flags = nullptr;
    uint32_t* msg_flags = nullptr;
    uint32_t* d_launch_count = nullptr;
    const uint32_t* launch_count = nullptr;   // count the last launch ran over: d_launch_count or the list's own count
    bool msg_count_in_kernel = false;        // the last launch wrote the message count itself
    std::vector<VarBuf> newbuf;     // output agent variables (cur only)
    void* death_scratch = nullptr;
    size_t death_scratch_bytes = 0;
    void* births_scratch = nullptr;
    size_t births_scratch_bytes = 0;
    void* append_scratch = nullptr;
    size_t append_scratch_bytes = 0;
    void* trans_scratch = nullptr;
    size_t trans_scratch_bytes = 0;
    // function condition: flags (!condition), disabled count, work word, partition scratch
    uint32_t* cond_flags = nullptr;
    uint32_t* d_first = nullptr;
    uint32_t* d_work = nullptr;
    uint32_t* d_id_slots = nullptr;   // agent output: set when a parent read its child's ID in-kernel
    void* cond_scratch = nullptr;
    size_t cond_scratch_bytes = 0;
    void* stream = nullptr;
};

struct KernelTime {
    double ms = 0.0;
    uint64_t count = 0;
};

}  // namespace detail

// ------------------------------------------------------------------ host API
class HostAgentAPI {
 public:
    HostAgentAPI(CUDASimulation& s, detail::AgentRT& a, detail::StateList& l) : s_(s), a_(a), l_(l) {}
    // include/flamegpu/runtime/agent/HostAgentAPI.cuh:174
    unsigned int count();
    // include/flamegpu/runtime/agent/HostAgentAPI.cuh:689 (deterministic fixed-shape tree on the device)
    template <typename InT, typename OutT = InT>
    OutT sum(const std::string& variable);
    template <typename T>
    T min(const std::string& variable);
    template <typename T>
    T max(const std::string& variable);
    // include/flamegpu/runtime/agent/HostAgentAPI.cuh: meanStandardDeviation (population standard deviation)
    template <typename T>
    std::pair<double, double> meanStandardDeviation(const std::string& variable);
    template <typename T>
    std::vector<T> getVariableValues(const std::string& variable);
    // include/flamegpu/runtime/agent/HostAgentAPI.cuh: newAgent (HostNewAgentAPI).  The agent
    // starts from the variable defaults with the next free ID, and joins its state list when
    // the host function returns.
    AgentVector::Agent newAgent();

 private:
    template <typename T>
    const detail::VarBuf& var(const std::string& name);
    // sum, min, max and squared deviation of a variable on the device (fgx_reduce_stats)
    template <typename T>
    void stats(const std::string& name, char out[32]);
    CUDASimulation& s_;
    detail::AgentRT& a_;
    detail::StateList& l_;
};

class HostEnvironment {
 public:
    explicit HostEnvironment(CUDASimulation& s) : s_(s) {}
    template <typename T>
    T getProperty(const std::string& name) const;
    template <typename T>
    T setProperty(const std::string& name, T value);
    // include/flamegpu/runtime/environment/HostEnvironment.cuh: array properties
    template <typename T, unsigned int N>
    std::array<T, N> getProperty(const std::string& name) const;
    template <typename T, unsigned int N>
    std::array<T, N> setProperty(const std::string& name, const std::array<T, N>& value);

 private:
    CUDASimulation& s_;
};

class HostRandom {
 public:
    explicit HostRandom(uint64_t seed) : g_(seed), seed_(seed) {}
    // include/flamegpu/runtime/random/HostRandom.cuh: getSeed
    uint64_t getSeed() const { return seed_; }
    template <typename T>
    T uniform() {
        return std::uniform_real_distribution<T>(0, 1)(g_);
    }
    template <typename T>
    T uniform(T mn, T mx) {
        if constexpr (std::is_integral<T>::value)
            return std::uniform_int_distribution<T>(mn, mx)(g_);
        else
            return std::uniform_real_distribution<T>(mn, mx)(g_);
    }
    template <typename T>
    T normal() {
        return std::normal_distribution<T>(0, 1)(g_);
    }
    void setSeed(uint64_t s) {
        g_.seed(s);
        seed_ = s;
    }

 private:
    std::mt19937_64 g_;
    uint64_t seed_ = 0;
};

class HostAPI {
 public:
    HostAPI(CUDASimulation& s, uint64_t seed) : environment(s), random(seed), s_(s) {}
    HostAgentAPI agent(const std::string& name, const std::string& state = "");
    unsigned int getStepCounter() const;
    HostEnvironment environment;
    HostRandom random;

 private:
    CUDASimulation& s_;
};

// include/flamegpu/runtime/HostAPI_macros.h
#define FLAMEGPU_INIT_FUNCTION(funcName)                      \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU);        \
    flamegpu::FLAMEGPU_INIT_FUNCTION_POINTER funcName = funcName##_impl; \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU)
#define FLAMEGPU_STEP_FUNCTION(funcName)                      \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU);        \
    flamegpu::FLAMEGPU_STEP_FUNCTION_POINTER funcName = funcName##_impl; \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU)
#define FLAMEGPU_EXIT_FUNCTION(funcName)                      \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU);        \
    flamegpu::FLAMEGPU_EXIT_FUNCTION_POINTER funcName = funcName##_impl; \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU)
#define FLAMEGPU_HOST_FUNCTION(funcName)                      \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU);        \
    flamegpu::FLAMEGPU_HOST_FUNCTION_POINTER funcName = funcName##_impl; \
    void funcName##_impl(flamegpu::HostAPI* FLAMEGPU)
#define FLAMEGPU_EXIT_CONDITION(funcName)                     \
    flamegpu::CONDITION_RESULT funcName##_impl(flamegpu::HostAPI* FLAMEGPU); \
    flamegpu::FLAMEGPU_EXIT_CONDITION_POINTER funcName = funcName##_impl; \
    flamegpu::CONDITION_RESULT funcName##_impl(flamegpu::HostAPI* FLAMEGPU)

// ------------------------------------------------------------------ run plans
// include/flamegpu/simulation/RunPlan.h
class RunPlan {
 public:
    explicit RunPlan(const ModelDescription& model) : m_(model.data()) {}
    void setRandomSimulationSeed(uint64_t s) { seed_ = s; }
    void setSteps(unsigned int s) { steps_ = s; }
    void setOutputSubdirectory(const std::string& d) { subdir_ = d; }
    template <typename T>
    void setProperty(const std::string& name, T value) {
        auto it = m_->env.props.find(name);
        if (it == m_->env.props.end()) throw exception::InvalidEnvProperty("RunPlan: unknown environment property '" + name + "'");
        if (it->second.type != std::type_index(typeid(T)))
            throw exception::InvalidEnvPropertyType("RunPlan: property '" + name + "' type mismatch");
        std::vector<char> b(sizeof(T));
        std::memcpy(b.data(), &value, sizeof(T));
        props_[name] = b;
    }
    template <typename T>
    T getProperty(const std::string& name) const {
        auto it = props_.find(name);
        if (it != props_.end()) {
            T v;
            std::memcpy(&v, it->second.data(), sizeof(T));
            return v;
        }
        return EnvironmentDescription(m_.get()).getProperty<T>(name);
    }
    uint64_t getRandomSimulationSeed() const { return seed_; }
    unsigned int getSteps() const { return steps_; }
    const std::shared_ptr<ModelData>& model() const { return m_; }
    std::string getOutputSubdirectory() const { return subdir_; }
    const std::map<std::string, std::vector<char>>& properties() const { return props_; }

 private:
    std::shared_ptr<ModelData> m_;
    uint64_t seed_ = 0;
    unsigned int steps_ = 1;
    std::string subdir_;
    std::map<std::string, std::vector<char>> props_;
};

// include/flamegpu/simulation/RunPlanVector.h:36-139
class RunPlanVector : public std::vector<RunPlan> {
 public:
    RunPlanVector(const ModelDescription& model, unsigned int n) : std::vector<RunPlan>(n, RunPlan(model)), rng_(0) {}
    void setRandomSimulationSeed(uint64_t initial, unsigned int step = 0) {
        uint64_t s = initial;
        for (auto& p : *this) {
            p.setRandomSimulationSeed(s);
            s += step;
        }
    }
    void setSteps(unsigned int s) {
        for (auto& p : *this) p.setSteps(s);
    }
    void setOutputSubdirectory(const std::string& d) {
        for (auto& p : *this) p.setOutputSubdirectory(d);
    }
    template <typename T>
    void setProperty(const std::string& name, T value) {
        for (auto& p : *this) p.setProperty<T>(name, value);
    }
    template <typename T>
    void setPropertyLerpRange(const std::string& name, T mn, T mx) {
        const size_t n = size();
        for (size_t i = 0; i < n; ++i) {
            const double a = n > 1 ? (double)i / (double)(n - 1) : 0.0;
            const double v = (double)mn * (1.0 - a) + (double)mx * a;
            (*this)[i].template setProperty<T>(name, std::is_integral<T>::value ? (T)std::llround(v) : (T)v);
        }
    }
    void setRandomPropertySeed(uint64_t s) { rng_.seed(s); }
    template <typename T>
    void setPropertyUniformRandom(const std::string& name, T mn, T mx) {
        for (auto& p : *this) {
            if constexpr (std::is_integral<T>::value)
                p.template setProperty<T>(name, std::uniform_int_distribution<T>(mn, mx)(rng_));
            else
                p.template setProperty<T>(name, std::uniform_real_distribution<T>(mn, mx)(rng_));
        }
    }

 private:
    std::mt19937_64 rng_;
};

// ------------------------------------------------------------------ simulation
class CUDASimulation {
 public:
    // include/flamegpu/simulation/Simulation.h:37-70
    struct Config {
        std::string input_file;
        uint64_t random_seed = 0;
        unsigned int steps = 1;
        bool verbose = false;
        bool timing = false;
        bool truncate_log_files = true;
    };
    // include/flamegpu/simulation/CUDASimulation.h:100-129 (+ graph capture switch)
    struct CUDAConfig {
        int device_id = 0;
        bool inLayerConcurrency = true;
        bool graphs = true;          // capture whole steps as hipGraphs where the model allows
        bool is_ensemble = false;
    };

    explicit CUDASimulation(const ModelDescription& model, int argc = 0, const char** argv = nullptr)
        : m_(model.data()) {
        env_host_ = m_->env.buffer;
        config_.random_seed = (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count();
        if (argc && argv) initialise(argc, argv);
    }
    CUDASimulation(const CUDASimulation&) = delete;
    CUDASimulation& operator=(const CUDASimulation&) = delete;
    ~CUDASimulation() { release(); }

    Config& SimulationConfig() { return config_; }
    const Config& getSimulationConfig() const { return config_; }
    CUDAConfig& CUDAConfig() { return cuda_config_; }
    const struct CUDAConfig& getCUDAConfig() const { return cuda_config_; }

    // src/flamegpu/simulation/Simulation.cu:228-350 (the step/seed/device/timing subset)
    void initialise(int argc, const char** argv) {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            auto next = [&](const char* what) -> std::string {
                if (i + 1 >= argc) throw exception::InvalidArgument(std::string("missing value for ") + what);
                return argv[++i];
            };
            if (a == "-s" || a == "--steps") config_.steps = (unsigned)std::stoul(next("--steps"));
            else if (a == "-r" || a == "--random") config_.random_seed = std::stoull(next("--random"));
            else if (a == "-d" || a == "--device") cuda_config_.device_id = std::stoi(next("--device"));
            else if (a == "-i" || a == "--in") {
                // loaded here so that later arguments override the file's configuration
                // (src/flamegpu/simulation/Simulation.cu: checkArgs)
                config_.input_file = next("--in");
                load_state(config_.input_file);
            }
            else if (a == "-t" || a == "--timing") config_.timing = true;
            else if (a == "-v" || a == "--verbose") config_.verbose = true;
            else if (a == "-q" || a == "--quiet") config_.verbose = false;
            else if (a == "--no-graphs") cuda_config_.graphs = false;
        }
    }
    // src/flamegpu/simulation/Simulation.cu:55-83: loads SimulationConfig().input_file once
    void applyConfig() {
        if (!config_.input_file.empty() && config_.input_file != loaded_input_file_) load_state(config_.input_file);
        ensure_init();
    }
    // src/flamegpu/simulation/Simulation.cu exportData: configuration, step count,
    // environment and every agent population to an .xml or .json state file.
    void exportData(const std::string& path, bool prettyPrint = true);

    // src/flamegpu/simulation/CUDASimulation.cu:575-642
    bool step();
    // src/flamegpu/simulation/CUDASimulation.cu:1220-1344
    void simulate() {
        const auto t0 = std::chrono::steady_clock::now();
        ensure_init();
        reset_log();
        run_init_functions();
        // the initial state is step log frame 0 (src/flamegpu/simulation/CUDASimulation.cu:1262-1263)
        process_step_log(elapsed_init_);
        for (unsigned int s = 0; s < config_.steps || config_.steps == 0; ++s) {
            if (!step()) break;
            if (config_.steps == 0 && s > 1000000000u) break;
        }
        run_exit_functions();
        synchronize();
        check_device_error();
        elapsed_sim_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        process_exit_log();
    }
    // The two host phases of simulate() around steps run elsewhere (an ensemble's batch):
    // init functions, then, once the run's final populations are set, exit functions and
    // the exit log as after `steps` steps.
    void runInitPhase() {
        ensure_init();
        reset_log();
        run_init_functions();
        synchronize();
        check_device_error();
    }
    void runExitPhase(unsigned int steps) {
        ensure_init();
        step_ = steps;
        detail::fgx_check(fgx_fill_u32(d_step_, steps, 1, stream_), "step counter");
        run_exit_functions();
        synchronize();
        check_device_error();
        process_exit_log();
    }
    void simulate(const RunPlan& plan) {
        config_.random_seed = plan.getRandomSimulationSeed();
        config_.steps = plan.getSteps();
        for (auto& kv : plan.properties()) set_env_raw(kv.first, kv.second);
        simulate();
    }
    // src/flamegpu/simulation/CUDASimulation.cu:1476-1491
    void setStepLog(const StepLoggingConfig& cfg) {
        check_log_model(cfg, "setStepLog");
        flush_step_log();   // frames already taken keep the configuration they were taken with
        step_log_cfg_ = std::make_shared<StepLoggingConfig>(cfg);
        step_plan_.reset();
        run_log_.step_log_frequency = cfg.frequency();
    }
    void setExitLog(const LoggingConfig& cfg) {
        check_log_model(cfg, "setExitLog");
        exit_log_cfg_ = std::make_shared<LoggingConfig>(cfg);
        exit_plan_.reset();
    }
    // src/flamegpu/simulation/CUDASimulation.cu:2023-2025.  Reads back logged steps
    // still held in device memory first.
    const RunLog& getRunLog() const {
        const_cast<CUDASimulation*>(this)->flush_step_log();
        return run_log_;
    }
    // src/flamegpu/simulation/Simulation.cu exportLog + src/flamegpu/io/JSONLogger.cu (JSON only)
    void exportLog(const std::string& path, bool steps, bool exit, bool stepTime, bool exitTime) {
        const std::string j = detail::run_log_json(getRunLog(), steps, exit, stepTime, exitTime);
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f) throw exception::InvalidArgument("Unable to open file '" + path + "' for writing");
        const size_t w = std::fwrite(j.data(), 1, j.size(), f);
        std::fclose(f);
        if (w != j.size()) throw exception::InvalidArgument("Unable to write file '" + path + "'");
    }
    double getElapsedTimeInitFunctions() const { return elapsed_init_; }
    double getElapsedTimeExitFunctions() const { return elapsed_exit_; }
    // src/flamegpu/simulation/CUDASimulation.cu:1418-1452
    void setPopulationData(AgentVector& pop, const std::string& state = "");
    void getPopulationData(AgentVector& pop, const std::string& state = "");
    void synchronize() {
        if (stream_) detail::fgx_check(fgx_stream_synchronize(stream_), "stream synchronize");
    }
    unsigned int getStepCounter() const { return step_; }
    void resetStepCounter() { step_ = 0; if (d_step_) detail::fgx_check(fgx_fill_u32(d_step_, 0, 1, stream_), "reset step"); }
    std::vector<double> getElapsedTimeSteps() const { return step_times_; }
    double getElapsedTimeStep(unsigned int s) const { return step_times_.at(s); }
    double getElapsedTimeSimulation() const { return elapsed_sim_; }
    template <typename T>
    void setEnvironmentProperty(const std::string& name, T value) {
        const EnvProp& p = env_prop<T>(name);
        std::memcpy(env_host_.data() + p.offset, &value, sizeof(T));
        env_dirty_ = true;
    }
    template <typename T>
    T getEnvironmentProperty(const std::string& name) const {
        const EnvProp& p = env_prop<T>(name);
        T v;
        std::memcpy(&v, env_host_.data() + p.offset, sizeof(T));
        return v;
    }
    template <typename T, unsigned int N>
    std::array<T, N> getEnvironmentProperty(const std::string& name) const {
        const EnvProp& p = env_prop<T>(name);
        if (p.elements != N)
            throw exception::OutOfBoundsException("Environment property '" + name + "' does not have " + std::to_string(N) + " elements");
        std::array<T, N> v;
        std::memcpy(v.data(), env_host_.data() + p.offset, sizeof(T) * N);
        return v;
    }
    template <typename T, unsigned int N>
    void setEnvironmentProperty(const std::string& name, const std::array<T, N>& value) {
        const EnvProp& p = env_prop<T>(name);
        if (p.elements != N)
            throw exception::OutOfBoundsException("Environment property '" + name + "' does not have " + std::to_string(N) + " elements");
        std::memcpy(env_host_.data() + p.offset, value.data(), sizeof(T) * N);
        env_dirty_ = true;
    }

    // ---- extensions used by the benchmark and the model C ABI
    void setKernelTiming(bool on) { timing_ = on; }
    // Stage spatial message neighbourhoods in LDS before iterating them (default off: on
    // MI355X the L2-resident HBM path is faster for the bundled models); off reads HBM directly.
    void setNeighbourhoodStaging(bool on) {
        stage_enabled_ = on;
        clear_graphs();
    }
    // Spatial message iteration reads each wave's bin-table ranges from a table it copies to
    // LDS when the iterator is built (default on); off reads the bin table from HBM at every
    // unit change.  Results are identical either way.
    void setUnitTables(bool on) {
        unit_tables_ = on;
        clear_graphs();
    }
    // The layer kernel of the model's translation unit (set by the model library): layers
    // of several independent functions then run as one launch when in-layer concurrency
    // is on.  setLayerFusion(false) keeps one launch per function, on forked streams.
    void setLayerKernel(const detail::LayerKernel& k) {
        layer_kernel_ = k;
        for (auto& f : fns_) f.fused_fn = -2;
        clear_graphs();
    }
    void setLayerFusion(bool on) {
        layer_fusion_ = on;
        clear_graphs();
    }
    // Fused layer launches enqueued so far (graph replays not counted).
    uint64_t fusedLayerLaunches() const { return fused_launches_; }
    uint64_t layerImages() {
        ensure_init();
        uint32_t n = 0;
        detail::fgx_check(fgx_layer_space_info(ctx_, &n, nullptr), "layer images");
        return n;
    }
    // Spatial message bins of width radius x mult, mult = 1 (the reference's grid) or 0.5.
    // With half-width bins the iterators search 2 bins either side: the same in-radius
    // neighbours from 5^3 bins of 1/8 the volume (125/216 of the candidates), visited in
    // the reference's order over the finer grid.  No counterpart in the reference; set
    // before the first step.
    void setBinWidthMultiplier(float mult);
    float getBinWidthMultiplier() const { return bin_mult_; }
    const std::map<std::string, detail::KernelTime>& getKernelTimes() {
        collect_timings();
        return ktimes_;
    }
    void resetKernelTimes() {
        collect_timings();
        ktimes_.clear();
    }
    uint32_t messageBins(const std::string& msg) {
        ensure_init();
        auto& m = msgs_.at(msg);
        return m.spatial ? m.meta.bins : 0u;
    }
    void getPBM(const std::string& msg, uint32_t* host) {
        auto& m = msgs_.at(msg);
        if (!m.spatial) throw exception::InvalidMessageType("message '" + msg + "' has no spatial index");
        synchronize();
        detail::fgx_check(fgx_memcpy_d2h(host, m.pbm, (size_t)(m.meta.bins + 1) * 4, stream_), "PBM readback");
        synchronize();
    }
    void* stream() { ensure_init(); return stream_; }
    bool graphActive() const { return !graphs_.empty(); }
    uint64_t seed() const { return config_.random_seed; }

    // internal (HostAPI)
    detail::AgentRT& agent_rt(const std::string& name) {
        ensure_init();
        auto it = agents_.find(name);
        if (it == agents_.end()) throw exception::InvalidAgentName("Agent '" + name + "' was not found");
        return it->second;
    }
    fgx_ctx* ctx() { return ctx_; }
    void* reduce_scratch(size_t bytes) {
        if (bytes > reduce_scratch_bytes_) {
            if (reduce_scratch_) fgx_free(ctx_, reduce_scratch_);
            detail::fgx_check(fgx_malloc(ctx_, bytes, &reduce_scratch_), "reduction scratch");
            detail::fgx_check(fgx_scratch_clear(reduce_scratch_, bytes, stream_), "reduction scratch clear");
            reduce_scratch_bytes_ = bytes;
        }
        return reduce_scratch_;
    }
    void* reduce_out() { return d_reduce_out_; }
    void check_device_error();
    void mark_env_dirty() { env_dirty_ = true; }
    // agents created by host functions (HostAgentAPI::newAgent), appended when the function returns
    AgentVector& host_new_agents(detail::AgentRT& a, detail::StateList& l) {
        auto& p = host_new_[&l];
        if (!p.v) {
            p.a = &a;
            p.v.reset(new AgentVector(AgentDescription(a.d)));
        }
        return *p.v;
    }
    bool host_agents_pending(const detail::AgentRT* a = nullptr) const {
        for (auto& kv : host_new_)
            if (kv.second.v && !kv.second.v->empty() && (!a || kv.second.a == a)) return true;
        return false;
    }
    void flush_host_agents();
    std::vector<char>& env_host() { return env_host_; }
    const std::shared_ptr<ModelData>& model() const { return m_; }
    void sync_counts();

 private:
    template <typename T>
    const EnvProp& env_prop(const std::string& name) const {
        auto it = m_->env.props.find(name);
        if (it == m_->env.props.end()) throw exception::InvalidEnvProperty("Environment property '" + name + "' was not found");
        if (it->second.type != std::type_index(typeid(T)))
            throw exception::InvalidEnvPropertyType("Environment property '" + name + "' type mismatch");
        return it->second;
    }
    void load_state(const std::string& path);
    std::string loaded_input_file_;
    void set_env_raw(const std::string& name, const std::vector<char>& b) {
        auto it = m_->env.props.find(name);
        if (it == m_->env.props.end()) throw exception::InvalidEnvProperty("Environment property '" + name + "' was not found");
        std::memcpy(env_host_.data() + it->second.offset, b.data(), std::min(b.size(), it->second.type_size * it->second.elements));
        env_dirty_ = true;
    }

    void ensure_init();
    void release();
    void alloc(void** p, size_t bytes) { detail::fgx_check(fgx_malloc(ctx_, bytes ? bytes : 4, p), "device allocation"); }
    void free_dev(void*& p) {
        if (p) fgx_free(ctx_, p);
        p = nullptr;
    }
    void* new_scratch(int op, uint32_t n, uint32_t bins, size_t* bytes) {
        detail::fgx_check(fgx_scratch_bytes(op, n, bins, bytes), "scratch sizing");
        void* p = nullptr;
        alloc(&p, *bytes);
        detail::fgx_check(fgx_scratch_clear(p, *bytes, stream_), "scratch clear");
        sorted_scratch_.erase(p);
        return p;
    }
    void grow_list(detail::StateList& l, uint32_t cap);
    void grow_message(detail::MessageRT& m, uint32_t cap);
    void size_function(detail::FnRT& f, uint32_t cap);
    bool ensure_capacity();
    bool model_can_grow() const;
    uint32_t prepare_function(detail::FnRT& f, void* st, detail::FnArgs& a);
    bool launch_layer_fused(const LayerData& layer);
    static constexpr uint32_t kFnBlock = 256;   // threads per block of every agent-function launch
    void plan_transitions(bool counted);
    std::vector<uint8_t> swap_plan() const {
        std::vector<uint8_t> p;
        for (auto& f : fns_) p.push_back(f.swap_transition ? 1 : 0);
        return p;
    }
    void enqueue_step();
    void enqueue_layer(const LayerData& layer);
    void launch_function(detail::FnRT& f, void* st);
    void fill_condition_args(detail::FnRT& f, detail::StateList& l, detail::FnArgs& a);
    void stage_neighbourhood(detail::FnRT& f, detail::StateList& l, detail::FnArgs& a);
    static uint32_t stage_smem_bytes(const detail::FnArgs& a) {
        return a.stage_ok ? (2u * (a.stage_max_entries + 1u) + a.stage_nvars * a.stage_cap) * 4u : 0u;
    }
    void unmap_function(detail::FnRT& f, void* st);
    void enqueue_end_of_step();
    void run_init_functions();
    void run_exit_functions();
    void upload_env();
    // step / exit logs
    struct LogVarPlan {
        std::string name;
        int type;               // fgx_reduce_stats type code
        std::type_index vtype;  // variable type
        size_t off;             // byte offset of its 4 words in a frame record
    };
    struct LogAgentPlan {
        std::string agent, state;
        bool count;
        size_t count_off;
        std::vector<LogVarPlan> vars;
        std::set<LoggingConfig::NameReduction> reductions;
    };
    struct LogPlan {
        std::vector<LogAgentPlan> agents;
        size_t record_bytes = 0;
    };
    struct PendingLogFrame {
        unsigned int step;
        double step_time;
        std::map<std::string, detail::LogValue> env;
    };
    void check_log_model(const LoggingConfig& cfg, const char* where) const {
        if (cfg.model().get() != m_.get())
            throw exception::InvalidArgument(std::string("Model descriptions attached to LoggingConfig and CUDASimulation do not match, in CUDASimulation::") +
                                             where + "()");
    }
    std::unique_ptr<LogPlan> make_log_plan(const LoggingConfig& cfg);
    void enqueue_log_frame(const LogPlan& plan, char* record);
    void decode_log_frame(const LoggingConfig& cfg, const LogPlan& plan, const char* record, LogFrame& f) const;
    std::map<std::string, detail::LogValue> log_env(const LoggingConfig& cfg) const;
    void reset_log();
    void process_step_log(double step_time);
    void flush_step_log();
    void process_exit_log();
    void tick(const std::string& label, void* st, bool begin);
    void collect_timings();

    std::shared_ptr<ModelData> m_;
    Config config_;
    struct CUDAConfig cuda_config_;
    bool initialised_ = false;
    fgx_ctx* ctx_ = nullptr;
    void* stream_ = nullptr;
    std::map<std::string, detail::AgentRT> agents_;
    std::map<std::string, detail::MessageRT> msgs_;
    std::vector<detail::FnRT> fns_;
    std::map<const AgentFunctionData*, size_t> fn_index_;
    std::vector<char> env_host_;
    char* d_env_ = nullptr;
    bool env_dirty_ = true;
    uint32_t* d_small_ = nullptr;    // [0..2] error, [3] step, [4] dummy counter, [8..] counts
    uint32_t* d_error_ = nullptr;
    uint32_t* d_step_ = nullptr;
    uint32_t* d_dummy_ = nullptr;
    uint32_t* d_overflow_ = nullptr;   // set when an append would pass a list's capacity (an internal bound error)
    uint32_t small_used_ = 8;
    uint32_t* h_small_ = nullptr;    // pinned mirror
    void* d_reduce_out_ = nullptr;
    void* d_pages_ = nullptr;        // [0, 128): zero page, [128, 256): sink page
    void* reduce_scratch_ = nullptr;
    size_t reduce_scratch_bytes_ = 0;
    unsigned int step_ = 0;
    std::vector<double> step_times_;
    double elapsed_sim_ = 0.0;
    double elapsed_init_ = 0.0;
    double elapsed_exit_ = 0.0;
    // logs
    std::shared_ptr<StepLoggingConfig> step_log_cfg_;
    std::shared_ptr<LoggingConfig> exit_log_cfg_;
    std::unique_ptr<LogPlan> step_plan_, exit_plan_;
    RunLog run_log_;
    char* d_log_ = nullptr;          // frame records of logged steps not yet read back
    size_t log_cap_frames_ = 0;
    std::vector<PendingLogFrame> log_pending_;
    char* d_exit_log_ = nullptr;
    size_t exit_log_bytes_ = 0;
    void* log_scratch_ = nullptr;
    std::unique_ptr<HostAPI> host_api_;
    bool init_functions_ran_ = false;
    // graphs
    struct GraphEntry {
        std::vector<void*> start, end;
        std::vector<uint8_t> swaps;   // the step's transition plan (which transitions swap lists)
        bool stale = false;           // recorded a layer image upload: captured again next time
        bool concurrent = true;       // captured with in-layer concurrency on
        fgx_graph* g = nullptr;
        uint64_t last_use = 0;
    };
    // Step graphs kept per simulation: one per double-buffer arrangement at step start.
    // Models alternate between a few arrangements; beyond this many the least recently
    // used graph is replaced (and a warning is printed once).
    static constexpr size_t kMaxStepGraphs = 8;
    std::vector<GraphEntry> graphs_;
    uint64_t graph_uses_ = 0;
    bool graph_evict_warned_ = false;
    bool graph_eligible_ = false;
    bool populations_set_ = false;   // setPopulationData since the last step: re-check capacities
    struct HostNew {
        detail::AgentRT* a = nullptr;
        std::unique_ptr<AgentVector> v;
    };
    std::map<detail::StateList*, HostNew> host_new_;   // pending host-created agents, by state list
    void clear_graphs() {
        for (auto& g : graphs_) fgx_graph_destroy(g.g);
        graphs_.clear();
    }
    std::vector<void*> buffer_state() const;
    void set_buffer_state(const std::vector<void*>& st);
    std::vector<void*> join_events_;
    void* fork_event_ = nullptr;
    // index builds beside the agent sorts of a layer
    void* build_stream_ = nullptr;
    void* build_fork_event_ = nullptr;
    void* build_join_event_ = nullptr;
    // kernel timing
    bool timing_ = false;
    bool stage_enabled_ = false;
    bool unit_tables_ = true;
    detail::LayerKernel layer_kernel_{};
    bool layer_fusion_ = true;
    bool layer_space_warned_ = false;
    bool new_layer_images_ = false;   // the enqueue being captured uploaded a layer image
    // sort scratch buffers that a sort has used: a capture that holds a buffer's first sort
    // also holds its clearing launch, so the graph is captured again after one run
    std::unordered_set<const void*> sorted_scratch_;
    bool first_sorts_ = false;        // the enqueue being captured holds a buffer's first sort
    void note_sort(const void* scratch) {
        if (sorted_scratch_.insert(scratch).second) first_sorts_ = true;
    }
    uint64_t fused_launches_ = 0;
    float bin_mult_ = 1.0f;
    struct PendingTime {
        std::string label;
        void* start;
        void* stop;
    };
    std::vector<PendingTime> pending_;
    std::vector<void*> free_events_;
    std::map<std::string, void*> open_events_;
    std::map<std::string, detail::KernelTime> ktimes_;
    friend class HostAgentAPI;
};

// ------------------------------------------------------------------ implementation
inline void CUDASimulation::setBinWidthMultiplier(float mult) {
    if (!(mult == 1.0f || mult == 0.5f))
        throw exception::InvalidArgument("bin-width multiplier must be 1 or 0.5");
    if (step_ > 0) throw exception::InvalidOperation("setBinWidthMultiplier() must be called before the first step");
    bin_mult_ = mult;
    if (!initialised_) return;
    // re-grid every spatial list: new bin tables, and build scratch sized for the new bin count
    for (auto& kv : msgs_) {
        detail::MessageRT& m = kv.second;
        if (!m.spatial) continue;
        const fgx_spatial_meta old = m.meta;
        detail::fgx_check(fgx_spatial_meta_init_mult(old.dims, old.min, old.max, old.radius, mult, &m.meta),
                          "spatial message metadata");
        void* old_pbm = m.pbm;
        free_dev(old_pbm);
        alloc((void**)&m.pbm, (size_t)(m.meta.bins + 1) * 4);
        detail::fgx_check(fgx_fill_u32(m.pbm, 0, m.meta.bins + 1, stream_), "PBM init");
        if (m.cap) {
            free_dev(m.build_scratch);
            m.build_scratch = new_scratch(FGX_OP_SPATIAL_BUILD, m.cap, m.meta.bins, &m.build_scratch_bytes);
        }
    }
    clear_graphs();
}

inline void CUDASimulation::ensure_init() {
    if (initialised_) return;
    detail::fgx_check(fgx_ctx_create(cuda_config_.device_id, &ctx_), "fgx_ctx_create");
    // every build and sort scratch buffer comes from new_scratch (zeroed) and serves only builds/sorts
    detail::fgx_check(fgx_ctx_set_option(ctx_, FGX_OPT_KEEP_SCRATCH_ZERO, 1), "context option");
    detail::fgx_check(fgx_stream_create(&stream_), "stream create");
    alloc((void**)&d_small_, 4096 * 4);
    detail::fgx_check(fgx_fill_u32(d_small_, 0, 4096, stream_), "zero counters");
    detail::fgx_check(fgx_host_alloc(4096 * 4, (void**)&h_small_), "pinned counters");
    d_error_ = d_small_;
    d_step_ = d_small_ + 3;
    d_dummy_ = d_small_ + 4;
    d_overflow_ = d_small_ + 5;
    alloc(&d_reduce_out_, 64);
    alloc(&d_pages_, 256);
    detail::fgx_check(fgx_fill_u32((uint32_t*)d_pages_, 0, 64, stream_), "zero page");
    auto counter = [&]() {
        if (small_used_ >= 4096) throw exception::InvalidOperation("too many agent states and message lists");
        return d_small_ + small_used_++;
    };
    // agents and their state lists
    for (auto& kv : m_->agents) {
        detail::AgentRT a;
        a.d = kv.second;
        a.d_next_id = counter();
        for (auto& s : kv.second->states) {
            detail::StateList l;
            l.state = s;
            for (auto& v : kv.second->variables) {
                detail::VarBuf b;
                b.name = v.first;
                b.bytes = (uint32_t)v.second.bytes();
                b.def = v.second.default_value;
                l.vars.push_back(b);
            }
            l.id_slot = l.var_index("_id");
            l.d_count = counter();
            a.states.emplace(s, std::move(l));
        }
        agents_.emplace(kv.first, std::move(a));
    }
    // message lists
    for (auto& kv : m_->messages) {
        detail::MessageRT mr;
        mr.d = kv.second;
        for (auto& v : kv.second->variables) {
            detail::VarBuf b;
            b.name = v.first;
            b.bytes = (uint32_t)v.second.bytes();
            b.def = v.second.default_value;
            mr.vars.push_back(b);
        }
        if (mr.vars.size() > detail::MAX_MSG_VARS)
            throw exception::InvalidMessageVar("Message '" + kv.first + "' has too many variables");
        mr.d_count = counter();
        mr.spatial = kv.second->kind == MsgKind::Spatial2D || kv.second->kind == MsgKind::Spatial3D;
        if (mr.spatial) {
            const int dims = kv.second->kind == MsgKind::Spatial3D ? 3 : 2;
            if (!kv.second->radius_set || !kv.second->min_set || !kv.second->max_set)
                throw exception::InvalidMessageType("Spatial message '" + kv.first + "' needs a radius, min and max");
            detail::fgx_check(fgx_spatial_meta_init_mult(dims, kv.second->min, kv.second->max, kv.second->radius,
                                                         bin_mult_, &mr.meta),
                              "spatial message metadata");
            alloc((void**)&mr.pbm, (size_t)(mr.meta.bins + 1) * 4);
            detail::fgx_check(fgx_fill_u32(mr.pbm, 0, mr.meta.bins + 1, stream_), "PBM init");
        }
        mr.kind = kv.second->kind;
        if (mr.kind == MsgKind::Bucket) {
            // src/flamegpu/runtime/messaging/MessageBucket.cu:38-46
            if (!kv.second->upper_set)
                throw exception::InvalidMessage("Bucket message '" + kv.first + "' needs an upper bound (setUpperBound / setBounds)");
            mr.key_min = kv.second->lower;
            const int64_t keys = (int64_t)kv.second->upper - kv.second->lower + 1;
            if (keys <= 0 || keys > (1ll << 30)) throw exception::InvalidMessage("Bucket message '" + kv.first + "' key range is too large");
            mr.keys = (uint32_t)keys;
            alloc((void**)&mr.pbm, (size_t)(mr.keys + 1) * 4);
            detail::fgx_check(fgx_fill_u32(mr.pbm, 0, mr.keys + 1, stream_), "bucket table init");
        } else if (is_array_kind(mr.kind)) {
            // src/flamegpu/runtime/messaging/MessageArray*.cu: Data() (zero length is an error)
            uint64_t len = 1;
            const int nd = mr.kind == MsgKind::Array ? 1 : (mr.kind == MsgKind::Array2D ? 2 : 3);
            for (int a = 0; a < 3; ++a) {
                mr.dims[a] = a < nd ? kv.second->dims[a] : 1u;
                len *= mr.dims[a];
            }
            if (len == 0) throw exception::InvalidMessage("Length must not be zero in array message '" + kv.first + "'");
            if (len > 0x7FFFFFFFull) throw exception::InvalidMessage("Array message '" + kv.first + "' is too large");
            mr.length = (uint32_t)len;
            alloc((void**)&mr.counts, (size_t)mr.length * 4);
        }
        msgs_.emplace(kv.first, std::move(mr));
    }
    // functions, in layer order
    for (auto& layer : m_->layers) {
        for (auto& f : layer->agent_functions) {
            if (fn_index_.count(f.get())) continue;
            detail::FnRT r;
            r.d = f;
            r.agent = &agents_.at(f->parent->name);
            r.in_state = f->initial_state;
            r.end_state = f->end_state;
            if (auto mi = f->message_input.lock()) r.min = &msgs_.at(mi->name);
            if (auto mo = f->message_output.lock()) r.mout = &msgs_.at(mo->name);
            if (f->in_kind != MsgKind::None && !r.min)
                throw exception::InvalidAgentFunc("Agent function '" + f->name + "' reads messages but has no message input set");
            if (f->out_kind != MsgKind::None && !r.mout)
                throw exception::InvalidAgentFunc("Agent function '" + f->name + "' outputs messages but has no message output set");
            if (auto ao = f->agent_output.lock()) {
                r.out_agent = &agents_.at(ao->name);
                r.out_state = f->agent_output_state;
            }
            if (r.min && r.min->spatial && f->parent->sortPeriod != 0) {
                const auto& vars = f->parent->variables;
                auto isf = [&](const char* n) {
                    auto it = vars.find(n);
                    return it != vars.end() && it->second.type == std::type_index(typeid(float)) && it->second.elements == 1;
                };
                if (r.min->meta.dims == 3 && isf("x") && isf("y") && isf("z")) r.sort_dims = 3;
                if (r.min->meta.dims == 2 && isf("x") && isf("y")) r.sort_dims = 2;
            }
            r.d_launch_count = counter();
            if (r.out_agent) r.d_id_slots = counter();
            if (f->condition) {
                r.d_first = counter();
                r.d_work = counter();
            }
            fn_index_[f.get()] = fns_.size();
            fns_.push_back(std::move(r));
        }
    }
    // environment
    alloc((void**)&d_env_, std::max<size_t>(env_host_.size(), 8));
    env_dirty_ = true;
    host_api_.reset(new HostAPI(*this, config_.random_seed));
    graph_eligible_ = true;
    for (auto& l : m_->layers)
        if (!l->host_functions.empty()) graph_eligible_ = false;
    for (auto& a : m_->agents)
        if (a.second->sortPeriod > 1) graph_eligible_ = false;
    for (auto& mm : m_->messages)
        if (mm.second->persistent) graph_eligible_ = false;
    detail::fgx_check(fgx_event_create(&fork_event_), "event create");
    detail::fgx_check(fgx_event_create(&build_fork_event_), "event create");
    detail::fgx_check(fgx_event_create(&build_join_event_), "event create");
    initialised_ = true;
}

inline void CUDASimulation::release() {
    if (!initialised_) return;
    fgx_stream_synchronize(stream_);
    collect_timings();
    clear_graphs();
    for (auto& kv : agents_)
        for (auto& s : kv.second.states) {
            for (auto& v : s.second.vars) {
                free_dev(v.cur);
                free_dev(v.alt);
            }
            free_dev(s.second.sort_scratch);
        }
    for (auto& kv : msgs_) {
        for (auto& v : kv.second.vars) {
            free_dev(v.cur);
            free_dev(v.alt);
        }
        void* p = kv.second.pbm;
        free_dev(p);
        p = kv.second.counts;
        free_dev(p);
        free_dev(kv.second.build_scratch);
        free_dev(kv.second.packed);
    }
    for (auto& f : fns_) {
        void* p = f.death_flags;
        free_dev(p);
        p = f.birth_flags;
        free_dev(p);
        p = f.msg_flags;
        free_dev(p);
        for (auto& v : f.newbuf) free_dev(v.cur);
        free_dev(f.death_scratch);
        free_dev(f.births_scratch);
        free_dev(f.append_scratch);
        free_dev(f.trans_scratch);
        p = f.cond_flags;
        free_dev(p);
        free_dev(f.cond_scratch);
        if (f.stream) fgx_stream_destroy(f.stream);
    }
    for (auto e : join_events_) fgx_event_destroy(e);
    if (fork_event_) fgx_event_destroy(fork_event_);
    if (build_fork_event_) fgx_event_destroy(build_fork_event_);
    if (build_join_event_) fgx_event_destroy(build_join_event_);
    if (build_stream_) fgx_stream_destroy(build_stream_);
    for (auto e : free_events_) fgx_event_destroy(e);
    void* p = d_env_;
    free_dev(p);
    p = d_small_;
    free_dev(p);
    free_dev(d_reduce_out_);
    free_dev(d_pages_);
    free_dev(reduce_scratch_);
    p = d_log_;
    free_dev(p);
    d_log_ = nullptr;
    log_cap_frames_ = 0;
    p = d_exit_log_;
    free_dev(p);
    d_exit_log_ = nullptr;
    exit_log_bytes_ = 0;
    free_dev(log_scratch_);
    if (h_small_) fgx_host_free(h_small_);
    fgx_stream_destroy(stream_);
    fgx_ctx_destroy(ctx_);
    initialised_ = false;
}

inline void CUDASimulation::grow_list(detail::StateList& l, uint32_t cap) {
    if (cap <= l.cap) return;
    cap = std::max<uint32_t>(cap, 256);
    for (auto& v : l.vars) {
        void *nc = nullptr, *na = nullptr;
        alloc(&nc, (size_t)cap * v.bytes);
        alloc(&na, (size_t)cap * v.bytes);
        if (v.cur && l.cap) detail::fgx_check(fgx_memcpy_d2d(nc, v.cur, (size_t)l.cap * v.bytes, stream_), "grow copy");
        // Keep both buffers valid before the old ones go.
        detail::fgx_check(fgx_stream_synchronize(stream_), "grow sync");
        free_dev(v.cur);
        free_dev(v.alt);
        v.cur = nc;
        v.alt = na;
    }
    free_dev(l.sort_scratch);
    l.cap = cap;
}

inline void CUDASimulation::grow_message(detail::MessageRT& m, uint32_t cap) {
    if (cap <= m.cap) return;
    cap = std::max<uint32_t>(cap, 256);
    for (auto& v : m.vars) {
        void *nc = nullptr, *na = nullptr;
        alloc(&nc, (size_t)cap * v.bytes);
        alloc(&na, (size_t)cap * v.bytes);
        if (v.cur && m.cap) detail::fgx_check(fgx_memcpy_d2d(nc, v.cur, (size_t)m.cap * v.bytes, stream_), "grow copy");
        detail::fgx_check(fgx_stream_synchronize(stream_), "grow sync");
        free_dev(v.cur);
        free_dev(v.alt);
        v.cur = nc;
        v.alt = na;
    }
    free_dev(m.build_scratch);
    if (m.spatial) {
        m.build_scratch = new_scratch(FGX_OP_SPATIAL_BUILD, cap, m.meta.bins, &m.build_scratch_bytes);
        free_dev(m.packed);
        alloc(&m.packed, (size_t)cap * 16);
    } else if (m.kind == MsgKind::Bucket) {
        m.build_scratch = new_scratch(FGX_OP_SPATIAL_BUILD, cap, m.keys, &m.build_scratch_bytes);
    }
    m.cap = cap;
}

inline void CUDASimulation::size_function(detail::FnRT& f, uint32_t cap) {
    if (cap <= f.cap) return;
    cap = std::max<uint32_t>(cap, 256);
    auto realloc_u32 = [&](uint32_t*& p) {
        void* q = p;
        free_dev(q);
        alloc(&q, (size_t)cap * 4);
        p = (uint32_t*)q;
    };
    detail::fgx_check(fgx_stream_synchronize(stream_), "resize sync");
    if (f.d->has_agent_death) {
        realloc_u32(f.death_flags);
        free_dev(f.death_scratch);
        f.death_scratch = new_scratch(FGX_OP_COMPACT, cap, 0, &f.death_scratch_bytes);
    }
    if (f.mout) {
        realloc_u32(f.msg_flags);
        free_dev(f.append_scratch);
        f.append_scratch = new_scratch(FGX_OP_BIRTHS, cap, 0, &f.append_scratch_bytes);
    }
    if (f.out_agent) {
        realloc_u32(f.birth_flags);
        free_dev(f.births_scratch);
        f.births_scratch = new_scratch(FGX_OP_BIRTHS, cap, 0, &f.births_scratch_bytes);
        for (auto& v : f.newbuf) free_dev(v.cur);
        f.newbuf.clear();
        for (auto& kv : f.out_agent->d->variables) {
            detail::VarBuf b;
            b.name = kv.first;
            b.bytes = (uint32_t)kv.second.bytes();
            b.def = kv.second.default_value;
            alloc(&b.cur, (size_t)cap * b.bytes);
            f.newbuf.push_back(b);
        }
    }
    if (f.in_state != f.end_state) {
        free_dev(f.trans_scratch);
        f.trans_scratch = new_scratch(FGX_OP_BIRTHS, cap, 0, &f.trans_scratch_bytes);
    }
    if (f.d->condition) {
        realloc_u32(f.cond_flags);
        free_dev(f.cond_scratch);
        f.cond_scratch = new_scratch(FGX_OP_COMPACT, cap, 0, &f.cond_scratch_bytes);
    }
    f.cap = cap;
}

inline bool CUDASimulation::model_can_grow() const {
    for (auto& f : fns_)
        if (f.out_agent || f.in_state != f.end_state) return true;
    return false;
}

inline void CUDASimulation::sync_counts() {
    synchronize();
    size_t k = 0;
    std::vector<uint32_t*> order;
    for (auto& kv : agents_)
        for (auto& s : kv.second.states) order.push_back(s.second.d_count);
    // counts live in one block: copy the whole used prefix
    detail::fgx_check(fgx_memcpy_d2h(h_small_, d_small_, (size_t)small_used_ * 4, stream_), "count readback");
    synchronize();
    for (auto& kv : agents_) {
        // IDs given to host-created agents not yet uploaded stay taken
        const uint32_t dev_next = h_small_[kv.second.d_next_id - d_small_];
        kv.second.h_next_id = host_agents_pending(&kv.second) ? std::max(dev_next, kv.second.h_next_id) : dev_next;
        for (auto& s : kv.second.states) s.second.h_count = h_small_[s.second.d_count - d_small_];
    }
    (void)k;
    (void)order;
    if (h_small_[5] != 0)
        throw exception::InvalidOperation("an agent list append passed its capacity: the step's population bound was too low");
    if (h_small_[0] != 0) check_device_error();
}

// Which state transitions of the coming step swap lists instead of copying
// (CUDAFatAgent::transitionState swaps when the end state holds no agent and no function
// condition left agents behind). The end state is known empty when its upper bound at that
// point of the step is zero: the bounds start from the counts read back at step start and
// grow as births and transitions are enqueued. A transition swaps only when no other
// function of its layer touches either list (they may run concurrently) and the two lists
// have equal capacities; nothing is known after a host layer function, which may add agents.
inline void CUDASimulation::plan_transitions(bool counted) {
    for (auto& f : fns_) f.swap_transition = false;
    if (!counted) return;
    std::map<const detail::StateList*, uint64_t> ub;
    for (auto& kv : agents_)
        for (auto& s : kv.second.states) ub[&s.second] = s.second.h_count;
    for (auto& layer : m_->layers) {
        if (!layer->host_functions.empty()) return;
        std::vector<detail::FnRT*> fl;
        for (auto& fd : layer->agent_functions) fl.push_back(&fns_[fn_index_.at(fd.get())]);
        for (auto* f : fl) {
            const auto* in = &f->agent->states.at(f->in_state);
            const uint64_t n = ub[in];
            if (f->in_state != f->end_state) {
                const auto* dst = &f->agent->states.at(f->end_state);
                bool alone = true;
                for (auto* g : fl) {
                    if (g == f) continue;
                    for (const std::string* s : {&f->in_state, &f->end_state}) {
                        if (g->agent == f->agent && (g->in_state == *s || g->end_state == *s)) alone = false;
                        if (g->out_agent == f->agent && g->out_state == *s) alone = false;
                    }
                }
                f->swap_transition = alone && !f->d->condition && ub[dst] == 0 && in->cap == dst->cap;
            }
            if (f->out_agent) ub[&f->out_agent->states.at(f->out_state)] += n;
            if (f->in_state != f->end_state) {
                ub[&f->agent->states.at(f->end_state)] += n;
                if (!f->d->condition) ub[in] = 0;
            }
        }
    }
}

// Upper bounds of every list for the coming step (births append, transitions move).
inline bool CUDASimulation::ensure_capacity() {
    std::map<detail::StateList*, uint64_t> ub, peak;
    for (auto& kv : agents_)
        for (auto& s : kv.second.states) {
            ub[&s.second] = s.second.h_count;
            peak[&s.second] = s.second.h_count;
        }
    std::map<detail::MessageRT*, uint64_t> mub;
    std::map<detail::FnRT*, uint64_t> fub;
    for (auto& layer : m_->layers) {
        for (auto& fd : layer->agent_functions) {
            auto& f = fns_[fn_index_.at(fd.get())];
            auto* in = &f.agent->states.at(f.in_state);
            const uint64_t n = ub[in];
            fub[&f] = std::max(fub[&f], n);
            if (f.mout) mub[f.mout] += n;
            if (f.out_agent) {
                auto* dst = &f.out_agent->states.at(f.out_state);
                ub[dst] += n;
                peak[dst] = std::max(peak[dst], ub[dst]);
            }
            if (f.in_state != f.end_state) {
                auto* dst = &f.agent->states.at(f.end_state);
                ub[dst] += n;
                peak[dst] = std::max(peak[dst], ub[dst]);
                // with a function condition the disabled agents stay behind
                if (!f.d->condition) ub[in] = 0;
            }
        }
    }
    bool changed = false;
    auto need = [](uint64_t want, uint32_t have) -> uint32_t {
        if (want <= have) return 0;
        uint64_t c = std::max<uint64_t>(want, (uint64_t)have + have / 2);
        if (c > 0xFFFFFFFFull / 2) throw exception::InvalidOperation("population exceeds the 32-bit index range");
        return (uint32_t)c;
    };
    std::map<detail::StateList*, uint32_t> target;
    for (auto& kv : peak) {
        const uint32_t c = need(kv.second, kv.first->cap);
        target[kv.first] = c ? c : kv.first->cap;
    }
    // the two lists of an unconditional transition keep equal capacities, so that the
    // transition can swap them when its end state is empty
    for (bool again = true; again;) {
        again = false;
        for (auto& f : fns_) {
            if (f.in_state == f.end_state || f.d->condition) continue;
            auto* a = &f.agent->states.at(f.in_state);
            auto* b = &f.agent->states.at(f.end_state);
            const uint32_t m = std::max(target[a], target[b]);
            if (target[a] != m || target[b] != m) {
                target[a] = target[b] = m;
                again = true;
            }
        }
    }
    for (auto& kv : target) {
        if (kv.second > kv.first->cap) {
            grow_list(*kv.first, kv.second);
            changed = true;
        }
    }
    for (auto& kv : msgs_) {
        auto& m = kv.second;
        uint64_t want = mub[&m] + (m.d->persistent ? m.cap : 0);
        want = std::max<uint64_t>(want, m.length);   // an array list holds every element
        if (uint32_t c = need(std::max<uint64_t>(want, 1), m.cap)) {
            grow_message(m, c);
            changed = true;
        }
    }
    // Per-function flags, new-agent buffers and scratch cover the whole input list
    // capacity: launches and compactions are shaped by that capacity.
    for (auto& f : fns_) {
        const uint64_t list_cap = f.agent->states.at(f.in_state).cap;
        const uint64_t want = std::max<uint64_t>(std::max<uint64_t>(fub[&f], list_cap), 1);
        if (want > f.cap) {
            size_function(f, (uint32_t)want);
            changed = true;
        }
    }
    // sort scratch depends on list capacity and on the largest key range of the functions
    // that sort the list (they may read different spatial messages)
    for (auto& f : fns_) {
        if (!f.sort_dims) continue;
        auto& l = f.agent->states.at(f.in_state);
        uint32_t bits = 0;
        detail::fgx_check(fgx_agent_sort_key_bits(f.sort_dims, f.min->meta.min, f.min->meta.max, f.min->meta.radius, &bits),
                          "sort key bits");
        if (!l.sort_scratch || l.sort_scratch_keys < (1u << bits)) {
            free_dev(l.sort_scratch);
            l.sort_scratch_keys = std::max(l.sort_scratch_keys, 1u << bits);
            l.sort_scratch = new_scratch(FGX_OP_AGENT_SORT, l.cap, l.sort_scratch_keys, &l.sort_scratch_bytes);
            changed = true;
        }
    }
    return changed;
}

inline void CUDASimulation::setPopulationData(AgentVector& pop, const std::string& state_in) {
    ensure_init();
    auto& a = agent_rt(pop.getAgentName());
    const std::string state = state_in.empty() ? a.d->initial_state : state_in;
    auto it = a.states.find(state);
    if (it == a.states.end()) throw exception::InvalidAgentState("Agent '" + a.d->name + "' has no state '" + state + "'");
    auto& l = it->second;
    if (step_ > 0 || initialised_) sync_counts();
    const uint32_t n = (uint32_t)pop.size();
    // IDs: CUDAFatAgent::assignIDs (src/flamegpu/simulation/detail/CUDAFatAgent.cu:347-391)
    auto& ids = pop.raw("_id");
    id_t mx = 0;
    for (uint32_t i = 0; i < n; ++i) {
        id_t v;
        std::memcpy(&v, ids.data() + (size_t)i * 4, 4);
        mx = std::max(mx, v);
    }
    uint32_t next = std::max<uint32_t>(a.h_next_id, mx + 1);
    for (uint32_t i = 0; i < n; ++i) {
        id_t v;
        std::memcpy(&v, ids.data() + (size_t)i * 4, 4);
        if (v == ID_NOT_SET) {
            v = next + i;
            std::memcpy(ids.data() + (size_t)i * 4, &v, 4);
        }
    }
    next += n;
    a.h_next_id = next;
    grow_list(l, std::max<uint32_t>(n, 1));
    for (auto& v : l.vars)
        detail::fgx_check(fgx_memcpy_h2d(v.cur, pop.raw(v.name).data(), (size_t)n * v.bytes, stream_), "population upload");
    l.h_count = n;
    detail::fgx_check(fgx_memcpy_h2d(l.d_count, &n, 4, stream_), "count upload");
    detail::fgx_check(fgx_memcpy_h2d(a.d_next_id, &next, 4, stream_), "next id upload");
    synchronize();
    clear_graphs();
    populations_set_ = true;
}

// src/flamegpu/simulation/Simulation.cu:55-83 + src/flamegpu/io/StateReader.cu
// getFullModelState: the file's configuration, then environment values, then every
// (agent, state) population of the model (empty when the file has none).
inline void CUDASimulation::load_state(const std::string& path) {
    io::StateData d = io::read_state(path);
    auto as_bool = [](std::string v) {
        for (auto& c : v) c = (char)std::tolower((unsigned char)c);
        return v == "true" || (v != "false" && std::stoll(v) != 0);
    };
    for (auto& kv : d.simulation) {
        if (kv.first == "random_seed") config_.random_seed = std::stoull(kv.second);
        else if (kv.first == "steps") config_.steps = (unsigned int)std::stoul(kv.second);
        else if (kv.first == "timing") config_.timing = as_bool(kv.second);
        else if (kv.first == "truncate_log_files") config_.truncate_log_files = as_bool(kv.second);
    }
    for (auto& kv : d.cuda) {
        if (kv.first == "device_id" && !initialised_) cuda_config_.device_id = std::stoi(kv.second);
        else if (kv.first == "inLayerConcurrency") cuda_config_.inLayerConcurrency = as_bool(kv.second);
    }
    for (auto& kv : d.environment) {
        auto it = m_->env.props.find(kv.first);
        if (it == m_->env.props.end())
            throw exception::InvalidEnvProperty("Input file contains unrecognised environment property '" + kv.first + "'");
        if (kv.second.size() > it->second.elements)
            throw exception::OutOfBoundsException("Input file contains environment property '" + kv.first + "' with too many elements");
        for (size_t e = 0; e < kv.second.size(); ++e)
            io::detail::parse_token(it->second.type, kv.second[e], env_host_.data() + it->second.offset, e);
        env_dirty_ = true;
    }
    for (auto& kv : d.agents) {
        auto a = m_->agents.find(kv.first.first);
        const std::string st = a == m_->agents.end() ? kv.first.second
                             : kv.first.second.empty() ? a->second->initial_state : kv.first.second;
        if (a == m_->agents.end() || !a->second->states.count(st))
            throw exception::InvalidAgentState("Agent '" + kv.first.first + "' with state '" + st + "', found in input file '" +
                                               path + "', is not part of the model description hierarchy");
    }
    ensure_init();
    for (auto& ap : m_->agents) {
        std::set<id_t> seen;
        for (auto& st : ap.second->states) {
            AgentVector pop(AgentDescription(ap.second));
            const std::vector<io::StateData::AgentRecord>* recs = nullptr;
            auto it = d.agents.find({ap.first, st});
            if (it != d.agents.end()) recs = &it->second;
            if (st == ap.second->initial_state) {
                auto it0 = d.agents.find({ap.first, std::string()});
                if (it0 != d.agents.end()) {
                    if (recs) throw exception::InvalidInputFile("Input file gives agent '" + ap.first + "' state '" + st + "' twice");
                    recs = &it0->second;
                }
            }
            const size_t n = recs ? recs->size() : 0;
            pop.resize(n);
            for (size_t i = 0; i < n; ++i) {
                for (auto& v : (*recs)[i].vars) {
                    auto vit = ap.second->variables.find(v.first);
                    if (vit == ap.second->variables.end())
                        throw exception::InvalidAgentVar("Input file agent '" + ap.first + "' has unrecognised variable '" + v.first + "'");
                    if (v.second.size() > vit->second.elements)
                        throw exception::OutOfBoundsException("Input file agent variable '" + v.first + "' has too many elements");
                    char* dst = pop.raw(v.first).data() + i * vit->second.bytes();
                    for (size_t e = 0; e < v.second.size(); ++e) io::detail::parse_token(vit->second.type, v.second[e], dst, e);
                }
                id_t id;
                std::memcpy(&id, pop.raw("_id").data() + i * sizeof(id_t), sizeof(id_t));
                // tests/test_cases/io/test_io.cu AgentID_FileInput_IDCollision
                if (id != ID_NOT_SET && !seen.insert(id).second)
                    throw exception::AgentIDCollision("Input file contains agent '" + ap.first + "' ID " + std::to_string(id) +
                                                      " more than once");
            }
            setPopulationData(pop, st);
        }
    }
    // Next agent IDs as exported (this build's stats/next_id), so births after a resume
    // get the IDs they would have had; never below the loaded IDs.
    for (auto& kv : d.next_id) {
        auto it = agents_.find(kv.first);
        if (it == agents_.end()) continue;
        sync_counts();
        uint32_t mx = 0;
        for (auto& st : it->second.states) {
            auto& l = st.second;
            if (!l.h_count) continue;
            std::vector<uint32_t> ids(l.h_count);
            detail::fgx_check(fgx_memcpy_d2h(ids.data(), l.vars[l.id_slot].cur, (size_t)l.h_count * 4, stream_), "id readback");
            synchronize();
            for (auto v : ids) mx = std::max(mx, v);
        }
        it->second.h_next_id = std::max(kv.second, mx + 1);
        detail::fgx_check(fgx_memcpy_h2d(it->second.d_next_id, &it->second.h_next_id, 4, stream_), "next id upload");
        synchronize();
    }
    // The step counter resumes from the file's step count, so that an exported state
    // continues bit-exactly (random streams are keyed by agent ID and step); the
    // reference restarts a loaded state at step 0.
    if (d.has_step) {
        step_ = d.step_count;
        detail::fgx_check(fgx_fill_u32(d_step_, step_, 1, stream_), "step counter");
        synchronize();
    }
    loaded_input_file_ = path;
}

inline void CUDASimulation::exportData(const std::string& path, bool prettyPrint) {
    ensure_init();
    sync_counts();
    io::StateOut s;
    s.simulation = {{"input_file", config_.input_file, true},
                    {"random_seed", std::to_string(config_.random_seed), false},
                    {"steps", std::to_string(config_.steps), false},
                    {"timing", config_.timing ? "true" : "false", false},
                    {"truncate_log_files", config_.truncate_log_files ? "true" : "false", false}};
    s.cuda = {{"device_id", std::to_string(cuda_config_.device_id), false},
              {"inLayerConcurrency", cuda_config_.inLayerConcurrency ? "true" : "false", false}};
    s.step_count = step_;
    for (auto& kv : agents_) s.next_id[kv.first] = kv.second.h_next_id;   // read by sync_counts above
    for (auto& kv : m_->env.props) {
        io::StateOut::Env e;
        e.name = kv.first;
        e.type = kv.second.type;
        e.elements = (unsigned int)kv.second.elements;
        e.bytes.assign(env_host_.data() + kv.second.offset,
                       env_host_.data() + kv.second.offset + kv.second.type_size * kv.second.elements);
        s.environment.push_back(std::move(e));
    }
    std::vector<std::unique_ptr<AgentVector>> keep;
    for (auto& ap : m_->agents)
        for (auto& st : ap.second->states) {
            keep.emplace_back(new AgentVector(AgentDescription(ap.second)));
            getPopulationData(*keep.back(), st);
            io::StateOut::Pop p;
            p.agent = ap.first;
            p.state = st;
            p.n = (uint32_t)keep.back()->size();
            for (auto& v : ap.second->variables)
                p.vars.push_back(io::StateOut::Pop::Var{v.first, v.second.type, (unsigned int)v.second.elements,
                                                        &keep.back()->raw(v.first)});
            s.pops.push_back(std::move(p));
        }
    io::write_state(path, s, prettyPrint);
}

inline void CUDASimulation::getPopulationData(AgentVector& pop, const std::string& state_in) {
    ensure_init();
    auto& a = agent_rt(pop.getAgentName());
    const std::string state = state_in.empty() ? a.d->initial_state : state_in;
    auto it = a.states.find(state);
    if (it == a.states.end()) throw exception::InvalidAgentState("Agent '" + a.d->name + "' has no state '" + state + "'");
    sync_counts();
    check_device_error();
    auto& l = it->second;
    pop.resize(l.h_count);
    for (auto& v : l.vars)
        if (l.h_count)
            detail::fgx_check(fgx_memcpy_d2h(pop.raw(v.name).data(), v.cur, (size_t)l.h_count * v.bytes, stream_),
                              "population download");
    synchronize();
}

inline void CUDASimulation::check_device_error() {
    synchronize();
    uint32_t e[3] = {0, 0, 0};
    detail::fgx_check(fgx_memcpy_d2h(e, d_error_, 12, stream_), "error readback");
    synchronize();
    if (e[0] != detail::DERR_NONE) {
        detail::fgx_check(fgx_fill_u32(d_error_, 0, 3, stream_), "error reset");
        synchronize();
        // src/flamegpu/simulation/detail/CUDAScatter.cu arrayMessageReorder raises this type
        if (e[0] == detail::DERR_ARRAY_CONFLICT)
            throw exception::ArrayMessageWriteConflict(detail::format("Device error %u (%s), array index %u", e[0],
                                                                      detail::device_error_text(e[0]), e[1]));
        throw exception::DeviceError(detail::format("Device error %u (%s), name hash 0x%08x, agent index %u", e[0],
                                                    detail::device_error_text(e[0]), e[1], e[2]));
    }
}

inline void CUDASimulation::upload_env() {
    if (!env_dirty_) return;
    if (!env_host_.empty())
        detail::fgx_check(fgx_memcpy_h2d(d_env_, env_host_.data(), env_host_.size(), stream_), "environment upload");
    synchronize();
    env_dirty_ = false;
}

inline void CUDASimulation::tick(const std::string& label, void* st, bool begin) {
    if (!timing_) return;
    if (begin) {
        void* e = nullptr;
        if (!free_events_.empty()) {
            e = free_events_.back();
            free_events_.pop_back();
        } else {
            detail::fgx_check(fgx_event_create(&e), "event create");
        }
        fgx_event_record(e, st);
        open_events_[label] = e;
    } else {
        void* e = nullptr;
        if (!free_events_.empty()) {
            e = free_events_.back();
            free_events_.pop_back();
        } else {
            detail::fgx_check(fgx_event_create(&e), "event create");
        }
        fgx_event_record(e, st);
        pending_.push_back({label, open_events_.at(label), e});
        open_events_.erase(label);
    }
}

inline void CUDASimulation::collect_timings() {
    if (pending_.empty()) return;
    for (auto& p : pending_) {
        float ms = 0.0f;
        fgx_event_elapsed_ms(p.start, p.stop, &ms);
        auto& k = ktimes_[p.label];
        k.ms += ms;
        k.count++;
        free_events_.push_back(p.start);
        free_events_.push_back(p.stop);
    }
    pending_.clear();
}

inline void CUDASimulation::launch_function(detail::FnRT& f, void* st) {
    detail::FnArgs a;
    const uint32_t grid = prepare_function(f, st, a);
    const std::string label = "fn:" + f.agent->d->name + "::" + f.d->name;
    tick(label, st, true);
    f.d->launch(a, grid, kFnBlock, stage_smem_bytes(a), st);
    tick(label, st, false);
    if (hipGetLastError() != hipSuccess) throw exception::CUDAError("agent function launch failed: " + f.d->name);
}

// Launch record and grid of one agent function; the device-side preparation (new-agent
// defaults, the launch count) is enqueued on `st`.
inline uint32_t CUDASimulation::prepare_function(detail::FnRT& f, void* st, detail::FnArgs& a) {
    auto& l = f.agent->states.at(f.in_state);
    if (f.cap < l.cap)
        throw exception::InvalidOperation("function '" + f.d->name + "' buffers are smaller than its agent list");
    std::memset(&a, 0, sizeof(a));
    a.n_cap = l.cap;
    a.zero_page = (const char*)d_pages_;
    a.sink_page = (char*)d_pages_ + 128;
    a.stream_id = f.d->stream_id;
    a.unit_tables = unit_tables_ ? 1 : 0;
    a.d_n = l.d_count;
    a.d_step = d_step_;
    a.seed = config_.random_seed;
    a.d_error = d_error_;
    a.env = d_env_;
    a.ids = l.id_slot >= 0 ? (const uint32_t*)l.vars[l.id_slot].cur : nullptr;
    auto fill_table = [](detail::VarBinding* tab, uint32_t* mul, const std::vector<detail::VarBuf>& vars, bool use_alt) {
        std::vector<uint32_t> hs;
        for (auto& v : vars) hs.push_back(detail::name_hash(v.name));
        *mul = detail::perfect_hash_mul(hs.data(), (unsigned)hs.size(), detail::VAR_SLOT_BITS);
        for (size_t i = 0; i < vars.size(); ++i) {
            const uint32_t s = detail::table_slot(hs[i], *mul, detail::VAR_SLOT_BITS);
            tab[s] = {hs[i], (uint16_t)vars[i].bytes, 0, (char*)(use_alt ? vars[i].alt : vars[i].cur)};
        }
    };
    fill_table(a.agent, &a.agent_mul, l.vars, false);
    if (f.min) {
        fill_table(a.msg_in, &a.msg_in_mul, f.min->vars, false);
        a.d_msg_in_count = f.min->d_count;
        if (f.min->spatial) {
            const auto& m = f.min->meta;
            for (int k = 0; k < 3; ++k) {
                a.in_meta.min[k] = m.min[k];
                a.in_meta.max[k] = m.max[k];
                a.in_meta.width[k] = m.width[k];
                a.in_meta.grid[k] = m.grid[k];
            }
            a.in_meta.radius = m.radius;
            a.in_meta.cell = m.cell;
            a.in_meta.span = m.span;
            a.in_meta.wrap_ok = m.wrap_ok;
            a.in_meta.pbm = f.min->pbm;
            a.in_meta.packed = f.min->packed;
            stage_neighbourhood(f, l, a);
        } else if (f.min->kind == MsgKind::Bucket) {
            a.in_meta.key_min = f.min->key_min;
            a.in_meta.keys = f.min->keys;
            a.in_meta.pbm = f.min->pbm;
        } else if (is_array_kind(f.min->kind)) {
            for (int k = 0; k < 3; ++k) a.in_meta.grid[k] = f.min->dims[k];
        }
    }
    if (f.mout) {
        for (int k = 0; k < 3; ++k) a.out_dims[k] = f.mout->dims[k];
        a.out_key_min = f.mout->key_min;
        a.out_keys = f.mout->keys;
        fill_table(a.msg_out, &a.msg_out_mul, f.mout->vars, true);
        a.msg_out_flags = f.msg_flags;
        a.msg_out_optional = f.d->message_output_optional ? 1 : 0;
    }
    if (f.out_agent) {
        fill_table(a.new_agent, &a.new_mul, f.newbuf, false);
        std::vector<void*> ptrs;
        std::vector<uint32_t> bytes, offs;
        std::vector<char> defs;
        for (size_t i = 0; i < f.newbuf.size(); ++i) {
            ptrs.push_back(f.newbuf[i].cur);
            bytes.push_back(f.newbuf[i].bytes);
            offs.push_back((uint32_t)defs.size());
            defs.insert(defs.end(), f.newbuf[i].def.begin(), f.newbuf[i].def.end());
        }
        a.birth_flags = f.birth_flags;
        a.d_new_next_id = f.out_agent->d_next_id;
        a.d_id_slots = f.d_id_slots;
        // broadcastInit: new agents start from the variable defaults.
        detail::fgx_check(fgx_fill_vars((uint32_t)ptrs.size(), ptrs.data(), bytes.data(), defs.data(), offs.data(), l.cap, st),
                          "new-agent defaults");
    }
    if (f.d->has_agent_death) a.death_flags = f.death_flags;
    {
        std::vector<uint32_t> hs;
        for (auto& kv : m_->env.props) hs.push_back(detail::name_hash(kv.first));
        a.env_mul = detail::perfect_hash_mul(hs.data(), (unsigned)hs.size(), detail::ENV_SLOT_BITS);
        size_t i = 0;
        for (auto& kv : m_->env.props) {
            const uint32_t s = detail::table_slot(hs[i], a.env_mul, detail::ENV_SLOT_BITS);
            a.env_b[s] = {hs[i], (uint16_t)kv.second.offset, (uint16_t)(kv.second.type_size * kv.second.elements)};
            ++i;
        }
    }
    // The count the launch runs over, for the post-processing in unmap_function: a
    // function that neither removes, adds nor moves agents leaves its list's count as
    // it is, so that count is read directly instead of being copied.
    const bool keeps_count = !f.d->has_agent_death && !f.out_agent && f.end_state == f.in_state;
    if (f.d->condition) {
        a.d_first = f.d_first;
        a.stage_ok = 0;
        detail::fgx_check(fgx_sub_u32(f.d_launch_count, l.d_count, f.d_first, st), "launch count");
        f.launch_count = f.d_launch_count;
    } else if (keeps_count) {
        f.launch_count = l.d_count;
    } else {
        detail::fgx_check(fgx_memcpy_d2d(f.d_launch_count, l.d_count, 4, st), "launch count");
        f.launch_count = f.d_launch_count;
    }
    // mandatory output into an emptied list: the kernel sets the message count
    f.msg_count_in_kernel = f.mout && f.mout->truncate && !f.d->message_output_optional && !f.d->condition;
    a.d_msg_count_out = f.msg_count_in_kernel ? f.mout->d_count : nullptr;
    return std::max<uint32_t>(1, (l.cap + kFnBlock - 1) / kFnBlock);
}

// In-layer concurrency as one launch: the layer kernel of the model's translation unit
// runs every function of the layer, each over its own blocks, from launch records held
// in device memory (fgx_layer_launch).  Layers it cannot take (a function outside the
// kernel, LDS staging) return false and run on forked streams instead.
inline bool CUDASimulation::launch_layer_fused(const LayerData& layer) {
    if (!layer_fusion_ || !layer_kernel_.kernel || layer.agent_functions.size() > FGX_LAYER_MAX_FNS) return false;
    std::vector<detail::FnRT*> fl;
    for (auto& fd : layer.agent_functions) {
        auto& f = fns_[fn_index_.at(fd.get())];
        if (f.fused_fn == -2) f.fused_fn = layer_kernel_.index_of(f.d->launch);
        if (f.fused_fn < 0) return false;
        if (stage_enabled_ && f.min && f.min->spatial) return false;
        fl.push_back(&f);
    }
    std::vector<detail::FnArgs> args(fl.size());
    std::vector<fgx_fn_launch> L(fl.size());
    std::vector<uint32_t> grids(fl.size());
    for (size_t k = 0; k < fl.size(); ++k) {
        grids[k] = prepare_function(*fl[k], stream_, args[k]);
        L[k].args = &args[k];
        L[k].fn = (uint32_t)fl[k]->fused_fn;
        L[k].blocks = grids[k];
        L[k].lds_bytes = layer_kernel_.lds_bytes(fl[k]->fused_fn, kFnBlock);
    }
    uint32_t before = 0;
    detail::fgx_check(fgx_layer_space_info(ctx_, &before, nullptr), "layer images");
    const int rc = fgx_layer_launch(ctx_, layer_kernel_.kernel, kFnBlock, L.data(), (uint32_t)L.size(),
                                    (uint32_t)sizeof(detail::FnArgs), stream_);
    if (rc == FGX_ERR_CAPACITY) {
        if (!layer_space_warned_) {
            std::fprintf(stderr, "flamegpu: fused layer images fill their device space; further new layers launch per function\n");
            layer_space_warned_ = true;
        }
        for (size_t k = 0; k < fl.size(); ++k) fl[k]->d->launch(args[k], grids[k], kFnBlock, 0, stream_);
        if (hipGetLastError() != hipSuccess) throw exception::CUDAError("agent function launch failed");
    } else {
        detail::fgx_check(rc, "fused layer launch");
        uint32_t after = 0;
        detail::fgx_check(fgx_layer_space_info(ctx_, &after, nullptr), "layer images");
        if (after != before) new_layer_images_ = true;   // this enqueue uploaded an image
        ++fused_launches_;
    }
    for (auto* f : fl) unmap_function(*f, stream_);
    return true;
}


// Launch arguments of a function condition: the agent's variables (read-only), environment, RNG.
inline void CUDASimulation::fill_condition_args(detail::FnRT& f, detail::StateList& l, detail::FnArgs& a) {
    a.n_cap = l.cap;
    a.zero_page = (const char*)d_pages_;
    a.sink_page = (char*)d_pages_ + 128;
    a.stream_id = f.d->stream_id;
    a.d_n = l.d_count;
    a.d_step = d_step_;
    a.seed = config_.random_seed;
    a.d_error = d_error_;
    a.env = d_env_;
    a.ids = l.id_slot >= 0 ? (const uint32_t*)l.vars[l.id_slot].cur : nullptr;
    std::vector<uint32_t> hs;
    for (auto& v : l.vars) hs.push_back(detail::name_hash(v.name));
    a.agent_mul = detail::perfect_hash_mul(hs.data(), (unsigned)hs.size(), detail::VAR_SLOT_BITS);
    for (size_t i = 0; i < l.vars.size(); ++i) {
        const uint32_t s = detail::table_slot(hs[i], a.agent_mul, detail::VAR_SLOT_BITS);
        a.agent[s] = {hs[i], (uint16_t)l.vars[i].bytes, 0, (char*)l.vars[i].cur};
    }
    std::vector<uint32_t> he;
    for (auto& kv : m_->env.props) he.push_back(detail::name_hash(kv.first));
    a.env_mul = detail::perfect_hash_mul(he.data(), (unsigned)he.size(), detail::ENV_SLOT_BITS);
    size_t i = 0;
    for (auto& kv : m_->env.props) {
        const uint32_t s = detail::table_slot(he[i], a.env_mul, detail::ENV_SLOT_BITS);
        a.env_b[s] = {he[i], (uint16_t)kv.second.offset, (uint16_t)(kv.second.type_size * kv.second.elements)};
        ++i;
    }
}

// LDS staging plan for a spatial-input function: the agent's own x/y(/z) is the
// assumed search origin, and up to four 4-byte message variables are staged
// (x, y, z first, then "id", then the rest in name order).  The budget keeps
// three 256-agent blocks resident per CU (52 KB of dynamic LDS each).
inline void CUDASimulation::stage_neighbourhood(detail::FnRT& f, detail::StateList& l, detail::FnArgs& a) {
    a.stage_ok = 0;
    if (!stage_enabled_ || f.min->meta.span != 1) return;   // the staged rows cover one bin either side
    const int dims = f.min->meta.dims;
    const char* axes[3] = {"x", "y", "z"};
    for (int k = 0; k < dims; ++k) {
        const int vi = l.var_index(axes[k]);
        if (vi < 0) return;
        const auto& var = f.agent->d->variables.at(axes[k]);
        if (var.type != std::type_index(typeid(float)) || var.elements != 1) return;
        a.stage_pos[k] = (const float*)l.vars[vi].cur;
    }
    // staged: the message position, then "id" when it is a 4-byte scalar (the variables
    // the iterator reads ahead of the loop body)
    std::vector<int> pick;
    for (int k = 0; k < dims; ++k) pick.push_back(f.min->var_index(axes[k]));
    const int vid = f.min->var_index("id");
    if (vid >= 0) {
        const auto& var = f.min->d->variables.at("id");
        if (var.elements == 1 && var.type_size == 4) pick.push_back(vid);
    }
    const uint32_t budget_words = 13312;   // 52 KB
    a.stage_max_entries = detail::STAGE_MAX_ENTRIES;
    a.stage_nvars = (uint32_t)pick.size();
    a.stage_cap = (budget_words - 2u * (a.stage_max_entries + 1u)) / a.stage_nvars;
    for (size_t k = 0; k < pick.size(); ++k) a.stage_src[k] = (const uint32_t*)f.min->vars[pick[k]].cur;
    a.stage_ok = 1;
}

inline void CUDASimulation::unmap_function(detail::FnRT& f, void* st) {
    auto& l = f.agent->states.at(f.in_state);
    // 1. message output: swap (first mandatory output this step) or append
    if (f.mout) {
        auto& m = *f.mout;
        tick("msgout:" + f.d->name, st, true);
        if (m.truncate && !f.d->message_output_optional) {
            for (auto& v : m.vars) std::swap(v.cur, v.alt);
            if (!f.msg_count_in_kernel)
                detail::fgx_check(fgx_memcpy_d2d(m.d_count, f.launch_count, 4, st), "message count");
        } else {
            if (m.truncate) detail::fgx_check(fgx_fill_u32(m.d_count, 0, 1, st), "message truncate");
            std::vector<const void*> src;
            std::vector<void*> dst;
            std::vector<uint32_t> by;
            for (auto& v : m.vars) {
                src.push_back(v.alt);
                dst.push_back(v.cur);
                by.push_back(v.bytes);
            }
            detail::fgx_check(fgx_append_births(ctx_, f.msg_flags, f.launch_count, l.cap, (uint32_t)m.vars.size(),
                                                src.data(), dst.data(), by.data(), -1, m.d_count, m.cap, d_dummy_,
                                                nullptr, d_overflow_, f.append_scratch, f.append_scratch_bytes, st),
                              "message append");
        }
        tick("msgout:" + f.d->name, st, false);
        m.truncate = false;
        m.pbm_dirty = true;
    }
    // 2. death (CUDAFatAgent::processDeath): stable compaction of every variable
    if (f.d->has_agent_death) {
        std::vector<const void*> src;
        std::vector<void*> dst;
        std::vector<uint32_t> by;
        for (auto& v : l.vars) {
            src.push_back(v.cur);
            dst.push_back(v.alt);
            by.push_back(v.bytes);
        }
        tick("death:" + f.d->name, st, true);
        if (f.d->condition) {
            // disabled agents stay in front (CUDAFatAgentStateList::scatterDeath with disabledAgents)
            detail::fgx_check(fgx_compact_prefix(ctx_, f.death_flags, 1, f.d_first, l.d_count, l.cap, (uint32_t)l.vars.size(),
                                                 src.data(), dst.data(), by.data(), l.d_count, f.death_scratch,
                                                 f.death_scratch_bytes, st),
                              "death compaction");
        } else {
            detail::fgx_check(fgx_compact(ctx_, f.death_flags, 1, l.d_count, l.cap, (uint32_t)l.vars.size(), src.data(),
                                          dst.data(), by.data(), l.d_count, f.death_scratch, f.death_scratch_bytes, st),
                              "death compaction");
        }
        tick("death:" + f.d->name, st, false);
        for (auto& v : l.vars) std::swap(v.cur, v.alt);
    }
    // 3. state transition (CUDAFatAgent::transitionState): into a list known to be empty
    // the two lists swap buffers (equal capacities) and only the count moves; otherwise
    // every agent is appended to the end state
    if (f.in_state != f.end_state && f.swap_transition) {
        auto& dl = f.agent->states.at(f.end_state);
        tick("transition-swap:" + f.d->name, st, true);
        for (size_t i = 0; i < l.vars.size(); ++i) {
            std::swap(l.vars[i].cur, dl.vars[i].cur);
            std::swap(l.vars[i].alt, dl.vars[i].alt);
        }
        detail::fgx_check(fgx_memcpy_d2d(dl.d_count, l.d_count, 4, st), "state transition count");
        detail::fgx_check(fgx_fill_u32(l.d_count, 0, 1, st), "state transition count");
        tick("transition-swap:" + f.d->name, st, false);
    } else if (f.in_state != f.end_state) {
        auto& dl = f.agent->states.at(f.end_state);
        tick("transition:" + f.d->name, st, true);
        std::vector<const void*> src;
        std::vector<void*> dst;
        std::vector<uint32_t> by;
        for (size_t i = 0; i < l.vars.size(); ++i) {
            src.push_back(l.vars[i].cur);
            dst.push_back(dl.vars[i].cur);
            by.push_back(l.vars[i].bytes);
        }
        if (f.d->condition) {
            // only the agents that ran move; the disabled ones stay in the start state
            detail::fgx_check(fgx_append_range(ctx_, f.d_first, l.d_count, l.cap, (uint32_t)l.vars.size(), src.data(),
                                               dst.data(), by.data(), dl.d_count, dl.cap, d_overflow_, f.trans_scratch,
                                               f.trans_scratch_bytes, st),
                              "state transition");
            detail::fgx_check(fgx_memcpy_d2d(l.d_count, f.d_first, 4, st), "state transition count");
        } else {
            detail::fgx_check(fgx_append_births(ctx_, nullptr, l.d_count, l.cap, (uint32_t)l.vars.size(), src.data(),
                                                dst.data(), by.data(), -1, dl.d_count, dl.cap, d_dummy_, nullptr,
                                                d_overflow_, f.trans_scratch, f.trans_scratch_bytes, st),
                              "state transition");
            detail::fgx_check(fgx_fill_u32(l.d_count, 0, 1, st), "state transition count");
        }
        tick("transition:" + f.d->name, st, false);
    }
    // 4. births (CUDAAgentStateList::scatterNew): appended after the survivors, ordered IDs
    if (f.out_agent) {
        auto& dl = f.out_agent->states.at(f.out_state);
        std::vector<const void*> src;
        std::vector<void*> dst;
        std::vector<uint32_t> by;
        for (size_t i = 0; i < dl.vars.size(); ++i) {
            src.push_back(f.newbuf[i].cur);
            dst.push_back(dl.vars[i].cur);
            by.push_back(dl.vars[i].bytes);
        }
        tick("births:" + f.d->name, st, true);
        detail::fgx_check(fgx_append_births(ctx_, f.birth_flags, f.launch_count, f.cap, (uint32_t)dl.vars.size(), src.data(),
                                            dst.data(), by.data(), dl.id_slot, dl.d_count, dl.cap, f.out_agent->d_next_id,
                                            f.d_id_slots, d_overflow_, f.births_scratch, f.births_scratch_bytes, st),
                          "agent births");
        tick("births:" + f.d->name, st, false);
    }
}

inline void CUDASimulation::enqueue_layer(const LayerData& layer) {
    if (!layer.host_functions.empty()) {
        synchronize();
        for (auto fn : layer.host_functions) {
            fn(host_api_.get());
            flush_host_agents();
        }
        upload_env();
        return;
    }
    // auto spatial sort (CUDASimulation.cu:671-682) and lazy index builds.  Sorts move agent
    // lists and builds move message lists, so with in-layer concurrency on the builds run on a
    // stream of their own beside the sorts: both are chains of small latency-bound kernels.
    bool sorts = false, builds = false;
    for (auto& fd : layer.agent_functions) {
        auto& f = fns_[fn_index_.at(fd.get())];
        const unsigned period = f.d->parent->sortPeriod;
        sorts = sorts || (f.sort_dims && period && step_ % period == 0);
        builds = builds || (f.min && f.min->pbm_dirty &&
                            (f.min->spatial || f.min->kind == MsgKind::Bucket || is_array_kind(f.min->kind)));
    }
    const bool overlap = sorts && builds && cuda_config_.inLayerConcurrency && !timing_;
    void* bst = stream_;
    if (overlap) {
        if (!build_stream_) detail::fgx_check(fgx_stream_create(&build_stream_), "stream create");
        detail::fgx_check(fgx_event_record(build_fork_event_, stream_), "fork");
        detail::fgx_check(fgx_stream_wait_event(build_stream_, build_fork_event_), "fork wait");
        bst = build_stream_;
    }
    for (auto& fd : layer.agent_functions) {
        auto& f = fns_[fn_index_.at(fd.get())];
        const unsigned period = f.d->parent->sortPeriod;
        if (f.sort_dims && period && (step_ % period == 0)) {
            auto& l = f.agent->states.at(f.in_state);
            const int ix = l.var_index("x"), iy = l.var_index("y"), iz = l.var_index("z");
            std::vector<const void*> src;
            std::vector<void*> dst;
            std::vector<uint32_t> by;
            for (auto& v : l.vars) {
                src.push_back(v.cur);
                dst.push_back(v.alt);
                by.push_back(v.bytes);
            }
            tick("sort:" + f.agent->d->name, stream_, true);
            note_sort(l.sort_scratch);
            detail::fgx_check(fgx_agent_sort(ctx_, f.sort_dims, f.min->meta.min, f.min->meta.max, f.min->meta.radius, l.d_count,
                                             l.cap, (const float*)l.vars[ix].cur, (const float*)l.vars[iy].cur,
                                             f.sort_dims == 3 ? (const float*)l.vars[iz].cur : nullptr,
                                             (uint32_t)l.vars.size(), src.data(), dst.data(), by.data(), l.sort_scratch,
                                             l.sort_scratch_bytes, stream_),
                              "agent spatial sort");
            tick("sort:" + f.agent->d->name, stream_, false);
            for (auto& v : l.vars) std::swap(v.cur, v.alt);
        }
    }
    for (auto& fd : layer.agent_functions) {
        auto& f = fns_[fn_index_.at(fd.get())];
        if (f.min && f.min->spatial && f.min->pbm_dirty) {
            auto& m = *f.min;
            const int ix = m.var_index("x"), iy = m.var_index("y"), iz = m.var_index("z");
            std::vector<const void*> src;
            std::vector<void*> dst;
            std::vector<uint32_t> by;
            for (auto& v : m.vars) {
                src.push_back(v.cur);
                dst.push_back(v.alt);
                by.push_back(v.bytes);
            }
            tick("build:" + m.d->name, bst, true);
            note_sort(m.build_scratch);
            // the walk's {x, y, z, id} records are written by the build's scatter
            const int ii = m.var_index("id");
            const int32_t rec_vars[4] = {ix, iy, m.meta.dims == 3 ? iz : -1,
                                         ii >= 0 && m.vars[ii].bytes == 4 ? ii : -1};
            detail::fgx_check(fgx_spatial_build_records(ctx_, &m.meta, m.d_count, m.cap, (const float*)m.vars[ix].cur,
                                                        (const float*)m.vars[iy].cur,
                                                        m.meta.dims == 3 ? (const float*)m.vars[iz].cur : nullptr,
                                                        (uint32_t)m.vars.size(), src.data(), dst.data(), by.data(),
                                                        rec_vars, m.packed, m.pbm, m.build_scratch,
                                                        m.build_scratch_bytes, bst),
                              "message index build");
            for (auto& v : m.vars) std::swap(v.cur, v.alt);
            tick("build:" + m.d->name, bst, false);
            m.pbm_dirty = false;
        } else if (f.min && (f.min->kind == MsgKind::Bucket || is_array_kind(f.min->kind)) && f.min->pbm_dirty) {
            auto& m = *f.min;
            std::vector<const void*> src;
            std::vector<void*> dst;
            std::vector<uint32_t> by;
            for (auto& v : m.vars) {
                src.push_back(v.cur);
                dst.push_back(v.alt);
                by.push_back(v.bytes);
            }
            tick("build:" + m.d->name, bst, true);
            if (m.kind == MsgKind::Bucket) {
                const int ik = m.var_index("_key");
                note_sort(m.build_scratch);
                detail::fgx_check(fgx_bucket_build(ctx_, m.key_min, m.keys, m.d_count, m.cap, (const int32_t*)m.vars[ik].cur,
                                                   (uint32_t)m.vars.size(), src.data(), dst.data(), by.data(), m.pbm,
                                                   m.build_scratch, m.build_scratch_bytes, bst),
                                  "bucket message index build");
            } else {
                const int ii = m.var_index("___INDEX");
                detail::fgx_check(fgx_array_build(ctx_, m.length, m.d_count, m.cap, (const uint32_t*)m.vars[ii].cur,
                                                  (uint32_t)m.vars.size(), src.data(), dst.data(), by.data(), m.counts,
                                                  d_error_, bst),
                                  "array message build");
                // the list now holds every element (MessageArray.cu buildIndex: count = length)
                detail::fgx_check(fgx_fill_u32(m.d_count, m.length, 1, bst), "array message count");
            }
            tick("build:" + m.d->name, bst, false);
            for (auto& v : m.vars) std::swap(v.cur, v.alt);
            m.pbm_dirty = false;
        }
    }
    if (overlap) {
        detail::fgx_check(fgx_event_record(build_join_event_, build_stream_), "join");
        detail::fgx_check(fgx_stream_wait_event(stream_, build_join_event_), "join wait");
    }
    // function conditions (CUDASimulation.cu:687-835): condition kernel, then a stable
    // partition that puts the disabled agents in front; the function runs on the rest
    for (auto& fd : layer.agent_functions) {
        auto& f = fns_[fn_index_.at(fd.get())];
        if (!f.d->condition) continue;
        auto& l = f.agent->states.at(f.in_state);
        detail::FnArgs a;
        std::memset(&a, 0, sizeof(a));
        fill_condition_args(f, l, a);
        const uint32_t grid = std::max<uint32_t>(1, (l.cap + 255) / 256);
        tick("condition:" + f.d->name, stream_, true);
        f.d->condition(a, f.cond_flags, grid, stream_);
        if (hipGetLastError() != hipSuccess) throw exception::CUDAError("agent function condition launch failed: " + f.d->name);
        std::vector<const void*> src;
        std::vector<void*> dst;
        std::vector<uint32_t> by;
        for (auto& v : l.vars) {
            src.push_back(v.cur);
            dst.push_back(v.alt);
            by.push_back(v.bytes);
        }
        detail::fgx_check(fgx_partition(ctx_, f.cond_flags, l.d_count, l.cap, (uint32_t)l.vars.size(), src.data(), dst.data(),
                                        by.data(), f.d_first, f.d_work, f.cond_scratch, f.cond_scratch_bytes, stream_),
                          "function condition partition");
        tick("condition:" + f.d->name, stream_, false);
        for (auto& v : l.vars) std::swap(v.cur, v.alt);
    }
    const bool concurrent = cuda_config_.inLayerConcurrency && layer.agent_functions.size() > 1 && !timing_;
    if (!concurrent) {
        for (auto& fd : layer.agent_functions) {
            auto& f = fns_[fn_index_.at(fd.get())];
            launch_function(f, stream_);
            unmap_function(f, stream_);
        }
        return;
    }
    if (launch_layer_fused(layer)) return;
    // fork: each function of the layer on its own stream, joined back into the main stream
    detail::fgx_check(fgx_event_record(fork_event_, stream_), "fork");
    size_t k = 0;
    for (auto& fd : layer.agent_functions) {
        auto& f = fns_[fn_index_.at(fd.get())];
        if (!f.stream) detail::fgx_check(fgx_stream_create(&f.stream), "stream create");
        detail::fgx_check(fgx_stream_wait_event(f.stream, fork_event_), "fork wait");
        launch_function(f, f.stream);
        unmap_function(f, f.stream);
        while (join_events_.size() <= k) {
            void* e = nullptr;
            detail::fgx_check(fgx_event_create(&e), "event create");
            join_events_.push_back(e);
        }
        detail::fgx_check(fgx_event_record(join_events_[k], f.stream), "join");
        ++k;
    }
    for (size_t j = 0; j < k; ++j) detail::fgx_check(fgx_stream_wait_event(stream_, join_events_[j]), "join wait");
}

inline void CUDASimulation::enqueue_end_of_step() {
    // CUDASimulation.cu:619-624: non-persistent message lists are emptied
    std::vector<uint32_t*> zero;
    for (auto& kv : msgs_) {
        auto& m = kv.second;
        if (!m.d->persistent) {
            zero.push_back(m.d_count);
            m.truncate = true;
            m.pbm_dirty = true;
        }
    }
    // one launch: message counts to zero and the step counter incremented
    size_t k = 0;
    do {
        const uint32_t nz = (uint32_t)std::min<size_t>(zero.size() - k, FGX_SET_COUNTS_MAX);
        const std::vector<uint32_t> vals(nz, 0u);
        const bool last = k + nz >= zero.size();
        detail::fgx_check(fgx_set_counts(nz, nz ? zero.data() + k : nullptr, vals.data(), last ? d_step_ : nullptr, stream_),
                          "end of step counters");
        k += nz;
    } while (k < zero.size());
}

inline void CUDASimulation::enqueue_step() {
    for (auto& layer : m_->layers) enqueue_layer(*layer);
    enqueue_end_of_step();
}

inline std::vector<void*> CUDASimulation::buffer_state() const {
    std::vector<void*> st;
    for (auto& kv : agents_)
        for (auto& s : kv.second.states)
            for (auto& v : s.second.vars) {
                st.push_back(v.cur);
                st.push_back(v.alt);
            }
    for (auto& kv : msgs_)
        for (auto& v : kv.second.vars) {
            st.push_back(v.cur);
            st.push_back(v.alt);
        }
    return st;
}

inline void CUDASimulation::set_buffer_state(const std::vector<void*>& st) {
    size_t k = 0;
    for (auto& kv : agents_)
        for (auto& s : kv.second.states)
            for (auto& v : s.second.vars) {
                v.cur = st[k++];
                v.alt = st[k++];
            }
    for (auto& kv : msgs_)
        for (auto& v : kv.second.vars) {
            v.cur = st[k++];
            v.alt = st[k++];
        }
}

inline bool CUDASimulation::step() {
    const auto t0 = std::chrono::steady_clock::now();
    ensure_init();
    if (!init_functions_ran_) run_init_functions();
    bool changed = false;
    bool counted = false;
    if (model_can_grow() || step_ == 0 || populations_set_) {
        populations_set_ = false;
        sync_counts();
        changed = ensure_capacity();
        counted = true;
    }
    plan_transitions(counted);
    upload_env();
    const bool use_graph = cuda_config_.graphs && graph_eligible_ && !timing_;
    if (use_graph) {
        if (changed) clear_graphs();
        // One single-step graph per double-buffer arrangement at step start; a step
        // swaps buffers, so a model alternates between a small number of arrangements.
        // a graph whose capture uploaded layer images repeats the upload on every replay:
        // drop it once it has run, so the arrangement is captured again without it
        bool stale = false;
        for (auto& g : graphs_) stale = stale || g.stale;
        if (stale) {
            synchronize();
            for (size_t k = graphs_.size(); k-- > 0;)
                if (graphs_[k].stale) {
                    fgx_graph_destroy(graphs_[k].g);
                    graphs_.erase(graphs_.begin() + (std::ptrdiff_t)k);
                }
        }
        const auto start = buffer_state();
        const auto swaps = swap_plan();
        GraphEntry* hit = nullptr;
        for (auto& g : graphs_)
            if (g.start == start && g.swaps == swaps && g.concurrent == cuda_config_.inLayerConcurrency) hit = &g;
        if (!hit && graphs_.size() >= kMaxStepGraphs) {
            size_t lru = 0;
            for (size_t k = 1; k < graphs_.size(); ++k)
                if (graphs_[k].last_use < graphs_[lru].last_use) lru = k;
            if (!graph_evict_warned_) {
                std::fprintf(stderr,
                             "flamegpu: more than %zu buffer arrangements at step start; replacing the least "
                             "recently used step graph (each new arrangement is captured once)\n",
                             kMaxStepGraphs);
                graph_evict_warned_ = true;
            }
            fgx_graph_destroy(graphs_[lru].g);
            graphs_.erase(graphs_.begin() + (std::ptrdiff_t)lru);
        }
        if (!hit) {
            GraphEntry e;
            e.start = start;
            e.swaps = swaps;
            e.concurrent = cuda_config_.inLayerConcurrency;
            new_layer_images_ = false;
            first_sorts_ = false;
            detail::fgx_check(fgx_graph_begin(stream_), "graph begin");
            enqueue_step();
            detail::fgx_check(fgx_graph_end(stream_, &e.g), "graph end");
            e.stale = new_layer_images_ || first_sorts_;
            e.end = buffer_state();
            set_buffer_state(start);   // capture recorded the work; nothing has run yet
            graphs_.push_back(std::move(e));
            hit = &graphs_.back();
        }
        hit->last_use = ++graph_uses_;
        detail::fgx_check(fgx_graph_launch(hit->g, stream_), "graph launch");
        set_buffer_state(hit->end);
    } else {
        enqueue_step();
    }
    ++step_;
    for (auto fn : m_->step_functions) {
        synchronize();
        fn(host_api_.get());
        flush_host_agents();
        upload_env();
    }
    bool cont = true;
    for (auto fn : m_->exit_conditions) {
        synchronize();
        if (fn(host_api_.get()) == EXIT) cont = false;
    }
    if (config_.timing) synchronize();
    step_times_.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    process_step_log(step_times_.back());
    return cont;
}

inline void CUDASimulation::run_init_functions() {
    if (init_functions_ran_) return;
    init_functions_ran_ = true;
    const auto t0 = std::chrono::steady_clock::now();
    ensure_init();
    for (auto fn : m_->init_functions) {
        synchronize();
        fn(host_api_.get());
        flush_host_agents();
    }
    upload_env();
    elapsed_init_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Host-created agents join their lists after the survivors, in creation order, with the
// IDs they were given at creation (HostAgentAPI::newAgent).
inline void CUDASimulation::flush_host_agents() {
    if (!host_agents_pending()) return;
    sync_counts();
    for (auto& kv : host_new_) {
        if (!kv.second.v || kv.second.v->empty()) continue;
        detail::StateList& l = *kv.first;
        AgentVector& v = *kv.second.v;
        const uint32_t k = (uint32_t)v.size(), n = l.h_count;
        if ((uint64_t)n + k > 0x7FFFFFFFull) throw exception::InvalidOperation("population exceeds the 32-bit index range");
        if (n + k > l.cap) grow_list(l, std::max<uint32_t>(n + k, l.cap + l.cap / 2));
        for (auto& var : l.vars)
            detail::fgx_check(fgx_memcpy_h2d((char*)var.cur + (size_t)n * var.bytes, v.raw(var.name).data(),
                                             (size_t)k * var.bytes, stream_),
                              "host agent upload");
        l.h_count = n + k;
        detail::fgx_check(fgx_memcpy_h2d(l.d_count, &l.h_count, 4, stream_), "count upload");
        synchronize();
        v.clear();
    }
    for (auto& kv : agents_)
        detail::fgx_check(fgx_memcpy_h2d(kv.second.d_next_id, &kv.second.h_next_id, 4, stream_), "next id upload");
    synchronize();
    clear_graphs();
    populations_set_ = true;
    ensure_capacity();   // a host layer function may add agents in the middle of a step
}

inline void CUDASimulation::run_exit_functions() {
    const auto t0 = std::chrono::steady_clock::now();
    for (auto fn : m_->exit_functions) {
        synchronize();
        fn(host_api_.get());
        flush_host_agents();
    }
    elapsed_exit_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------ step / exit logs
// Record layout of one logged frame: per (agent, state) an 8-byte count slot when
// the count is logged, then 32 bytes per logged variable (fgx_reduce_stats output).
inline std::unique_ptr<CUDASimulation::LogPlan> CUDASimulation::make_log_plan(const LoggingConfig& cfg) {
    std::unique_ptr<LogPlan> p(new LogPlan());
    size_t off = 0;
    for (auto& kv : cfg.agents()) {
        LogAgentPlan a;
        a.agent = kv.first.first;
        a.state = kv.first.second;
        a.count = kv.second->count;
        a.reductions = kv.second->reductions;
        a.count_off = off;   // captured even when not logged: mean and deviation divide by it
        off += 8;
        const auto& ad = m_->agents.at(a.agent);
        std::set<std::string> names;
        for (auto& r : a.reductions) names.insert(r.name);
        for (auto& n : names) {
            const auto& v = ad->variables.at(n);
            a.vars.push_back(LogVarPlan{n, detail::stats_type_code(v.type), v.type, off});
            off += 32;
        }
        p->agents.push_back(std::move(a));
    }
    p->record_bytes = std::max<size_t>(off, 8);
    return p;
}

inline void CUDASimulation::enqueue_log_frame(const LogPlan& plan, char* record) {
    size_t sb = 0;
    detail::fgx_check(fgx_scratch_bytes(FGX_OP_STATS, 0, 0, &sb), "log scratch sizing");
    if (!log_scratch_) alloc(&log_scratch_, sb);
    for (auto& a : plan.agents) {
        auto& ar = agent_rt(a.agent);
        auto it = ar.states.find(a.state);
        if (it == ar.states.end()) throw exception::InvalidAgentState("Agent '" + a.agent + "' has no state '" + a.state + "'");
        detail::StateList& l = it->second;
        detail::fgx_check(fgx_memcpy_d2d(record + a.count_off, l.d_count, 4, stream_), "log count");
        for (auto& v : a.vars) {
            const int vi = l.var_index(v.name);
            detail::fgx_check(fgx_reduce_stats(ctx_, v.type, l.vars[vi].cur, l.d_count, l.cap, record + v.off, log_scratch_, sb,
                                               stream_),
                              "log reduction");
        }
    }
}


inline void CUDASimulation::decode_log_frame(const LoggingConfig& cfg, const LogPlan& plan, const char* record,
                                             LogFrame& f) const {
    (void)cfg;
    for (auto& a : plan.agents) {
        auto& entry = f.agents[{a.agent, a.state}];
        uint32_t n = 0;
        std::memcpy(&n, record + a.count_off, 4);
        entry.second = a.count ? n : UINT_MAX;
        for (auto& r : a.reductions) {
            const LogVarPlan* vp = nullptr;
            for (auto& v : a.vars)
                if (v.name == r.name) vp = &v;
            const char* w = record + vp->off;
            const uint32_t cnt = n;
            detail::LogValue v;
            switch (vp->type) {
                case 0: detail::decode_stats<float, double>(w, cnt, r.reduction, v); break;
                case 1: detail::decode_stats<double, double>(w, cnt, r.reduction, v); break;
                case 2: detail::decode_stats<int32_t, long long>(w, cnt, r.reduction, v); break;
                case 3: detail::decode_stats<uint32_t, unsigned long long>(w, cnt, r.reduction, v); break;
                case 4: detail::decode_stats<int64_t, long long>(w, cnt, r.reduction, v); break;
                default: detail::decode_stats<uint64_t, unsigned long long>(w, cnt, r.reduction, v); break;
            }
            entry.first.emplace(r, std::move(v));
        }
    }
}

inline std::map<std::string, detail::LogValue> CUDASimulation::log_env(const LoggingConfig& cfg) const {
    std::map<std::string, detail::LogValue> out;
    for (auto& name : cfg.environment()) {
        const auto& p = m_->env.props.at(name);
        detail::LogValue v;
        v.type = p.type;
        v.elements = (unsigned int)p.elements;
        v.bytes.assign(env_host_.data() + p.offset, env_host_.data() + p.offset + p.type_size * p.elements);
        out.emplace(name, std::move(v));
    }
    return out;
}

// src/flamegpu/simulation/CUDASimulation.cu: resetLog()
inline void CUDASimulation::reset_log() {
    log_pending_.clear();
    run_log_ = RunLog();
    run_log_.random_seed = config_.random_seed;
    run_log_.step_log_frequency = step_log_cfg_ ? step_log_cfg_->frequency() : 0;
    run_log_.performance_specs.device_name = "AMD Instinct MI355X (gfx950)";
    run_log_.performance_specs.device_cc_major = 9;
    run_log_.performance_specs.device_cc_minor = 5;
}

// src/flamegpu/simulation/CUDASimulation.cu:1949-1983: a frame every `frequency` steps.
inline void CUDASimulation::process_step_log(double step_time) {
    if (!step_log_cfg_) return;
    const unsigned int freq = step_log_cfg_->frequency();
    if (freq == 0 || step_ % freq != 0) return;
    if (!step_plan_) {
        flush_step_log();
        step_plan_ = make_log_plan(*step_log_cfg_);
        void* p = d_log_;
        free_dev(p);
        d_log_ = nullptr;
        log_cap_frames_ = 0;
    }
    if (!d_log_) {
        log_cap_frames_ = 256;
        void* p = nullptr;
        alloc(&p, log_cap_frames_ * step_plan_->record_bytes);
        d_log_ = static_cast<char*>(p);
    }
    if (log_pending_.size() == log_cap_frames_) flush_step_log();
    enqueue_log_frame(*step_plan_, d_log_ + log_pending_.size() * step_plan_->record_bytes);
    log_pending_.push_back(PendingLogFrame{step_, step_time, log_env(*step_log_cfg_)});
}

inline void CUDASimulation::flush_step_log() {
    if (log_pending_.empty()) return;
    std::vector<char> host(log_pending_.size() * step_plan_->record_bytes);
    detail::fgx_check(fgx_memcpy_d2h(host.data(), d_log_, host.size(), stream_), "log readback");
    synchronize();
    for (size_t i = 0; i < log_pending_.size(); ++i) {
        StepLogFrame f;
        f.step_count = log_pending_[i].step;
        f.step_time = log_pending_[i].step_time;
        f.environment = std::move(log_pending_[i].env);
        decode_log_frame(*step_log_cfg_, *step_plan_, host.data() + i * step_plan_->record_bytes, f);
        run_log_.step.push_back(std::move(f));
    }
    log_pending_.clear();
}

// src/flamegpu/simulation/CUDASimulation.cu:1985-2021
inline void CUDASimulation::process_exit_log() {
    if (!exit_log_cfg_) return;
    if (!exit_plan_) exit_plan_ = make_log_plan(*exit_log_cfg_);
    if (exit_log_bytes_ < exit_plan_->record_bytes) {
        void* p = d_exit_log_;
        free_dev(p);
        alloc(&p, exit_plan_->record_bytes);
        d_exit_log_ = static_cast<char*>(p);
        exit_log_bytes_ = exit_plan_->record_bytes;
    }
    enqueue_log_frame(*exit_plan_, d_exit_log_);
    std::vector<char> host(exit_plan_->record_bytes);
    detail::fgx_check(fgx_memcpy_d2h(host.data(), d_exit_log_, host.size(), stream_), "exit log readback");
    synchronize();
    ExitLogFrame f;
    f.step_count = step_;
    f.environment = log_env(*exit_log_cfg_);
    decode_log_frame(*exit_log_cfg_, *exit_plan_, host.data(), f);
    f.init_time = elapsed_init_;
    f.exit_time = elapsed_exit_;
    f.total_time = elapsed_sim_;
    run_log_.exit = std::move(f);
}

// ------------------------------------------------------------------ HostAPI implementation
inline HostAgentAPI HostAPI::agent(const std::string& name, const std::string& state) {
    auto& a = s_.agent_rt(name);
    const std::string st = state.empty() ? a.d->initial_state : state;
    auto it = a.states.find(st);
    if (it == a.states.end()) throw exception::InvalidAgentState("Agent '" + name + "' has no state '" + st + "'");
    return HostAgentAPI(s_, a, it->second);
}
inline unsigned int HostAPI::getStepCounter() const { return s_.getStepCounter(); }

inline unsigned int HostAgentAPI::count() {
    s_.sync_counts();
    return l_.h_count;
}

inline AgentVector::Agent HostAgentAPI::newAgent() {
    if (!s_.host_agents_pending()) s_.sync_counts();   // the device's next ID, advanced by births since the last read
    AgentVector& v = s_.host_new_agents(a_, l_);
    v.push_back();
    AgentVector::Agent ag = v.back();
    ag.setVariable<id_t>("_id", a_.h_next_id++);
    return ag;
}

template <typename T>
inline const detail::VarBuf& HostAgentAPI::var(const std::string& name) {
    auto vit = a_.d->variables.find(name);
    if (vit == a_.d->variables.end())
        throw exception::InvalidAgentVar("Agent '" + a_.d->name + "' does not contain variable '" + name + "'");
    if (vit->second.type != std::type_index(typeid(T)) || vit->second.elements != 1)
        throw exception::InvalidVarType("Variable '" + name + "' is not of the requested type");
    return l_.vars[l_.var_index(name)];
}

template <typename InT, typename OutT>
inline OutT HostAgentAPI::sum(const std::string& name) {
    const auto& v = var<InT>(name);
    s_.sync_counts();
    int type;
    if (std::is_same<InT, float>::value) type = 0;
    else if (std::is_same<InT, double>::value) type = 1;
    else if (std::is_same<InT, int32_t>::value) type = 2;
    else if (std::is_same<InT, uint32_t>::value) type = 3;
    else if (std::is_same<InT, int64_t>::value || std::is_same<InT, long long>::value) type = 4;
    else if (std::is_same<InT, uint64_t>::value || std::is_same<InT, unsigned long long>::value) type = 5;
    else {
        auto vals = getVariableValues<InT>(name);
        OutT r = 0;
        for (auto x : vals) r += (OutT)x;
        return r;
    }
    size_t sb = 0;
    detail::fgx_check(fgx_scratch_bytes(FGX_OP_SCAN, std::max<uint32_t>(l_.cap, 1), 0, &sb), "reduction scratch");
    void* scratch = s_.reduce_scratch(sb);
    detail::fgx_check(fgx_reduce_sum(s_.ctx(), type, v.cur, l_.d_count, l_.cap, s_.reduce_out(), scratch, sb, s_.stream()),
                      "reduction");
    double dres = 0;
    long long ires = 0;
    if (type <= 1) {
        detail::fgx_check(fgx_memcpy_d2h(&dres, s_.reduce_out(), 8, s_.stream()), "reduction readback");
        s_.synchronize();
        return (OutT)dres;
    }
    detail::fgx_check(fgx_memcpy_d2h(&ires, s_.reduce_out(), 8, s_.stream()), "reduction readback");
    s_.synchronize();
    if (type == 5 || type == 3) return (OutT)(unsigned long long)ires;
    return (OutT)ires;
}

template <typename T>
inline std::vector<T> HostAgentAPI::getVariableValues(const std::string& name) {
    const auto& v = var<T>(name);
    s_.sync_counts();
    std::vector<T> out(l_.h_count);
    if (l_.h_count)
        detail::fgx_check(fgx_memcpy_d2h(out.data(), v.cur, (size_t)l_.h_count * sizeof(T), s_.stream()), "variable readback");
    s_.synchronize();
    return out;
}

template <typename T>
inline void HostAgentAPI::stats(const std::string& name, char out[32]) {
    const auto& v = var<T>(name);
    const int type = detail::stats_type_code(std::type_index(typeid(T)));
    size_t sb = 0;
    detail::fgx_check(fgx_scratch_bytes(FGX_OP_STATS, 0, 0, &sb), "statistics scratch");
    void* scratch = s_.reduce_scratch(sb);
    detail::fgx_check(fgx_reduce_stats(s_.ctx(), type, v.cur, l_.d_count, l_.cap, s_.reduce_out(), scratch, sb, s_.stream()),
                      "statistics");
    detail::fgx_check(fgx_memcpy_d2h(out, s_.reduce_out(), 32, s_.stream()), "statistics readback");
    s_.synchronize();
}
template <typename T>
inline T HostAgentAPI::min(const std::string& name) {
    if (detail::stats_type_code(std::type_index(typeid(T))) < 0) {
        auto v = getVariableValues<T>(name);
        return v.empty() ? std::numeric_limits<T>::max() : *std::min_element(v.begin(), v.end());
    }
    char w[32];
    stats<T>(name, w);
    detail::LogValue r;
    if (std::is_floating_point<T>::value) detail::decode_stats<T, double>(w, 0, LoggingConfig::Min, r);
    else if (std::is_signed<T>::value) detail::decode_stats<T, long long>(w, 0, LoggingConfig::Min, r);
    else detail::decode_stats<T, unsigned long long>(w, 0, LoggingConfig::Min, r);
    return r.get<T>();
}
template <typename T>
inline T HostAgentAPI::max(const std::string& name) {
    if (detail::stats_type_code(std::type_index(typeid(T))) < 0) {
        auto v = getVariableValues<T>(name);
        return v.empty() ? std::numeric_limits<T>::lowest() : *std::max_element(v.begin(), v.end());
    }
    char w[32];
    stats<T>(name, w);
    detail::LogValue r;
    if (std::is_floating_point<T>::value) detail::decode_stats<T, double>(w, 0, LoggingConfig::Max, r);
    else if (std::is_signed<T>::value) detail::decode_stats<T, long long>(w, 0, LoggingConfig::Max, r);
    else detail::decode_stats<T, unsigned long long>(w, 0, LoggingConfig::Max, r);
    return r.get<T>();
}
template <typename T>
inline std::pair<double, double> HostAgentAPI::meanStandardDeviation(const std::string& name) {
    if (detail::stats_type_code(std::type_index(typeid(T))) < 0)
        throw exception::UnsupportedVarType("meanStandardDeviation supports float, double and 32/64-bit integers");
    const unsigned int n = count();
    char w[32];
    stats<T>(name, w);
    detail::LogValue mean, sd;
    if (std::is_floating_point<T>::value) {
        detail::decode_stats<T, double>(w, n, LoggingConfig::Mean, mean);
        detail::decode_stats<T, double>(w, n, LoggingConfig::StandardDev, sd);
    } else if (std::is_signed<T>::value) {
        detail::decode_stats<T, long long>(w, n, LoggingConfig::Mean, mean);
        detail::decode_stats<T, long long>(w, n, LoggingConfig::StandardDev, sd);
    } else {
        detail::decode_stats<T, unsigned long long>(w, n, LoggingConfig::Mean, mean);
        detail::decode_stats<T, unsigned long long>(w, n, LoggingConfig::StandardDev, sd);
    }
    return {mean.get<double>(), sd.get<double>()};
}

template <typename T>
inline T HostEnvironment::getProperty(const std::string& name) const {
    return s_.getEnvironmentProperty<T>(name);
}
template <typename T>
inline T HostEnvironment::setProperty(const std::string& name, T value) {
    const T old = s_.getEnvironmentProperty<T>(name);
    s_.setEnvironmentProperty<T>(name, value);
    return old;
}
template <typename T, unsigned int N>
inline std::array<T, N> HostEnvironment::getProperty(const std::string& name) const {
    return s_.getEnvironmentProperty<T, N>(name);
}
template <typename T, unsigned int N>
inline std::array<T, N> HostEnvironment::setProperty(const std::string& name, const std::array<T, N>& value) {
    const std::array<T, N> old = s_.getEnvironmentProperty<T, N>(name);
    s_.setEnvironmentProperty<T, N>(name, value);
    return old;
}

// ------------------------------------------------------------------ ensembles
// Per-run bookkeeping of an ensemble (plan, seed, wall time, final counts, error).
struct RunRecord {
    bool started = false;         // the run began (Fast error handling stops the others)
    unsigned int plan_index = 0;
    uint64_t seed = 0;
    unsigned int steps = 0;
    double wall_seconds = 0.0;
    std::map<std::string, unsigned int> final_counts;
    std::string error;
};

namespace util {
inline void cleanup() { fgx_device_synchronize(); }
}  // namespace util

}  // namespace flamegpu
