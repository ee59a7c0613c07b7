// Batched simulations: many independent simulations of one model co-scheduled in the
// same launches (one segmented launch per stage), with simulations joining and
// leaving between steps.
//
// The reference runs every simulation of an ensemble with its own launch sequence
// (CUDAEnsemble::simulate, src/flamegpu/simulation/CUDAEnsemble.cu; SimRunner::main,
// src/flamegpu/simulation/detail/SimRunner.cu:51-90), so small simulations are bound
// by launch latency.  Here the lists of all resident simulations are packed one
// after another in HBM and every stage (agent sort, message-list build, each agent
// function) is one launch over the batch: the sort/build use the segmented
// primitives of libfgx (fgx_agent_sort_seg, fgx_spatial_build_seg) and the agent
// functions the segmented kernel (agent_function_kernel_seg), which gives each
// simulation its own bin table, environment, step counter, seed and error words.
// A simulation's results do not depend on what it shares the batch with: IDs,
// RNG streams (keyed by seed, ID, step), sort order and message order are those of
// the same simulation run alone.
//
// Populations may change: agent death and agent output (births into an agent's initial
// state) run over the whole batch.  A stable compaction of the packed list keeps each
// simulation's survivors together and in order, each simulation's new agents are placed
// after its own agents (IDs from its own counter), and per-simulation counts stay on the
// device, where fgx_seg_layout patches every segment table after each change.  The host
// reads the totals back once per step to keep capacities ahead of the step's bound.
//
// Supported models: agents in their initial state (no state transitions), mandatory
// message output, every message list written by one agent function before it is read
// in the step, no function conditions, no host functions inside layers and no step
// functions or exit conditions (they need per-simulation host logic).  AgentOut::getID()
// reports a device error in a batch: children are numbered per simulation after the launch.
#pragma once

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "flamegpu/simulation.h"

namespace flamegpu {

// Geometry of a spatial message list for one simulation of a batch (per-simulation
// radius and bounds; the reference fixes them in the model description).
struct SpatialBounds {
    float min[3] = {0.0f, 0.0f, 0.0f};
    float max[3] = {0.0f, 0.0f, 0.0f};
    float radius = 0.0f;
};

class BatchSimulation {
 public:
    // Launch shape of the agent-function launches (the autotuner's search space).
    struct Tuning {
        uint32_t block = 256;             // threads per block, a multiple of 64 up to 1024
        uint32_t agents_per_thread = 1;   // agents each thread runs
        uint32_t blocks_per_cu = 0;       // 0: one block per tile; otherwise a persistent grid
    };
    // One simulation.
    struct Plan {
        uint64_t seed = 0;
        unsigned int steps = 1;
        std::vector<char> env;                          // environment buffer (empty: model defaults)
        std::map<std::string, SpatialBounds> messages;  // spatial geometry overrides by message name
        std::map<std::string, std::shared_ptr<AgentVector>> populations;   // by agent name (initial state)
    };

    // host_phases_outside: the caller runs the model's init and exit functions itself
    // (CUDAEnsemble does, on a simulation of its own per run), so the model may have them.
    explicit BatchSimulation(const ModelDescription& model, int device = 0, bool host_phases_outside = false)
        : m_(model.data()), device_(device) {
        check_model(*m_, host_phases_outside);
    }
    // Throws InvalidOperation naming the first feature of the model a batch cannot run.
    static void check_model(const ModelData& m, bool host_phases_outside);
    BatchSimulation(const BatchSimulation&) = delete;
    BatchSimulation& operator=(const BatchSimulation&) = delete;
    ~BatchSimulation() { release(); }

    // Queues a simulation; it joins the batch at the next step.  Returns its handle.
    int add(const Plan& p);
    // HBM bytes a simulation occupies once resident (lists, bin tables, scratch share).
    size_t estimateBytes(const Plan& p) const;
    // One step of every resident simulation that has steps left.
    void step();
    // Handles whose steps are done (their state stays resident until remove()).
    std::vector<int> finished() const;
    unsigned int stepsDone(int h) const { return sim(h).steps_done; }
    void getPopulationData(int h, AgentVector& pop);
    void remove(int h);
    void setTuning(const Tuning& t);
    const Tuning& tuning() const { return tuning_; }
    void synchronize() {
        if (stream_) detail::fgx_check(fgx_stream_synchronize(stream_), "stream synchronize");
    }
    size_t residentBytes() const { return resident_bytes_; }
    // Resident simulations, not counting those removed since the last step.
    uint32_t residentCount() const { return (uint32_t)(order_.size() - removing_.size()); }
    uint32_t pendingCount() const { return (uint32_t)pending_.size(); }
    // Agents over all resident simulations (agent-steps per step).
    uint64_t residentAgents() const;
    // Per-wave unit tables for spatial iteration (CUDASimulation::setUnitTables).
    void setUnitTables(bool on) {
        unit_tables_ = on;
        clear_graphs();
    }
    void setGraphs(bool on) {
        graphs_on_ = on;
        clear_graphs();
    }
    // Bin-width multiplier (1 or 0.5; CUDASimulation::setBinWidthMultiplier) of the
    // simulations added from now on.
    void setBinWidthMultiplier(float mult) {
        if (!(mult == 1.0f || mult == 0.5f)) throw exception::InvalidArgument("bin-width multiplier must be 1 or 0.5");
        bin_mult_ = mult;
    }
    float binWidthMultiplier() const { return bin_mult_; }
    void* stream() {
        init();
        return stream_;
    }

 private:
    struct Sim {
        int handle = -1;
        uint64_t seed = 0;
        unsigned int steps_target = 0, steps_done = 0;
        std::vector<char> env;
        std::map<std::string, fgx_spatial_meta> meta;            // spatial message lists
        std::map<std::string, uint32_t> n;                        // agent name -> population (as of the last repack)
        std::map<std::string, uint32_t> next_id;                  // agent name -> next agent ID (as of the last repack)
        std::map<std::string, std::shared_ptr<AgentVector>> pop;  // until uploaded
        bool resident = false;
    };
    struct List {   // one agent list (initial state) or message list, packed over the batch
        std::vector<detail::VarBuf> vars;
        uint32_t cap = 0;
        uint32_t total = 0;
        std::vector<uint32_t> off;   // per resident simulation (batch order), host copy as of the last repack
        std::vector<uint32_t> cnt;
        // device copies, kept by fgx_seg_layout while births and deaths change the populations
        uint32_t* d_off = nullptr;
        uint32_t* d_cnt = nullptr;
        uint32_t* d_total = nullptr;
        int id_slot = -1;
        int var_index(const std::string& n) const {
            for (size_t i = 0; i < vars.size(); ++i)
                if (vars[i].name == n) return (int)i;
            return -1;
        }
    };
    struct MsgB {
        std::shared_ptr<MessageData> d;
        List list;
        std::string src_agent;        // the agent whose function writes the list
        bool spatial = false;
        int dims = 0;
        uint32_t* pbm = nullptr;
        uint32_t pbm_cap = 0;
        fgx_seg_key* d_keys = nullptr;
        uint32_t keys_cap = 0;
        uint32_t keys_total = 0, tiles = 0;
        std::vector<fgx_seg_key> keys;
        void* scratch = nullptr;
        size_t scratch_bytes = 0;
        void* packed = nullptr;       // spatial: 16-byte {x, y, z, id} record per message
        uint32_t packed_cap = 0;
        bool dirty = true;
    };
    struct AgentB {
        std::shared_ptr<AgentData> d;
        List list;
        std::string sort_msg;         // spatial message that drives the agent sort ("" = none)
        fgx_seg_key* d_keys = nullptr;
        uint32_t keys_cap = 0;
        uint32_t keys_total = 0, tiles = 0;
        void* scratch = nullptr;
        size_t scratch_bytes = 0;
        // changing populations: per-simulation next IDs, new agents and append work space
        uint32_t* d_next_id = nullptr;
        uint32_t* d_born = nullptr;
        uint32_t* d_work = nullptr;
    };
    struct FnB {
        std::shared_ptr<AgentFunctionData> d;
        AgentB* agent = nullptr;
        MsgB* in = nullptr;
        MsgB* out = nullptr;
        detail::SegDev* d_segs = nullptr;
        uint32_t segs_cap = 0;
        uint32_t tiles = 0;
        uint32_t* msg_flags = nullptr;   // message-output flags, one per agent of the list
        uint32_t flags_cap = 0;
        // agent death and agent output
        AgentB* out_agent = nullptr;
        uint32_t* death_flags = nullptr;
        uint32_t* birth_flags = nullptr;
        std::vector<detail::VarBuf> newbuf;   // new agents by parent index (cur) and compacted (alt)
        uint32_t pop_cap = 0;                 // capacity of the buffers above
        void* death_scratch = nullptr;
        size_t death_scratch_bytes = 0;
        void* birth_scratch = nullptr;
        size_t birth_scratch_bytes = 0;
    };

    Sim& sim(int h) {
        auto it = sims_.find(h);
        if (it == sims_.end()) throw exception::InvalidArgument("BatchSimulation: unknown simulation handle");
        return it->second;
    }
    const Sim& sim(int h) const {
        auto it = sims_.find(h);
        if (it == sims_.end()) throw exception::InvalidArgument("BatchSimulation: unknown simulation handle");
        return it->second;
    }
    void init();
    void release();
    void alloc(void** p, size_t bytes) { detail::fgx_check(fgx_malloc(ctx_, bytes ? bytes : 4, p), "device allocation"); }
    void free_dev(void*& p) {
        if (p) fgx_free(ctx_, p);
        p = nullptr;
    }
    template <typename T>
    void free_typed(T*& p) {
        void* q = p;
        free_dev(q);
        p = nullptr;
    }
    void rebuild();
    void repack(List& l, const std::vector<uint32_t>& keep_old_slot, const std::vector<uint32_t>& new_cnt);
    // changing populations
    void read_counts();
    void ensure_step_capacity();
    void grow_agent_list(AgentB& a, uint32_t cap);
    void size_pop_buffers(FnB& f);
    void size_sort_scratch(AgentB& a);
    void size_records(MsgB& m);
    void size_build_scratch(MsgB& m);
    void upload_layout(List& l, uint32_t S);
    std::vector<fgx_seg_target> agent_targets(AgentB& a);
    std::vector<fgx_seg_target> message_targets(MsgB& m);
    void relayout(List& l, const std::vector<fgx_seg_target>& t, const uint32_t* d_add);
    uint32_t tile_bound(uint32_t cap, uint32_t tile) const { return (cap + tile - 1) / tile + (uint32_t)order_.size(); }
    void after_launch(FnB& f);
    void enqueue_step();
    void launch(FnB& f);
    void clear_graphs() {
        for (auto& g : graphs_) fgx_graph_destroy(g.g);
        graphs_.clear();
    }
    std::vector<void*> buffer_state() const;
    void set_buffer_state(const std::vector<void*>& st);
    void check_errors(const Sim& s, uint32_t slot);

    std::shared_ptr<ModelData> m_;
    int device_ = 0;
    bool initialised_ = false;
    fgx_ctx* ctx_ = nullptr;
    void* stream_ = nullptr;
    // index builds beside the agent sorts of a layer
    void* build_stream_ = nullptr;
    void* build_fork_event_ = nullptr;
    void* build_join_event_ = nullptr;
    Tuning tuning_;
    int cu_count_ = 256;
    std::map<std::string, AgentB> agents_;
    std::map<std::string, MsgB> msgs_;
    std::vector<FnB> fns_;
    std::map<const AgentFunctionData*, size_t> fn_index_;
    size_t env_stride_ = 64;
    std::map<int, Sim> sims_;
    std::vector<int> order_;                  // resident simulations in batch order
    std::vector<int> pending_;                // queued simulations
    std::set<int> removing_;
    int next_handle_ = 1;
    bool dirty_ = true;
    // per-simulation device data (batch order)
    char* d_env_ = nullptr;
    uint32_t* d_steps_ = nullptr;
    uint32_t* d_errors_ = nullptr;
    uint32_t sim_cap_ = 0;
    void* d_pages_ = nullptr;
    uint32_t* d_ranges_ = nullptr;
    uint32_t ranges_cap_ = 0;
    size_t resident_bytes_ = 0;
    // graphs keyed by the buffer arrangement at step start
    struct GraphEntry {
        std::vector<void*> start, end;
        fgx_graph* g = nullptr;
    };
    std::vector<GraphEntry> graphs_;
    float bin_mult_ = 1.0f;
    bool graphs_on_ = true;
    bool unit_tables_ = true;
    uint32_t stable_steps_ = 0;   // steps since the last repack
    bool dynamic_ = false;        // some function removes or adds agents
    uint32_t* d_word_ = nullptr;  // a count nobody reads (the births compaction's total)
};

// ------------------------------------------------------------------ implementation
inline void BatchSimulation::check_model(const ModelData& m, bool host_phases_outside) {
    if (!m.step_functions.empty() || !m.exit_conditions.empty())
        throw exception::InvalidOperation("BatchSimulation: models with host step functions or exit conditions cannot be batched");
    if (!host_phases_outside && (!m.init_functions.empty() || !m.exit_functions.empty()))
        throw exception::InvalidOperation("BatchSimulation: models with host init/exit functions cannot be batched");
    std::map<std::string, int> writers;
    std::set<std::string> written;
    for (auto& layer : m.layers) {
        if (!layer->host_functions.empty())
            throw exception::InvalidOperation("BatchSimulation: layers with host functions cannot be batched");
        for (auto& f : layer->agent_functions) {
            if (f->initial_state != f->end_state || f->initial_state != f->parent->initial_state)
                throw exception::InvalidOperation("BatchSimulation: agent function '" + f->name +
                                                  "' changes agent state; batched agents stay in their initial state");
            if (auto oa = f->agent_output.lock())
                if (f->agent_output_state != oa->initial_state)
                    throw exception::InvalidOperation("BatchSimulation: agent function '" + f->name +
                                                      "' creates agents outside their initial state");
            if (f->message_output_optional)
                throw exception::InvalidOperation("BatchSimulation: optional message output ('" + f->name + "') cannot be batched");
            if (f->condition)
                throw exception::InvalidOperation("BatchSimulation: agent function conditions ('" + f->name + "') cannot be batched");
            if (auto mi = f->message_input.lock())
                if (!written.count(mi->name))
                    throw exception::InvalidOperation("BatchSimulation: message list '" + mi->name +
                                                      "' is read before it is written in the step");
            if (auto mo = f->message_output.lock()) {
                if (++writers[mo->name] > 1)
                    throw exception::InvalidOperation("BatchSimulation: message list '" + mo->name + "' has several writers");
                written.insert(mo->name);
            }
        }
    }
    for (auto& mm : m.messages)
        if (mm.second->kind == MsgKind::Bucket || is_array_kind(mm.second->kind))
            throw exception::InvalidOperation("BatchSimulation: bucket and array message lists ('" + mm.first +
                                              "') cannot be batched");
    for (auto& mm : m.messages)
        if (mm.second->persistent)
            throw exception::InvalidOperation("BatchSimulation: persistent message lists cannot be batched");
    for (auto& a : m.agents)
        if (a.second->sortPeriod > 1)
            throw exception::InvalidOperation("BatchSimulation: agent '" + a.first +
                                              "' sorts every few steps; batched simulations sort every step or never");
}

inline void BatchSimulation::init() {
    if (initialised_) return;
    detail::fgx_check(fgx_ctx_create(device_, &ctx_), "fgx_ctx_create");
    detail::fgx_check(fgx_stream_create(&stream_), "stream create");
    detail::fgx_check(fgx_event_create(&build_fork_event_), "event create");
    detail::fgx_check(fgx_event_create(&build_join_event_), "event create");
    int cus = 0, xcds = 0;
    size_t hbm = 0;
    if (fgx_device_info(ctx_, &cus, &xcds, &hbm) == FGX_OK && cus > 0) cu_count_ = cus;
    alloc(&d_pages_, 256);
    detail::fgx_check(fgx_fill_u32((uint32_t*)d_pages_, 0, 64, stream_), "zero page");
    for (auto& kv : m_->agents) {
        AgentB a;
        a.d = kv.second;
        for (auto& v : kv.second->variables) {
            detail::VarBuf b;
            b.name = v.first;
            b.bytes = (uint32_t)v.second.bytes();
            b.def = v.second.default_value;
            a.list.vars.push_back(b);
        }
        a.list.id_slot = a.list.var_index("_id");
        agents_.emplace(kv.first, std::move(a));
    }
    for (auto& kv : m_->messages) {
        MsgB mb;
        mb.d = kv.second;
        for (auto& v : kv.second->variables) {
            detail::VarBuf b;
            b.name = v.first;
            b.bytes = (uint32_t)v.second.bytes();
            b.def = v.second.default_value;
            mb.list.vars.push_back(b);
        }
        mb.spatial = kv.second->kind == MsgKind::Spatial2D || kv.second->kind == MsgKind::Spatial3D;
        mb.dims = kv.second->kind == MsgKind::Spatial3D ? 3 : 2;
        msgs_.emplace(kv.first, std::move(mb));
    }
    for (auto& layer : m_->layers)
        for (auto& f : layer->agent_functions) {
            if (fn_index_.count(f.get())) continue;
            FnB r;
            r.d = f;
            r.agent = &agents_.at(f->parent->name);
            if (auto mi = f->message_input.lock()) r.in = &msgs_.at(mi->name);
            if (auto mo = f->message_output.lock()) {
                r.out = &msgs_.at(mo->name);
                r.out->src_agent = f->parent->name;
            }
            if (auto oa = f->agent_output.lock()) r.out_agent = &agents_.at(oa->name);
            if (f->has_agent_death || r.out_agent) dynamic_ = true;
            for (auto& kv : r.out_agent ? r.out_agent->d->variables : std::map<std::string, Variable>()) {
                detail::VarBuf b;
                b.name = kv.first;
                b.bytes = (uint32_t)kv.second.bytes();
                b.def = kv.second.default_value;
                r.newbuf.push_back(b);
            }
            if (r.in && r.in->spatial && f->parent->sortPeriod != 0) {
                const auto& vars = f->parent->variables;
                auto isf = [&](const char* n) {
                    auto it = vars.find(n);
                    return it != vars.end() && it->second.type == std::type_index(typeid(float)) && it->second.elements == 1;
                };
                if (isf("x") && isf("y") && (r.in->dims == 2 || isf("z"))) r.agent->sort_msg = r.in->d->name;
            }
            fn_index_[f.get()] = fns_.size();
            fns_.push_back(std::move(r));
        }
    env_stride_ = std::max<size_t>((m_->env.buffer.size() + 63) / 64 * 64, 64);
    {
        void* p = nullptr;
        alloc(&p, 4);
        d_word_ = (uint32_t*)p;
        for (auto& kv : agents_) {
            alloc(&p, 4);
            kv.second.list.d_total = (uint32_t*)p;
        }
        for (auto& kv : msgs_) {
            alloc(&p, 4);
            kv.second.list.d_total = (uint32_t*)p;
        }
    }
    initialised_ = true;
}

inline void BatchSimulation::release() {
    if (!initialised_) return;
    fgx_stream_synchronize(stream_);
    clear_graphs();
    for (auto& kv : agents_) {
        for (auto& v : kv.second.list.vars) {
            free_dev(v.cur);
            free_dev(v.alt);
        }
        free_typed(kv.second.d_keys);
        free_dev(kv.second.scratch);
        free_typed(kv.second.list.d_off);
        free_typed(kv.second.list.d_cnt);
        free_typed(kv.second.list.d_total);
        free_typed(kv.second.d_next_id);
        free_typed(kv.second.d_born);
        free_typed(kv.second.d_work);
    }
    for (auto& kv : msgs_) {
        for (auto& v : kv.second.list.vars) {
            free_dev(v.cur);
            free_dev(v.alt);
        }
        free_typed(kv.second.pbm);
        free_typed(kv.second.d_keys);
        free_dev(kv.second.scratch);
        free_dev(kv.second.packed);
        kv.second.packed_cap = 0;
        free_typed(kv.second.list.d_off);
        free_typed(kv.second.list.d_cnt);
        free_typed(kv.second.list.d_total);
    }
    for (auto& f : fns_) {
        free_typed(f.d_segs);
        free_typed(f.msg_flags);
        free_typed(f.death_flags);
        free_typed(f.birth_flags);
        for (auto& v : f.newbuf) {
            free_dev(v.cur);
            free_dev(v.alt);
        }
        free_dev(f.death_scratch);
        free_dev(f.birth_scratch);
    }
    free_typed(d_word_);
    free_typed(d_env_);
    free_typed(d_steps_);
    free_typed(d_errors_);
    free_dev(d_pages_);
    free_typed(d_ranges_);
    fgx_event_destroy(build_fork_event_);
    fgx_event_destroy(build_join_event_);
    build_fork_event_ = build_join_event_ = nullptr;
    if (build_stream_) fgx_stream_destroy(build_stream_);
    build_stream_ = nullptr;
    fgx_stream_destroy(stream_);
    fgx_ctx_destroy(ctx_);
    initialised_ = false;
}

inline int BatchSimulation::add(const Plan& p) {
    init();
    Sim s;
    s.handle = next_handle_++;
    s.seed = p.seed;
    s.steps_target = p.steps;
    s.env = p.env.empty() ? m_->env.buffer : p.env;
    if (s.env.size() != m_->env.buffer.size())
        throw exception::InvalidArgument("BatchSimulation::add: environment buffer size does not match the model");
    for (auto& kv : msgs_) {
        if (!kv.second.spatial) continue;
        const auto& md = kv.second.d;
        SpatialBounds b;
        auto it = p.messages.find(kv.first);
        if (it != p.messages.end()) {
            b = it->second;
        } else {
            for (int a = 0; a < 3; ++a) {
                b.min[a] = md->min[a];
                b.max[a] = md->max[a];
            }
            b.radius = md->radius;
        }
        fgx_spatial_meta m;
        detail::fgx_check(fgx_spatial_meta_init_mult(kv.second.dims, b.min, b.max, b.radius, bin_mult_, &m),
                          "spatial message metadata");
        s.meta[kv.first] = m;
    }
    for (auto& kv : agents_) {
        auto it = p.populations.find(kv.first);
        std::shared_ptr<AgentVector> pop;
        if (it != p.populations.end() && it->second) {
            if (it->second->getAgentName() != kv.first)
                throw exception::InvalidPopulationData("BatchSimulation::add: population of the wrong agent type");
            pop = std::make_shared<AgentVector>(*it->second);
        } else {
            pop = std::make_shared<AgentVector>(AgentDescription(kv.second.d), 0);
        }
        // IDs as a simulation run alone assigns them (CUDAFatAgent::assignIDs,
        // src/flamegpu/simulation/detail/CUDAFatAgent.cu:347-391): unset IDs follow the largest set one.
        auto& ids = pop->raw("_id");
        const size_t n = pop->size();
        id_t mx = 0;
        for (size_t i = 0; i < n; ++i) {
            id_t v;
            std::memcpy(&v, ids.data() + i * 4, 4);
            mx = std::max(mx, v);
        }
        const id_t next = std::max<id_t>(1u, mx + 1);
        for (size_t i = 0; i < n; ++i) {
            id_t v;
            std::memcpy(&v, ids.data() + i * 4, 4);
            if (v == ID_NOT_SET) {
                v = next + (id_t)i;
                std::memcpy(ids.data() + i * 4, &v, 4);
            }
        }
        // the next ID after the population (CUDASimulation::setPopulationData reserves n from `next`)
        s.next_id[kv.first] = next + (id_t)n;
        s.n[kv.first] = (uint32_t)n;
        s.pop[kv.first] = pop;
    }
    pending_.push_back(s.handle);
    sims_.emplace(s.handle, std::move(s));
    dirty_ = true;
    return pending_.back();
}

inline size_t BatchSimulation::estimateBytes(const Plan& p) const {
    size_t bytes = 0;
    std::map<std::string, uint64_t> n;
    for (auto& kv : m_->agents) {
        auto it = p.populations.find(kv.first);
        n[kv.first] = (it != p.populations.end() && it->second) ? it->second->size() : 0;
        size_t ab = 0;
        for (auto& v : kv.second->variables) ab += v.second.bytes();
        bytes += 2 * n[kv.first] * ab + 16 * n[kv.first];   // double buffer + sort scratch
    }
    for (auto& kv : m_->messages) {
        uint64_t cnt = 0;
        for (auto& layer : m_->layers)
            for (auto& f : layer->agent_functions)
                if (auto mo = f->message_output.lock())
                    if (mo->name == kv.first) cnt += n[f->parent->name];
        size_t mb = 0;
        for (auto& v : kv.second->variables) mb += v.second.bytes();
        bytes += 2 * cnt * mb + 16 * cnt;
        if (kv.second->kind == MsgKind::Spatial2D || kv.second->kind == MsgKind::Spatial3D) {
            SpatialBounds b;
            auto it = p.messages.find(kv.first);
            if (it != p.messages.end()) b = it->second;
            else {
                for (int a = 0; a < 3; ++a) {
                    b.min[a] = kv.second->min[a];
                    b.max[a] = kv.second->max[a];
                }
                b.radius = kv.second->radius;
            }
            uint64_t bins = 1;
            const int dims = kv.second->kind == MsgKind::Spatial3D ? 3 : 2;
            for (int a = 0; a < dims; ++a) bins *= (uint64_t)std::ceil((b.max[a] - b.min[a]) / b.radius);
            bytes += 3 * 4 * (bins + 1) + 3 * 8 * bins;   // bin table, histogram, sort keys (2x bins upper bound)
        }
    }
    return bytes + env_stride_ + 64;
}

inline uint64_t BatchSimulation::residentAgents() const {
    uint64_t t = 0;
    for (int h : order_) {
        auto& s = sims_.at(h);
        if (s.steps_done >= s.steps_target || removing_.count(h)) continue;
        for (auto& kv : s.n) t += kv.second;
    }
    return t;
}

inline std::vector<int> BatchSimulation::finished() const {
    std::vector<int> r;
    for (int h : order_) {
        auto& s = sims_.at(h);
        if (s.steps_done >= s.steps_target && !removing_.count(h)) r.push_back(h);
    }
    return r;
}

// A removed simulation leaves at the next repack: it keeps its place (and is still
// computed, harmlessly) until a simulation joins or the removed ones hold more than
// a quarter of the batch's agents, so a stream of completions costs no repacks.
inline void BatchSimulation::remove(int h) {
    Sim& s = sim(h);
    if (!s.resident) {
        pending_.erase(std::remove(pending_.begin(), pending_.end(), h), pending_.end());
        sims_.erase(h);
        return;
    }
    if (removing_.count(h)) return;
    removing_.insert(h);
    uint64_t dead = 0, all = 0;
    for (int k : order_) {
        uint64_t n = 0;
        for (auto& kv : sims_.at(k).n) n += kv.second;
        all += n;
        if (removing_.count(k)) dead += n;
    }
    if (dead * 4 > all || removing_.size() == order_.size()) dirty_ = true;
}

inline void BatchSimulation::setTuning(const Tuning& t) {
    if (t.block == 0 || t.block > 1024 || t.block % 64 != 0)
        throw exception::InvalidArgument("BatchSimulation: block size must be a multiple of 64 up to 1024");
    if (t.agents_per_thread == 0 || t.agents_per_thread > 16)
        throw exception::InvalidArgument("BatchSimulation: agents per thread must be 1..16");
    tuning_ = t;
    dirty_ = true;   // tile tables depend on the launch shape
}

// Repacks list l for the new batch order: kept simulations (old slot per new slot,
// or UINT32_MAX for a joining one) move from the current into the alternate buffer
// (both at least the new total), then the buffers swap.  Kept simulations come
// first in the new order, so their destination ranges are contiguous.
inline void BatchSimulation::repack(List& l, const std::vector<uint32_t>& old_slot, const std::vector<uint32_t>& new_cnt) {
    const uint32_t S = (uint32_t)new_cnt.size();
    std::vector<uint32_t> off(S);
    uint64_t total = 0;
    for (uint32_t s = 0; s < S; ++s) {
        off[s] = (uint32_t)total;
        total += new_cnt[s];
    }
    if (total >= (1ull << 31)) throw exception::InvalidOperation("BatchSimulation: batch exceeds the 32-bit index range");
    std::vector<uint32_t> ranges;
    for (uint32_t s = 0; s < S; ++s) {
        if (old_slot[s] == 0xFFFFFFFFu || !new_cnt[s]) continue;
        ranges.push_back(l.off[old_slot[s]]);
        ranges.push_back(off[s]);
        ranges.push_back(new_cnt[s]);
    }
    if (total > l.cap) throw exception::InvalidOperation("BatchSimulation: list capacity below the packed batch");
    // gather kept ranges from cur into alt
    if (!ranges.empty()) {
        const uint32_t nr = (uint32_t)ranges.size() / 3;
        if (ranges.size() > ranges_cap_) {
            free_typed(d_ranges_);
            ranges_cap_ = (uint32_t)ranges.size() * 2;
            void* p = nullptr;
            alloc(&p, (size_t)ranges_cap_ * 4);
            d_ranges_ = (uint32_t*)p;
        }
        detail::fgx_check(fgx_memcpy_h2d(d_ranges_, ranges.data(), ranges.size() * 4, stream_), "range table upload");
        std::vector<const void*> src;
        std::vector<void*> dst;
        std::vector<uint32_t> by;
        for (auto& v : l.vars) {
            src.push_back(v.cur);
            dst.push_back(v.alt);
            by.push_back(v.bytes);
        }
        uint32_t kept_total = 0;
        for (uint32_t r = 0; r < nr; ++r) kept_total += ranges[3 * r + 2];
        detail::fgx_check(fgx_copy_ranges((uint32_t)l.vars.size(), src.data(), dst.data(), by.data(), d_ranges_, nr,
                                          kept_total, stream_),
                          "batch repack");
        synchronize();
    }
    for (auto& v : l.vars) std::swap(v.cur, v.alt);
    l.off = off;
    l.cnt = new_cnt;
    l.total = (uint32_t)total;
}

inline void BatchSimulation::rebuild() {
    init();
    synchronize();
    clear_graphs();
    if (dynamic_) read_counts();
    // new batch order: kept simulations in their order, then joining ones
    std::vector<int> neworder;
    std::vector<uint32_t> old_slot;
    for (uint32_t i = 0; i < order_.size(); ++i) {
        if (removing_.count(order_[i])) continue;
        neworder.push_back(order_[i]);
        old_slot.push_back(i);
    }
    for (int h : pending_) {
        neworder.push_back(h);
        old_slot.push_back(0xFFFFFFFFu);
    }
    // error words of kept simulations survive the repack
    std::vector<uint32_t> old_err(order_.size() * 4, 0);
    if (!order_.empty() && d_errors_)
        detail::fgx_check(fgx_memcpy_d2h(old_err.data(), d_errors_, old_err.size() * 4, stream_), "error readback");
    synchronize();
    for (int h : removing_) sims_.erase(h);
    removing_.clear();
    pending_.clear();
    const uint32_t S = (uint32_t)neworder.size();
    // agent lists
    for (auto& kv : agents_) {
        AgentB& a = kv.second;
        std::vector<uint32_t> cnt(S);
        for (uint32_t s = 0; s < S; ++s) cnt[s] = sims_.at(neworder[s]).n.at(kv.first);
        // allocate both buffers up front when growing (gather source = old cur)
        uint64_t total = 0;
        for (auto c : cnt) total += c;
        if (total > a.list.cap) {
            const uint32_t cap = std::max<uint32_t>((uint32_t)std::max<uint64_t>(total, 1), a.list.cap + a.list.cap / 2);
            for (auto& v : a.list.vars) {
                void* na = nullptr;
                alloc(&na, (size_t)cap * v.bytes);
                free_dev(v.alt);
                v.alt = na;
            }
            // after the gather (into alt) the old cur is replaced by a new spare
            std::vector<void*> old_cur;
            for (auto& v : a.list.vars) old_cur.push_back(v.cur);
            const uint32_t old_cap = a.list.cap;
            a.list.cap = cap;   // repack must not grow again
            repack(a.list, old_slot, cnt);
            for (size_t k = 0; k < a.list.vars.size(); ++k) {
                auto& v = a.list.vars[k];
                free_dev(v.alt);   // == old cur after the swap
                void* nb = nullptr;
                alloc(&nb, (size_t)cap * v.bytes);
                v.alt = nb;
            }
            (void)old_cur;
            (void)old_cap;
        } else {
            repack(a.list, old_slot, cnt);
        }
        // joining simulations: upload their populations
        for (uint32_t s = 0; s < S; ++s) {
            Sim& sm = sims_.at(neworder[s]);
            if (old_slot[s] != 0xFFFFFFFFu) continue;
            auto& pop = *sm.pop.at(kv.first);
            if (!cnt[s]) continue;
            for (auto& v : a.list.vars)
                detail::fgx_check(fgx_memcpy_h2d((char*)v.cur + (size_t)a.list.off[s] * v.bytes, pop.raw(v.name).data(),
                                                 (size_t)cnt[s] * v.bytes, stream_),
                                  "population upload");
        }
    }
    for (uint32_t s = 0; s < S; ++s) {
        Sim& sm = sims_.at(neworder[s]);
        sm.pop.clear();
        sm.resident = true;
    }
    // message lists: one element per agent of the writing agent
    for (auto& kv : msgs_) {
        MsgB& mb = kv.second;
        std::vector<uint32_t> cnt(S, 0);
        if (!mb.src_agent.empty())
            for (uint32_t s = 0; s < S; ++s) cnt[s] = sims_.at(neworder[s]).n.at(mb.src_agent);
        uint64_t total = 0;
        for (auto c : cnt) total += c;
        const uint32_t need = (uint32_t)std::max<uint64_t>(total, 1);
        if (need > mb.list.cap) {
            const uint32_t cap = std::max<uint32_t>(need, mb.list.cap + mb.list.cap / 2);
            for (auto& v : mb.list.vars) {
                free_dev(v.cur);
                free_dev(v.alt);
                alloc(&v.cur, (size_t)cap * v.bytes);
                alloc(&v.alt, (size_t)cap * v.bytes);
            }
            mb.list.cap = cap;
        }
        size_records(mb);
        mb.list.off.assign(S, 0);
        uint32_t o = 0;
        for (uint32_t s = 0; s < S; ++s) {
            mb.list.off[s] = o;
            o += cnt[s];
        }
        mb.list.cnt = cnt;
        mb.list.total = (uint32_t)total;
        if (mb.spatial) {
            mb.keys.assign(S, fgx_seg_key{});
            for (uint32_t s = 0; s < S; ++s) {
                const fgx_spatial_meta& m = sims_.at(neworder[s]).meta.at(kv.first);
                detail::fgx_check(fgx_seg_key_init(FGX_SEG_MESSAGE_BINS, mb.dims, m.min, m.max, m.cell, &mb.keys[s]),
                                  "message segment");
                mb.keys[s].n = cnt[s];
            }
            uint32_t nt = 0;
            detail::fgx_check(fgx_segments_pack(mb.keys.data(), S, &nt, &mb.keys_total, &mb.tiles), "segment packing");
            if (S > mb.keys_cap) {
                free_typed(mb.d_keys);
                mb.keys_cap = S * 2;
                void* p = nullptr;
                alloc(&p, (size_t)mb.keys_cap * sizeof(fgx_seg_key));
                mb.d_keys = (fgx_seg_key*)p;
            }
            if (S)
                detail::fgx_check(fgx_memcpy_h2d(mb.d_keys, mb.keys.data(), (size_t)S * sizeof(fgx_seg_key), stream_),
                                  "segment table upload");
            if (mb.keys_total + 1 > mb.pbm_cap) {
                free_typed(mb.pbm);
                mb.pbm_cap = (mb.keys_total + 1) + (mb.keys_total + 1) / 2;
                void* p = nullptr;
                alloc(&p, (size_t)mb.pbm_cap * 4);
                mb.pbm = (uint32_t*)p;
            }
            size_t sb = 0;
            detail::fgx_check(fgx_scratch_bytes(FGX_OP_SPATIAL_BUILD, std::max<uint32_t>(mb.list.cap, 1), mb.keys_total, &sb),
                              "scratch sizing");
            if (sb > mb.scratch_bytes) {
                free_dev(mb.scratch);
                sb += sb / 2;
                alloc(&mb.scratch, sb);
                detail::fgx_check(fgx_scratch_clear(mb.scratch, sb, stream_), "scratch clear");
                mb.scratch_bytes = sb;
            }
        }
        mb.dirty = true;
    }
    // agent sort segments (geometry of the message list that drives the sort)
    for (auto& kv : agents_) {
        AgentB& a = kv.second;
        if (a.sort_msg.empty()) continue;
        const MsgB& mb = msgs_.at(a.sort_msg);
        std::vector<fgx_seg_key> keys(S);
        for (uint32_t s = 0; s < S; ++s) {
            const fgx_spatial_meta& m = sims_.at(neworder[s]).meta.at(a.sort_msg);
            detail::fgx_check(fgx_seg_key_init(FGX_SEG_AGENT_SORT, mb.dims, m.min, m.max, m.radius, &keys[s]),
                              "agent sort segment");
            keys[s].n = a.list.cnt[s];
        }
        uint32_t nt = 0;
        detail::fgx_check(fgx_segments_pack(keys.data(), S, &nt, &a.keys_total, &a.tiles), "segment packing");
        if (S > a.keys_cap) {
            free_typed(a.d_keys);
            a.keys_cap = S * 2;
            void* p = nullptr;
            alloc(&p, (size_t)a.keys_cap * sizeof(fgx_seg_key));
            a.d_keys = (fgx_seg_key*)p;
        }
        if (S) detail::fgx_check(fgx_memcpy_h2d(a.d_keys, keys.data(), (size_t)S * sizeof(fgx_seg_key), stream_), "sort table upload");
        size_t sb = 0;
        detail::fgx_check(fgx_scratch_bytes(FGX_OP_AGENT_SORT, std::max<uint32_t>(a.list.cap, 1), a.keys_total, &sb),
                          "scratch sizing");
        if (sb > a.scratch_bytes) {
            free_dev(a.scratch);
            sb += sb / 2;
            alloc(&a.scratch, sb);
            detail::fgx_check(fgx_scratch_clear(a.scratch, sb, stream_), "scratch clear");
            a.scratch_bytes = sb;
        }
    }
    // per-simulation environment, step counters and error words
    if (S > sim_cap_) {
        free_typed(d_env_);
        free_typed(d_steps_);
        free_typed(d_errors_);
        sim_cap_ = S * 2;
        void* p = nullptr;
        alloc(&p, (size_t)sim_cap_ * env_stride_);
        d_env_ = (char*)p;
        alloc(&p, (size_t)sim_cap_ * 4);
        d_steps_ = (uint32_t*)p;
        alloc(&p, (size_t)sim_cap_ * 16);
        d_errors_ = (uint32_t*)p;
        auto per_sim = [&](uint32_t*& q, size_t words) {
            free_typed(q);
            void* r = nullptr;
            alloc(&r, words * 4);
            q = (uint32_t*)r;
        };
        for (auto& kv : agents_) {
            per_sim(kv.second.list.d_off, sim_cap_);
            per_sim(kv.second.list.d_cnt, sim_cap_);
            per_sim(kv.second.d_next_id, sim_cap_);
            per_sim(kv.second.d_born, sim_cap_);
            per_sim(kv.second.d_work, 2 * (size_t)sim_cap_ + 2);
        }
        for (auto& kv : msgs_) {
            per_sim(kv.second.list.d_off, sim_cap_);
            per_sim(kv.second.list.d_cnt, sim_cap_);
        }
    }
    // device copies of every list's layout and of the agents' next IDs
    for (auto& kv : agents_) {
        upload_layout(kv.second.list, S);
        std::vector<uint32_t> ids(S);
        for (uint32_t k = 0; k < S; ++k) ids[k] = sims_.at(neworder[k]).next_id.at(kv.first);
        if (S) detail::fgx_check(fgx_memcpy_h2d(kv.second.d_next_id, ids.data(), (size_t)S * 4, stream_), "next IDs");
    }
    for (auto& kv : msgs_) upload_layout(kv.second.list, S);
    if (S) {
        std::vector<char> env((size_t)S * env_stride_, 0);
        std::vector<uint32_t> steps(S);
        for (uint32_t s = 0; s < S; ++s) {
            const Sim& sm = sims_.at(neworder[s]);
            std::memcpy(env.data() + (size_t)s * env_stride_, sm.env.data(), sm.env.size());
            steps[s] = sm.steps_done;
        }
        detail::fgx_check(fgx_memcpy_h2d(d_env_, env.data(), env.size(), stream_), "environment upload");
        detail::fgx_check(fgx_memcpy_h2d(d_steps_, steps.data(), (size_t)S * 4, stream_), "step counters upload");
        std::vector<uint32_t> err((size_t)S * 4, 0);
        for (uint32_t s = 0; s < S; ++s)
            if (old_slot[s] != 0xFFFFFFFFu)
                for (int k = 0; k < 4; ++k) err[4 * s + k] = old_err[4 * old_slot[s] + k];
        detail::fgx_check(fgx_memcpy_h2d(d_errors_, err.data(), err.size() * 4, stream_), "error words");
    }
    // per-function segment tables
    const uint32_t tile = tuning_.block * tuning_.agents_per_thread;
    for (auto& f : fns_) {
        std::vector<detail::SegDev> segs(S + 1);
        uint32_t t = 0;
        for (uint32_t s = 0; s < S; ++s) {
            const Sim& sm = sims_.at(neworder[s]);
            detail::SegDev& d = segs[s];
            std::memset(&d, 0, sizeof(d));
            d.agent_off = f.agent->list.off[s];
            d.n = f.agent->list.cnt[s];
            d.tile_off = t;
            t += (d.n + tile - 1) / tile;
            d.seed = sm.seed;
            d.d_step = d_steps_ + s;
            d.env = d_env_ + (size_t)s * env_stride_;
            d.d_error = d_errors_ + 4 * s;
            if (f.in) {
                d.msg_off = f.in->list.off[s];
                d.msg_n = f.in->list.cnt[s];
                if (f.in->spatial) {
                    const fgx_spatial_meta& m = sm.meta.at(f.in->d->name);
                    for (int k = 0; k < 3; ++k) {
                        d.meta.min[k] = m.min[k];
                        d.meta.max[k] = m.max[k];
                        d.meta.width[k] = m.width[k];
                        d.meta.grid[k] = m.grid[k] ? m.grid[k] : 1u;
                    }
                    d.meta.radius = m.radius;
                    d.meta.cell = m.cell;
                    d.meta.span = m.span;
                    d.meta.wrap_ok = m.wrap_ok;
                    d.meta.pbm = f.in->pbm + f.in->keys[s].key_off;
                }
            }
        }
        f.tiles = t;
        {
            // sentinel record: the batch's tile count (the kernel's loop bound)
            detail::SegDev& d = segs[S];
            std::memset(&d, 0, sizeof(d));
            d.agent_off = f.agent->list.total;
            d.tile_off = t;
        }
        if (f.out && f.agent->list.cap > f.flags_cap) {
            free_typed(f.msg_flags);
            f.flags_cap = f.agent->list.cap;
            void* p = nullptr;
            alloc(&p, (size_t)f.flags_cap * 4);
            f.msg_flags = (uint32_t*)p;
        }
        if (S + 1 > f.segs_cap) {
            free_typed(f.d_segs);
            f.segs_cap = (S + 1) * 2;
            void* p = nullptr;
            alloc(&p, (size_t)f.segs_cap * sizeof(detail::SegDev));
            f.d_segs = (detail::SegDev*)p;
        }
        detail::fgx_check(fgx_memcpy_h2d(f.d_segs, segs.data(), (size_t)(S + 1) * sizeof(detail::SegDev), stream_),
                          "function segments");
        if (dynamic_) size_pop_buffers(f);
    }
    order_ = neworder;
    // resident bytes: lists (both buffers), bin tables, scratch, per-simulation tables
    size_t rb = 0;
    for (auto& kv : agents_) {
        for (auto& v : kv.second.list.vars) rb += 2ull * kv.second.list.cap * v.bytes;
        rb += kv.second.scratch_bytes;
    }
    for (auto& kv : msgs_) {
        for (auto& v : kv.second.list.vars) rb += 2ull * kv.second.list.cap * v.bytes;
        rb += kv.second.scratch_bytes + (size_t)kv.second.pbm_cap * 4;
    }
    resident_bytes_ = rb + (size_t)sim_cap_ * (env_stride_ + 20);
    synchronize();
    dirty_ = false;
    stable_steps_ = 0;
}

inline void BatchSimulation::launch(FnB& f) {
    detail::FnArgs a;
    std::memset(&a, 0, sizeof(a));
    List& l = f.agent->list;
    a.n_cap = l.total;
    a.unit_tables = unit_tables_ ? 1 : 0;
    a.zero_page = (const char*)d_pages_;
    a.sink_page = (char*)d_pages_ + 128;
    a.stream_id = f.d->stream_id;
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
    if (f.in) {
        fill_table(a.msg_in, &a.msg_in_mul, f.in->list.vars, false);
        a.in_meta.packed = f.in->packed;   // covers the whole batch, indexed like the message arrays
    }
    if (f.out) {
        fill_table(a.msg_out, &a.msg_out_mul, f.out->list.vars, true);
        if (f.flags_cap < (dynamic_ ? l.cap : l.total)) throw exception::InvalidOperation("BatchSimulation: message flags below the list size");
        a.msg_out_flags = f.msg_flags;   // written 1 per agent (mandatory output)
        a.msg_out_optional = 0;
    }
    if (f.d->has_agent_death) a.death_flags = f.death_flags;
    if (f.out_agent) {
        // new agents start from the variable defaults (broadcastInit), one slot per parent
        fill_table(a.new_agent, &a.new_mul, f.newbuf, false);
        a.birth_flags = f.birth_flags;
        std::vector<void*> ptrs;
        std::vector<uint32_t> bytes, offs;
        std::vector<char> defs;
        for (auto& v : f.newbuf) {
            ptrs.push_back(v.cur);
            bytes.push_back(v.bytes);
            offs.push_back((uint32_t)defs.size());
            defs.insert(defs.end(), v.def.begin(), v.def.end());
        }
        detail::fgx_check(fgx_fill_vars((uint32_t)ptrs.size(), ptrs.data(), bytes.data(), defs.data(), offs.data(), l.cap,
                                        stream_),
                          "new-agent defaults");
    }
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
    const uint32_t S = (uint32_t)order_.size();
    // changing populations: the tile count is on the device (sentinel record); the grid
    // covers the list's capacity
    const uint32_t tiles = dynamic_ ? tile_bound(l.cap, tuning_.block * tuning_.agents_per_thread) : f.tiles;
    if (!S || !tiles) return;
    uint32_t grid = tiles;
    if (tuning_.blocks_per_cu) grid = std::min<uint32_t>(grid, tuning_.blocks_per_cu * (uint32_t)cu_count_);
    f.d->launch_seg(a, f.d_segs, S, tuning_.agents_per_thread, grid, tuning_.block, stream_);
    if (hipGetLastError() != hipSuccess) throw exception::CUDAError("batched agent function launch failed: " + f.d->name);
}

inline void BatchSimulation::enqueue_step() {
    for (auto& layer : m_->layers) {
        // agent sorts and message-list builds for the layer's spatial inputs.  Sorts move agent
        // lists and builds message lists: when a layer has both, the builds run on a stream of
        // their own beside the sorts.
        std::vector<AgentB*> sorts;
        std::vector<MsgB*> builds;
        for (auto& fd : layer->agent_functions) {
            FnB& f = fns_[fn_index_.at(fd.get())];
            if (!f.in || !f.in->spatial) continue;
            AgentB& a = *f.agent;
            if (!a.sort_msg.empty() && a.d->sortPeriod == 1 && (dynamic_ ? a.list.cap : a.list.total) &&
                std::find(sorts.begin(), sorts.end(), &a) == sorts.end())
                sorts.push_back(&a);
            if (f.in->dirty && std::find(builds.begin(), builds.end(), f.in) == builds.end()) builds.push_back(f.in);
        }
        for (AgentB* ap : sorts) {
            AgentB& a = *ap;
            List& l = a.list;
            const int ix = l.var_index("x"), iy = l.var_index("y"), iz = l.var_index("z");
            std::vector<const void*> src;
            std::vector<void*> dst;
            std::vector<uint32_t> by;
            for (auto& v : l.vars) {
                src.push_back(v.cur);
                dst.push_back(v.alt);
                by.push_back(v.bytes);
            }
            const MsgB& sm = msgs_.at(a.sort_msg);
            detail::fgx_check(fgx_agent_sort_seg(ctx_, a.d_keys, (uint32_t)order_.size(),
                                                 dynamic_ ? tile_bound(l.cap, 2048) : a.tiles, dynamic_ ? l.cap : l.total,
                                                 dynamic_ ? l.d_total : nullptr, a.keys_total,
                                                 (const float*)l.vars[ix].cur, (const float*)l.vars[iy].cur,
                                                 sm.dims == 3 ? (const float*)l.vars[iz].cur : nullptr,
                                                 (uint32_t)l.vars.size(), src.data(), dst.data(), by.data(), a.scratch,
                                                 a.scratch_bytes, stream_),
                              "batched agent sort");
            for (auto& v : l.vars) std::swap(v.cur, v.alt);
        }
        const bool overlap = !sorts.empty() && !builds.empty();
        void* bst = stream_;
        if (overlap) {
            if (!build_stream_) detail::fgx_check(fgx_stream_create(&build_stream_), "stream create");
            detail::fgx_check(fgx_event_record(build_fork_event_, stream_), "fork");
            detail::fgx_check(fgx_stream_wait_event(build_stream_, build_fork_event_), "fork wait");
            bst = build_stream_;
        }
        for (MsgB* mp : builds) {
            MsgB& m = *mp;
            List& l = m.list;
            const int ix = l.var_index("x"), iy = l.var_index("y"), iz = l.var_index("z");
            std::vector<const void*> src;
            std::vector<void*> dst;
            std::vector<uint32_t> by;
            for (auto& v : l.vars) {
                src.push_back(v.cur);
                dst.push_back(v.alt);
                by.push_back(v.bytes);
            }
            detail::fgx_check(fgx_spatial_build_seg(ctx_, m.d_keys, (uint32_t)order_.size(),
                                                    dynamic_ ? tile_bound(l.cap, 2048) : m.tiles,
                                                    dynamic_ ? l.cap : l.total, dynamic_ ? l.d_total : nullptr,
                                                    m.keys_total,
                                                    (const float*)l.vars[ix].cur, (const float*)l.vars[iy].cur,
                                                    m.dims == 3 ? (const float*)l.vars[iz].cur : nullptr,
                                                    (uint32_t)l.vars.size(), src.data(), dst.data(), by.data(), m.pbm,
                                                    m.scratch, m.scratch_bytes, bst),
                              "batched message-list build");
            for (auto& v : l.vars) std::swap(v.cur, v.alt);
            {
                const int ii = l.var_index("id");
                const void* idp = ii >= 0 && l.vars[ii].bytes == 4 ? l.vars[ii].cur : nullptr;
                detail::fgx_check(fgx_pack_records(dynamic_ ? l.d_total : nullptr, dynamic_ ? l.cap : l.total,
                                                   l.vars[ix].cur, l.vars[iy].cur, m.dims == 3 ? l.vars[iz].cur : nullptr,
                                                   idp, m.packed, bst),
                                  "batched message records");
            }
            m.dirty = false;
        }
        if (overlap) {
            detail::fgx_check(fgx_event_record(build_join_event_, build_stream_), "join");
            detail::fgx_check(fgx_stream_wait_event(stream_, build_join_event_), "join wait");
        }
        for (auto& fd : layer->agent_functions) {
            FnB& f = fns_[fn_index_.at(fd.get())];
            launch(f);
            after_launch(f);
        }
    }
    for (auto& kv : msgs_) kv.second.dirty = true;
    detail::fgx_check(fgx_add_u32_n(d_steps_, (uint32_t)order_.size(), 1, stream_), "step counters");
}

inline std::vector<void*> BatchSimulation::buffer_state() const {
    std::vector<void*> st;
    for (auto& kv : agents_)
        for (auto& v : kv.second.list.vars) {
            st.push_back(v.cur);
            st.push_back(v.alt);
        }
    for (auto& kv : msgs_)
        for (auto& v : kv.second.list.vars) {
            st.push_back(v.cur);
            st.push_back(v.alt);
        }
    return st;
}

inline void BatchSimulation::set_buffer_state(const std::vector<void*>& st) {
    size_t k = 0;
    for (auto& kv : agents_)
        for (auto& v : kv.second.list.vars) {
            v.cur = st[k++];
            v.alt = st[k++];
        }
    for (auto& kv : msgs_)
        for (auto& v : kv.second.list.vars) {
            v.cur = st[k++];
            v.alt = st[k++];
        }
}

inline void BatchSimulation::step() {
    init();
    // simulations whose steps are done leave before the step (their results stay
    // readable until remove()); joiners and leavers trigger one repack.
    if (dirty_) rebuild();
    if (order_.empty()) return;
    bool all_active = true;
    for (int h : order_) {
        const Sim& s = sims_.at(h);
        if (s.steps_done >= s.steps_target && !removing_.count(h)) all_active = false;
    }
    if (!all_active)
        throw exception::InvalidOperation("BatchSimulation::step: finished simulations are still resident; remove() them first");
    if (dynamic_) ensure_step_capacity();
    // Graphs only once the batch has been stable for two steps: under churn the
    // capture would be thrown away at the next repack.
    if (graphs_on_ && stable_steps_ >= 2) {
        const auto start = buffer_state();
        GraphEntry* hit = nullptr;
        for (auto& g : graphs_)
            if (g.start == start) hit = &g;
        if (!hit && graphs_.size() < 4) {
            GraphEntry e;
            e.start = start;
            detail::fgx_check(fgx_graph_begin(stream_), "graph begin");
            enqueue_step();
            detail::fgx_check(fgx_graph_end(stream_, &e.g), "graph end");
            e.end = buffer_state();
            set_buffer_state(start);
            graphs_.push_back(std::move(e));
            hit = &graphs_.back();
        }
        if (hit) {
            detail::fgx_check(fgx_graph_launch(hit->g, stream_), "graph launch");
            set_buffer_state(hit->end);
        } else {
            enqueue_step();
        }
    } else {
        enqueue_step();
    }
    for (int h : order_) sims_.at(h).steps_done++;
    ++stable_steps_;
}

// ------------------------------------------------------------------ changing populations
inline void BatchSimulation::upload_layout(List& l, uint32_t S) {
    if (S) {
        detail::fgx_check(fgx_memcpy_h2d(l.d_cnt, l.cnt.data(), (size_t)S * 4, stream_), "list counts");
        detail::fgx_check(fgx_memcpy_h2d(l.d_off, l.off.data(), (size_t)S * 4, stream_), "list offsets");
    }
    detail::fgx_check(fgx_memcpy_h2d(l.d_total, &l.total, 4, stream_), "list total");
    synchronize();   // the host vectors may change before the copies run otherwise
}

// Per-simulation counts, offsets and next IDs as the device holds them (before a repack).
inline void BatchSimulation::read_counts() {
    const uint32_t S = (uint32_t)order_.size();
    if (!S) return;
    for (auto& kv : agents_) {
        List& l = kv.second.list;
        std::vector<uint32_t> cnt(S), off(S), ids(S);
        detail::fgx_check(fgx_memcpy_d2h(cnt.data(), l.d_cnt, (size_t)S * 4, stream_), "count readback");
        detail::fgx_check(fgx_memcpy_d2h(off.data(), l.d_off, (size_t)S * 4, stream_), "offset readback");
        detail::fgx_check(fgx_memcpy_d2h(ids.data(), kv.second.d_next_id, (size_t)S * 4, stream_), "next ID readback");
        synchronize();
        uint64_t t = 0;
        for (uint32_t k = 0; k < S; ++k) {
            Sim& sm = sims_.at(order_[k]);
            sm.n[kv.first] = cnt[k];
            sm.next_id[kv.first] = ids[k];
            t += cnt[k];
        }
        l.cnt = cnt;
        l.off = off;
        l.total = (uint32_t)t;
    }
}

inline std::vector<fgx_seg_target> BatchSimulation::agent_targets(AgentB& a) {
    std::vector<fgx_seg_target> t;
    const uint32_t tile = tuning_.block * tuning_.agents_per_thread;
    for (auto& f : fns_) {
        if (f.agent != &a || !f.d_segs) continue;
        t.push_back({f.d_segs, (uint32_t)sizeof(detail::SegDev), (int32_t)offsetof(detail::SegDev, agent_off),
                     (int32_t)offsetof(detail::SegDev, n), (int32_t)offsetof(detail::SegDev, tile_off), tile, 1});
    }
    if (!a.sort_msg.empty() && a.d_keys)
        t.push_back({a.d_keys, (uint32_t)sizeof(fgx_seg_key), (int32_t)offsetof(fgx_seg_key, elem_off),
                     (int32_t)offsetof(fgx_seg_key, n), (int32_t)offsetof(fgx_seg_key, tile_off), 2048u, 0});
    return t;
}

inline std::vector<fgx_seg_target> BatchSimulation::message_targets(MsgB& m) {
    std::vector<fgx_seg_target> t;
    if (m.spatial && m.d_keys)
        t.push_back({m.d_keys, (uint32_t)sizeof(fgx_seg_key), (int32_t)offsetof(fgx_seg_key, elem_off),
                     (int32_t)offsetof(fgx_seg_key, n), (int32_t)offsetof(fgx_seg_key, tile_off), 2048u, 0});
    for (auto& f : fns_) {
        if (f.in != &m || !f.d_segs) continue;
        t.push_back({f.d_segs, (uint32_t)sizeof(detail::SegDev), (int32_t)offsetof(detail::SegDev, msg_off),
                     (int32_t)offsetof(detail::SegDev, msg_n), -1, 0u, 0});
    }
    return t;
}

// New per-simulation counts (plus d_add) -> offsets, total and every table of the list.
inline void BatchSimulation::relayout(List& l, const std::vector<fgx_seg_target>& t, const uint32_t* d_add) {
    const uint32_t S = (uint32_t)order_.size();
    size_t k = 0;
    do {
        const uint32_t nt = (uint32_t)std::min<size_t>(t.size() - k, FGX_SEG_MAX_TARGETS);
        detail::fgx_check(fgx_seg_layout(S, l.d_cnt, k == 0 ? d_add : nullptr, l.d_off, l.d_total, nt ? t.data() + k : nullptr,
                                         nt, stream_),
                          "batch layout");
        k += nt;
    } while (k < t.size());
}

// After an agent function of a batch whose populations change (CUDASimulation's unmap
// order: message output, death, births): the message list takes the layout the agents had
// during the launch; births are counted and compacted against the parents' layout before
// the deaths change it; then each simulation's survivors and new agents are packed.
inline void BatchSimulation::after_launch(FnB& f) {
    const uint32_t S = (uint32_t)order_.size();
    List& A = f.agent->list;
    if (f.out) {
        for (auto& v : f.out->list.vars) std::swap(v.cur, v.alt);
        f.out->dirty = true;
        if (dynamic_ && S) {
            detail::fgx_check(fgx_memcpy_d2d(f.out->list.d_cnt, A.d_cnt, (size_t)S * 4, stream_), "message layout");
            relayout(f.out->list, message_targets(*f.out), nullptr);
        }
    }
    if (!dynamic_ || !S) return;
    auto ptrs = [](std::vector<detail::VarBuf>& vars, bool alt) {
        std::vector<void*> p;
        for (auto& v : vars) p.push_back(alt ? v.alt : v.cur);
        return p;
    };
    auto bytes = [](const std::vector<detail::VarBuf>& vars) {
        std::vector<uint32_t> b;
        for (auto& v : vars) b.push_back(v.bytes);
        return b;
    };
    if (f.out_agent) {
        AgentB& B = *f.out_agent;
        detail::fgx_check(fgx_seg_count(S, A.d_off, A.d_cnt, f.birth_flags, 1u, B.d_born, stream_), "birth counts");
        const auto src = ptrs(f.newbuf, false);
        const auto dst = ptrs(f.newbuf, true);
        const auto by = bytes(f.newbuf);
        std::vector<const void*> csrc(src.begin(), src.end());
        detail::fgx_check(fgx_compact(ctx_, f.birth_flags, 1, A.d_total, A.cap, (uint32_t)by.size(), csrc.data(), dst.data(),
                                      by.data(), d_word_, f.birth_scratch, f.birth_scratch_bytes, stream_),
                          "new agents");
    }
    if (f.d->has_agent_death) {
        detail::fgx_check(fgx_seg_count(S, A.d_off, A.d_cnt, f.death_flags, 1u, A.d_cnt, stream_), "survivor counts");
        const auto src = ptrs(A.vars, false);
        const auto dst = ptrs(A.vars, true);
        const auto by = bytes(A.vars);
        std::vector<const void*> csrc(src.begin(), src.end());
        detail::fgx_check(fgx_compact(ctx_, f.death_flags, 1, A.d_total, A.cap, (uint32_t)by.size(), csrc.data(), dst.data(),
                                      by.data(), A.d_total, f.death_scratch, f.death_scratch_bytes, stream_),
                          "death compaction");
        for (auto& v : A.vars) std::swap(v.cur, v.alt);
        relayout(A, agent_targets(*f.agent), nullptr);
    }
    if (f.out_agent) {
        AgentB& B = *f.out_agent;
        List& L = B.list;
        const auto cur = ptrs(L.vars, false);
        const auto alt = ptrs(L.vars, true);
        const auto born = ptrs(f.newbuf, true);
        const auto by = bytes(L.vars);
        std::vector<const void*> ccur(cur.begin(), cur.end()), cborn(born.begin(), born.end());
        detail::fgx_check(fgx_seg_append(S, L.d_off, L.d_cnt, B.d_born, L.cap, (uint32_t)by.size(), ccur.data(), cborn.data(),
                                         alt.data(), by.data(), L.id_slot, B.d_next_id, B.d_work, stream_),
                          "birth append");
        for (auto& v : L.vars) std::swap(v.cur, v.alt);
        relayout(L, agent_targets(B), B.d_born);
    }
}

inline void BatchSimulation::grow_agent_list(AgentB& a, uint32_t cap) {
    List& l = a.list;
    if (cap <= l.cap) return;
    synchronize();
    for (auto& v : l.vars) {
        void *nc = nullptr, *na = nullptr;
        alloc(&nc, (size_t)cap * v.bytes);
        alloc(&na, (size_t)cap * v.bytes);
        if (v.cur && l.cap) detail::fgx_check(fgx_memcpy_d2d(nc, v.cur, (size_t)l.cap * v.bytes, stream_), "list growth");
        synchronize();
        free_dev(v.cur);
        free_dev(v.alt);
        v.cur = nc;
        v.alt = na;
    }
    l.cap = cap;
}

inline void BatchSimulation::size_pop_buffers(FnB& f) {
    const uint32_t cap = std::max<uint32_t>(f.agent->list.cap, 1);
    if (f.out && cap > f.flags_cap) {
        free_typed(f.msg_flags);
        f.flags_cap = cap;
        void* p = nullptr;
        alloc(&p, (size_t)cap * 4);
        f.msg_flags = (uint32_t*)p;
    }
    if (cap <= f.pop_cap) return;
    void* p = nullptr;
    size_t sb = 0;
    detail::fgx_check(fgx_scratch_bytes(FGX_OP_COMPACT, cap, 0, &sb), "compaction scratch");
    if (f.d->has_agent_death) {
        free_typed(f.death_flags);
        alloc(&p, (size_t)cap * 4);
        f.death_flags = (uint32_t*)p;
        free_dev(f.death_scratch);
        alloc(&f.death_scratch, sb);
        f.death_scratch_bytes = sb;
    }
    if (f.out_agent) {
        free_typed(f.birth_flags);
        alloc(&p, (size_t)cap * 4);
        f.birth_flags = (uint32_t*)p;
        for (auto& v : f.newbuf) {
            free_dev(v.cur);
            free_dev(v.alt);
            alloc(&v.cur, (size_t)cap * v.bytes);
            alloc(&v.alt, (size_t)cap * v.bytes);
        }
        free_dev(f.birth_scratch);
        alloc(&f.birth_scratch, sb);
        f.birth_scratch_bytes = sb;
    }
    f.pop_cap = cap;
}

inline void BatchSimulation::size_sort_scratch(AgentB& a) {
    if (a.sort_msg.empty()) return;
    size_t sb = 0;
    detail::fgx_check(fgx_scratch_bytes(FGX_OP_AGENT_SORT, std::max<uint32_t>(a.list.cap, 1), a.keys_total, &sb),
                      "scratch sizing");
    if (sb <= a.scratch_bytes) return;
    free_dev(a.scratch);
    sb += sb / 2;
    alloc(&a.scratch, sb);
    detail::fgx_check(fgx_scratch_clear(a.scratch, sb, stream_), "scratch clear");
    a.scratch_bytes = sb;
}

// The packed {x, y, z, id} records of a spatial list cover its capacity.  A capacity
// change also rebuilds the launch arguments (graphs are cleared), which carry the pointer.
inline void BatchSimulation::size_records(MsgB& m) {
    if (!m.spatial || m.list.cap <= m.packed_cap) return;
    free_dev(m.packed);
    alloc(&m.packed, (size_t)m.list.cap * 16);
    m.packed_cap = m.list.cap;
}

inline void BatchSimulation::size_build_scratch(MsgB& m) {
    if (!m.spatial) return;
    size_t sb = 0;
    detail::fgx_check(fgx_scratch_bytes(FGX_OP_SPATIAL_BUILD, std::max<uint32_t>(m.list.cap, 1), m.keys_total, &sb),
                      "scratch sizing");
    if (sb <= m.scratch_bytes) return;
    free_dev(m.scratch);
    sb += sb / 2;
    alloc(&m.scratch, sb);
    detail::fgx_check(fgx_scratch_clear(m.scratch, sb, stream_), "scratch clear");
    m.scratch_bytes = sb;
}

// Before a step of a batch whose populations change: read the totals back and grow every
// list to the step's bound (births add at most one agent per parent and function), as
// CUDASimulation::ensure_capacity does for one simulation.
inline void BatchSimulation::ensure_step_capacity() {
    std::map<AgentB*, uint64_t> ub, peak;
    for (auto& kv : agents_) {
        uint32_t t = 0;
        detail::fgx_check(fgx_memcpy_d2h(&t, kv.second.list.d_total, 4, stream_), "total readback");
        synchronize();
        kv.second.list.total = t;
        ub[&kv.second] = t;
        peak[&kv.second] = t;
    }
    std::map<MsgB*, uint64_t> mub;
    for (auto& layer : m_->layers)
        for (auto& fd : layer->agent_functions) {
            FnB& f = fns_[fn_index_.at(fd.get())];
            const uint64_t n = ub[f.agent];
            if (f.out) mub[f.out] = std::max(mub[f.out], n);
            if (f.out_agent) {
                ub[f.out_agent] += n;
                peak[f.out_agent] = std::max(peak[f.out_agent], ub[f.out_agent]);
            }
        }
    bool changed = false;
    for (auto& kv : peak) {
        AgentB& a = *kv.first;
        if (kv.second >= (1ull << 31)) throw exception::InvalidOperation("BatchSimulation: batch exceeds the 32-bit index range");
        if (kv.second > a.list.cap) {
            grow_agent_list(a, (uint32_t)std::max<uint64_t>(kv.second, (uint64_t)a.list.cap + a.list.cap / 2));
            changed = true;
        }
        size_sort_scratch(a);
    }
    for (auto& kv : mub) {
        MsgB& m = *kv.first;
        if (kv.second > m.list.cap) {
            synchronize();
            const uint32_t cap = (uint32_t)std::max<uint64_t>(kv.second, (uint64_t)m.list.cap + m.list.cap / 2);
            for (auto& v : m.list.vars) {
                free_dev(v.cur);
                free_dev(v.alt);
                alloc(&v.cur, (size_t)cap * v.bytes);
                alloc(&v.alt, (size_t)cap * v.bytes);
            }
            m.list.cap = cap;
            changed = true;
        }
        size_records(m);
        size_build_scratch(m);
    }
    for (auto& f : fns_) {
        const uint32_t before = f.pop_cap;
        size_pop_buffers(f);
        changed = changed || f.pop_cap != before;
    }
    if (changed) clear_graphs();
}

inline void BatchSimulation::check_errors(const Sim& s, uint32_t slot) {
    uint32_t e[4] = {0, 0, 0, 0};
    detail::fgx_check(fgx_memcpy_d2h(e, d_errors_ + 4 * slot, 16, stream_), "error readback");
    synchronize();
    if (e[0] != detail::DERR_NONE)
        throw exception::DeviceError(detail::format("Device error %u (%s) in batched simulation %d, name hash 0x%08x, agent index %u",
                                                    e[0], detail::device_error_text(e[0]), s.handle, e[1], e[2]));
}

inline void BatchSimulation::getPopulationData(int h, AgentVector& pop) {
    Sim& s = sim(h);
    if (!s.resident) {
        auto it = s.pop.find(pop.getAgentName());
        if (it == s.pop.end()) throw exception::InvalidAgentName("BatchSimulation: unknown agent");
        pop = *it->second;
        return;
    }
    synchronize();
    uint32_t slot = 0;
    while (slot < order_.size() && order_[slot] != h) ++slot;
    check_errors(s, slot);
    auto it = agents_.find(pop.getAgentName());
    if (it == agents_.end()) throw exception::InvalidAgentName("BatchSimulation: unknown agent");
    const List& l = it->second.list;
    uint32_t n = l.cnt[slot], o = l.off[slot];
    if (dynamic_) {
        detail::fgx_check(fgx_memcpy_d2h(&n, l.d_cnt + slot, 4, stream_), "count readback");
        detail::fgx_check(fgx_memcpy_d2h(&o, l.d_off + slot, 4, stream_), "offset readback");
        synchronize();
    }
    pop.resize(n);
    for (auto& v : l.vars)
        if (n)
            detail::fgx_check(fgx_memcpy_d2h(pop.raw(v.name).data(), (const char*)v.cur + (size_t)o * v.bytes,
                                             (size_t)n * v.bytes, stream_),
                              "population download");
    synchronize();
}


// include/flamegpu/simulation/CUDAEnsemble.h:40-129.  simulate() runs the plans in one
// BatchSimulation when the model can be batched (the default; Config().batched = false
// opts out): every run's init functions run first on a simulation of its own, the
// batch steps all runs together, and each run's exit functions and exit log run on a
// simulation holding its final populations.  Other models, and ensembles with a step
// log, run each plan as its own CUDASimulation on `concurrent_runs` host threads
// sharing the device (src/flamegpu/simulation/detail/SimRunner.cu:51-90).
class CUDAEnsemble {
 public:
    struct EnsembleConfig {
        std::string out_directory;
        std::string out_format = "json";
        unsigned int concurrent_runs = 4;
        std::set<int> devices;
        bool quiet = false;
        bool verbose = false;
        bool timing = false;
        enum ErrorLevel { Off = 0, Slow = 1, Fast = 2 };
        ErrorLevel error_level = Slow;
        bool truncate_log_files = false;
        bool batched = true;          // run batchable models as one batch (no counterpart in the reference)
        size_t batch_bytes = 0;       // HBM the batch may hold at once (0: three quarters of the device's)
    };
    explicit CUDAEnsemble(const ModelDescription& model, int argc = 0, const char** argv = nullptr) : model_(model) {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if ((a == "-c" || a == "--concurrent") && i + 1 < argc) config_.concurrent_runs = (unsigned)std::stoul(argv[++i]);
            else if (a == "-q" || a == "--quiet") config_.quiet = true;
            else if (a == "-t" || a == "--timing") config_.timing = true;
        }
    }
    EnsembleConfig& Config() { return config_; }
    const EnsembleConfig& getConfig() const { return config_; }
    // Population initialiser called for every run before it starts.
    void setPopulationInit(std::function<void(CUDASimulation&, const RunPlan&, unsigned int)> f) { init_ = std::move(f); }
    // src/flamegpu/simulation/CUDAEnsemble.cu: setStepLog / setExitLog
    void setStepLog(const StepLoggingConfig& cfg) {
        if (cfg.model().get() != model_.data().get())
            throw exception::InvalidArgument("Model descriptions attached to LoggingConfig and CUDAEnsemble do not match, in CUDAEnsemble::setStepLog()");
        step_log_ = std::make_shared<StepLoggingConfig>(cfg);
    }
    void setExitLog(const LoggingConfig& cfg) {
        if (cfg.model().get() != model_.data().get())
            throw exception::InvalidArgument("Model descriptions attached to LoggingConfig and CUDAEnsemble do not match, in CUDAEnsemble::setExitLog()");
        exit_log_ = std::make_shared<LoggingConfig>(cfg);
    }
    unsigned int simulate(const RunPlanVector& plans) {
        for (auto& p : plans)
            if (p.model() != model_.data())
                throw exception::InvalidArgument("CUDAEnsemble::simulate: the run plans are for a different model");
        logs_.clear();
        logs_.resize(plans.size());
        run_logs_.clear();
        const auto t0 = std::chrono::steady_clock::now();
        unsigned int errors = 0;
        batched_ = config_.batched && batchable(plans);
        if (batched_) {
            try {
                errors = simulate_batched(plans);
            } catch (const NotBatchable&) {
                // an init function left agents outside their initial state: run every plan alone
                logs_.assign(plans.size(), RunRecord());
                run_logs_.clear();
                batched_ = false;
            }
        }
        if (!batched_) errors = simulate_pool(plans);
        elapsed_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (errors && config_.error_level != EnsembleConfig::Off) {
            throw exception::EnsembleError(detail::format("%u/%zu ensemble runs failed: %s", errors, plans.size(),
                                                          first_error().c_str()));
        }
        return errors;
    }
    // Whether the last simulate() ran its plans as one batch.
    bool lastRunBatched() const { return batched_; }
    // include/flamegpu/simulation/CUDAEnsemble.h: getLogs(), keyed by plan index
    const std::map<unsigned int, RunLog>& getLogs() const { return run_logs_; }
    const std::vector<RunRecord>& getRunRecords() const { return logs_; }
    double getEnsembleElapsedTime() const { return elapsed_; }

 private:

    struct NotBatchable {};
    bool batchable(const RunPlanVector& plans) const {
        if (step_log_) return false;
        for (auto& p : plans)
            if (p.getSteps() == 0) return false;   // runs until an exit condition: not a batch
        try {
            BatchSimulation::check_model(*model_.data(), true);
        } catch (const exception::InvalidOperation&) {
            return false;
        }
        return true;
    }
    int device() const { return config_.devices.empty() ? 0 : *config_.devices.begin(); }
    static void plan_env(CUDASimulation& sim, const RunPlan& p) {
        for (auto& kv : p.properties()) sim_set_env(sim, kv.first, kv.second);
    }
    // A run's batch plan: its populations and environment once its init functions ran.
    BatchSimulation::Plan prologue(const RunPlan& p, unsigned int i) {
        BatchSimulation::Plan bp;
        bp.seed = p.getRandomSimulationSeed();
        bp.steps = p.getSteps();
        const auto& m = *model_.data();
        if (!init_ && m.init_functions.empty()) {
            bp.env = m.env.buffer;
            for (auto& kv : p.properties()) {
                const auto& prop = m.env.props.at(kv.first);
                std::memcpy(bp.env.data() + prop.offset, kv.second.data(), kv.second.size());
            }
            return bp;
        }
        CUDASimulation sim(model_);
        sim.CUDAConfig().is_ensemble = true;
        sim.CUDAConfig().device_id = device();
        sim.SimulationConfig().random_seed = bp.seed;
        sim.SimulationConfig().steps = bp.steps;
        plan_env(sim, p);
        if (init_) init_(sim, p, i);
        sim.runInitPhase();
        bp.env = sim.env_host();
        for (auto& a : m.agents) {
            for (auto& st : a.second->states)
                if (st != a.second->initial_state && sim_count(sim, a.first, st) != 0) throw NotBatchable();
            auto pop = std::make_shared<AgentVector>(AgentDescription(a.second));
            sim.getPopulationData(*pop, a.second->initial_state);
            bp.populations[a.first] = pop;
        }
        return bp;
    }
    // A finished run: final populations, then exit functions and the exit log.
    void epilogue(BatchSimulation& batch, int h, const RunPlan& p, const std::vector<char>& env, RunRecord& lg,
                  unsigned int i) {
        const auto& m = *model_.data();
        std::map<std::string, std::shared_ptr<AgentVector>> pops;
        for (auto& a : m.agents) {
            auto pop = std::make_shared<AgentVector>(AgentDescription(a.second));
            batch.getPopulationData(h, *pop);
            lg.final_counts[a.first] = (unsigned int)pop->size();
            pops[a.first] = pop;
        }
        RunLog rl;
        if (!m.exit_functions.empty() || exit_log_) {
            CUDASimulation sim(model_);
            sim.CUDAConfig().is_ensemble = true;
            sim.CUDAConfig().device_id = device();
            sim.SimulationConfig().random_seed = p.getRandomSimulationSeed();
            sim.SimulationConfig().steps = p.getSteps();
            sim.env_host() = env;
            sim.mark_env_dirty();
            if (exit_log_) sim.setExitLog(*exit_log_);
            for (auto& a : m.agents)
                if (pops[a.first]->size()) sim.setPopulationData(*pops[a.first], a.second->initial_state);
            sim.runExitPhase(p.getSteps());
            rl = sim.getRunLog();
        }
        run_logs_[i] = rl;
    }
    // Runs join the batch while its HBM estimate stays within the budget; finished runs
    // leave after every step, and queued runs take their place.
    unsigned int simulate_batched(const RunPlanVector& plans) {
        BatchSimulation batch(model_, device(), true);
        size_t budget = config_.batch_bytes;
        if (!budget) {
            size_t free_b = 0, total_b = 0;
            budget = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? free_b / 4 * 3 : (size_t)64 << 30;
        }
        struct Running {
            size_t plan;
            size_t bytes;
            std::vector<char> env;
            std::chrono::steady_clock::time_point start;
        };
        std::map<int, Running> running;
        size_t used = 0, next = 0;
        unsigned int errors = 0;
        bool stop = false;
        while (!stop && (next < plans.size() || !running.empty())) {
            while (!stop && next < plans.size()) {
                const size_t i = next;
                RunRecord& lg = logs_[i];
                lg.started = true;
                lg.plan_index = (unsigned int)i;
                lg.seed = plans[i].getRandomSimulationSeed();
                lg.steps = plans[i].getSteps();
                const auto r0 = std::chrono::steady_clock::now();
                BatchSimulation::Plan bp;
                try {
                    bp = prologue(plans[i], (unsigned int)i);
                } catch (const NotBatchable&) {
                    throw;
                } catch (const std::exception& e) {
                    lg.error = e.what();
                    ++errors;
                    lg.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - r0).count();
                    if (config_.error_level == EnsembleConfig::Fast) stop = true;
                    ++next;
                    continue;
                }
                const size_t est = batch.estimateBytes(bp);
                if (!running.empty() && used + est > budget) break;   // waits for runs to finish
                const int h = batch.add(bp);
                running[h] = Running{i, est, bp.env, r0};
                used += est;
                ++next;
            }
            if (stop || running.empty()) continue;
            try {
                batch.step();
            } catch (const std::exception& e) {
                // a launch or host error of the batch: every resident run fails with it
                for (auto& kv : running) {
                    RunRecord& lg = logs_[kv.second.plan];
                    lg.error = e.what();
                    ++errors;
                }
                running.clear();
                if (config_.error_level == EnsembleConfig::Fast) stop = true;
                batch.synchronize();
                continue;
            }
            for (int h : batch.finished()) {
                Running& r = running.at(h);
                RunRecord& lg = logs_[r.plan];
                try {
                    epilogue(batch, h, plans[r.plan], r.env, lg, (unsigned int)r.plan);
                } catch (const std::exception& e) {
                    lg.error = e.what();
                    ++errors;
                    if (config_.error_level == EnsembleConfig::Fast) stop = true;
                }
                lg.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - r.start).count();
                used -= r.bytes;
                batch.remove(h);
                running.erase(h);
            }
        }
        return errors;
    }
    unsigned int simulate_pool(const RunPlanVector& plans) {
        std::mutex log_mutex;
        std::atomic<unsigned> next{0}, errors{0};
        std::atomic<bool> stop{false};
        auto worker = [&]() {
            while (!stop) {
                const unsigned i = next++;
                if (i >= plans.size()) break;
                RunRecord& lg = logs_[i];
                lg.started = true;
                lg.plan_index = i;
                lg.seed = plans[i].getRandomSimulationSeed();
                lg.steps = plans[i].getSteps();
                const auto r0 = std::chrono::steady_clock::now();
                try {
                    CUDASimulation sim(model_);
                    sim.CUDAConfig().is_ensemble = true;
                    sim.CUDAConfig().device_id = device();
                    sim.SimulationConfig().random_seed = plans[i].getRandomSimulationSeed();
                    sim.SimulationConfig().steps = plans[i].getSteps();
                    for (auto& kv : plans[i].properties()) sim_set_env(sim, kv.first, kv.second);
                    if (init_) init_(sim, plans[i], i);
                    if (step_log_) sim.setStepLog(*step_log_);
                    if (exit_log_) sim.setExitLog(*exit_log_);
                    sim.simulate();
                    for (auto& a : model_.data()->agents) lg.final_counts[a.first] = sim_count(sim, a.first);
                    const RunLog& rl = sim.getRunLog();
                    std::lock_guard<std::mutex> lock(log_mutex);
                    run_logs_[i] = rl;
                } catch (const std::exception& e) {
                    lg.error = e.what();
                    ++errors;
                    if (config_.error_level == EnsembleConfig::Fast) stop = true;
                }
                lg.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - r0).count();
            }
        };
        std::vector<std::thread> th;
        const unsigned nt = std::max(1u, std::min<unsigned>(config_.concurrent_runs, (unsigned)plans.size()));
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(worker);
        for (auto& t : th) t.join();
        return errors;
    }
    std::string first_error() const {
        for (auto& l : logs_)
            if (!l.error.empty()) return l.error;
        return "";
    }
    static void sim_set_env(CUDASimulation& s, const std::string& name, const std::vector<char>& b) {
        auto& props = s.model()->env.props;
        auto it = props.find(name);
        if (it == props.end()) throw exception::InvalidEnvProperty("unknown property " + name);
        std::memcpy(s.env_host().data() + it->second.offset, b.data(), b.size());
        s.mark_env_dirty();
    }
    static unsigned int sim_count(CUDASimulation& s, const std::string& agent, const std::string& state = "") {
        HostAPI api(s, 0);
        return api.agent(agent, state).count();
    }
    ModelDescription model_;
    EnsembleConfig config_;
    std::vector<RunRecord> logs_;
    std::map<unsigned int, RunLog> run_logs_;
    std::shared_ptr<StepLoggingConfig> step_log_;
    std::shared_ptr<LoggingConfig> exit_log_;
    std::function<void(CUDASimulation&, const RunPlan&, unsigned int)> init_;
    double elapsed_ = 0.0;
    bool batched_ = false;
};


}  // namespace flamegpu
