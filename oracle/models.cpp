// Serial CPU oracle: whole-model simulations.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Step semantics follow CUDASimulation::step/stepLayer
// (src/flamegpu/simulation/CUDASimulation.cu:575-1109):
//   per layer: auto spatial sort of agents whose function reads spatial messages
//   (sortPeriod 1, :671-682) -> lazy message index build (:864) -> functions ->
//   unmap: message count, death compaction, births appended (:1042-1094);
//   message lists are truncated at the end of the step (:619-624).
//
// Workloads:
//   model 0  circles spatial-3D: examples/cpp/circles_spatial3D/src/main.cu:5-54
//            (force k = sinf((d/r)*3.141*-2)*repulse), plus a clamp to [min,max]^3.
//            params {max, radius, repulse}; min = 0.
//   model 1  boids spatial-3D, wrapped: examples/python_native/boids_spatial3D_wrapped/
//            boids_spatial3D.py:79-240 (wrap() iterator, virtual coordinates, position wrap),
//            with the speed bounds 1.0/0.5 made parameters.
//            params {min, max, interaction_r, separation_r, time_scale, global_scale,
//                    steer_scale, collision_scale, match_scale, min_speed, max_speed}
//   model 2  predator-prey spatial-2D with birth/death (no reference example; the model
//            is specified here and mirrored by the library's example).
//            params {width, radius, eat_r, p_prey_death, p_prey_repro, p_pred_repro,
//                    starve_steps, jitter, flee_scale, hunt_scale}
//   model 3  circles brute-force: examples/cpp/circles_bruteforce/src/main.cu (same force).
//            params {max, radius, repulse}
// All FP32 arithmetic is written in the same order as the library's agent functions;
// build with -ffp-contract=off so that no FMA contraction changes rounding.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <memory>

namespace {

// Generic SoA agent list.
struct AgentList {
    std::vector<uint32_t> bytes;           // per-variable element size
    std::vector<std::vector<char>> data;   // per-variable storage
    std::vector<uint32_t> id;
    uint32_t n = 0;
    uint32_t next_id = 1;                   // ID_NOT_SET + 1
    template <typename T> T* v(size_t k) { return reinterpret_cast<T*>(data[k].data()); }
    void resize(uint32_t cap) {
        for (size_t k = 0; k < data.size(); ++k) data[k].resize((size_t)cap * bytes[k]);
        id.resize(cap);
    }
    void permute(const std::vector<uint32_t>& perm) {
        for (size_t k = 0; k < data.size(); ++k) {
            std::vector<char> t(data[k].size());
            const uint32_t b = bytes[k];
            for (uint32_t i = 0; i < n; ++i) std::memcpy(&t[(size_t)i * b], &data[k][(size_t)perm[i] * b], b);
            data[k].swap(t);
        }
        std::vector<uint32_t> t(id.size());
        for (uint32_t i = 0; i < n; ++i) t[i] = id[perm[i]];
        id.swap(t);
    }
    // Death compaction (stable, keep flag==1) over every variable and _id.
    void compact(const std::vector<uint32_t>& alive) {
        uint32_t p = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!alive[i]) continue;
            if (p != i) {
                for (size_t k = 0; k < data.size(); ++k)
                    std::memmove(&data[k][(size_t)p * bytes[k]], &data[k][(size_t)i * bytes[k]], bytes[k]);
                id[p] = id[i];
            }
            ++p;
        }
        n = p;
    }
};

// Message list with its spatial index.
struct MsgList {
    orc_spatial_meta meta{};
    bool spatial = true;
    uint32_t n = 0;
    std::vector<uint32_t> id;
    std::vector<float> x, y, z;
    std::vector<uint32_t> pbm;
    double last_candidates = 0.0;
    void build() {
        if (!spatial) return;
        std::vector<uint32_t> perm(n);
        pbm.assign(meta.bins + 1, 0);
        std::vector<uint32_t> id2(n);
        std::vector<float> x2(n), y2(n), z2(meta.dims == 3 ? n : 0);
        const void* in[4] = {id.data(), x.data(), y.data(), z.data()};
        void* out[4] = {id2.data(), x2.data(), y2.data(), z2.data()};
        const uint32_t b[4] = {4, 4, 4, 4};
        orc_spatial_build(&meta, n, x.data(), y.data(), meta.dims == 3 ? z.data() : nullptr,
                          meta.dims == 3 ? 4 : 3, in, b, out, pbm.data(), perm.data());
        id.swap(id2); x.swap(x2); y.swap(y2);
        if (meta.dims == 3) z.swap(z2);
        extra_permute(perm);
    }
    // Extra (non-location) message variables, permuted alongside.
    std::vector<std::vector<float>> extra;
    void extra_permute(const std::vector<uint32_t>& perm) {
        for (auto& e : extra) {
            std::vector<float> t(n);
            for (uint32_t i = 0; i < n; ++i) t[i] = e[perm[i]];
            e.swap(t);
        }
    }
};

template <typename Visit>
void for_candidates(const MsgList& m, int mode, float x, float y, float z, Visit&& visit) {
    if (!m.spatial) {
        for (uint32_t i = 0; i < m.n; ++i) visit(i);
        return;
    }
    // Reuse the primitive's enumeration order via a small buffer.
    static thread_local std::vector<uint32_t> buf;
    uint32_t c = orc_spatial_candidates(&m.meta, m.pbm.data(), mode, x, y, z, nullptr, 0);
    if (buf.size() < c) buf.resize(c);
    orc_spatial_candidates(&m.meta, m.pbm.data(), mode, x, y, z, buf.data(), c);
    for (uint32_t k = 0; k < c; ++k) visit(buf[k]);
}

void sort_agents(AgentList& a, const orc_spatial_meta& m, int xv, int yv, int zv) {
    std::vector<uint32_t> keys(a.n), perm(a.n);
    orc_agent_sort_keys(m.dims, m.min, m.max, m.radius, a.n, a.v<float>(xv), a.v<float>(yv),
                        m.dims == 3 ? a.v<float>(zv) : nullptr, keys.data(), nullptr);
    orc_stable_sort_perm(a.n, keys.data(), perm.data());
    a.permute(perm);
}

inline float vec3Length(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

struct Sim {
    int model = 0;
    std::vector<float> p;
    uint64_t seed = 0;
    uint32_t step = 0;
    std::vector<AgentList> agents;
    std::vector<MsgList> msgs;
    virtual ~Sim() = default;
    virtual void do_step() = 0;
};

// ---------------------------------------------------------------- circles
struct Circles : Sim {
    // agent vars: x, y, z, drift
    bool brute = false;
    void init() {
        agents.resize(1);
        agents[0].bytes = {4, 4, 4, 4};
        agents[0].data.resize(4);
        msgs.resize(1);
        const float mn[3] = {0, 0, 0};
        const float mx[3] = {p[0], p[0], p[0]};
        orc_spatial_meta_init(3, mn, mx, p[1], &msgs[0].meta);
        msgs[0].spatial = !brute;
    }
    void do_step() override {
        AgentList& a = agents[0];
        MsgList& m = msgs[0];
        const float MAXP = p[0], RADIUS = p[1], REPULSE = p[2];
        // Layer 1: output_message (setLocation / setVariable x,y,z; id)
        m.n = a.n;
        m.id.assign(a.id.begin(), a.id.begin() + a.n);
        m.x.assign(a.v<float>(0), a.v<float>(0) + a.n);
        m.y.assign(a.v<float>(1), a.v<float>(1) + a.n);
        m.z.assign(a.v<float>(2), a.v<float>(2) + a.n);
        // Layer 2: move; agent auto-sort then index build.
        if (!brute) sort_agents(a, m.meta, 0, 1, 2);
        m.build();
        float* X = a.v<float>(0); float* Y = a.v<float>(1); float* Z = a.v<float>(2); float* D = a.v<float>(3);
        uint64_t cand = 0;
        for (uint32_t i = 0; i < a.n; ++i) {
            const uint32_t ID = a.id[i];
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
            const float x1 = X[i], y1 = Y[i], z1 = Z[i];
            int count = 0;
            for_candidates(m, 0, x1, y1, z1, [&](uint32_t j) {
                ++cand;
                if (m.id[j] != ID) {
                    const float x2 = m.x[j], y2 = m.y[j], z2 = m.z[j];
                    float x21 = x2 - x1, y21 = y2 - y1, z21 = z2 - z1;
                    const float separation = std::sqrt(x21 * x21 + y21 * y21 + z21 * z21);
                    if (separation < RADIUS && separation > 0.0f) {
                        const float k = std::sin((separation / RADIUS) * 3.141f * -2.0f) * REPULSE;
                        x21 /= separation; y21 /= separation; z21 /= separation;
                        fx += k * x21; fy += k * y21; fz += k * z21;
                        count++;
                    }
                }
            });
            fx /= (float)(count > 0 ? count : 1);
            fy /= (float)(count > 0 ? count : 1);
            fz /= (float)(count > 0 ? count : 1);
            float nx = x1 + fx, ny = y1 + fy, nz = z1 + fz;
            nx = nx < 0.0f ? 0.0f : (nx > MAXP ? MAXP : nx);
            ny = ny < 0.0f ? 0.0f : (ny > MAXP ? MAXP : ny);
            nz = nz < 0.0f ? 0.0f : (nz > MAXP ? MAXP : nz);
            X[i] = nx; Y[i] = ny; Z[i] = nz;
            D[i] = std::sqrt(fx * fx + fy * fy + fz * fz);
        }
        m.last_candidates = a.n ? (double)cand / a.n : 0.0;
    }
};

// ---------------------------------------------------------------- boids (wrapped)
struct Boids : Sim {
    // agent vars: x, y, z, fx, fy, fz; message: id, x, y, z, fx, fy, fz
    void init() {
        agents.resize(1);
        agents[0].bytes = {4, 4, 4, 4, 4, 4};
        agents[0].data.resize(6);
        msgs.resize(1);
        const float mn[3] = {p[0], p[0], p[0]};
        const float mx[3] = {p[1], p[1], p[1]};
        orc_spatial_meta_init(3, mn, mx, p[2], &msgs[0].meta);
        msgs[0].extra.resize(3);
    }
    static float virt(float x2, float x1, float width) {
        const float x21 = x2 - x1;
        return std::fabs(x21) > width / 2.0f ? x2 - (x21 / std::fabs(x21) * width) : x2;
    }
    void do_step() override {
        AgentList& a = agents[0];
        MsgList& m = msgs[0];
        const float MINP = p[0], MAXP = p[1], INTERACTION_RADIUS = p[2], SEPARATION_RADIUS = p[3];
        const float TIME_SCALE = p[4], GLOBAL_SCALE = p[5], STEER_SCALE = p[6];
        const float COLLISION_SCALE = p[7], MATCH_SCALE = p[8], MIN_SPEED = p[9], MAX_SPEED = p[10];
        // Layer 1: outputdata
        m.n = a.n;
        m.id.assign(a.id.begin(), a.id.begin() + a.n);
        m.x.assign(a.v<float>(0), a.v<float>(0) + a.n);
        m.y.assign(a.v<float>(1), a.v<float>(1) + a.n);
        m.z.assign(a.v<float>(2), a.v<float>(2) + a.n);
        for (int k = 0; k < 3; ++k) m.extra[k].assign(a.v<float>(3 + k), a.v<float>(3 + k) + a.n);
        // Layer 2: inputdata
        sort_agents(a, m.meta, 0, 1, 2);
        m.build();
        const float WX = m.meta.width[0], WY = m.meta.width[1], WZ = m.meta.width[2];
        float* X = a.v<float>(0); float* Y = a.v<float>(1); float* Z = a.v<float>(2);
        float* FX = a.v<float>(3); float* FY = a.v<float>(4); float* FZ = a.v<float>(5);
        const float* MFX = m.extra[0].data(); const float* MFY = m.extra[1].data(); const float* MFZ = m.extra[2].data();
        uint64_t cand = 0;
        for (uint32_t i = 0; i < a.n; ++i) {
            const uint32_t id = a.id[i];
            float agent_x = X[i], agent_y = Y[i], agent_z = Z[i];
            float agent_fx = FX[i], agent_fy = FY[i], agent_fz = FZ[i];
            float pcx = 0.0f, pcy = 0.0f, pcz = 0.0f;
            int perceived_count = 0;
            float gvx = 0.0f, gvy = 0.0f, gvz = 0.0f;
            float vcx = 0.0f, vcy = 0.0f, vcz = 0.0f;
            for_candidates(m, 1, agent_x, agent_y, agent_z, [&](uint32_t j) {
                ++cand;
                if (m.id[j] != id) {
                    const float message_x = virt(m.x[j], agent_x, WX);
                    const float message_y = virt(m.y[j], agent_y, WY);
                    const float message_z = virt(m.z[j], agent_z, WZ);
                    const float separation = vec3Length(agent_x - message_x, agent_y - message_y, agent_z - message_z);
                    if (separation < INTERACTION_RADIUS) {
                        pcx += message_x; pcy += message_y; pcz += message_z;
                        perceived_count++;
                        gvx += MFX[j]; gvy += MFY[j]; gvz += MFZ[j];
                        if (separation < SEPARATION_RADIUS) {
                            const float normalizedSeparation = separation / SEPARATION_RADIUS;
                            const float invNormSep = 1.0f - normalizedSeparation;
                            const float invSqSep = invNormSep * invNormSep;
                            vcx += COLLISION_SCALE * (agent_x - message_x) * invSqSep;
                            vcy += COLLISION_SCALE * (agent_y - message_y) * invSqSep;
                            vcz += COLLISION_SCALE * (agent_z - message_z) * invSqSep;
                        }
                    }
                }
            });
            if (perceived_count) {
                const float c = (float)perceived_count;
                pcx /= c; pcy /= c; pcz /= c;
                gvx /= c; gvy /= c; gvz /= c;
                vcx += (pcx - agent_x) * STEER_SCALE;
                vcy += (pcy - agent_y) * STEER_SCALE;
                vcz += (pcz - agent_z) * STEER_SCALE;
                vcx += gvx * MATCH_SCALE - agent_fx;
                vcy += gvy * MATCH_SCALE - agent_fy;
                vcz += gvz * MATCH_SCALE - agent_fz;
            }
            vcx *= GLOBAL_SCALE; vcy *= GLOBAL_SCALE; vcz *= GLOBAL_SCALE;
            agent_fx += vcx; agent_fy += vcy; agent_fz += vcz;
            const float agent_fscale = vec3Length(agent_fx, agent_fy, agent_fz);
            if (agent_fscale > MAX_SPEED) {
                agent_fx /= agent_fscale; agent_fy /= agent_fscale; agent_fz /= agent_fscale;
                agent_fx *= MAX_SPEED; agent_fy *= MAX_SPEED; agent_fz *= MAX_SPEED;
            }
            if (agent_fscale < MIN_SPEED) {
                agent_fx /= agent_fscale; agent_fy /= agent_fscale; agent_fz /= agent_fscale;
                agent_fx *= MIN_SPEED; agent_fy *= MIN_SPEED; agent_fz *= MIN_SPEED;
            }
            agent_x += agent_fx * TIME_SCALE;
            agent_y += agent_fy * TIME_SCALE;
            agent_z += agent_fz * TIME_SCALE;
            const float width = MAXP - MINP;
            if (agent_x < MINP) agent_x += width;
            if (agent_y < MINP) agent_y += width;
            if (agent_z < MINP) agent_z += width;
            if (agent_x > MAXP) agent_x -= width;
            if (agent_y > MAXP) agent_y -= width;
            if (agent_z > MAXP) agent_z -= width;
            X[i] = agent_x; Y[i] = agent_y; Z[i] = agent_z;
            FX[i] = agent_fx; FY[i] = agent_fy; FZ[i] = agent_fz;
        }
        m.last_candidates = a.n ? (double)cand / a.n : 0.0;
    }
};

// ---------------------------------------------------------------- predator-prey
// Prey vars: x, y, vx, vy.  Predator vars: x, y, vx, vy, life (u32, steps since fed).
// Function streams (RNG) follow definition order:
//   prey: output 0, flee 1, move 2, reproduce 3; predator: output 4, hunt 5, move 6,
//   reproduce 7, starve 8.
// Layers: [prey_output, pred_output] [prey_flee, pred_hunt] [prey_move, pred_move]
//         [prey_reproduce, pred_reproduce] [pred_starve]
struct PredPrey : Sim {
    void init() {
        agents.resize(2);
        agents[0].bytes = {4, 4, 4, 4};
        agents[0].data.resize(4);
        agents[1].bytes = {4, 4, 4, 4, 4};
        agents[1].data.resize(5);
        msgs.resize(2);  // 0 prey_location, 1 predator_location
        const float mn[2] = {0, 0};
        const float mx[2] = {p[0], p[0]};
        orc_spatial_meta_init(2, mn, mx, p[1], &msgs[0].meta);
        orc_spatial_meta_init(2, mn, mx, p[1], &msgs[1].meta);
    }
    static float wrapc(float v, float w) {
        if (v < 0.0f) v += w;
        if (v >= w) v -= w;
        return v;
    }
    void output(AgentList& a, MsgList& m) {
        m.n = a.n;
        m.id.assign(a.id.begin(), a.id.begin() + a.n);
        m.x.assign(a.v<float>(0), a.v<float>(0) + a.n);
        m.y.assign(a.v<float>(1), a.v<float>(1) + a.n);
    }
    void do_step() override {
        const float W = p[0], EAT_R = p[2], P_PREY_DEATH = p[3], P_PREY_REPRO = p[4];
        const float P_PRED_REPRO = p[5];
        const uint32_t STARVE = (uint32_t)p[6];
        const float JITTER = p[7], FLEE = p[8], HUNT = p[9];
        AgentList& prey = agents[0];
        AgentList& pred = agents[1];
        MsgList& mprey = msgs[0];
        MsgList& mpred = msgs[1];
        // L1
        output(prey, mprey);
        output(pred, mpred);
        // L2: sort both populations (each function reads spatial-2D messages), then build.
        sort_agents(prey, mpred.meta, 0, 1, -1);
        sort_agents(pred, mprey.meta, 0, 1, -1);
        mpred.build();
        mprey.build();
        {   // prey_flee: MessageSpatial2D in, death
            const float R = mpred.meta.radius;
            std::vector<uint32_t> alive(prey.n, 1);
            float* X = prey.v<float>(0); float* Y = prey.v<float>(1);
            float* VX = prey.v<float>(2); float* VY = prey.v<float>(3);
            uint64_t cand = 0;
            for (uint32_t i = 0; i < prey.n; ++i) {
                const float x1 = X[i], y1 = Y[i];
                float fx = 0.0f, fy = 0.0f;
                int eaten = 0;
                for_candidates(mpred, 0, x1, y1, 0.0f, [&](uint32_t j) {
                    ++cand;
                    const float dx = x1 - mpred.x[j];
                    const float dy = y1 - mpred.y[j];
                    const float d = std::sqrt(dx * dx + dy * dy);
                    if (d < R && d > 0.0f) {
                        fx += dx / d;
                        fy += dy / d;
                    }
                    if (d < EAT_R) eaten = 1;
                });
                float vx = VX[i] + fx * FLEE;
                float vy = VY[i] + fy * FLEE;
                const float s = std::sqrt(vx * vx + vy * vy);
                if (s > 1.0f) { vx /= s; vy /= s; }
                VX[i] = vx; VY[i] = vy;
                if (eaten) alive[i] = 0;
                else if (orc_rng_uniform(seed, prey.id[i], step, 1, 0) < P_PREY_DEATH) alive[i] = 0;
            }
            mpred.last_candidates = prey.n ? (double)cand / prey.n : 0.0;
            // pred_hunt runs in the same layer on the other population; prey death is
            // processed in the unmap phase after both functions ran.
            {
                const float RR = mprey.meta.radius;
                float* PX = pred.v<float>(0); float* PY = pred.v<float>(1);
                float* PVX = pred.v<float>(2); float* PVY = pred.v<float>(3);
                uint32_t* LIFE = pred.v<uint32_t>(4);
                uint64_t cand2 = 0;
                for (uint32_t i = 0; i < pred.n; ++i) {
                    const float x1 = PX[i], y1 = PY[i];
                    float best = RR, bx = 0.0f, by = 0.0f;
                    int found = 0;
                    for_candidates(mprey, 0, x1, y1, 0.0f, [&](uint32_t j) {
                        ++cand2;
                        const float dx = mprey.x[j] - x1;
                        const float dy = mprey.y[j] - y1;
                        const float d = std::sqrt(dx * dx + dy * dy);
                        if (d < best) { best = d; bx = dx; by = dy; found = 1; }
                    });
                    float vx = PVX[i], vy = PVY[i];
                    if (found && best > 0.0f) {
                        vx += bx / best * HUNT;
                        vy += by / best * HUNT;
                    }
                    const float s = std::sqrt(vx * vx + vy * vy);
                    if (s > 1.0f) { vx /= s; vy /= s; }
                    PVX[i] = vx; PVY[i] = vy;
                    LIFE[i] = (found && best < EAT_R) ? 0u : LIFE[i] + 1u;
                }
                mprey.last_candidates = pred.n ? (double)cand2 / pred.n : 0.0;
            }
            prey.compact(alive);
        }
        // L3: move
        for (int t = 0; t < 2; ++t) {
            AgentList& a = agents[t];
            float* X = a.v<float>(0); float* Y = a.v<float>(1);
            float* VX = a.v<float>(2); float* VY = a.v<float>(3);
            for (uint32_t i = 0; i < a.n; ++i) {
                X[i] = wrapc(X[i] + VX[i], W);
                Y[i] = wrapc(Y[i] + VY[i], W);
            }
        }
        // L4: reproduce (agent output into the same agent type), both in one layer.
        {
            const uint32_t np = prey.n;
            std::vector<uint32_t> flag(np, 0);
            std::vector<float> nx(np, 0.0f), ny(np, 0.0f), nvx(np, 0.0f), nvy(np, 0.0f);
            float* X = prey.v<float>(0); float* Y = prey.v<float>(1);
            for (uint32_t i = 0; i < np; ++i) {
                const uint32_t id = prey.id[i];
                if (orc_rng_uniform(seed, id, step, 3, 0) < P_PREY_REPRO) {
                    flag[i] = 1;
                    nx[i] = wrapc(X[i] + (orc_rng_uniform(seed, id, step, 3, 1) - 0.5f) * JITTER, W);
                    ny[i] = wrapc(Y[i] + (orc_rng_uniform(seed, id, step, 3, 2) - 0.5f) * JITTER, W);
                    nvx[i] = orc_rng_uniform(seed, id, step, 3, 3) * 2.0f - 1.0f;
                    nvy[i] = orc_rng_uniform(seed, id, step, 3, 4) * 2.0f - 1.0f;
                }
            }
            const uint32_t npd = pred.n;
            std::vector<uint32_t> pflag(npd, 0), nlife(npd, 0);
            std::vector<float> px(npd, 0.0f), py(npd, 0.0f), pvx(npd, 0.0f), pvy(npd, 0.0f);
            float* PX = pred.v<float>(0); float* PY = pred.v<float>(1);
            const uint32_t* LIFE = pred.v<uint32_t>(4);
            for (uint32_t i = 0; i < npd; ++i) {
                const uint32_t id = pred.id[i];
                if (LIFE[i] == 0u && orc_rng_uniform(seed, id, step, 7, 0) < P_PRED_REPRO) {
                    pflag[i] = 1;
                    px[i] = wrapc(PX[i] + (orc_rng_uniform(seed, id, step, 7, 1) - 0.5f) * JITTER, W);
                    py[i] = wrapc(PY[i] + (orc_rng_uniform(seed, id, step, 7, 2) - 0.5f) * JITTER, W);
                    pvx[i] = orc_rng_uniform(seed, id, step, 7, 3) * 2.0f - 1.0f;
                    pvy[i] = orc_rng_uniform(seed, id, step, 7, 4) * 2.0f - 1.0f;
                    nlife[i] = 0u;
                }
            }
            {
                uint32_t births = 0;
                for (uint32_t i = 0; i < np; ++i) births += flag[i];
                prey.resize(prey.n + births);
                const void* nb[4] = {nx.data(), ny.data(), nvx.data(), nvy.data()};
                void* dst[5] = {prey.data[0].data(), prey.data[1].data(), prey.data[2].data(), prey.data[3].data(), prey.id.data()};
                const uint32_t by[5] = {4, 4, 4, 4, 4};
                const void* nb5[5] = {nb[0], nb[1], nb[2], nb[3], nullptr};
                prey.n = orc_births(np, flag.data(), 5, nb5, by, dst, prey.n, 4, &prey.next_id);
            }
            {
                uint32_t births = 0;
                for (uint32_t i = 0; i < npd; ++i) births += pflag[i];
                pred.resize(pred.n + births);
                void* dst[6] = {pred.data[0].data(), pred.data[1].data(), pred.data[2].data(), pred.data[3].data(),
                                pred.data[4].data(), pred.id.data()};
                const uint32_t by[6] = {4, 4, 4, 4, 4, 4};
                const void* nb[6] = {px.data(), py.data(), pvx.data(), pvy.data(), nlife.data(), nullptr};
                pred.n = orc_births(npd, pflag.data(), 6, nb, by, dst, pred.n, 5, &pred.next_id);
            }
        }
        // L5: pred_starve
        {
            std::vector<uint32_t> alive(pred.n, 1);
            const uint32_t* LIFE = pred.v<uint32_t>(4);
            for (uint32_t i = 0; i < pred.n; ++i) alive[i] = LIFE[i] >= STARVE ? 0u : 1u;
            pred.compact(alive);
        }
    }
};

}  // namespace

struct orc_sim {
    std::unique_ptr<Sim> s;
};

extern "C" orc_sim* orc_sim_create(int model, const float* params, uint32_t nparams, uint64_t seed) {
    std::unique_ptr<Sim> s;
    if (model == 0 || model == 3) {
        if (nparams < 3) return nullptr;
        auto c = new Circles();
        c->brute = model == 3;
        c->p.assign(params, params + nparams);
        c->init();
        s.reset(c);
    } else if (model == 1) {
        if (nparams < 11) return nullptr;
        auto b = new Boids();
        b->p.assign(params, params + nparams);
        b->init();
        s.reset(b);
    } else if (model == 2) {
        if (nparams < 10) return nullptr;
        auto pp = new PredPrey();
        pp->p.assign(params, params + nparams);
        pp->init();
        s.reset(pp);
    } else {
        return nullptr;
    }
    s->model = model;
    s->seed = seed;
    auto* r = new orc_sim();
    r->s = std::move(s);
    return r;
}

extern "C" void orc_sim_destroy(orc_sim* s) { delete s; }

extern "C" uint32_t orc_sim_var_count(const orc_sim* s, int t) {
    if (t < 0 || t >= (int)s->s->agents.size()) return 0;
    return (uint32_t)s->s->agents[t].bytes.size();
}
extern "C" uint32_t orc_sim_var_bytes(const orc_sim* s, int t, uint32_t v) {
    if (t < 0 || t >= (int)s->s->agents.size()) return 0;
    const auto& a = s->s->agents[t];
    return v < a.bytes.size() ? a.bytes[v] : 0;
}

extern "C" int orc_sim_set_agents(orc_sim* s, int t, uint32_t n, const void* const* vars, const uint32_t* ids) {
    if (t < 0 || t >= (int)s->s->agents.size()) return -1;
    AgentList& a = s->s->agents[t];
    a.n = n;
    a.resize(n);
    for (size_t k = 0; k < a.data.size(); ++k) std::memcpy(a.data[k].data(), vars[k], (size_t)n * a.bytes[k]);
    if (ids) {
        uint32_t mx = 0;
        for (uint32_t i = 0; i < n; ++i) { a.id[i] = ids[i]; mx = ids[i] > mx ? ids[i] : mx; }
        a.next_id = mx + 1;
    } else {
        // assignIDs (CUDAFatAgent.cu:347-391): id = nextID + index, nextID += count.
        for (uint32_t i = 0; i < n; ++i) a.id[i] = a.next_id + i;
        a.next_id += n;
    }
    return 0;
}

extern "C" uint32_t orc_sim_count(const orc_sim* s, int t) {
    if (t < 0 || t >= (int)s->s->agents.size()) return 0;
    return s->s->agents[t].n;
}

extern "C" int orc_sim_get_agents(const orc_sim* s, int t, void* const* vars, uint32_t* ids) {
    if (t < 0 || t >= (int)s->s->agents.size()) return -1;
    const AgentList& a = s->s->agents[t];
    for (size_t k = 0; k < a.data.size(); ++k)
        if (vars[k]) std::memcpy(vars[k], a.data[k].data(), (size_t)a.n * a.bytes[k]);
    if (ids) std::memcpy(ids, a.id.data(), (size_t)a.n * 4);
    return 0;
}

extern "C" int orc_sim_step(orc_sim* s, uint32_t nsteps) {
    for (uint32_t k = 0; k < nsteps; ++k) {
        s->s->do_step();
        s->s->step++;
    }
    return 0;
}

extern "C" uint32_t orc_sim_bins(const orc_sim* s, int msg) {
    if (msg < 0 || msg >= (int)s->s->msgs.size()) return 0;
    return s->s->msgs[msg].meta.bins;
}

extern "C" int orc_sim_get_pbm(const orc_sim* s, int msg, uint32_t* pbm) {
    if (msg < 0 || msg >= (int)s->s->msgs.size()) return -1;
    const MsgList& m = s->s->msgs[msg];
    if (m.pbm.empty()) return -1;
    std::memcpy(pbm, m.pbm.data(), m.pbm.size() * 4);
    return 0;
}

extern "C" double orc_sim_last_candidates(const orc_sim* s, int msg) {
    if (msg < 0 || msg >= (int)s->s->msgs.size()) return 0.0;
    return s->s->msgs[msg].last_candidates;
}

extern "C" int orc_sim_set_bin_mult(orc_sim* s, float mult) {
    for (auto& m : s->s->msgs) {
        if (!m.spatial) continue;
        const orc_spatial_meta old = m.meta;
        if (orc_spatial_meta_init_mult(old.dims, old.min, old.max, old.radius, mult, &m.meta)) return -1;
    }
    return 0;
}
