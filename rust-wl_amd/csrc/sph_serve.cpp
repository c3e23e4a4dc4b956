// Continuous-batching server (see sph_serve.h).
#include "sph_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace vxl {

namespace {

using Clock = std::chrono::steady_clock;

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double r = q * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(r), hi = (size_t)std::ceil(r);
    return v[lo] + (v[hi] - v[lo]) * (r - (double)lo);
}

uint64_t fnv_fluid(const std::vector<float> &pos, const std::vector<float> &vel, uint32_t nf) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        h ^= (uint64_t)u;
        h *= 0x00000100000001b3ull;
    };
    for (uint32_t i = 0; i < nf; ++i) {
        mix(pos[3 * i]); mix(pos[3 * i + 1]); mix(pos[3 * i + 2]);
        mix(vel[3 * i]); mix(vel[3 * i + 1]); mix(vel[3 * i + 2]);
    }
    return h;
}

struct Built {
    uint32_t index;
    SceneConfig sc;
    std::vector<float> pos, vel, pmass;
    std::vector<Span> spans;
    uint32_t nf = 0;
    uint32_t cells_hint = 0;
    // coupled requests (dynamic boxes)
    std::vector<BoxBody> bodies;
    std::vector<float> density;
    RigidConfig rc;
};

}  // namespace

void request_bodies(const SceneSample &req, std::vector<BoxBody> &bodies, std::vector<float> &density,
                    RigidConfig &rc) {
    bodies = req.dam.walls ? tank_bodies(req.dam.tank, req.dam.d) : std::vector<BoxBody>{};
    density.assign(bodies.size(), 0.0f);
    for (uint32_t b = 0; b < req.nboxes && b < 4; ++b) {
        bodies.push_back(req.boxes[b]);
        density.push_back(req.box_density[b]);
    }
    rc = RigidConfig{};
    for (int a = 0; a < 3; ++a) rc.tank[a] = req.dam.tank[a];
}

void build_request(const SceneSample &req, const ServeConfig &cfg, SceneConfig &sc, std::vector<float> &pos,
                   std::vector<float> &vel, std::vector<float> &pmass, std::vector<Span> &spans, uint32_t &nf,
                   std::vector<BoxBody> *bodies, std::vector<float> *density, RigidConfig *rc) {
    const float d = req.dam.d;
    sc = SceneConfig{};
    sc.smoothing_radius = 2.0f * d;
    sc.sound_speed = req.sound_speed;
    sc.artificial_viscosity = req.alpha;
    sc.xsph_viscosity = 0.5f;
    sc.substeps = 1;
    sc.boundary_layers = 1;
    sc.particle_mass = ((sc.rest_density * d) * d) * d;
    sc.spacing = d;
    sc.max_bins = cfg.max_bins;
    sc.dt_tick = cfg.dt;
    if (!req.nboxes) {
        dam_break(req.dam, sc, pos, vel, pmass, spans, nf);
        return;
    }
    if (!bodies || !density || !rc) throw std::invalid_argument("a request with boxes needs its body list");
    DamBreakSpec fluid_only = req.dam;
    fluid_only.walls = false;
    dam_break(fluid_only, sc, pos, vel, pmass, spans, nf);
    request_bodies(req, *bodies, *density, *rc);
    append_boundary_boxes(sc, *bodies, pos, vel, pmass, spans);
}

uint32_t scene_cells_hint(const std::vector<float> &pos, size_t n, float h, uint64_t max_bins) {
    float lo[3], hi[3], gmin[3], bin;
    uint32_t dims[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = pos[a];
    for (size_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], pos[3 * i + a]);
            hi[a] = fmaxf(hi[a], pos[3 * i + a]);
        }
    grid_box(lo, hi, h, max_bins, gmin, bin, dims);
    return dims[0] * dims[1] * dims[2];
}

uint64_t request_solo_bytes(const SceneSample &req, const ServeConfig &cfg) {
    SceneConfig sc;
    std::vector<float> pos, vel, pmass, density;
    std::vector<Span> spans;
    std::vector<BoxBody> bodies;
    RigidConfig rc;
    uint32_t nf = 0;
    build_request(req, cfg, sc, pos, vel, pmass, spans, nf, &bodies, &density, &rc);
    const uint32_t n = (uint32_t)pmass.size();
    return Engine::solo_peak(n, scene_cells_hint(pos, n, sc.smoothing_radius, sc.max_bins), sc.max_bins,
                             !bodies.empty());
}

ServeStats serve(Engine &eng, const std::vector<SceneSample> &reqs, const ServeConfig &cfg,
                 std::vector<ServeResult> &results) {
    ServeStats st;
    st.requests = (uint32_t)reqs.size();
    results.assign(reqs.size(), ServeResult{});
    std::vector<uint32_t> order(reqs.size());
    for (uint32_t k = 0; k < order.size(); ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });

    const auto t0 = Clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    auto sched_now = [&] { return now_s() * cfg.time_scale; };

    // Builder threads: build arrived requests in arrival order.
    std::mutex mu;
    std::condition_variable cv;
    size_t next_build = 0;  // position in `order`
    std::map<uint32_t, Built> ready;  // by arrival position
    std::string build_error;
    bool stop = false;
    auto builder = [&] {
        for (;;) {
            size_t pos_k;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] {
                    return stop || (next_build < order.size() && reqs[order[next_build]].arrival_s <= sched_now());
                });
                if (stop) return;
                pos_k = next_build++;
            }
            Built b;
            b.index = order[pos_k];
            try {
                build_request(reqs[b.index], cfg, b.sc, b.pos, b.vel, b.pmass, b.spans, b.nf, &b.bodies, &b.density,
                              &b.rc);
                b.cells_hint = scene_cells_hint(b.pos, b.pmass.size(), b.sc.smoothing_radius, b.sc.max_bins);
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> lk(mu);
                build_error = e.what();
                stop = true;
                cv.notify_all();
                return;
            }
            std::lock_guard<std::mutex> lk(mu);
            ready.emplace(pos_k, std::move(b));
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t k = 0; k < std::max<uint32_t>(cfg.builders, 1); ++k) pool.emplace_back(builder);
    // Wake builders as arrivals pass (they wait on the schedule clock).
    std::thread ticker([&] {
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (stop || next_build >= order.size()) return;
            }
            cv.notify_all();
            std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
    });

    struct Resident {
        uint32_t index, scene, steps_left, nf, n;
    };
    std::vector<Resident> resident;
    size_t next_admit = 0;  // position in `order`
    std::vector<double> step_ms, loop_ms, lat, wait;
    std::string error;
    // Pool growth stays inside the budget; admission checks the engine's
    // projected peak (old and new pool alive together while scenes move,
    // readback staging included), so device bytes never exceed the budget.
    eng.set_memory_limit(cfg.budget_bytes);
    try {
        while (st.completed + st.rejected < st.requests) {
            if (now_s() > cfg.max_wall_s) break;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!build_error.empty()) throw std::runtime_error("scene build failed: " + build_error);
            }
            const auto loop_t0 = Clock::now();
            // Admit built requests in arrival order while the budget allows.
            for (;;) {
                if (next_admit >= order.size()) break;
                if (cfg.max_resident && resident.size() >= cfg.max_resident) break;
                Built b;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    auto it = ready.find((uint32_t)next_admit);
                    if (it == ready.end()) break;
                    b = std::move(it->second);
                    ready.erase(it);
                }
                const uint32_t n = (uint32_t)b.pmass.size();
                if (resident.empty()) eng.release_if_empty();
                const uint64_t peak = eng.projected_peak(n, b.cells_hint, cfg.max_bins, !b.bodies.empty());
                if (peak > cfg.budget_bytes) {
                    if (!resident.empty()) {
                        // wait for retirements; keep the built scene for the next try
                        std::lock_guard<std::mutex> lk(mu);
                        ready.emplace((uint32_t)next_admit, std::move(b));
                        break;
                    }
                    // Too large even alone: rejected (steps = 0 in its result).
                    ServeResult &res = results[b.index];
                    res.index = b.index;
                    res.n = n;
                    res.nf = b.nf;
                    res.steps = 0;
                    res.arrival_s = reqs[b.index].arrival_s;
                    res.admit_s = res.finish_s = sched_now();
                    ++st.rejected;
                    ++next_admit;
                    continue;
                }
                const uint32_t sid =
                    b.bodies.empty()
                        ? eng.add_scene(b.sc, n, b.nf, b.pos.data(), b.vel.data(), b.pmass.data(), b.spans.data(),
                                        (uint32_t)b.spans.size(), false)
                        : eng.add_coupled_scene(b.sc, b.nf, b.pos.data(), b.vel.data(), b.bodies, b.density, b.rc);
                const SceneSample &rq = reqs[b.index];
                resident.push_back(Resident{b.index, sid, std::max<uint32_t>(rq.steps, 1), b.nf, n});
                ServeResult &res = results[b.index];
                res.index = b.index;
                res.n = n;
                res.nf = b.nf;
                res.steps = std::max<uint32_t>(rq.steps, 1);
                res.arrival_s = rq.arrival_s;
                res.admit_s = sched_now();
                wait.push_back(res.admit_s - rq.arrival_s);
                ++next_admit;
            }
            if (resident.empty()) {
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                continue;
            }
            st.peak_resident = std::max<uint32_t>(st.peak_resident, (uint32_t)resident.size());
            // One pool step for every resident scene.
            const std::vector<float> ms = eng.step_timed(1);
            step_ms.push_back(ms.empty() ? 0.0 : ms[0]);
            ++st.pool_steps;
            // Retire finished scenes.
            for (size_t k = 0; k < resident.size();) {
                Resident &r = resident[k];
                st.fluid_updates += r.nf;
                if (--r.steps_left == 0) {
                    std::vector<float> p(3ull * r.n), v(3ull * r.n);
                    eng.read_state(r.scene, p.data(), v.data());
                    ServeResult &res = results[r.index];
                    res.state_hash = fnv_fluid(p, v, r.nf);
                    res.finish_s = sched_now();
                    lat.push_back(res.finish_s - res.arrival_s);
                    eng.remove_scene(r.scene);
                    ++st.completed;
                    resident[k] = resident.back();
                    resident.pop_back();
                } else {
                    ++k;
                }
            }
            loop_ms.push_back(std::chrono::duration<double, std::milli>(Clock::now() - loop_t0).count());
        }
    } catch (const std::exception &e) {
        error = e.what();
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto &t : pool) t.join();
    ticker.join();
    for (auto &r : resident) eng.remove_scene(r.scene);
    if (!error.empty()) throw std::runtime_error(error);

    st.wall_s = now_s();
    st.updates_per_s = st.wall_s > 0 ? (double)st.fluid_updates / st.wall_s : 0.0;
    st.step_ms_p50 = pct(step_ms, 0.50);
    st.step_ms_p99 = pct(step_ms, 0.99);
    double sum = 0.0;
    for (double v : step_ms) sum += v;
    st.step_ms_mean = step_ms.empty() ? 0.0 : sum / (double)step_ms.size();
    st.loop_ms_p50 = pct(loop_ms, 0.50);
    st.loop_ms_p99 = pct(loop_ms, 0.99);
    st.latency_s_p50 = pct(lat, 0.50);
    st.latency_s_p99 = pct(lat, 0.99);
    st.wait_s_p50 = pct(wait, 0.50);
    st.wait_s_p99 = pct(wait, 0.99);
    st.layouts = eng.layouts();
    st.peak_bytes = eng.peak_bytes();
    return st;
}

}  // namespace vxl
