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

uint64_t fnv_fluid(const float *pos, const float *vel, uint32_t nf) {
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

// A request's scene, prepared by a builder thread (records packed, bodies
// and lattices built), so that admitting it only moves it into the pool.
struct Built {
    uint32_t index = 0;
    SceneSlot scene;
    uint32_t n = 0, nf = 0;
    bool coupled = false;
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
                             !bodies.empty(), true);
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
                const SceneSample &rq = reqs[b.index];
                SceneConfig sc;
                std::vector<float> pos, vel, pmass;
                std::vector<Span> spans;
                if (rq.nboxes) {
                    // fluid only here: prepare_coupled samples the bodies
                    SceneSample fluid_only = rq;
                    fluid_only.nboxes = 0;
                    fluid_only.dam.walls = false;
                    build_request(fluid_only, cfg, sc, pos, vel, pmass, spans, b.nf);
                    std::vector<BoxBody> bodies;
                    std::vector<float> density;
                    RigidConfig rc;
                    request_bodies(rq, bodies, density, rc);
                    b.scene = Engine::prepare_coupled(sc, b.nf, pos.data(), vel.data(), bodies, density, rc);
                    b.coupled = true;
                } else {
                    build_request(rq, cfg, sc, pos, vel, pmass, spans, b.nf);
                    b.scene = Engine::prepare_scene(sc, (uint32_t)pmass.size(), b.nf, pos.data(), vel.data(),
                                                    pmass.data(), spans.data(), (uint32_t)spans.size(), false);
                }
                b.n = b.scene.params.n;
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
    // A pool step in flight: device events around its tick.
    struct Step {
        hipEvent_t e0, e1;
    };
    // A retired request whose state is on its way to pinned host memory.
    struct Pinned {
        float *p = nullptr;
        size_t floats = 0;
    };
    struct Retire {
        uint32_t index, nf;
        Pinned buf;
        hipEvent_t done;
    };
    std::vector<Resident> resident;
    std::deque<Step> steps;
    std::deque<Retire> retiring;
    std::vector<hipEvent_t> spare_events;
    std::vector<Pinned> spare_pinned;
    hipEvent_t last_end = nullptr;  // end event of the previous completed step (interval clock)
    size_t next_admit = 0;          // position in `order`
    uint32_t steps_since_layout = 0;
    uint64_t resident_particles = 0;
    std::vector<double> step_ms, loop_ms, lat, wait;
    std::string error;
    auto ok = [](hipError_t e, const char *what) {
        if (e != hipSuccess) throw DeviceError(e, std::string(what) + ": " + hipGetErrorString(e));
    };
    auto event = [&]() {
        if (!spare_events.empty()) {
            hipEvent_t e = spare_events.back();
            spare_events.pop_back();
            return e;
        }
        hipEvent_t e;
        ok(hipEventCreate(&e), "hipEventCreate");
        return e;
    };
    auto pinned = [&](size_t floats) {
        for (size_t k = 0; k < spare_pinned.size(); ++k)
            if (spare_pinned[k].floats >= floats) {
                Pinned b = spare_pinned[k];
                spare_pinned.erase(spare_pinned.begin() + (long)k);
                return b;
            }
        Pinned b;
        b.floats = floats + floats / 2;
        ok(hipHostMalloc((void **)&b.p, 4 * b.floats, hipHostMallocDefault), "hipHostMalloc");
        return b;
    };
    // Completed steps give their device time and the interval since the
    // previous step's end (the step plus any gap the loop left on the device).
    auto harvest_steps = [&](bool wait_oldest) {
        while (!steps.empty()) {
            Step &q = steps.front();
            if (wait_oldest) {
                ok(hipEventSynchronize(q.e1), "hipEventSynchronize");
                wait_oldest = false;
            } else if (hipEventQuery(q.e1) != hipSuccess) {
                break;
            }
            float ms = 0.0f;
            ok(hipEventElapsedTime(&ms, q.e0, q.e1), "hipEventElapsedTime");
            step_ms.push_back(ms);
            if (last_end) {
                float iv = 0.0f;
                ok(hipEventElapsedTime(&iv, last_end, q.e1), "hipEventElapsedTime");
                loop_ms.push_back(iv);
                spare_events.push_back(last_end);
            }
            last_end = q.e1;
            spare_events.push_back(q.e0);
            steps.pop_front();
        }
    };
    // A worker hashes the state of retired requests off the loop thread (a
    // large request's hash takes milliseconds, which would leave the device
    // idle); the pinned buffer comes back to the loop afterwards.
    std::mutex hmu;
    std::condition_variable hcv;
    std::deque<Retire> to_hash;
    std::vector<Pinned> hashed_bufs;
    uint32_t hashed = 0, handed = 0;  // states hashed / handed to the worker
    bool hstop = false;
    std::thread hasher([&] {
        for (;;) {
            Retire r;
            {
                std::unique_lock<std::mutex> lk(hmu);
                hcv.wait(lk, [&] { return hstop || !to_hash.empty(); });
                if (to_hash.empty()) return;
                r = to_hash.front();
                to_hash.pop_front();
            }
            ServeResult &res = results[r.index];
            res.state_hash = fnv_fluid(r.buf.p, r.buf.p + 3ull * res.n, r.nf);
            res.finish_s = sched_now();
            std::lock_guard<std::mutex> lk(hmu);
            lat.push_back(res.finish_s - res.arrival_s);
            hashed_bufs.push_back(r.buf);
            ++hashed;
        }
    });
    auto harvest_retired = [&](bool all) {
        while (!retiring.empty()) {
            Retire &r = retiring.front();
            if (all)
                ok(hipEventSynchronize(r.done), "hipEventSynchronize");
            else if (hipEventQuery(r.done) != hipSuccess)
                break;
            spare_events.push_back(r.done);
            {
                std::lock_guard<std::mutex> lk(hmu);
                to_hash.push_back(r);
                ++handed;
            }
            hcv.notify_one();
            retiring.pop_front();
        }
        std::lock_guard<std::mutex> lk(hmu);
        st.completed = hashed;
        for (const Pinned &b : hashed_bufs) spare_pinned.push_back(b);
        hashed_bufs.clear();
    };
    auto wait_hashed = [&] {
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(hmu);
                if (retiring.empty() && hashed == handed) {
                    st.completed = hashed;
                    for (const Pinned &b : hashed_bufs) spare_pinned.push_back(b);
                    hashed_bufs.clear();
                    return;
                }
            }
            harvest_retired(true);
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    };
    // Pool growth stays inside the budget; admission checks the engine's
    // projected peak (old and new pool alive together while scenes move,
    // readback staging included), so device bytes never exceed the budget.
    // Every scene reserves cells for its whole grid budget, so no substep is
    // ever abandoned for capacity and readbacks need no check.
    eng.set_memory_limit(cfg.budget_bytes);
    eng.set_full_cell_reserve(true);
    constexpr uint32_t kInFlight = 4;  // pool steps enqueued ahead of the device
    const uint32_t layout_every = std::max<uint32_t>(cfg.admit_every, 1u);
    try {
        while (next_admit < order.size() || !resident.empty()) {
            if (now_s() > cfg.max_wall_s) break;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!build_error.empty()) throw std::runtime_error("scene build failed: " + build_error);
            }
            harvest_steps(false);
            harvest_retired(false);
            // Admit built requests in arrival order while the budget allows,
            // at most once every admit_every pool steps (one layout for all).
            const bool may_admit = resident.empty() || steps_since_layout >= layout_every;
            bool admitted = false;
            while (may_admit) {
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
                const uint32_t n = b.n;
                if (cfg.max_particles && !resident.empty() && resident_particles + n > cfg.max_particles) {
                    // the pool step is as large as allowed: wait for retirements
                    std::lock_guard<std::mutex> lk(mu);
                    ready.emplace((uint32_t)next_admit, std::move(b));
                    break;
                }
                if (resident.empty() && steps.empty() && retiring.empty()) eng.release_if_empty();
                const uint64_t peak = eng.projected_peak(n, b.scene.cells_hint, cfg.max_bins, b.coupled);
                if (peak > cfg.budget_bytes) {
                    if (!resident.empty() || !retiring.empty()) {
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
                const uint32_t sid = eng.add_prepared(std::move(b.scene));
                const SceneSample &rq = reqs[b.index];
                resident.push_back(Resident{b.index, sid, std::max<uint32_t>(rq.steps, 1), b.nf, n});
                resident_particles += n;
                ServeResult &res = results[b.index];
                res.index = b.index;
                res.n = n;
                res.nf = b.nf;
                res.steps = std::max<uint32_t>(rq.steps, 1);
                res.arrival_s = rq.arrival_s;
                res.admit_s = sched_now();
                wait.push_back(res.admit_s - rq.arrival_s);
                ++next_admit;
                admitted = true;
            }
            if (resident.empty()) {
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                continue;
            }
            if (admitted) {
                eng.prepare();  // one layout for every admission and retirement since the last
                steps_since_layout = 0;
            }
            st.peak_resident = std::max<uint32_t>(st.peak_resident, (uint32_t)resident.size());
            // One pool step for every resident scene, enqueued with its events;
            // the host stays at most kInFlight steps ahead of the device.
            if (steps.size() >= kInFlight) harvest_steps(true);
            Step q{event(), event()};
            ok(hipEventRecord(q.e0, eng.stream()), "hipEventRecord");
            eng.step(1);
            ok(hipEventRecord(q.e1, eng.stream()), "hipEventRecord");
            steps.push_back(q);
            ++st.pool_steps;
            ++steps_since_layout;
            // Requests that just had their last step: read back asynchronously
            // and retired in place (no relayout until the next admission).
            for (size_t k = 0; k < resident.size();) {
                Resident &r = resident[k];
                st.fluid_updates += r.nf;
                if (--r.steps_left == 0) {
                    Retire t{r.index, r.nf, pinned(6ull * r.n), event()};
                    eng.read_state_async(r.scene, t.buf.p, t.buf.p + 3ull * r.n);
                    ok(hipEventRecord(t.done, eng.stream()), "hipEventRecord");
                    eng.freeze_scene(r.scene);
                    resident_particles -= r.n;
                    retiring.push_back(t);
                    resident[k] = resident.back();
                    resident.pop_back();
                } else {
                    ++k;
                }
            }
            if (resident.empty()) {
                // Nothing left to step: let the readbacks finish, and do not
                // count the idle time before the next step as an interval.
                ok(hipStreamSynchronize(eng.stream()), "drain");
                harvest_steps(false);
                harvest_retired(true);
                if (last_end) spare_events.push_back(last_end);
                last_end = nullptr;
            }
        }
        ok(hipStreamSynchronize(eng.stream()), "drain");
        harvest_steps(false);
        wait_hashed();
    } catch (const std::exception &e) {
        error = e.what();
    }
    {
        std::lock_guard<std::mutex> lk(hmu);
        hstop = true;
    }
    hcv.notify_all();
    hasher.join();
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto &t : pool) t.join();
    ticker.join();
    (void)hipStreamSynchronize(eng.stream());
    for (auto &q : steps) {
        (void)hipEventDestroy(q.e0);
        (void)hipEventDestroy(q.e1);
    }
    for (auto &r : retiring) {
        (void)hipEventDestroy(r.done);
        (void)hipHostFree(r.buf.p);
    }
    for (auto e : spare_events) (void)hipEventDestroy(e);
    if (last_end) (void)hipEventDestroy(last_end);
    for (auto &b : spare_pinned) (void)hipHostFree(b.p);
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
    {
        std::vector<double> lm(eng.layout_ms().begin(), eng.layout_ms().end());
        st.layout_ms_p50 = pct(lm, 0.50);
        st.layout_ms_p99 = pct(lm, 0.99);
        for (int k = 0; k < 5; ++k) st.layout_phase_ms[k] = eng.layout_phase_ms()[k];
    }
    return st;
}

}  // namespace vxl
