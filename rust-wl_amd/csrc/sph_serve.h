// Continuous-batching server for dam-break step requests on one device.
//
// Requests arrive over time (a recorded arrival schedule, optionally time
// compressed). Their scenes are built by host worker threads once they have
// arrived, admitted in arrival order as soon as the pool's HBM budget allows,
// and co-stepped with every other resident scene in one launch per stage.
// Finished scenes are retired (state hash recorded) at step boundaries;
// admissions and retirements between two steps share one pool relayout.
#ifndef VXL_SPH_SERVE_H
#define VXL_SPH_SERVE_H
#include <cstdint>
#include <vector>

#include "sph_engine.h"

namespace vxl {

struct ServeConfig {
    uint64_t budget_bytes = 200ull << 30;  // device bytes the resident scenes may use
    double time_scale = 1.0;               // arrival schedule speed-up (1 = real time)
    double max_wall_s = 3600.0;            // stop admitting/stepping after this
    uint32_t max_resident = 0;             // 0 = limited by the budget only
    uint64_t max_particles = 0;            // resident particles (fluid + boundary) at most; 0 = no limit
    uint32_t admit_every = 128;            // pool steps between admissions (one layout for all of them)
    uint32_t builders = 4;                 // host threads building arrived scenes
    uint64_t max_bins = 1ull << 21;        // grid budget of every scene
    float dt = 5e-4f;
    float d = 0.02f;
};

struct ServeResult {
    uint32_t index;
    uint32_t n, nf, steps;
    double arrival_s;   // schedule time (s)
    double admit_s;     // wall time since start, scaled to schedule time
    double finish_s;
    uint64_t state_hash;
};

struct ServeStats {
    uint32_t requests = 0, completed = 0;
    uint64_t pool_steps = 0;
    uint64_t fluid_updates = 0;
    double wall_s = 0.0;
    double updates_per_s = 0.0;
    double step_ms_p50 = 0.0, step_ms_p99 = 0.0, step_ms_mean = 0.0;   // device time of a pool step
    double loop_ms_p50 = 0.0, loop_ms_p99 = 0.0;                        // wall time per step incl. admission
    double latency_s_p50 = 0.0, latency_s_p99 = 0.0;                    // arrival -> finish (schedule s)
    double wait_s_p50 = 0.0, wait_s_p99 = 0.0;                          // arrival -> admission
    uint32_t peak_resident = 0;
    uint32_t rejected = 0;    // requests larger than the budget on their own (not served)
    uint64_t peak_bytes = 0;  // highest device bytes the engine held
    uint64_t layouts = 0;
    double layout_ms_p50 = 0.0, layout_ms_p99 = 0.0;  // host time of a layout (admissions)
    double layout_phase_ms[5] = {0, 0, 0, 0, 0};      // summed: wait, move, upload, tables, graph
};

// Builds the scene of one request: the dam-break fluid and, with its tank,
// all particles (pos/vel/pmass/spans). A request with dynamic boxes is a
// coupled scene: `bodies`/`density` then hold the tank's static Box bodies
// followed by the boxes (density > 0), and `rc` its rigid configuration; the
// particle arrays still hold every particle (for the footprint and the grid
// hint), the fluid prefix being the first nf.
void build_request(const SceneSample &req, const ServeConfig &cfg, SceneConfig &sc, std::vector<float> &pos,
                   std::vector<float> &vel, std::vector<float> &pmass, std::vector<Span> &spans, uint32_t &nf,
                   std::vector<BoxBody> *bodies = nullptr, std::vector<float> *density = nullptr,
                   RigidConfig *rc = nullptr);
// The Box bodies of a request with dynamic boxes (tank, then boxes) and its
// rigid configuration (PhysConfig defaults, contacts with the request's tank).
void request_bodies(const SceneSample &req, std::vector<BoxBody> &bodies, std::vector<float> &density,
                    RigidConfig &rc);

// Grid cells of the initial box of a scene's positions (sizes its cell tables).
uint32_t scene_cells_hint(const std::vector<float> &pos, size_t n, float h, uint64_t max_bins);
// Peak device bytes of serving this request alone (Engine::solo_peak of its scene).
uint64_t request_solo_bytes(const SceneSample &req, const ServeConfig &cfg);

ServeStats serve(Engine &eng, const std::vector<SceneSample> &reqs, const ServeConfig &cfg,
                 std::vector<ServeResult> &results);

}  // namespace vxl
#endif
