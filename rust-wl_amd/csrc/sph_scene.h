// Host-side scene construction for the WCSPH path: fluid constants
// (FluidSystem::new), Akinci boundary lattices (boundary::lattice) and the
// seeded dam-break scenes the step is benchmarked on.
#ifndef VXL_SPH_SCENE_H
#define VXL_SPH_SCENE_H
#include <cstdint>
#include <vector>

#include "sph_layout.h"

namespace vxl {

// FluidConfig (crates/vxl-phys-fluid/src/config.rs:6-69) plus what the
// restated scenes need: an explicit particle mass, the lattice spacing, the
// grid budget (crates/vxl-phys-core/src/grid.rs:20) and the tick length.
struct SceneConfig {
    float rest_density = 1000.0f;
    float smoothing_radius = 0.1f;
    float xsph_viscosity = 0.1f;
    float sound_speed = 10.0f;
    float gamma_tait = 7.0f;
    uint32_t substeps = 4;
    uint32_t tensile_clamp = 1;
    uint32_t boundary_layers = 2;
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    float max_speed_frac = 0.4f;
    float artificial_viscosity = 1.0f;
    float particle_mass = 0.0f;  // 0 = lattice calibration of FluidSystem::new
    float spacing = 0.05f;
    uint64_t max_bins = 1ull << 21;
    float dt_tick = 1.0f / 60.0f;
};

// Boundary span (body, origin, [start, end)) of FluidSystem::spans.
struct Span {
    uint32_t body;
    float origin[3];
    uint32_t start, end;
};

struct FluidConstants {
    float h, h2, k6, ks, w0, b_tait, mass, alpha_c;
};

// fluid_access.rs:8-42: kernel constants and, when cfg.particle_mass == 0,
// the lattice mass calibration m = rho0 / sum_lattice W.
FluidConstants fluid_constants(const SceneConfig &cfg);

// Device parameter block of a scene (slot range filled in by the layout).
SceneParams scene_params(const SceneConfig &cfg, uint32_t n, uint32_t nf);

// boundary.rs:67-88 for a Box: surface lattice, layer offsets and the
// self-consistent volume V_b = 1 / mean(sum_b W). Returns the point count.
uint32_t box_lattice(const float half[3], float s, float h, uint32_t layers, std::vector<float> &pts,
                     uint32_t &n0, float &volume);

// boundary.rs:106-129 evaluated through a uniform grid with neighbour lists
// sorted by index (same summation order as the O(n^2) loop).
float boundary_self_sum(const float *pts, uint32_t n, float h);

// A rigid Box body sampled into Akinci particles (fluid_step.rs:13-53).
struct BoxBody {
    uint32_t id;
    float half[3];
    float pos[3];
    float rot[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float linvel[3] = {0.0f, 0.0f, 0.0f};
    float angvel[3] = {0.0f, 0.0f, 0.0f};
};

// mass_props for Shape::Box (crates/vxl-phys-core/src/mass.rs:35-48):
// m = density ex ey ez, I = m/12 (e_j^2 + e_k^2); density <= 0 falls back to 1.
void box_mass_props(const float half[3], float density, float &inv_mass, float inv_inertia[3]);

// Append the boundary particles of `bodies` after the fluid prefix
// (body order x layer x face order), pmass = rho0 * V_b, spans per body.
void append_boundary_boxes(const SceneConfig &cfg, const std::vector<BoxBody> &bodies,
                           std::vector<float> &pos, std::vector<float> &vel, std::vector<float> &pmass,
                           std::vector<Span> &spans);

// Seeded dam break: fluid block of bx*by*bz particles on a lattice of spacing
// d in the corner of an open tank [0,L]^3, jitter U(-j*d, j*d) per coordinate,
// tank = 5 Box bodies (floor + 4 walls) of thickness 2d sampled with
// `cfg.boundary_layers` Akinci layers.
struct DamBreakSpec {
    float tank[3] = {2.0f, 1.0f, 1.0f};
    uint32_t block[3] = {32, 32, 32};
    float d = 0.02f;
    float jitter = 0.05f;
    uint64_t seed = 1;
    bool walls = true;
};
void dam_break(const DamBreakSpec &spec, const SceneConfig &cfg, std::vector<float> &pos,
               std::vector<float> &vel, std::vector<float> &pmass, std::vector<Span> &spans,
               uint32_t &nf);
std::vector<BoxBody> tank_bodies(const float tank[3], float d);

// A sampled dam-break scene: geometry, physics and (served form) its request.
struct SceneSample {
    DamBreakSpec dam;
    float sound_speed = 20.0f;
    float alpha = 0.02f;
    uint32_t steps = 1000;
    double arrival_s = 0.0;
    uint32_t nboxes = 0;
    BoxBody boxes[4];
    float box_density[4] = {0, 0, 0, 0};
    double fill = 0.0;  // realised fluid volume / tank volume
};

// Batched form: scene `index` of a seeded set. Fluid count log-uniform in
// [nmin, nmax], tank x, z uniform in [1, 4] m with y = 1 m, fill fraction
// uniform in [0.1, 0.5] (block height = fill * 1 m, footprint from the count
// with the tank's aspect, clamped to the tank), c0 uniform in [15, 30] m/s,
// alpha uniform in [0.01, 0.05], d = 0.02 m.
SceneSample sample_batch_scene(uint64_t seed, uint32_t index, uint32_t nmin = 8192, uint32_t nmax = 524288);

// Served form: request `index` of a seeded Poisson stream of `rate` requests/s.
// Fluid count log-uniform in [nmin, nmax], steps uniform in [smin, smax],
// 0..max_boxes dynamic boxes of side 0.1-0.3 m and density 500-2000 kg/m^3
// dropped above the fluid. Arrival times are cumulative exponential gaps.
std::vector<SceneSample> sample_requests(uint64_t seed, uint32_t count, double rate, uint32_t nmin = 4096,
                                         uint32_t nmax = 2097152, uint32_t smin = 100, uint32_t smax = 2000,
                                         uint32_t max_boxes = 4);

// splitmix64 stream used by every seeded scene.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    // uniform f32 in [0, 1) with 24 random bits (exact)
    float unit() { return (float)(next() >> 40) * 5.9604644775390625e-08f; }
};

// Grid box rule (crates/vxl-phys-core/src/grid.rs:34-51) on the host.
void grid_box(const float lo[3], const float hi[3], float h, uint64_t max_bins, float out_min[3],
              float &out_bin, uint32_t out_dims[3]);

}  // namespace vxl
#endif
