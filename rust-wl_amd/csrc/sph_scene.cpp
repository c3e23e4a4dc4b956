// Host-side scene construction (strict f32, built with -ffp-contract=off).
#include "sph_scene.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace vxl {

namespace {

constexpr float kPi = 3.14159274101257324f;  // std::f32::consts::PI

// h.powi(n) as the reference binaries evaluate it: with the smoothing
// radius a compile-time constant, LLVM folds llvm.powi.f32 through pow() in
// double precision and rounds once to f32.
float powi(float x, int n) { return (float)std::pow((double)x, (double)n); }

float poly6_k6(float h) { return 315.0f / ((64.0f * kPi) * powi(h, 9)); }

// Rust `f as usize` (saturating).
uint64_t sat_usize(float f) {
    if (!(f > 0.0f)) return 0;
    if (f >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)f;
}

uint32_t axis_bin(float o, float v, float inv, uint32_t n) {
    float f = fmaxf(floorf((v - o) * inv), 0.0f);
    uint32_t lim = n - 1u;
    return f >= (float)lim ? lim : (uint32_t)f;
}

// boundary.rs:192-194
uint32_t lattice_count(float len, float s) {
    long long c = (long long)roundf(len / s);
    return (uint32_t)std::max<long long>(c, 1);
}

}  // namespace

void grid_box(const float lo[3], const float hi[3], float h, uint64_t max_bins, float out_min[3],
              float &out_bin, uint32_t out_dims[3]) {
    float bin = fmaxf(h, 1e-6f);
    const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    uint32_t dims[3] = {1, 1, 1};
    for (int it = 0; it < 256; ++it) {
        for (int a = 0; a < 3; ++a) {
            uint64_t d = sat_usize(floorf(ext[a] / bin)) + 1;
            dims[a] = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(d, 1), kGridDimMax);
        }
        if ((uint64_t)dims[0] * dims[1] * dims[2] <= max_bins) break;
        bin *= 2.0f;
    }
    for (int a = 0; a < 3; ++a) {
        out_min[a] = lo[a];
        out_dims[a] = dims[a];
    }
    out_bin = bin;
}

FluidConstants fluid_constants(const SceneConfig &cfg) {
    FluidConstants c{};
    const float h = fmaxf(cfg.smoothing_radius, 1e-5f);
    c.h = h;
    c.h2 = h * h;
    c.k6 = poly6_k6(h);
    c.ks = 45.0f / (kPi * powi(h, 6));
    c.w0 = ((c.k6 * c.h2) * c.h2) * c.h2;
    c.b_tait = ((cfg.sound_speed * cfg.sound_speed) * cfg.rest_density) / cfg.gamma_tait;
    c.alpha_c = cfg.artificial_viscosity * cfg.sound_speed;
    if (cfg.particle_mass > 0.0f) {
        c.mass = cfg.particle_mass;
    } else {
        // Lattice calibration: sum W over the spacing lattice within h.
        const float sp = cfg.spacing;
        float wsum = 0.0f;
        const int side = (int)ceilf(h / fmaxf(sp, 1e-6f));
        for (int i = -side; i <= side; ++i)
            for (int j = -side; j <= side; ++j)
                for (int k = -side; k <= side; ++k) {
                    const float dx = (float)i * sp, dy = (float)j * sp, dz = (float)k * sp;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 <= c.h2) {
                        const float t = c.h2 - r2;
                        wsum += ((c.k6 * t) * t) * t;
                    }
                }
        c.mass = cfg.rest_density / wsum;
    }
    return c;
}

SceneParams scene_params(const SceneConfig &cfg, uint32_t n, uint32_t nf) {
    const FluidConstants c = fluid_constants(cfg);
    SceneParams p{};
    p.n = n;
    p.nf = nf;
    p.max_bins = std::min<uint64_t>(cfg.max_bins ? cfg.max_bins : (1ull << 21), 1ull << 30);
    p.h = c.h;
    p.h2 = c.h2;
    p.k6 = c.k6;
    p.ks = c.ks;
    p.w0 = c.w0;
    p.mass = c.mass;
    p.rho0 = cfg.rest_density;
    p.b_tait = c.b_tait;
    p.alpha_c = c.alpha_c;
    p.eps = cfg.xsph_viscosity;
    p.gx = cfg.gravity[0];
    p.gy = cfg.gravity[1];
    p.gz = cfg.gravity[2];
    const uint32_t sub = std::max<uint32_t>(cfg.substeps, 1);
    p.dt = cfg.dt_tick / (float)sub;                 // fluid_step.rs:76
    p.vmax = (cfg.max_speed_frac * c.h) / p.dt;      // fluid_step.rs:110
    p.vmax2 = p.vmax * p.vmax;
    p.clamp_neg = cfg.tensile_clamp ? 1u : 0u;
    p.substeps = sub;
    return p;
}

float boundary_self_sum(const float *pts, uint32_t n, float h) {
    if (n == 0) return 0.0f;
    const float h2 = h * h;
    const float k6 = poly6_k6(h);
    const float w0 = ((k6 * h2) * h2) * h2;
    // Uniform grid (bin = h) over the points; cells hold ascending indices.
    float lo[3] = {pts[0], pts[1], pts[2]}, hi[3] = {pts[0], pts[1], pts[2]};
    for (uint32_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], pts[3 * i + a]);
            hi[a] = fmaxf(hi[a], pts[3 * i + a]);
        }
    float gmin[3], bin;
    uint32_t dims[3];
    grid_box(lo, hi, h, 1ull << 26, gmin, bin, dims);
    const float inv = 1.0f / bin;
    const uint64_t total = (uint64_t)dims[0] * dims[1] * dims[2];
    std::vector<uint32_t> start(total + 1, 0), items(n), cell(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = (axis_bin(gmin[0], pts[3 * i], inv, dims[0]) * dims[1] +
                            axis_bin(gmin[1], pts[3 * i + 1], inv, dims[1])) * dims[2] +
                           axis_bin(gmin[2], pts[3 * i + 2], inv, dims[2]);
        cell[i] = c;
        start[c + 1] += 1;
    }
    for (uint64_t c = 0; c < total; ++c) start[c + 1] += start[c];
    {
        std::vector<uint32_t> cur(start.begin(), start.end() - 1);
        for (uint32_t i = 0; i < n; ++i) items[cur[cell[i]]++] = i;
    }
    // Per-particle sums (index-ascending neighbour order), then the serial mean.
    std::vector<float> per(n);
    auto work = [&](uint32_t a, uint32_t b) {
        std::vector<std::pair<uint32_t, float>> nb;
        for (uint32_t i = a; i < b; ++i) {
            nb.clear();
            const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
            const int cx = (int)axis_bin(gmin[0], px, inv, dims[0]);
            const int cy = (int)axis_bin(gmin[1], py, inv, dims[1]);
            const int cz = (int)axis_bin(gmin[2], pz, inv, dims[2]);
            for (int x = cx - 1; x <= cx + 1; ++x) {
                if (x < 0 || x >= (int)dims[0]) continue;
                for (int y = cy - 1; y <= cy + 1; ++y) {
                    if (y < 0 || y >= (int)dims[1]) continue;
                    for (int z = cz - 1; z <= cz + 1; ++z) {
                        if (z < 0 || z >= (int)dims[2]) continue;
                        const uint64_t c = ((uint64_t)x * dims[1] + (uint64_t)y) * dims[2] + (uint64_t)z;
                        for (uint32_t k = start[c]; k < start[c + 1]; ++k) {
                            const uint32_t j = items[k];
                            if (j == i) continue;
                            const float dx = px - pts[3 * j], dy = py - pts[3 * j + 1], dz = pz - pts[3 * j + 2];
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            if (r2 <= h2) nb.push_back({j, r2});
                        }
                    }
                }
            }
            std::sort(nb.begin(), nb.end(), [](const auto &u, const auto &v) { return u.first < v.first; });
            float sum = w0;
            for (const auto &e : nb) {
                const float t = h2 - e.second;
                sum += ((k6 * t) * t) * t;
            }
            per[i] = sum;
        }
    };
    const uint32_t nt = n < 65536 ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nt == 1) {
        work(0, n);
    } else {
        std::vector<std::thread> ts;
        const uint32_t chunk = (n + nt - 1) / nt;
        for (uint32_t a = 0; a < n; a += chunk) ts.emplace_back(work, a, std::min(n, a + chunk));
        for (auto &t : ts) t.join();
    }
    float acc = 0.0f;
    for (uint32_t i = 0; i < n; ++i) acc += per[i];
    return acc / (float)n;
}

uint32_t box_lattice(const float half[3], float s_in, float h, uint32_t layers_in, std::vector<float> &pts,
                     uint32_t &n0, float &volume) {
    const float s = fmaxf(s_in, 1e-6f);
    const uint32_t layers = std::max<uint32_t>(layers_in, 1);
    // box_faces (boundary.rs:208-232): 3 axes x (+, -) faces x cell-centred grid.
    std::vector<float> sp, sn;
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        const float hb = half[b], hc = half[c];
        const uint32_t nb = lattice_count(2.0f * hb, s), nc = lattice_count(2.0f * hc, s);
        const float db = 2.0f * hb / (float)nb, dc = 2.0f * hc / (float)nc;
        for (float sign : {1.0f, -1.0f})
            for (uint32_t i = 0; i < nb; ++i)
                for (uint32_t j = 0; j < nc; ++j) {
                    float p[3] = {0, 0, 0}, q[3] = {0, 0, 0};
                    p[a] = sign * half[a];
                    q[a] = sign;
                    p[b] = -hb + ((float)i + 0.5f) * db;
                    p[c] = -hc + ((float)j + 0.5f) * dc;
                    sp.insert(sp.end(), p, p + 3);
                    sn.insert(sn.end(), q, q + 3);
                }
    }
    n0 = (uint32_t)(sp.size() / 3);
    const float hm = fminf(fminf(half[0], half[1]), half[2]);
    pts.clear();
    pts.reserve((size_t)n0 * layers * 3);
    for (uint32_t k = 0; k < layers; ++k) {
        const float want = ((float)k + 0.5f) * s;
        const float cap = (0.45f + 0.45f * (float)k) * hm;
        const float d = fmaxf(fminf(fminf(want, cap), 0.95f * hm), 1e-6f);
        for (uint32_t q = 0; q < n0; ++q)
            for (int a = 0; a < 3; ++a) pts.push_back(sp[3 * q + a] - sn[3 * q + a] * d);
    }
    const uint32_t n = (uint32_t)(pts.size() / 3);
    const float sum = boundary_self_sum(pts.data(), n, fmaxf(h, s));
    volume = sum > 1e-12f ? 1.0f / sum : 1.0f;
    return n;
}

void box_mass_props(const float half[3], float density, float &inv_mass, float inv_inertia[3]) {
    const float rho = density > 0.0f ? density : 1.0f;
    const float ex = 2.0f * half[0], ey = 2.0f * half[1], ez = 2.0f * half[2];
    const float m = rho * ex * ey * ez;
    const float ix = m / 12.0f * (ey * ey + ez * ez);
    const float iy = m / 12.0f * (ex * ex + ez * ez);
    const float iz = m / 12.0f * (ex * ex + ey * ey);
    inv_mass = 1.0f / m;
    inv_inertia[0] = 1.0f / ix;
    inv_inertia[1] = 1.0f / iy;
    inv_inertia[2] = 1.0f / iz;
}

void append_boundary_boxes(const SceneConfig &cfg, const std::vector<BoxBody> &bodies,
                           std::vector<float> &pos, std::vector<float> &vel, std::vector<float> &pmass,
                           std::vector<Span> &spans) {
    const float h = fmaxf(cfg.smoothing_radius, 1e-5f);
    for (const BoxBody &bd : bodies) {
        std::vector<float> lp;
        uint32_t n0 = 0;
        float vol = 0.0f;
        const uint32_t cnt = box_lattice(bd.half, cfg.spacing, h, cfg.boundary_layers, lp, n0, vol);
        if (cnt == 0) continue;
        const uint32_t start = (uint32_t)(pos.size() / 3);
        const float *q = bd.rot;
        for (uint32_t k = 0; k < cnt; ++k) {
            const float v[3] = {lp[3 * k], lp[3 * k + 1], lp[3 * k + 2]};
            // Quat::rotate_vec3 (math.rs:209-213): t = (qv x v)*2; v + qv x t + t*w.
            const float t[3] = {(q[1] * v[2] - q[2] * v[1]) * 2.0f, (q[2] * v[0] - q[0] * v[2]) * 2.0f,
                                (q[0] * v[1] - q[1] * v[0]) * 2.0f};
            const float c[3] = {q[1] * t[2] - q[2] * t[1], q[2] * t[0] - q[0] * t[2], q[0] * t[1] - q[1] * t[0]};
            float wp[3];
            for (int a = 0; a < 3; ++a) wp[a] = bd.pos[a] + ((v[a] + c[a]) + t[a] * q[3]);
            // v_lin + w x (wp - pos)
            const float r[3] = {wp[0] - bd.pos[0], wp[1] - bd.pos[1], wp[2] - bd.pos[2]};
            const float *w = bd.angvel;
            const float wv[3] = {bd.linvel[0] + (w[1] * r[2] - w[2] * r[1]),
                                 bd.linvel[1] + (w[2] * r[0] - w[0] * r[2]),
                                 bd.linvel[2] + (w[0] * r[1] - w[1] * r[0])};
            pos.insert(pos.end(), wp, wp + 3);
            vel.insert(vel.end(), wv, wv + 3);
            pmass.push_back(cfg.rest_density * vol);
        }
        const uint32_t end = (uint32_t)(pos.size() / 3);
        if (end > start) spans.push_back(Span{bd.id, {bd.pos[0], bd.pos[1], bd.pos[2]}, start, end});
    }
}

std::vector<BoxBody> tank_bodies(const float L[3], float d) {
    const float t = 2.0f * d;
    const float hx = L[0] * 0.5f, hy = L[1] * 0.5f, hz = L[2] * 0.5f;
    std::vector<BoxBody> b(5);
    auto set = [](BoxBody &o, uint32_t id, float h0, float h1, float h2, float p0, float p1, float p2) {
        o.id = id;
        o.half[0] = h0; o.half[1] = h1; o.half[2] = h2;
        o.pos[0] = p0; o.pos[1] = p1; o.pos[2] = p2;
    };
    set(b[0], 0, hx + t, t * 0.5f, hz + t, hx, -(t * 0.5f), hz);          // floor
    set(b[1], 1, t * 0.5f, hy, hz + t, -(t * 0.5f), hy, hz);              // wall x = 0
    set(b[2], 2, t * 0.5f, hy, hz + t, L[0] + t * 0.5f, hy, hz);          // wall x = Lx
    set(b[3], 3, hx + t, hy, t * 0.5f, hx, hy, -(t * 0.5f));              // wall z = 0
    set(b[4], 4, hx + t, hy, t * 0.5f, hx, hy, L[2] + t * 0.5f);          // wall z = Lz
    return b;
}

void dam_break(const DamBreakSpec &spec, const SceneConfig &cfg, std::vector<float> &pos, std::vector<float> &vel,
               std::vector<float> &pmass, std::vector<Span> &spans, uint32_t &nf) {
    const FluidConstants c = fluid_constants(cfg);
    const uint32_t bx = spec.block[0], by = spec.block[1], bz = spec.block[2];
    nf = bx * by * bz;
    pos.clear();
    vel.clear();
    pmass.clear();
    spans.clear();
    pos.reserve((size_t)nf * 3);
    SplitMix64 rng(spec.seed);
    const float d = spec.d;
    const float amp = spec.jitter * d;
    for (uint32_t i = 0; i < bx; ++i)
        for (uint32_t j = 0; j < by; ++j)
            for (uint32_t k = 0; k < bz; ++k) {
                const float base[3] = {((float)i + 0.5f) * d, ((float)j + 0.5f) * d, ((float)k + 0.5f) * d};
                for (int a = 0; a < 3; ++a) pos.push_back(base[a] + (2.0f * rng.unit() - 1.0f) * amp);
            }
    vel.assign((size_t)nf * 3, 0.0f);
    pmass.assign(nf, c.mass);
    if (spec.walls) append_boundary_boxes(cfg, tank_bodies(spec.tank, d), pos, vel, pmass, spans);
}


namespace {

// Block dimensions for ~n particles of spacing d with height `height` in a
// tank footprint lx x lz (aspect kept, clamped to the footprint).
void block_dims(double n, double lx, double lz, double height, double d, uint32_t out[3]) {
    const uint32_t ny = (uint32_t)std::max(1.0, std::round(height / d));
    const double nxz = std::max(1.0, n / ny);
    const uint32_t cx = (uint32_t)std::max(1.0, std::floor(lx / d)), cz = (uint32_t)std::max(1.0, std::floor(lz / d));
    uint32_t nx = (uint32_t)std::max(1.0, std::round(std::sqrt(nxz * lx / lz)));
    nx = std::min(nx, cx);
    uint32_t nz = (uint32_t)std::max(1.0, std::round(nxz / nx));
    nz = std::min(nz, cz);
    out[0] = nx;
    out[1] = ny;
    out[2] = nz;
}

double uniform(SplitMix64 &r, double a, double b) { return a + (b - a) * ((double)(r.next() >> 11) * 0x1.0p-53); }

}  // namespace

SceneSample sample_batch_scene(uint64_t seed, uint32_t index, uint32_t nmin, uint32_t nmax) {
    SplitMix64 r(seed * 0x9e3779b97f4a7c15ull + (uint64_t)index * 0xd1b54a32d192ed03ull + 1);
    SceneSample s;
    const double n = std::exp(uniform(r, std::log((double)nmin), std::log((double)nmax)));
    const float lx = (float)uniform(r, 1.0, 4.0), lz = (float)uniform(r, 1.0, 4.0), ly = 1.0f;
    const double fill = uniform(r, 0.1, 0.5);
    s.sound_speed = (float)uniform(r, 15.0, 30.0);
    s.alpha = (float)uniform(r, 0.01, 0.05);
    s.dam.tank[0] = lx;
    s.dam.tank[1] = ly;
    s.dam.tank[2] = lz;
    s.dam.d = 0.02f;
    s.dam.jitter = 0.05f;
    s.dam.seed = r.next();
    s.dam.walls = true;
    block_dims(n, lx, lz, fill * ly, s.dam.d, s.dam.block);
    s.fill = (double)s.dam.block[0] * s.dam.block[1] * s.dam.block[2] * 0.02 * 0.02 * 0.02 / ((double)lx * ly * lz);
    return s;
}

std::vector<SceneSample> sample_requests(uint64_t seed, uint32_t count, double rate, uint32_t nmin, uint32_t nmax,
                                         uint32_t smin, uint32_t smax, uint32_t max_boxes) {
    std::vector<SceneSample> out;
    SplitMix64 clock(seed ^ 0x5851f42d4c957f2dull);
    double t = 0.0;
    for (uint32_t k = 0; k < count; ++k) {
        t += -std::log(1.0 - uniform(clock, 0.0, 1.0)) / rate;
        SceneSample s = sample_batch_scene(seed ^ 0xa0761d6478bd642full, k, nmin, nmax);
        SplitMix64 r(seed * 0xe7037ed1a0b428dbull + (uint64_t)k * 0x8ebc6af09c88c6e3ull + 7);
        s.arrival_s = t;
        s.steps = (uint32_t)std::min<double>(smax, std::floor(uniform(r, smin, (double)smax + 1.0)));
        s.nboxes = max_boxes ? (uint32_t)std::min<double>(max_boxes, std::floor(uniform(r, 0.0, max_boxes + 1.0))) : 0u;
        // Boxes dropped above the fluid surface inside the tank.
        const float top = (float)s.dam.block[1] * s.dam.d;
        for (uint32_t b = 0; b < s.nboxes; ++b) {
            BoxBody &bx = s.boxes[b];
            const float side = (float)uniform(r, 0.1, 0.3);
            bx.id = 100u + b;
            bx.half[0] = bx.half[1] = bx.half[2] = 0.5f * side;
            bx.pos[0] = (float)uniform(r, 0.3, s.dam.tank[0] - 0.3);
            bx.pos[1] = top + 0.2f + side * (float)b;
            bx.pos[2] = (float)uniform(r, 0.3, s.dam.tank[2] - 0.3);
            s.box_density[b] = (float)uniform(r, 500.0, 2000.0);
        }
        out.push_back(s);
    }
    return out;
}

}  // namespace vxl
