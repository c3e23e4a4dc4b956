// WCSPH oracle — TEST INFRASTRUCTURE ONLY (see wcsph_oracle.h).
//
// Strict-f32 restatement of crates/vxl-phys-fluid (the reference CPU WCSPH).
// Build: g++ -O2 -ffp-contract=off -fno-fast-math (SSE f32, no x87).
// Every expression keeps the reference's left-to-right operation order; the
// comments cite the reference lines each block follows.
#include "wcsph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr float kPi = 3.14159274101257324f;  // std::f32::consts::PI
constexpr uint32_t kGridDimMax = 1u << 14;     // core/grid.rs:8

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
// math.rs:54-56: x*x' + y*y' + z*z' evaluated left to right.
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len2(V3 a) { return dot(a, a); }
// math.rs:59-65
inline V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Rust f32::min / f32::max (NaN-ignoring), as fminf/fmaxf.
inline V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

// Quat::rotate_vec3 (math.rs:209-213): t = (qv x v)*2; v + qv x t + t*w.
inline V3 rotate(const float q[4], V3 v) {
    V3 qv = v3(q[0], q[1], q[2]);
    V3 t = mul(cross(qv, v), 2.0f);
    return add(add(v, cross(qv, t)), mul(t, q[3]));
}

// h.powi(n) as the reference binaries evaluate it: the smoothing radius is a
// compile-time constant there, and LLVM folds llvm.powi.f32 through the host
// pow() in double precision, rounding once to f32. (At run time powi would
// be binary exponentiation in f32, which differs by an ulp for h = 0.1.)
// Pinned by the reference's recorded Akinci canary readout, y_min = -1.81726
// (docs/PLAN-gpu.md:2704), which only the folded constants reproduce.
inline float powi_rs(float x, int n) { return (float)std::pow((double)x, (double)n); }

// `f as u32` in Rust saturates: NaN -> 0, negative -> 0, >= 2^32 -> MAX.
inline uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
inline uint64_t sat_usize(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)f;
}

}  // namespace

// core/grid.rs:34-51
extern "C" void or_grid_box(const float lo[3], const float hi[3], float h, uint64_t max_bins,
                            float out_min[3], float *out_bin, uint32_t out_dims[3]) {
    float bin = fmaxf(h, 1e-6f);
    float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    uint32_t dims[3] = {1, 1, 1};
    for (;;) {
        for (int a = 0; a < 3; ++a) {
            uint64_t d = sat_usize(floorf(ext[a] / bin)) + 1;
            d = std::min<uint64_t>(std::max<uint64_t>(d, 1), kGridDimMax);
            dims[a] = (uint32_t)d;
        }
        if ((uint64_t)dims[0] * (uint64_t)dims[1] * (uint64_t)dims[2] <= max_bins) break;
        bin *= 2.0f;
    }
    out_min[0] = lo[0];
    out_min[1] = lo[1];
    out_min[2] = lo[2];
    *out_bin = bin;
    out_dims[0] = dims[0];
    out_dims[1] = dims[1];
    out_dims[2] = dims[2];
}

namespace {

// fluid/src/grid.rs:13-27
struct Grid {
    float bin = 0, inv = 0;
    V3 min{0, 0, 0};
    uint32_t nx = 0, ny = 0, nz = 0;
    std::vector<uint32_t> items, start, counts;

    // grid.rs:91-101 (per-axis clamp); also used by bin_index_of (:80-89).
    inline uint32_t axis(float o, float v, uint32_t n) const {
        uint32_t u = sat_u32(fmaxf(floorf((v - o) * inv), 0.0f));
        return std::min(u, n - 1);
    }
    inline uint32_t key(V3 p, uint64_t total) const {
        uint32_t idx = (axis(min.x, p.x, nx) * ny + axis(min.y, p.y, ny)) * nz + axis(min.z, p.z, nz);
        uint64_t lim = total ? total - 1 : 0;
        return (uint32_t)std::min<uint64_t>(idx, lim);
    }
    uint64_t total() const { return (uint64_t)nx * ny * nz; }

    // grid.rs:29-76
    void rebuild(const std::vector<V3> &pos, float h, uint64_t max_bins) {
        size_t n = pos.size();
        if (n == 0) {
            nx = ny = nz = 0;
            items.clear();
            start.assign(1, 0);
            return;
        }
        V3 lo = pos[0], hi = pos[0];
        for (size_t k = 1; k < n; ++k) {
            lo = vmin(lo, pos[k]);
            hi = vmax(hi, pos[k]);
        }
        float l3[3] = {lo.x, lo.y, lo.z}, h3[3] = {hi.x, hi.y, hi.z}, m3[3];
        uint32_t d3[3];
        or_grid_box(l3, h3, h, max_bins, m3, &bin, d3);
        inv = 1.0f / bin;
        min = v3(m3[0], m3[1], m3[2]);
        nx = d3[0];
        ny = d3[1];
        nz = d3[2];
        uint64_t tot = total();
        counts.assign(tot + 1, 0);
        for (size_t k = 0; k < n; ++k) counts[key(pos[k], tot) + 1] += 1;
        for (uint64_t c = 0; c < tot; ++c) counts[c + 1] += counts[c];
        start = counts;
        items.assign(n, 0);
        for (size_t k = 0; k < n; ++k) {
            uint32_t c = key(pos[k], tot);
            items[counts[c]] = (uint32_t)k;
            counts[c] += 1;
        }
    }

    // grid.rs:115-157: cells dz, dy, dx (x fastest in the loop nest), items in
    // ascending index order; j == i skipped; accept r2 <= h2.
    int fault = 0;  // test-only planted mistakes (see or_set_fault)
    template <class F>
    inline void for_neighbors(const std::vector<V3> &pos, float h2, size_t i, F &&f) const {
        V3 pi = pos[i];
        uint32_t cx = axis(min.x, pi.x, nx), cy = axis(min.y, pi.y, ny), cz = axis(min.z, pi.z, nz);
        for (int o0 = -1; o0 <= 1; ++o0) {
            const int dz = fault == 3 ? 0 : o0;
            const int dxo = fault == 3 ? o0 : 0;
            (void)dxo;
            int z = (int)cz + (fault == 3 ? 0 : dz);
            if (fault != 3 && (z < 0 || z >= (int)nz)) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int y = (int)cy + dy;
                if (y < 0 || y >= (int)ny) continue;
                for (int o2 = -1; o2 <= 1; ++o2) {
                    // fault 3: dx outer / dz inner (transposed enumeration)
                    const int dx = fault == 3 ? o0 : o2;
                    if (fault == 3) z = (int)cz + o2;
                    if (fault == 3 && (z < 0 || z >= (int)nz)) continue;
                    int x = (int)cx + dx;
                    if (x < 0 || x >= (int)nx) continue;
                    uint32_t idx = ((uint32_t)x * ny + (uint32_t)y) * nz + (uint32_t)z;
                    for (uint32_t k = start[idx]; k < start[idx + 1]; ++k) {
                        uint32_t j = items[k];
                        if (j == i) continue;
                        V3 d = sub(pi, pos[j]);
                        float r2 = len2(d);
                        if (r2 <= h2) f((size_t)j, d, r2);
                    }
                }
            }
        }
    }
};

}  // namespace

struct OrSystem {
    OrConfig cfg;
    int fault = 0;  // planted mistake for tests of the pins (0 = none)
    float h, h2, k6, ks, w0, b_tait, mass, spacing;
    std::vector<V3> pos, vel, acc, xsph, bforce;
    std::vector<float> dens, press, pmass;
    std::vector<uint32_t> ncount;
    size_t n_fluid;
    struct Span {
        uint32_t body;
        V3 origin;
        uint32_t start, end;
    };
    std::vector<Span> spans;
    struct React {
        uint32_t body;
        V3 f, tau;
    };
    std::vector<React> breact;
    Grid grid;
    // Provider walls. wall_mode 1: half-space providers queried every substep
    // (the CPU path: fluid_density.rs:17-51, fluid_boundary.rs:45-131);
    // wall_mode 2: host-gathered plane tables held for the tick (the GPU path:
    // wall_ghost.wgsl:61-178). The providers also serve the table gather.
    struct WallProv {
        uint32_t id;
        bool has_bounds;
        V3 lo, hi;
        uint32_t face0, nface;  // its faces, in contact order
    };
    struct WallFace {
        uint32_t axis;
        float plane, sign;
    };
    std::vector<WallProv> provs;
    std::vector<WallFace> faces;
    int wall_mode = 0;
    struct WallTable {
        std::vector<uint32_t> ids, start;
        std::vector<V3> pt, nrm;
    };
    WallTable tables[2];  // 0: mirror, 1: project
    float skin = 0.0f;    // fluid_access.rs:47
    uint32_t last_pushed = 0;
};

namespace {

// fluid_access.rs:8-13 (constants) — shared by both constructors.
void init_constants(OrSystem *s, const OrConfig *cfg) {
    s->cfg = *cfg;
    float h = fmaxf(cfg->smoothing_radius, 1e-5f);
    s->h = h;
    s->h2 = h * h;
    s->k6 = 315.0f / ((64.0f * kPi) * powi_rs(h, 9));
    s->ks = 45.0f / (kPi * powi_rs(h, 6));
    s->w0 = ((s->k6 * s->h2) * s->h2) * s->h2;
    s->skin = 0.15f * h;
    s->b_tait = ((cfg->sound_speed * cfg->sound_speed) * cfg->rest_density) / cfg->gamma_tait;
}

void alloc_fluid(OrSystem *s, size_t n) {
    s->n_fluid = n;
    s->vel.assign(n, v3(0, 0, 0));
    s->dens.assign(n, 0.0f);
    s->press.assign(n, 0.0f);
    s->acc.assign(n, v3(0, 0, 0));
    s->xsph.assign(n, v3(0, 0, 0));
    s->pmass.assign(n, s->mass);
    s->bforce.assign(n, v3(0, 0, 0));
    s->ncount.assign(n, 0);
}

void truncate_to_fluid(OrSystem *s) {  // fluid_step.rs:56-70
    size_t nf = s->n_fluid;
    s->pos.resize(nf);
    s->vel.resize(nf);
    s->dens.resize(nf);
    s->press.resize(nf);
    s->acc.resize(nf);
    s->xsph.resize(nf);
    s->pmass.resize(nf);
    s->bforce.resize(nf);
    s->ncount.resize(nf);
    s->spans.clear();
    s->breact.clear();
}

template <class F>
void parallel_chunks(size_t lo, size_t hi, uint32_t threads, F &&body) {
    size_t n = hi - lo;
    if (threads <= 1 || n == 0) {
        body(lo, hi);
        return;
    }
    size_t per = std::max<size_t>((n + threads - 1) / threads, 1);
    std::vector<std::thread> ts;
    for (size_t a = lo; a < hi; a += per) {
        size_t b = std::min(hi, a + per);
        ts.emplace_back([&, a, b] { body(a, b); });
    }
    for (auto &t : ts) t.join();
}

inline float &comp(V3 &v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// A half-space provider's point query (the reference's test doubles FloorY and
// Tank, crates/vxl-phys-fluid/src/tests.rs:11-110): face (axis, plane, sign)
// within 2 * probe of p yields the contact point p with p[axis] = plane and the
// outward normal sign * e_axis, in face order.
template <class F>
void provider_contacts(const OrSystem *s, const OrSystem::WallProv &pv, V3 p, float probe, F &&emit) {
    for (uint32_t k = 0; k < pv.nface; ++k) {
        const OrSystem::WallFace &f = s->faces[pv.face0 + k];
        V3 pt = p;
        const float sdf = (comp(pt, f.axis) - f.plane) * f.sign;
        if (sdf > 2.0f * probe) continue;
        comp(pt, f.axis) = f.plane;
        V3 n = v3(0.0f, 0.0f, 0.0f);
        comp(n, f.axis) = f.sign;
        emit(pt, n);
    }
}

// The AABB prefilter of every provider query, margin m (fluid_density.rs:88-99,
// fluid_boundary.rs:62-73, wall_gather.rs:63-73): true = skip the provider.
inline bool outside(const OrSystem::WallProv &pv, V3 p, float m) {
    if (!pv.has_bounds) return false;
    return p.x < pv.lo.x - m || p.x > pv.hi.x + m || p.y < pv.lo.y - m || p.y > pv.hi.y + m ||
           p.z < pv.lo.z - m || p.z > pv.hi.z + m;
}

// wall_planes_in (fluid_density.rs:76-121): the planes within the h band of
// pi (0 < sdf < h), at most 8, providers in boundary order.
uint32_t wall_planes(const OrSystem *s, V3 pi, V3 pt[8], V3 nrm[8]) {
    uint32_t np = 0;
    for (const auto &pv : s->provs) {
        if (outside(pv, pi, s->h)) continue;
        provider_contacts(s, pv, pi, s->h, [&](V3 point, V3 n) {
            const float sdf = dot(sub(pi, point), n);
            if (sdf > 0.0f && sdf < s->h && np < 8) {
                pt[np] = point;
                nrm[np] = n;
                ++np;
            }
        });
    }
    return np;
}

// fluid_density.rs:17-70. Fluid i uses two accumulators; boundary i takes
// fluid neighbours only. With half-space providers (wall_mode 1) the mirror
// ghosts of every neighbour are summed into `sum` right after it
// (fluid_density.rs:36-45).
void density_pass(OrSystem *s) {
    const size_t nf = s->n_fluid, nt = s->pos.size();
    const float h2 = s->h2, k6 = s->k6, w0 = s->w0, mass = s->mass;
    const Grid &g = s->grid;
    auto fluid = [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) {
            float sum = w0, sum_b = 0.0f;
            uint32_t cnt = 0;
            const V3 pi = s->pos[i];
            V3 pl_pt[8], pl_n[8];
            const uint32_t np = s->wall_mode == 1 ? wall_planes(s, pi, pl_pt, pl_n) : 0u;
            g.for_neighbors(s->pos, h2, i, [&](size_t j, V3 d, float r2) {
                float t = h2 - r2;
                float w = s->fault == 1 ? k6 * ((t * t) * t) : ((k6 * t) * t) * t;
                if (j < nf)
                    sum += w;
                else
                    sum_b += s->pmass[j] * w;
                ++cnt;
                for (uint32_t k = 0; k < np; ++k) {
                    const V3 pj = sub(pi, d);
                    const float dn = dot(sub(pj, pl_pt[k]), pl_n[k]);
                    const V3 gh = s->fault == 4 ? add(pj, mul(pl_n[k], 2.0f * dn)) : sub(pj, mul(pl_n[k], 2.0f * dn));
                    const float rg2 = len2(sub(pi, gh));
                    if (rg2 <= h2) {
                        const float tg = h2 - rg2;
                        sum += ((k6 * tg) * tg) * tg;
                    }
                }
            });
            s->dens[i] = mass * sum + sum_b;
            s->ncount[i] = cnt;
        }
    };
    auto bdry = [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) {
            float sum = 0.0f;
            uint32_t cnt = 0;
            g.for_neighbors(s->pos, h2, i, [&](size_t j, V3, float r2) {
                ++cnt;
                if (j >= nf) return;
                float t = h2 - r2;
                sum += s->pmass[j] * (((k6 * t) * t) * t);
            });
            s->dens[i] = s->pmass[i] * w0 + sum;
            s->ncount[i] = cnt;
        }
    };
    parallel_chunks(0, nf, s->cfg.threads, fluid);
    parallel_chunks(nf, nt, s->cfg.threads, bdry);
}

// fluid_force.rs:11-29 (gamma == 7 branch; clamp p < 0 -> 0).
void pressure_pass(OrSystem *s) {
    const float rho0 = s->cfg.rest_density, gam = s->cfg.gamma_tait;
    for (size_t i = 0; i < s->pos.size(); ++i) {
        float q = s->dens[i] / rho0;
        float p;
        if (fabsf(gam - 7.0f) < 1e-6f) {
            float q2 = q * q;
            float q4 = q2 * q2;
            p = s->b_tait * ((q * q2) * q4 - 1.0f);
        } else {
            p = s->b_tait * (powf(q, gam) - 1.0f);  // not bit-pinned: Rust powf
        }
        if (s->cfg.tensile_clamp && p < 0.0f) p = 0.0f;
        s->press[i] = p;
    }
}

// fluid_boundary.rs:10-31: per span, F = sum bforce[k], tau += (pos[k]-origin) x fk,
// accumulated into breact across substeps.
void aggregate_reactions(OrSystem *s) {
    for (size_t si = 0; si < s->spans.size(); ++si) {
        const auto &sp = s->spans[si];
        V3 f = v3(0, 0, 0), tau = v3(0, 0, 0);
        for (uint32_t k = sp.start; k < sp.end; ++k) {
            V3 fk = s->bforce[k];
            f = add(f, fk);
            tau = add(tau, cross(sub(s->pos[k], sp.origin), fk));
        }
        if (si < s->breact.size() && s->breact[si].body == sp.body) {
            s->breact[si].f = add(s->breact[si].f, f);
            s->breact[si].tau = add(s->breact[si].tau, tau);
        } else if (si < s->breact.size()) {
            s->breact[si] = {sp.body, f, tau};
        } else {
            s->breact.push_back({sp.body, f, tau});
        }
    }
}

// fluid_force.rs:165-217 (serial) and :52-162 (parallel: per-particle chunks,
// then a serial bforce pass in ascending fluid index) — bitwise equal.
void force_pass(OrSystem *s) {
    const size_t nf = s->n_fluid, nt = s->pos.size();
    const V3 grav = v3(s->cfg.gravity[0], s->cfg.gravity[1], s->cfg.gravity[2]);
    const float alpha_c = s->cfg.artificial_viscosity * s->cfg.sound_speed;
    const float h = s->h, h2 = s->h2, k6 = s->k6, ks = s->ks;
    const Grid &g = s->grid;
    for (size_t k = nf; k < nt; ++k) s->bforce[k] = v3(0, 0, 0);
    auto fluid = [&](size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) {
            V3 vi = s->vel[i];
            float rho_i = s->dens[i];
            float ci2 = s->press[i] / (rho_i * rho_i);
            V3 acc = grav, xs = v3(0, 0, 0);
            g.for_neighbors(s->pos, h2, i, [&](size_t j, V3 d, float r2) {
                float rho_j = s->dens[j];
                float cj2 = s->press[j] / (rho_j * rho_j);
                float r = sqrtf(r2);
                float t = h - r;
                float mj = s->pmass[j];
                float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
                float denom = fmaxf(r, 1e-9f);
                acc = add(acc, mul(d, coef / denom));
                V3 vij = sub(vi, s->vel[j]);
                float vdn = -dot(vij, d);
                if (vdn > 0.0f) {
                    float mu = (vdn * h) / (r2 + 0.01f * h2);
                    const float rho_bar = s->fault == 6 ? rho_i + rho_j : 0.5f * (rho_i + rho_j);
                    float c = (mj * ((alpha_c * mu) / rho_bar)) * ((ks * t) * t);
                    acc = add(acc, mul(d, c / denom));
                }
                float tt = h2 - r2;
                float w = ((k6 * tt) * tt) * tt;
                if (s->fault != 2) xs = add(xs, mul(sub(s->vel[j], vi), ((mj * 2.0f) / (rho_i + rho_j)) * w));
            });
            s->acc[i] = acc;
            s->xsph[i] = xs;
        }
    };
    parallel_chunks(0, nf, s->cfg.threads, fluid);
    if (nt > nf && !s->cfg.skip_reactions) {
        // Reaction on boundary particles: bforce[j] -= d * ((m_i*(coef+cv))/denom),
        // accumulated over fluid i ascending (fluid_force.rs:196-206, :111-145).
        for (size_t i = 0; i < nf; ++i) {
            V3 vi = s->vel[i];
            float rho_i = s->dens[i];
            float ci2 = s->press[i] / (rho_i * rho_i);
            float mi = s->pmass[i];
            g.for_neighbors(s->pos, h2, i, [&](size_t j, V3 d, float r2) {
                if (j < nf) return;
                float rho_j = s->dens[j];
                float cj2 = s->press[j] / (rho_j * rho_j);
                float r = sqrtf(r2);
                float t = h - r;
                float mj = s->pmass[j];
                float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
                float denom = fmaxf(r, 1e-9f);
                V3 vij = sub(vi, s->vel[j]);
                float vdn = -dot(vij, d);
                float cv = 0.0f;
                if (vdn > 0.0f) {
                    float mu = (vdn * h) / (r2 + 0.01f * h2);
                    cv = (mj * ((alpha_c * mu) / (0.5f * (rho_i + rho_j)))) * ((ks * t) * t);
                }
                s->bforce[j] = sub(s->bforce[j], mul(d, (mi * (coef + cv)) / denom));
            });
        }
    }
    aggregate_reactions(s);
}

// separate_coincident (fluid_boundary.rs:110-131): pushed particles in
// ascending index; the later of a pair closer than 0.2 h moves along its own
// contact normal sum by r_min - r.
void separate_coincident(OrSystem *s, std::vector<std::pair<size_t, V3>> pushed) {
    if (s->fault == 5) std::reverse(pushed.begin(), pushed.end());
    const float r_min = 0.2f * s->h;
    for (size_t a = 0; a < pushed.size(); ++a) {
        const size_t ia = pushed[a].first;
        const V3 na = pushed[a].second;
        const float nl = sqrtf(len2(na));
        if (nl <= 1e-9f) continue;
        const V3 dir = mul(na, 1.0f / nl);
        V3 pa = s->pos[ia];
        for (size_t b = 0; b < a; ++b) {
            const V3 d = sub(pa, s->pos[pushed[b].first]);
            const float r2 = len2(d);
            if (r2 < r_min * r_min) {
                const float r = sqrtf(r2);
                pa = add(pa, mul(dir, r_min - r));
            }
        }
        s->pos[ia] = pa;
    }
}

// boundary_pass (fluid_boundary.rs:45-104): every fluid particle queried at
// its integrated position pi; each contact with sdf(pi) < 0 pushes the
// particle to the rest line sdf = +skin (at most h per contact) and removes
// the inward normal velocity; then separate_coincident.
void boundary_pass(OrSystem *s) {
    if (s->provs.empty()) return;
    std::vector<std::pair<size_t, V3>> pushed;
    for (size_t i = 0; i < s->n_fluid; ++i) {
        const V3 pi = s->pos[i];
        V3 p = pi, v = s->vel[i], n_sum = v3(0.0f, 0.0f, 0.0f);
        bool hit = false;
        for (const auto &pv : s->provs) {
            if (outside(pv, pi, s->h)) continue;
            provider_contacts(s, pv, pi, s->skin, [&](V3 point, V3 n) {
                const float sdf = dot(sub(pi, point), n);
                const float pen = -sdf;
                if (pen <= 0.0f) return;
                p = add(p, mul(n, fminf(pen + s->skin, s->h)));
                n_sum = add(n_sum, n);
                hit = true;
                const float vn = dot(v, n);
                if (vn < 0.0f) v = sub(v, mul(n, vn));
            });
            s->pos[i] = p;
            s->vel[i] = v;
        }
        if (hit) pushed.emplace_back(i, n_sum);
    }
    s->last_pushed = (uint32_t)pushed.size();
    separate_coincident(s, pushed);
}

// wall_ghost (wall_ghost.wgsl:61-136): per mirror-table entry, the ghosts of
// every neighbour over the entry's planes whose band holds p_i now
// (0 < sdf < h), summed apart and added once: dens += mass * add.
void wall_ghost_tables(OrSystem *s) {
    const auto &T = s->tables[0];
    const float h = s->h, h2 = s->h2, k6 = s->k6;
    for (size_t e = 0; e < T.ids.size(); ++e) {
        const uint32_t i = T.ids[e], st = T.start[e], en = T.start[e + 1];
        if (st >= en) continue;
        const V3 pi = s->pos[i];
        float acc = 0.0f;
        s->grid.for_neighbors(s->pos, h2, i, [&](size_t, V3 d, float) {
            const V3 pj = sub(pi, d);
            for (uint32_t q = st; q < en; ++q) {
                const float sdf_i = dot(sub(pi, T.pt[q]), T.nrm[q]);
                if (sdf_i <= 0.0f || sdf_i >= h) continue;
                const float dn = dot(sub(pj, T.pt[q]), T.nrm[q]);
                const V3 gh = sub(pj, mul(T.nrm[q], 2.0f * dn));
                const float rg2 = len2(sub(pi, gh));
                if (rg2 <= h2) {
                    const float t = h2 - rg2;
                    acc = acc + ((k6 * t) * t) * t;
                }
            }
        });
        if (acc != 0.0f) s->dens[i] = s->dens[i] + s->mass * acc;
    }
}

// wall_project (wall_ghost.wgsl:137-178): per project-table entry, planes in
// order against the current (already pushed) position; no coincident
// separation.
void wall_project_tables(OrSystem *s) {
    const auto &T = s->tables[1];
    for (size_t e = 0; e < T.ids.size(); ++e) {
        const uint32_t i = T.ids[e], st = T.start[e], en = T.start[e + 1];
        if (st >= en) continue;
        V3 p = s->pos[i], v = s->vel[i];
        bool hit = false;
        for (uint32_t q = st; q < en; ++q) {
            const float sdf = dot(sub(p, T.pt[q]), T.nrm[q]);
            const float pen = -sdf;
            if (pen > 0.0f) {
                const float push = fminf(pen + s->skin, s->h);
                p = add(p, mul(T.nrm[q], push));
                const float vn = dot(v, T.nrm[q]);
                if (vn < 0.0f) v = sub(v, mul(T.nrm[q], vn));
                hit = true;
            }
        }
        if (hit) {
            s->pos[i] = p;
            s->vel[i] = v;
        }
    }
}

// fluid_step.rs:91-127; with plane tables the ghosts follow the density pass
// and the projection follows the integration, as the GPU pipeline encodes
// them (pipeline.rs:515-555).
void substep(OrSystem *s, float dt) {
    const size_t nf = s->n_fluid;
    if (nf == 0) return;
    s->grid.rebuild(s->pos, s->h, s->cfg.max_bins);
    density_pass(s);
    if (s->wall_mode == 2) wall_ghost_tables(s);
    pressure_pass(s);
    force_pass(s);
    const float vmax = (s->cfg.max_speed_frac * s->h) / dt;
    const float vmax2 = vmax * vmax;
    const float eps = s->cfg.xsph_viscosity;
    for (size_t i = 0; i < nf; ++i) {
        V3 v = add(add(s->vel[i], mul(s->acc[i], dt)), mul(s->xsph[i], eps));
        float s2 = len2(v);
        if (s2 > vmax2 && s->fault != 7) v = mul(v, vmax / sqrtf(s2));
        s->vel[i] = v;
        s->pos[i] = add(s->pos[i], mul(v, dt));
    }
    if (s->wall_mode == 1) boundary_pass(s);
    if (s->wall_mode == 2) wall_project_tables(s);
}

// boundary.rs:192-194: round half away from zero, at least 1.
inline uint32_t lat_count(float len, float s) {
    long long c = (long long)roundf(len / s);
    return (uint32_t)std::max<long long>(c, 1);
}

// boundary.rs:208-232
void box_faces(const float half[3], float s, std::vector<V3> &pts, std::vector<V3> &nrm) {
    for (int a = 0; a < 3; ++a) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        float hb = half[b], hc = half[c];
        uint32_t nb = lat_count(2.0f * hb, s);
        uint32_t nc = lat_count(2.0f * hc, s);
        float db = 2.0f * hb / (float)nb;
        float dc = 2.0f * hc / (float)nc;
        const float signs[2] = {1.0f, -1.0f};
        for (float sign : signs) {
            for (uint32_t i = 0; i < nb; ++i) {
                for (uint32_t j = 0; j < nc; ++j) {
                    float p[3] = {0, 0, 0}, n[3] = {0, 0, 0};
                    p[a] = sign * half[a];
                    n[a] = sign;
                    p[b] = -hb + ((float)i + 0.5f) * db;
                    p[c] = -hc + ((float)j + 0.5f) * dc;
                    pts.push_back(v3(p[0], p[1], p[2]));
                    nrm.push_back(v3(n[0], n[1], n[2]));
                }
            }
        }
    }
}

float kernel_k6(float h) { return 315.0f / ((64.0f * kPi) * powi_rs(h, 9)); }

}  // namespace

// boundary.rs:106-129, literally O(n^2).
extern "C" float or_boundary_self_sum_bruteforce(const float *p, uint32_t n, float h) {
    if (n == 0) return 0.0f;
    float h2 = h * h;
    float k6 = kernel_k6(h);
    float w0 = ((k6 * h2) * h2) * h2;
    float acc = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        V3 pi = v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        float sum = w0;
        for (uint32_t j = 0; j < n; ++j) {
            if (i == j) continue;
            V3 pj = v3(p[3 * j], p[3 * j + 1], p[3 * j + 2]);
            float r2 = len2(sub(pi, pj));
            if (r2 <= h2) {
                float t = h2 - r2;
                sum += ((k6 * t) * t) * t;
            }
        }
        acc += sum;
    }
    return acc / (float)n;
}

// The same summation through a grid: every accepted j of particle i is
// visited in ascending j (the neighbour list is sorted by index before the
// sum), so each per-particle sum, and therefore the outer sum, is
// bit-identical to the O(n^2) loop.
extern "C" float or_boundary_self_sum_grid(const float *p, uint32_t n, float h) {
    if (n == 0) return 0.0f;
    float h2 = h * h;
    float k6 = kernel_k6(h);
    float w0 = ((k6 * h2) * h2) * h2;
    std::vector<V3> pts(n);
    for (uint32_t i = 0; i < n; ++i) pts[i] = v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    Grid g;
    g.rebuild(pts, h, 1ull << 26);
    float acc = 0.0f;
    std::vector<std::pair<uint32_t, float>> nb;
    for (uint32_t i = 0; i < n; ++i) {
        nb.clear();
        g.for_neighbors(pts, h2, i, [&](size_t j, V3, float r2) { nb.push_back({(uint32_t)j, r2}); });
        std::sort(nb.begin(), nb.end(), [](auto &a, auto &b) { return a.first < b.first; });
        float sum = w0;
        for (auto &e : nb) {
            float t = h2 - e.second;
            sum += ((k6 * t) * t) * t;
        }
        acc += sum;
    }
    return acc / (float)n;
}

// boundary.rs:67-88 for Shape::Box.
extern "C" uint32_t or_lattice_box(const float half[3], float s_in, float h, uint32_t layers_in,
                                   float *out_pts, uint32_t cap, uint32_t *out_n0,
                                   float *out_volume) {
    float s = fmaxf(s_in, 1e-6f);
    uint32_t layers = std::max<uint32_t>(layers_in, 1);
    std::vector<V3> sp, sn;
    box_faces(half, s, sp, sn);
    size_t n0 = sp.size();
    float hm = fminf(fminf(half[0], half[1]), half[2]);
    std::vector<V3> pts;
    pts.reserve(n0 * layers);
    for (uint32_t k = 0; k < layers; ++k) {
        float want = ((float)k + 0.5f) * s;
        float capd = (0.45f + 0.45f * (float)k) * hm;
        float d = fmaxf(fminf(fminf(want, capd), 0.95f * hm), 1e-6f);
        for (size_t q = 0; q < n0; ++q) pts.push_back(sub(sp[q], mul(sn[q], d)));
    }
    // volume_from_self_sum (boundary.rs:95-102), radius h.max(s).
    float hh = fmaxf(h, s);
    std::vector<float> flat(pts.size() * 3);
    for (size_t q = 0; q < pts.size(); ++q) {
        flat[3 * q] = pts[q].x;
        flat[3 * q + 1] = pts[q].y;
        flat[3 * q + 2] = pts[q].z;
    }
    float sum = pts.size() <= 4096 ? or_boundary_self_sum_bruteforce(flat.data(), (uint32_t)pts.size(), hh)
                                   : or_boundary_self_sum_grid(flat.data(), (uint32_t)pts.size(), hh);
    float vol = sum > 1e-12f ? 1.0f / sum : 1.0f;
    uint32_t m = (uint32_t)std::min<size_t>(pts.size(), cap);
    if (out_pts)
        for (uint32_t q = 0; q < m; ++q) {
            out_pts[3 * q] = pts[q].x;
            out_pts[3 * q + 1] = pts[q].y;
            out_pts[3 * q + 2] = pts[q].z;
        }
    if (out_n0) *out_n0 = (uint32_t)n0;
    if (out_volume) *out_volume = vol;
    return (uint32_t)pts.size();
}

extern "C" void or_default_config(OrConfig *c) {  // config.rs:51-69
    c->rest_density = 1000.0f;
    c->smoothing_radius = 0.1f;
    c->xsph_viscosity = 0.1f;
    c->sound_speed = 10.0f;
    c->gamma_tait = 7.0f;
    c->substeps = 4;
    c->threads = 1;
    c->tensile_clamp = 1;
    c->boundary_layers = 2;
    c->gravity[0] = 0.0f;
    c->gravity[1] = -9.81f;
    c->gravity[2] = 0.0f;
    c->max_speed_frac = 0.4f;
    c->artificial_viscosity = 1.0f;
    c->max_bins = 1ull << 21;
    c->skip_reactions = 0;
    c->pad = 0;
}

// fluid_access.rs:7-76
extern "C" OrSystem *or_new_lattice(const OrConfig *cfg, const float origin[3],
                                    const uint32_t dims[3], float spacing) {
    OrSystem *s = new OrSystem();
    init_constants(s, cfg);
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    s->pos.reserve(n);
    for (uint32_t i = 0; i < dims[0]; ++i)
        for (uint32_t j = 0; j < dims[1]; ++j)
            for (uint32_t k = 0; k < dims[2]; ++k)
                s->pos.push_back(v3(origin[0] + ((float)i + 0.5f) * spacing,
                                    origin[1] + ((float)j + 0.5f) * spacing,
                                    origin[2] + ((float)k + 0.5f) * spacing));
    float wsum = 0.0f;
    int side = (int)ceilf(s->h / fmaxf(spacing, 1e-6f));
    for (int i = -side; i <= side; ++i)
        for (int j = -side; j <= side; ++j)
            for (int k = -side; k <= side; ++k) {
                V3 d = v3((float)i * spacing, (float)j * spacing, (float)k * spacing);
                float r2 = len2(d);
                if (r2 <= s->h2) {
                    float t = s->h2 - r2;
                    wsum += ((s->k6 * t) * t) * t;
                }
            }
    s->mass = cfg->rest_density / wsum;
    s->spacing = spacing;
    alloc_fluid(s, n);
    return s;
}

extern "C" OrSystem *or_new_particles(const OrConfig *cfg, uint32_t n_fluid, const float *pos,
                                      const float *vel, float mass, float spacing) {
    OrSystem *s = new OrSystem();
    init_constants(s, cfg);
    s->mass = mass;
    s->spacing = spacing;
    s->pos.resize(n_fluid);
    for (uint32_t i = 0; i < n_fluid; ++i) s->pos[i] = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    alloc_fluid(s, n_fluid);
    if (vel)
        for (uint32_t i = 0; i < n_fluid; ++i) s->vel[i] = v3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    return s;
}

extern "C" void or_free(OrSystem *s) { delete s; }

static void append_boundary(OrSystem *s, V3 p, V3 v, float m) {
    s->pos.push_back(p);
    s->vel.push_back(v);
    s->pmass.push_back(m);
    s->dens.push_back(0.0f);
    s->press.push_back(0.0f);
    s->acc.push_back(v3(0, 0, 0));
    s->xsph.push_back(v3(0, 0, 0));
    s->ncount.push_back(0);
}

// fluid_step.rs:13-53 (Box bodies; the per-shape cache only saves work, so
// the lattice is rebuilt per call here).
extern "C" uint32_t or_set_boundary_boxes(OrSystem *s, uint32_t nbodies, const OrBoxBody *bodies) {
    truncate_to_fluid(s);
    for (uint32_t b = 0; b < nbodies; ++b) {
        const OrBoxBody &bd = bodies[b];
        uint32_t total = or_lattice_box(bd.half, s->spacing, s->h, s->cfg.boundary_layers, nullptr, 0,
                                        nullptr, nullptr);
        std::vector<float> pts((size_t)total * 3);
        uint32_t n0 = 0;
        float vol = 0.0f;
        or_lattice_box(bd.half, s->spacing, s->h, s->cfg.boundary_layers, pts.data(), total, &n0, &vol);
        if (total == 0) continue;
        uint32_t start = (uint32_t)s->pos.size();
        V3 bp = v3(bd.pos[0], bd.pos[1], bd.pos[2]);
        V3 lv = v3(bd.linvel[0], bd.linvel[1], bd.linvel[2]);
        V3 av = v3(bd.angvel[0], bd.angvel[1], bd.angvel[2]);
        for (uint32_t q = 0; q < total; ++q) {
            V3 lp = v3(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]);
            V3 wp = add(bp, rotate(bd.rot, lp));
            V3 wv = add(lv, cross(av, sub(wp, bp)));
            append_boundary(s, wp, wv, s->cfg.rest_density * vol);
        }
        uint32_t end = (uint32_t)s->pos.size();
        if (end > start) s->spans.push_back({bd.id, bp, start, end});
    }
    s->bforce.assign(s->pos.size(), v3(0, 0, 0));
    s->breact.clear();
    return (uint32_t)(s->pos.size() - s->n_fluid);
}

extern "C" void or_set_boundary_raw(OrSystem *s, uint32_t nb, const float *pos, const float *vel,
                                    const float *pmass, uint32_t nspans, const uint32_t *span_body,
                                    const float *span_origin, const uint32_t *span_start,
                                    const uint32_t *span_end) {
    truncate_to_fluid(s);
    for (uint32_t q = 0; q < nb; ++q)
        append_boundary(s, v3(pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]),
                        vel ? v3(vel[3 * q], vel[3 * q + 1], vel[3 * q + 2]) : v3(0, 0, 0), pmass[q]);
    for (uint32_t k = 0; k < nspans; ++k)
        s->spans.push_back({span_body[k],
                            v3(span_origin[3 * k], span_origin[3 * k + 1], span_origin[3 * k + 2]),
                            span_start[k], span_end[k]});
    s->bforce.assign(s->pos.size(), v3(0, 0, 0));
    s->breact.clear();
}

extern "C" void or_set_velocities(OrSystem *s, uint32_t n, const float *vel) {
    uint32_t m = (uint32_t)std::min<size_t>(n, s->n_fluid);
    for (uint32_t i = 0; i < m; ++i) s->vel[i] = v3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
}

extern "C" void or_set_threads(OrSystem *s, uint32_t threads) { s->cfg.threads = threads; }

// Test-only: plant a plausible mistake so tests can show the pins catch it.
// 1 = poly6 product reassociated k6*((t*t)*t); 2 = XSPH term dropped;
// 3 = neighbour cells enumerated dx outer / dz inner; 4 = mirror ghost
// reflected the wrong way; 5 = coincident separation in descending order;
// 6 = artificial viscosity over the density sum instead of its mean;
// 7 = CFL speed clamp skipped.
extern "C" void or_set_fault(OrSystem *s, int fault) {
    s->fault = fault;
    s->grid.fault = fault;
}

extern "C" int or_set_halfspace_walls(OrSystem *s, uint32_t nprov, const OrWallProvider *provs,
                                      uint32_t nface, const OrWallFace *faces) {
    std::vector<OrSystem::WallProv> pv;
    std::vector<OrSystem::WallFace> fc;
    for (uint32_t k = 0; k < nprov; ++k) {
        OrSystem::WallProv p;
        p.id = provs[k].id;
        p.has_bounds = provs[k].has_bounds != 0;
        p.lo = v3(provs[k].lo[0], provs[k].lo[1], provs[k].lo[2]);
        p.hi = v3(provs[k].hi[0], provs[k].hi[1], provs[k].hi[2]);
        p.face0 = (uint32_t)fc.size();
        p.nface = 0;
        for (uint32_t f = 0; f < nface; ++f) {
            if (faces[f].provider != k) continue;
            if (faces[f].axis > 2) return -1;
            fc.push_back({faces[f].axis, faces[f].plane, faces[f].sign});
            ++p.nface;
        }
        pv.push_back(p);
    }
    if (fc.size() != nface) return -1;  // a face names no provider
    s->provs = std::move(pv);
    s->faces = std::move(fc);
    s->wall_mode = s->provs.empty() ? 0 : 1;
    return 0;
}

// gather_in (wall_gather.rs:52-100): per fluid particle at its current
// position, the planes of every provider within the h band (side 0 mirror:
// 0 < sdf < h; side 1 project: sdf < h), at most 8, as sparse rows.
extern "C" uint32_t or_gather_wall_tables(const OrSystem *s, uint32_t side, uint32_t cap_e, uint32_t *ids,
                                          uint32_t *start, uint32_t cap_p, float *planes, uint32_t *out_np) {
    uint32_t ne = 0, np_total = 0;
    if (start) start[0] = 0;
    for (size_t i = 0; i < s->n_fluid; ++i) {
        const V3 p = s->pos[i];
        uint32_t np = 0;
        for (const auto &pv : s->provs) {
            if (outside(pv, p, s->h)) continue;
            provider_contacts(s, pv, p, s->h, [&](V3 point, V3 n) {
                const float sdf = dot(sub(p, point), n);
                const bool keep = side == 0 ? (sdf > 0.0f && sdf < s->h) : (sdf < s->h);
                if (keep && np < 8) {
                    if (planes && np_total < cap_p) {
                        float *o = planes + 6ull * np_total;
                        o[0] = point.x; o[1] = point.y; o[2] = point.z;
                        o[3] = n.x; o[4] = n.y; o[5] = n.z;
                    }
                    ++np_total;
                    ++np;
                }
            });
        }
        if (np > 0) {
            if (ids && ne < cap_e) {
                ids[ne] = (uint32_t)i;
                start[ne + 1] = np_total;
            }
            ++ne;
        }
    }
    if (out_np) *out_np = np_total;
    return ne;
}

extern "C" int or_set_wall_tables(OrSystem *s, uint32_t side, uint32_t ne, const uint32_t *ids,
                                  const uint32_t *start, uint32_t np, const float *planes) {
    if (side > 1) return -1;
    OrSystem::WallTable t;
    t.ids.assign(ids, ids + ne);
    t.start.assign(start, start + ne + 1);
    if (t.start[0] != 0 || t.start[ne] != np) return -1;
    for (uint32_t e = 0; e < ne; ++e) {
        if (t.ids[e] >= s->n_fluid || (e && t.ids[e] <= t.ids[e - 1]) || t.start[e + 1] < t.start[e]) return -1;
    }
    for (uint32_t q = 0; q < np; ++q) {
        const float *o = planes + 6ull * q;
        t.pt.push_back(v3(o[0], o[1], o[2]));
        t.nrm.push_back(v3(o[3], o[4], o[5]));
    }
    s->tables[side] = std::move(t);
    s->wall_mode = 2;
    return 0;
}

extern "C" void or_clear_walls(OrSystem *s) {
    s->provs.clear();
    s->faces.clear();
    s->tables[0] = OrSystem::WallTable{};
    s->tables[1] = OrSystem::WallTable{};
    s->wall_mode = 0;
}

extern "C" uint32_t or_last_pushed(const OrSystem *s) { return s->last_pushed; }

// fluid_step.rs:75-89
extern "C" void or_step(OrSystem *s, float dt_tick) {
    uint32_t sub = std::max<uint32_t>(s->cfg.substeps, 1);
    float dt = dt_tick / (float)sub;
    s->breact.clear();
    for (uint32_t k = 0; k < sub; ++k) substep(s, dt);
    if (sub > 1) {
        float inv = 1.0f / (float)sub;
        for (auto &e : s->breact) {
            e.f = mul(e.f, inv);
            e.tau = mul(e.tau, inv);
        }
    }
}

extern "C" uint32_t or_n_fluid(const OrSystem *s) { return (uint32_t)s->n_fluid; }
extern "C" uint32_t or_n_total(const OrSystem *s) { return (uint32_t)s->pos.size(); }
extern "C" uint32_t or_n_spans(const OrSystem *s) { return (uint32_t)s->spans.size(); }
extern "C" float or_mass(const OrSystem *s) { return s->mass; }
extern "C" void or_constants(const OrSystem *s, float o[8]) {
    o[0] = s->h;
    o[1] = s->h2;
    o[2] = s->k6;
    o[3] = s->ks;
    o[4] = s->w0;
    o[5] = s->b_tait;
    o[6] = s->mass;
    o[7] = s->cfg.artificial_viscosity * s->cfg.sound_speed;
}

static void put3(const std::vector<V3> &v, size_t a, size_t b, float *out) {
    for (size_t k = a; k < b; ++k) {
        out[3 * (k - a)] = v[k].x;
        out[3 * (k - a) + 1] = v[k].y;
        out[3 * (k - a) + 2] = v[k].z;
    }
}
extern "C" void or_get_pos(const OrSystem *s, float *o) { put3(s->pos, 0, s->pos.size(), o); }
extern "C" void or_get_vel(const OrSystem *s, float *o) { put3(s->vel, 0, s->vel.size(), o); }
extern "C" void or_get_acc(const OrSystem *s, float *o) { put3(s->acc, 0, s->n_fluid, o); }
extern "C" void or_get_xsph(const OrSystem *s, float *o) { put3(s->xsph, 0, s->n_fluid, o); }
extern "C" void or_get_bforce(const OrSystem *s, float *o) { put3(s->bforce, s->n_fluid, s->bforce.size(), o); }
extern "C" void or_get_pmass(const OrSystem *s, float *o) { std::copy(s->pmass.begin(), s->pmass.end(), o); }
extern "C" void or_get_dens(const OrSystem *s, float *o) { std::copy(s->dens.begin(), s->dens.end(), o); }
extern "C" void or_get_press(const OrSystem *s, float *o) { std::copy(s->press.begin(), s->press.end(), o); }
extern "C" void or_get_ncount(const OrSystem *s, uint32_t *o) { std::copy(s->ncount.begin(), s->ncount.end(), o); }
extern "C" void or_get_reactions(const OrSystem *s, uint32_t *body, float *ft) {
    for (size_t k = 0; k < s->breact.size(); ++k) {
        body[k] = s->breact[k].body;
        const V3 &f = s->breact[k].f, &t = s->breact[k].tau;
        float *o = ft + 6 * k;
        o[0] = f.x; o[1] = f.y; o[2] = f.z;
        o[3] = t.x; o[4] = t.y; o[5] = t.z;
    }
}

extern "C" void or_grid_info(const OrSystem *s, float mn[3], float *inv, float *bin, uint32_t dims[3]) {
    mn[0] = s->grid.min.x;
    mn[1] = s->grid.min.y;
    mn[2] = s->grid.min.z;
    *inv = s->grid.inv;
    *bin = s->grid.bin;
    dims[0] = s->grid.nx;
    dims[1] = s->grid.ny;
    dims[2] = s->grid.nz;
}
extern "C" uint64_t or_grid_total(const OrSystem *s) { return s->grid.total(); }
extern "C" void or_grid_tables(const OrSystem *s, uint32_t *start, uint32_t *items, uint32_t *keys) {
    const Grid &g = s->grid;
    uint64_t tot = g.total();
    if (start) std::copy(g.start.begin(), g.start.begin() + (tot + 1), start);
    if (items) std::copy(g.items.begin(), g.items.end(), items);
    if (keys)  // per-particle cell index recovered from the table itself
        for (uint64_t c = 0; c < tot; ++c)
            for (uint32_t k = g.start[c]; k < g.start[c + 1]; ++k) keys[g.items[k]] = (uint32_t)c;
}

static inline void fnv(uint64_t &h, uint32_t v) {
    h ^= (uint64_t)v;
    h *= 0x00000100000001b3ull;
}

extern "C" uint64_t or_grid_hash(const OrSystem *s) {
    const Grid &g = s->grid;
    uint64_t tot = g.total();
    std::vector<uint32_t> bins(g.items.size(), 0xffffffffu);
    for (uint64_t c = 0; c < tot; ++c)
        for (uint32_t k = g.start[c]; k < g.start[c + 1]; ++k) bins[g.items[k]] = (uint32_t)c;
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t c = 0; c <= tot; ++c) fnv(h, g.start[c]);
    fnv(h, 0xffffffffu);
    for (uint32_t v : g.items) fnv(h, v);
    fnv(h, 0xffffffffu);
    for (uint32_t v : bins) fnv(h, v);
    return h;
}

static inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

extern "C" uint64_t or_state_hash(const OrSystem *s) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < s->n_fluid; ++i) {
        fnv(h, fbits(s->pos[i].x));
        fnv(h, fbits(s->pos[i].y));
        fnv(h, fbits(s->pos[i].z));
        fnv(h, fbits(s->vel[i].x));
        fnv(h, fbits(s->vel[i].y));
        fnv(h, fbits(s->vel[i].z));
    }
    return h;
}

extern "C" uint32_t or_neighbors(const OrSystem *s, uint32_t i, uint32_t *out, uint32_t cap) {
    uint32_t c = 0;
    s->grid.for_neighbors(s->pos, s->h2, i, [&](size_t j, V3, float) {
        if (c < cap) out[c] = (uint32_t)j;
        ++c;
    });
    return c;
}
