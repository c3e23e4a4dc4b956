// Rigid-box leg oracle — TEST INFRASTRUCTURE ONLY (see rigid_oracle.h).
// Strict f32, -ffp-contract=off; every expression keeps the reference's
// left-to-right operation order.
#include "rigid_oracle.h"

#include <cmath>

namespace {

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
// math.rs:54-56
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// math.rs:59-65
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// math.rs:73-75
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// math.rs:109-111
inline V3 mul_per_elem(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }

struct Q {
    float x, y, z, w;
};
// math.rs:209-213
inline V3 rotate(Q q, V3 v) {
    const V3 qv = v3(q.x, q.y, q.z);
    const V3 t = mul(cross(qv, v), 2.0f);
    return add(add(v, cross(qv, t)), mul(t, q.w));
}
// math.rs:216-224
inline Q normalize(Q q) {
    const float len2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (len2 > 1e-16f) {
        const float inv = 1.0f / std::sqrt(len2);
        return Q{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
    }
    return Q{0.0f, 0.0f, 0.0f, 1.0f};
}
// math.rs:245-252 (Hamilton product a * b)
inline Q qmul(Q a, Q b) {
    return Q{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// math.rs:228-240
inline Q integrate_angular(Q q, V3 w, float dt) {
    const Q wq{w.x, w.y, w.z, 0.0f};
    const Q dq = qmul(wq, q);
    const float hdt = 0.5f * dt;
    return normalize(Q{q.x + dq.x * hdt, q.y + dq.y * hdt, q.z + dq.z * hdt, q.w + dq.w * hdt});
}

// Mat3::from_quat (math.rs:275-292), row major.
struct M3 {
    float m[3][3];
};
inline M3 from_quat(Q q) {
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3 r;
    r.m[0][0] = 1.0f - 2.0f * (yy + zz);
    r.m[0][1] = 2.0f * (xy - wz);
    r.m[0][2] = 2.0f * (xz + wy);
    r.m[1][0] = 2.0f * (xy + wz);
    r.m[1][1] = 1.0f - 2.0f * (xx + zz);
    r.m[1][2] = 2.0f * (yz - wx);
    r.m[2][0] = 2.0f * (xz - wy);
    r.m[2][1] = 2.0f * (yz + wx);
    r.m[2][2] = 1.0f - 2.0f * (xx + yy);
    return r;
}
// math.rs:295-301
inline V3 mul_vec3(const M3 &r, V3 v) {
    return V3{r.m[0][0] * v.x + r.m[0][1] * v.y + r.m[0][2] * v.z, r.m[1][0] * v.x + r.m[1][1] * v.y + r.m[1][2] * v.z,
              r.m[2][0] * v.x + r.m[2][1] * v.y + r.m[2][2] * v.z};
}
// math.rs:305-311
inline V3 transpose_mul_vec3(const M3 &r, V3 v) {
    return V3{r.m[0][0] * v.x + r.m[1][0] * v.y + r.m[2][0] * v.z, r.m[0][1] * v.x + r.m[1][1] * v.y + r.m[2][1] * v.z,
              r.m[0][2] * v.x + r.m[1][2] * v.y + r.m[2][2] * v.z};
}
// BodySet::apply_world_inv_inertia (body.rs:172-176).
inline V3 world_inv_inertia(Q q, V3 d, V3 l) {
    const M3 r = from_quat(q);
    const V3 lt = transpose_mul_vec3(r, l);
    return mul_vec3(r, mul_per_elem(lt, d));
}

inline V3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }
inline void st3(float *p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// The tank contact rule (rigid_oracle.h).
void tank_contacts(const OrRigidConfig *cfg, V3 &pos, Q q, V3 &lv, V3 &av, float inv_mass, V3 d, V3 half) {
    const V3 nrm[5] = {v3(0.0f, 1.0f, 0.0f), v3(1.0f, 0.0f, 0.0f), v3(-1.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 1.0f),
                       v3(0.0f, 0.0f, -1.0f)};
    const float off[5] = {0.0f, 0.0f, -cfg->tank[0], 0.0f, -cfg->tank[2]};
    for (int p = 0; p < 5; ++p) {
        const V3 n = nrm[p];
        float best = 0.0f;
        V3 br = v3(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 8; ++k) {
            const V3 lp = v3((k & 1) ? half.x : -half.x, (k & 2) ? half.y : -half.y, (k & 4) ? half.z : -half.z);
            const V3 r = rotate(q, lp);
            const float depth = off[p] - dot(n, add(pos, r));
            if (depth > best) {
                best = depth;
                br = r;
            }
        }
        if (!(best > 0.0f)) continue;
        pos = add(pos, mul(n, best));
        const float vn = dot(add(lv, cross(av, br)), n);
        if (vn < 0.0f) {
            const V3 irn = world_inv_inertia(q, d, cross(br, n));
            const float k = inv_mass + dot(n, cross(irn, br));
            const float j = (-vn) / k;
            lv = add(lv, mul(n, j * inv_mass));
            av = add(av, world_inv_inertia(q, d, cross(br, mul(n, j))));
        }
    }
}

}  // namespace

extern "C" void or_box_mass_props(const float half[3], float density, float *inv_mass, float inv_inertia[3]) {
    const float dens = density > 0.0f ? density : 1.0f;
    const float ex = 2.0f * half[0], ey = 2.0f * half[1], ez = 2.0f * half[2];
    const float m = dens * ex * ey * ez;
    const float ix = m / 12.0f * (ey * ey + ez * ez);
    const float iy = m / 12.0f * (ex * ex + ez * ez);
    const float iz = m / 12.0f * (ex * ex + ey * ey);
    *inv_mass = 1.0f / m;
    inv_inertia[0] = 1.0f / ix;
    inv_inertia[1] = 1.0f / iy;
    inv_inertia[2] = 1.0f / iz;
}

extern "C" void or_rigid_tick(const OrRigidConfig *cfg, uint32_t n, OrRigidBody *bodies, const float *react) {
    const uint32_t sub = cfg->substeps ? cfg->substeps : 1u;
    const float dt = cfg->dt_tick / (float)sub;  // World::step (world_step.rs:13-15)
    const V3 g = ld3(cfg->gravity);
    const V3 zero = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t s = 0; s < sub; ++s) {
        for (uint32_t b = 0; b < n; ++b) {
            OrRigidBody &B = bodies[b];
            V3 pos = ld3(B.pos), lv = ld3(B.linvel), av = ld3(B.angvel);
            Q q{B.rot[0], B.rot[1], B.rot[2], B.rot[3]};
            const V3 d = ld3(B.inv_inertia), half = ld3(B.half);
            // Force accumulator: gravity field, then the fluid reaction.
            V3 force = zero, torque = zero;
            const float mass = 1.0f / B.inv_mass;
            force = add(force, mul(g, mass));
            force = add(force, ld3(react + 6 * b));
            torque = add(torque, ld3(react + 6 * b + 3));
            // integrate_velocities (gravity passed as zero: it is in the force).
            const V3 accel = add(zero, mul(force, B.inv_mass));
            lv = add(lv, mul(accel, dt));
            const V3 ang_accel = world_inv_inertia(q, d, torque);
            av = add(av, mul(ang_accel, dt));
            const float sp = length(lv);
            if (sp > cfg->max_linear) lv = mul(lv, cfg->max_linear / sp);
            const float ws = length(av);
            if (ws > cfg->max_angular) av = mul(av, cfg->max_angular / ws);
            // integrate_positions
            pos = add(pos, mul(lv, dt));
            q = integrate_angular(q, av, dt);
            if (cfg->tank[0] > 0.0f) tank_contacts(cfg, pos, q, lv, av, B.inv_mass, d, half);
            st3(B.pos, pos);
            st3(B.linvel, lv);
            st3(B.angvel, av);
            B.rot[0] = q.x;
            B.rot[1] = q.y;
            B.rot[2] = q.z;
            B.rot[3] = q.w;
        }
    }
}
