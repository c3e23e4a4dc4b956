// Rigid leg of the coupled step (dynamic Box bodies), once per tick:
//
//   k_rigid          before the tick's first fluid substep: per body and rigid
//                    substep, force = gravity + the previous tick's fluid
//                    reaction, semi-implicit Euler with speed clamps, then the
//                    tank contact rule; the body's reaction span takes the new
//                    position as its torque origin.
//   k_body_refresh   rebuilds the bodies' Akinci particles from the new pose
//                    (set_boundary_particles, fluid_step.rs:13-53) in place in
//                    the grid-ordered state, and the affected scenes' AABBs.
//   k_rigid_collect  after the tick's last substep: keeps the tick-averaged
//                    reaction of each body for the next tick.
//
// Arithmetic is strict f32 in the reference's operation order (math.rs,
// integrate/src/lib.rs, body.rs), so the coupled trajectory is reproducible
// bit for bit.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_kernels.h"

namespace vxl {

namespace {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vmul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
// Vec3::dot (math.rs:54-56), left to right.
__device__ __forceinline__ float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vec3::cross (math.rs:59-65).
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 velem(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 ld3(const float *p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void st3(float *p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

struct Quat {
    float x, y, z, w;
};
// Quat::rotate_vec3 (math.rs:209-213): t = (qv x v) 2; v + qv x t + t w.
__device__ __forceinline__ V3 qrot(Quat q, V3 v) {
    const V3 qv = v3(q.x, q.y, q.z);
    const V3 t = vmul(vcross(qv, v), 2.0f);
    return vadd(vadd(v, vcross(qv, t)), vmul(t, q.w));
}
// Quat::integrate_angular (math.rs:228-240): normalize(q + (0, w) q dt / 2),
// Hamilton product as math.rs:245-252, normalize as :216-224.
__device__ __forceinline__ Quat qintegrate(Quat q, V3 w, float dt) {
    const float dx = 0.0f * q.x + w.x * q.w + w.y * q.z - w.z * q.y;
    const float dy = 0.0f * q.y - w.x * q.z + w.y * q.w + w.z * q.x;
    const float dz = 0.0f * q.z + w.x * q.y - w.y * q.x + w.z * q.w;
    const float dw = 0.0f * q.w - w.x * q.x - w.y * q.y - w.z * q.z;
    const float hdt = 0.5f * dt;
    const Quat r{q.x + dx * hdt, q.y + dy * hdt, q.z + dz * hdt, q.w + dw * hdt};
    const float len2 = r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w;
    if (len2 > 1e-16f) {
        const float inv = 1.0f / sqrtf(len2);
        return Quat{r.x * inv, r.y * inv, r.z * inv, r.w * inv};
    }
    return Quat{0.0f, 0.0f, 0.0f, 1.0f};
}
// BodySet::apply_world_inv_inertia (body.rs:172-176): R (diag(d) (R^T l)) with
// R = Mat3::from_quat (math.rs:275-292) and the row/column products of
// math.rs:295-311.
__device__ __forceinline__ V3 world_inv_inertia(Quat q, V3 d, V3 l) {
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    const float m00 = 1.0f - 2.0f * (yy + zz), m01 = 2.0f * (xy - wz), m02 = 2.0f * (xz + wy);
    const float m10 = 2.0f * (xy + wz), m11 = 1.0f - 2.0f * (xx + zz), m12 = 2.0f * (yz - wx);
    const float m20 = 2.0f * (xz - wy), m21 = 2.0f * (yz + wx), m22 = 1.0f - 2.0f * (xx + yy);
    const V3 lt = v3(m00 * l.x + m10 * l.y + m20 * l.z, m01 * l.x + m11 * l.y + m21 * l.z,
                     m02 * l.x + m12 * l.y + m22 * l.z);
    const V3 s = velem(lt, d);
    return v3(m00 * s.x + m01 * s.y + m02 * s.z, m10 * s.x + m11 * s.y + m12 * s.z, m20 * s.x + m21 * s.y + m22 * s.z);
}

// Tank contact rule (no counterpart in the reference, whose particle-only
// tank has no rigid contact): for each plane in the order floor (y >= 0),
// x >= 0, x <= L0, z >= 0, z <= L2, the deepest corner (first maximum over
// k = 0..7; bit 0/1/2 of k picks +half on x/y/z) that penetrates by d > 0
// moves the box by d along the plane normal, and a frictionless, fully
// inelastic impulse removes that corner's approach velocity.
__device__ __forceinline__ void tank_contacts(const RigidDev &B, V3 &pos, Quat q, V3 &lv, V3 &av, V3 d) {
    const V3 half = ld3(B.half);
    for (int p = 0; p < 5; ++p) {
        const V3 n = p == 0 ? v3(0.0f, 1.0f, 0.0f)
                     : p == 1 ? v3(1.0f, 0.0f, 0.0f)
                     : p == 2 ? v3(-1.0f, 0.0f, 0.0f)
                     : p == 3 ? v3(0.0f, 0.0f, 1.0f)
                              : v3(0.0f, 0.0f, -1.0f);
        const float off = p == 2 ? -B.tank[0] : (p == 4 ? -B.tank[2] : 0.0f);
        float best = 0.0f;
        V3 br = v3(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 8; ++k) {
            const V3 lp = v3((k & 1) ? half.x : -half.x, (k & 2) ? half.y : -half.y, (k & 4) ? half.z : -half.z);
            const V3 r = qrot(q, lp);
            const float depth = off - vdot(n, vadd(pos, r));
            if (depth > best) {
                best = depth;
                br = r;
            }
        }
        if (!(best > 0.0f)) continue;
        pos = vadd(pos, vmul(n, best));
        const float vn = vdot(vadd(lv, vcross(av, br)), n);
        if (vn < 0.0f) {
            const V3 irn = world_inv_inertia(q, d, vcross(br, n));
            const float kk = B.inv_mass + vdot(n, vcross(irn, br));
            const float j = (-vn) / kk;
            lv = vadd(lv, vmul(n, j * B.inv_mass));
            av = vadd(av, world_inv_inertia(q, d, vcross(br, vmul(n, j))));
        }
    }
}

}  // namespace

__global__ __launch_bounds__(64) void k_rigid(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t k = blockIdx.x * 64 + threadIdx.x; k < P.nrigid; k += gridDim.x * 64) {
        RigidDev B = P.rigid[k];
        const uint32_t sub = B.substeps ? B.substeps : 1u;
        const float dt = B.dt_tick / (float)sub;  // World::step (world_step.rs:13-15)
        const V3 g = ld3(B.gravity), d = ld3(B.inv_inertia), zero = v3(0.0f, 0.0f, 0.0f);
        V3 pos = ld3(B.pos), lv = ld3(B.linvel), av = ld3(B.angvel);
        Quat q{B.rot[0], B.rot[1], B.rot[2], B.rot[3]};
        for (uint32_t s = 0; s < sub; ++s) {
            // Force accumulator: gravity field (field/src/lib.rs:22-31), then the
            // fluid reaction through the coupling applier (world_step.rs:235-250).
            V3 force = vadd(zero, vmul(g, 1.0f / B.inv_mass));
            force = vadd(force, ld3(B.react));
            const V3 torque = vadd(zero, ld3(B.react + 3));
            // Integrator::integrate_velocities (integrate/src/lib.rs:18-52),
            // gravity passed as zero (it is in the force).
            const V3 accel = vadd(zero, vmul(force, B.inv_mass));
            lv = vadd(lv, vmul(accel, dt));
            av = vadd(av, vmul(world_inv_inertia(q, d, torque), dt));
            const float sp = sqrtf(vdot(lv, lv));
            if (sp > B.max_linear) lv = vmul(lv, B.max_linear / sp);
            const float ws = sqrtf(vdot(av, av));
            if (ws > B.max_angular) av = vmul(av, B.max_angular / ws);
            // Integrator::integrate_positions (integrate/src/lib.rs:55-64)
            pos = vadd(pos, vmul(lv, dt));
            q = qintegrate(q, av, dt);
            if (B.tank[0] > 0.0f) tank_contacts(B, pos, q, lv, av, d);
        }
        RigidDev &O = P.rigid[k];
        st3(O.pos, pos);
        st3(O.linvel, lv);
        st3(O.angvel, av);
        O.rot[0] = q.x;
        O.rot[1] = q.y;
        O.rot[2] = q.z;
        O.rot[3] = q.w;
        st3(P.spans[B.span].origin, pos);  // reaction torque is taken about the body origin
    }
}

constexpr uint32_t kNoBody = 0xffffffffu;  // a slot that holds no body particle

// Rebuild the dynamic bodies' particles from their pose (fluid_step.rs:32-39:
// world = p + q lp, velocity = v + w x (world - p)) and recompute the AABB of
// every scene that has such bodies into box[parity], which this kernel resets
// for k_box_reduce to refill.
__global__ __launch_bounds__(kBlock) void k_body_refresh(Pool P) {
    if (P.ctl->err) return;
    const uint32_t par = P.ctl->step & 1u;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        if (sp.nrbody == 0u) continue;
        const uint32_t m = b * kBlock + threadIdx.x;
        const bool live = m < sp.pbase + sp.n;
        // the body whose particle this slot holds, if any
        uint32_t gid = 0, body = kNoBody;
        if (live) {
            gid = P.cg[m];
            if (gid >= sp.nf)
                for (uint32_t k = sp.rbody0; k < sp.rbody0 + sp.nrbody; ++k)
                    if (gid >= P.rigid[k].gstart && gid < P.rigid[k].gend) body = k;
        }
        if (b == sp.pbase / kBlock && threadIdx.x == 0) {
            SceneBox &o = P.box[par * P.nscenes + s];
            for (int ax = 0; ax < 3; ++ax) {
                o.lo[ax] = 0xffffffffu;
                o.hi[ax] = 0u;
            }
        }
        // A block without body particles keeps the partial box of its
        // positions that the last force pass (or the upload) wrote.
        if (!__syncthreads_or(body != kNoBody)) continue;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (live) a = P.cA[m];
        if (body != kNoBody) {
            const RigidDev &B = P.rigid[body];
            const float4 l = P.rlat[B.lat + (gid - B.gstart)];
            const V3 pos = ld3(B.pos);
            const Quat q{B.rot[0], B.rot[1], B.rot[2], B.rot[3]};
            const V3 wp = vadd(pos, qrot(q, v3(l.x, l.y, l.z)));
            const V3 wv = vadd(ld3(B.linvel), vcross(ld3(B.angvel), vsub(wp, pos)));
            a = make_float4(wp.x, wp.y, wp.z, a.w);
            P.cA[m] = a;
            const float4 v = P.cV[m];
            P.cV[m] = make_float4(wv.x, wv.y, wv.z, v.w);
        }
        dev::wave_box(live, a.x, a.y, a.z, &P.part[m / kPartSlots]);
    }
}

__global__ __launch_bounds__(64) void k_rigid_collect(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t k = blockIdx.x * 64 + threadIdx.x; k < P.nrigid; k += gridDim.x * 64) {
        RigidDev &B = P.rigid[k];
        const float *r = P.react + 6ull * B.span;
        for (int a = 0; a < 6; ++a) B.react[a] = r[a];
    }
}

}  // namespace vxl
