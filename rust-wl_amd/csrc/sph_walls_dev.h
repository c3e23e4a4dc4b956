// Device helpers of provider walls, shared by the density pass (mirror ghosts
// of half-space walls) and the wall kernels. Strict f32: every Vec3 operation
// written out in the reference's component order (math.rs:54-70).
#ifndef VXL_SPH_WALLS_DEV_H
#define VXL_SPH_WALLS_DEV_H
#include <hip/hip_runtime.h>

#include "sph_layout.h"

namespace vxl {
namespace wall {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }
// Component `axis` of v (selects, so v stays in registers).
__device__ __forceinline__ float get_axis(V3 v, uint32_t axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
__device__ __forceinline__ V3 with_axis(V3 v, uint32_t axis, float c) {
    return V3{axis == 0 ? c : v.x, axis == 1 ? c : v.y, axis == 2 ? c : v.z};
}

// Point query of one half-space provider (the reference's test doubles FloorY
// and Tank, crates/vxl-phys-fluid/src/tests.rs:11-110): face (axis, plane,
// sign) within 2 * probe of p yields the contact point p with p[axis] = plane
// and the outward normal sign * e_axis, faces in order; emit(face index,
// point, normal).
template <class F>
__device__ __forceinline__ void prov_contacts(const WallSceneDev &W, const WallProvDev &pv, V3 p, float probe,
                                              F &&emit) {
    for (uint32_t k = 0; k < pv.nface; ++k) {
        const WallFaceDev f = W.face[pv.face0 + k];
        const float sdf = (get_axis(p, f.axis) - f.plane) * f.sign;
        if (sdf > 2.0f * probe) continue;
        emit(pv.face0 + k, with_axis(p, f.axis, f.plane), with_axis(V3{0.0f, 0.0f, 0.0f}, f.axis, f.sign));
    }
}

// AABB prefilter of every provider query with margin m (fluid_density.rs:88-99,
// fluid_boundary.rs:62-73): true = the provider is skipped.
__device__ __forceinline__ bool outside(const WallProvDev &pv, V3 p, float m) {
    if (!pv.has_bounds) return false;
    return p.x < pv.lo[0] - m || p.x > pv.hi[0] + m || p.y < pv.lo[1] - m || p.y > pv.hi[1] + m ||
           p.z < pv.lo[2] - m || p.z > pv.hi[2] + m;
}

// Mirror image of neighbour j (at pi - d) in plane (pt, n), and its kernel
// weight at pi when within h (fluid_density.rs:36-45).
__device__ __forceinline__ bool ghost_w(V3 pi, V3 d, V3 pt, V3 n, float h2, float k6, float &w) {
    const V3 pj = sub(pi, d);
    const float dn = dot(sub(pj, pt), n);
    const V3 g = sub(pj, mul(n, 2.0f * dn));
    const V3 dr = sub(pi, g);
    const float rg2 = dot(dr, dr);
    if (!(rg2 <= h2)) return false;
    const float t = h2 - rg2;
    w = ((k6 * t) * t) * t;
    return true;
}

// The mirror planes of fluid particle pi (wall_planes_in, fluid_density.rs:
// 76-121): the faces within its h band (0 < sdf < h), at most 8, providers in
// order; bit k = face k of the scene (faces are stored provider by provider,
// so ascending bits are the reference's plane order).
__device__ __forceinline__ uint32_t band_faces(const WallSceneDev &W, V3 pi, float h) {
    uint32_t fm = 0, np = 0;
    for (uint32_t k = 0; k < W.nprov; ++k) {
        const WallProvDev &pv = W.prov[k];
        if (outside(pv, pi, h)) continue;
        prov_contacts(W, pv, pi, h, [&](uint32_t f, V3 point, V3 n) {
            const float sdf = dot(sub(pi, point), n);
            if (sdf > 0.0f && sdf < h && np < kWallMirrorPlanes) {
                fm |= 1u << f;
                ++np;
            }
        });
    }
    return fm;
}

// The ghost of neighbour d (= pi - p_j) in band face f of particle pi: the
// contact point of pi's query is pi with pi[axis] = plane.
__device__ __forceinline__ bool face_ghost(const WallFaceDev &f, V3 pi, V3 d, float h2, float k6, float &w) {
    return ghost_w(pi, d, with_axis(pi, f.axis, f.plane), with_axis(V3{0.0f, 0.0f, 0.0f}, f.axis, f.sign), h2, k6,
                   w);
}

}  // namespace wall
}  // namespace vxl
#endif
