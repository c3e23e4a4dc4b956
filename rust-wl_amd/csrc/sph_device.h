// Device helpers shared by the kernel translation units.
#ifndef VXL_SPH_DEVICE_H
#define VXL_SPH_DEVICE_H
#include <hip/hip_runtime.h>

#include "sph_layout.h"

namespace vxl {
namespace dev {

// Minimum / maximum over the wave's active lanes (the device library's DPP
// wave reductions; every caller has the whole wave active).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) { return __ockl_wfred_min_u32(v); }
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) { return __ockl_wfred_max_u32(v); }

// AABB of the positions handed in by the wave's live lanes, written by lane 0
// as the wave's partial box (ordered u32; the per-scene reduction runs after
// the launch). No atomics and no LDS: every wave of every workgroup size owns
// the partial box of its 64 slots. The whole wave must call it.
__device__ __forceinline__ void wave_box(bool live, float x, float y, float z, SceneBox *part) {
    uint32_t lx = live ? f2o(x) : 0xffffffffu, ly = live ? f2o(y) : 0xffffffffu,
             lz = live ? f2o(z) : 0xffffffffu;
    uint32_t hx = live ? f2o(x) : 0u, hy = live ? f2o(y) : 0u, hz = live ? f2o(z) : 0u;
    lx = wave_min_u32(lx);
    ly = wave_min_u32(ly);
    lz = wave_min_u32(lz);
    hx = wave_max_u32(hx);
    hy = wave_max_u32(hy);
    hz = wave_max_u32(hz);
    if ((threadIdx.x & 63u) == 0u) {
        part->lo[0] = lx; part->lo[1] = ly; part->lo[2] = lz;
        part->hi[0] = hx; part->hi[1] = hy; part->hi[2] = hz;
    }
}

// fluid grid.rs:91-101: u = min(max(floor((v - o) * inv), 0) as u32, n - 1);
// also returns the position inside the cell in cell units, clamped to [0, 1].
__device__ __forceinline__ uint32_t axis_bin(float o, float v, float inv, uint32_t n, float *frac) {
    const float t = (v - o) * inv;
    const float f = fmaxf(floorf(t), 0.0f);
    const uint32_t lim = n - 1u;
    const uint32_t u = f >= (float)lim ? lim : (uint32_t)f;
    *frac = fminf(fmaxf(t - (float)u, 0.0f), 1.0f);
    return u;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2. Neighbour kernels remap the hardware block id so that XCD x works on one
// contiguous run of pool blocks: spatial neighbours then share an L2 instead
// of being fetched into all eight. (A bijection for any grid size.)
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t hw, uint32_t nb) {
    const uint32_t q = nb / kXcds, r = nb % kXcds;
    const uint32_t x = hw % kXcds, k = hw / kXcds;
    return x * q + min(x, r) + k;
}

// Record `idx` of a per-slot or per-cell array through a 32-bit byte offset
// from the array base, which lets the load use a scalar base and a 32-bit
// vector offset. Callers keep idx * sizeof(T) below 2^32: a pool holds at
// most kMaxPoolSlots slots (density records: 2 x 16 bytes per slot) and
// kMaxCells cells (sph_layout.h).
template <class T>
__device__ __forceinline__ const T *at(const T *__restrict__ base, uint32_t idx) {
    return (const T *)((const char *)base + idx * (uint32_t)sizeof(T));
}

// Neighbour geometry of one particle: its cell and where it sits inside it.
struct Cursor {
    int cx, cy, cz;
    float fx, fy, fz;
    float cut;  // (h / bin)^2 with margin, in cell units
};

__device__ __forceinline__ Cursor make_cursor(const SceneGrid &g, float h, float px, float py, float pz) {
    Cursor c;
    c.cx = (int)axis_bin(g.gmin[0], px, g.inv, g.dims[0], &c.fx);
    c.cy = (int)axis_bin(g.gmin[1], py, g.inv, g.dims[1], &c.fy);
    c.cz = (int)axis_bin(g.gmin[2], pz, g.inv, g.dims[2], &c.fz);
    const float hc = h * g.inv;
    c.cut = hc * hc * 1.02f + 1e-6f;
    return c;
}

// Visit the 3x3 rows of the 27-cell neighbourhood in the reference order
// (dz outer, dy inner). Each row is handed over as its index
// r = 3 (dz + 1) + dy + 1, its first and last internal cell (the cells
// dx = -1, 0, +1 that survive the conservative distance test) and whether it
// starts at the cell dx = -1.
template <class F>
__device__ __forceinline__ void for_row_windows(const SceneGrid &g, const Cursor &c, F &&visit) {
    const int nx = (int)g.dims[0], ny = (int)g.dims[1], nz = (int)g.dims[2];
    const float lx = c.fx * c.fx, rx = (1.0f - c.fx) * (1.0f - c.fx);
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = c.cz + dz;
        if (z < 0 || z >= nz) continue;
        const float ez = dz < 0 ? c.fz : (dz > 0 ? 1.0f - c.fz : 0.0f);
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = c.cy + dy;
            if (y < 0 || y >= ny) continue;
            const float ey = dy < 0 ? c.fy : (dy > 0 ? 1.0f - c.fy : 0.0f);
            const float rd = ez * ez + ey * ey;
            if (rd > c.cut) continue;
            const int xl = (c.cx > 0 && rd + lx <= c.cut) ? c.cx - 1 : c.cx;
            const int xr = (c.cx + 1 < nx && rd + rx <= c.cut) ? c.cx + 1 : c.cx;
            const uint32_t row = g.cbase + ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
            visit((uint32_t)(3 * (dz + 1) + dy + 1), row + (uint32_t)xl, row + (uint32_t)xr, xl < c.cx);
        }
    }
}

// for_row_windows with each row as one slot range [j0, j1) (x-fastest cells
// make the row's cells contiguous).
template <class F>
__device__ __forceinline__ void for_rows(const SceneGrid &g, const uint32_t *__restrict__ start, const Cursor &c,
                                         F &&visit) {
    for_row_windows(g, c, [&](uint32_t r, uint32_t cf, uint32_t cl, bool left) {
        visit(r, left, start[cf], start[cl + 1u]);
    });
}

// Tait EOS, gamma = 7 (fluid_force.rs:11-29), and the per-particle pressure
// term p / (rho * rho) of the force sum.
__device__ __forceinline__ void tait(const SceneParams &sp, float rho, float &p, float &c2) {
    const float q = rho / sp.rho0;
    const float q2 = q * q;
    const float q4 = q2 * q2;
    p = sp.b_tait * ((q * q2) * q4 - 1.0f);
    if (sp.clamp_neg && p < 0.0f) p = 0.0f;
    c2 = p / (rho * rho);
}

}  // namespace dev
}  // namespace vxl
#endif
