// gfx950 kernels of the WCSPH substep (one launch per stage for every scene
// of the pool; a single scene is a pool of one).
//
// Substep = grid_box -> keys -> scan (3 passes) -> place -> canon_gather ->
// density(+Tait) -> force(+AV, XSPH, integrate, next AABB).
//
// Arithmetic contract: strict IEEE f32 (built with -ffp-contract=off,
// correctly rounded / and sqrt, denormals kept), every expression in the
// operation order of the reference CPU step, neighbours enumerated cell by
// cell (dz, dy, dx; within a cell ascending original index), so every sum is
// the reference's sum bit for bit.
#include <hip/hip_runtime.h>

#include "sph_kernels.h"

namespace vxl {

namespace {

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// fluid grid.rs:91-101: u = min(max(floor((v - o) * inv), 0) as u32, n - 1).
__device__ __forceinline__ uint32_t axis_bin(float o, float v, float inv, uint32_t n) {
    float f = fmaxf(floorf((v - o) * inv), 0.0f);
    uint32_t lim = n - 1u;
    return f >= (float)lim ? lim : (uint32_t)f;
}

// Block-wide AABB of the positions handed in by live lanes, merged into one
// scene's accumulator with 6 integer atomics (min/max are order free, exact).
__device__ __forceinline__ void block_box(bool live, float x, float y, float z, SceneBox *dst) {
    __shared__ uint32_t sh[6][kBlock / 64];
    uint32_t lx = live ? f2o(x) : 0xffffffffu, ly = live ? f2o(y) : 0xffffffffu,
             lz = live ? f2o(z) : 0xffffffffu;
    uint32_t hx = live ? f2o(x) : 0u, hy = live ? f2o(y) : 0u, hz = live ? f2o(z) : 0u;
    lx = wave_min_u32(lx);
    ly = wave_min_u32(ly);
    lz = wave_min_u32(lz);
    hx = wave_max_u32(hx);
    hy = wave_max_u32(hy);
    hz = wave_max_u32(hz);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sh[0][w] = lx; sh[1][w] = ly; sh[2][w] = lz;
        sh[3][w] = hx; sh[4][w] = hy; sh[5][w] = hz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(kBlock / 64); ++k) {
            lx = min(lx, sh[0][k]); ly = min(ly, sh[1][k]); lz = min(lz, sh[2][k]);
            hx = max(hx, sh[3][k]); hy = max(hy, sh[4][k]); hz = max(hz, sh[5][k]);
        }
        if (lx != 0xffffffffu) {
            atomicMin(&dst->lo[0], lx); atomicMin(&dst->lo[1], ly); atomicMin(&dst->lo[2], lz);
            atomicMax(&dst->hi[0], hx); atomicMax(&dst->hi[1], hy); atomicMax(&dst->hi[2], hz);
        }
    }
    __syncthreads();
}

}  // namespace

// AABB of the cur positions into box[parity(step)] (used at upload).
__global__ __launch_bounds__(kBlock) void k_box(Pool P) {
    const uint32_t par = P.ctl->step & 1u;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t m = b * kBlock + threadIdx.x;
        const bool live = m < sp.pbase + sp.n;
        float x = 0, y = 0, z = 0;
        if (live) { x = P.cx[m]; y = P.cy[m]; z = P.cz[m]; }
        block_box(live, x, y, z, &P.box[par * P.nscenes + s]);
    }
}

// grid_box rule per scene (core/grid.rs:34-51), cell bases, pad-cell counts,
// AABB ping-pong reset, capacity check. One block.
__global__ __launch_bounds__(1024) void k_grid_box(Pool P) {
    __shared__ uint32_t scan[1024];
    __shared__ uint32_t carry;
    __shared__ uint32_t bad;
    if (P.ctl->err) return;  // an earlier substep failed: leave everything as is
    const uint32_t par = P.ctl->step & 1u;
    if (threadIdx.x == 0) {
        carry = 0;
        bad = 0;
    }
    __syncthreads();
    for (uint32_t s0 = 0; s0 < P.nscenes; s0 += 1024) {
        const uint32_t s = s0 + threadIdx.x;
        uint32_t need = 0;
        if (s < P.nscenes) {
            const SceneParams &sp = P.params[s];
            SceneBox *bx = &P.box[par * P.nscenes + s];
            SceneBox *nx = &P.box[(par ^ 1u) * P.nscenes + s];
            float lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = o2f(bx->lo[a]);
                hi[a] = o2f(bx->hi[a]);
                nx->lo[a] = 0xffffffffu;
                nx->hi[a] = 0u;
            }
            float bin = fmaxf(sp.h, 1e-6f);
            const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
            uint32_t dims[3] = {1, 1, 1};
            for (int it = 0; it < 256; ++it) {
                for (int a = 0; a < 3; ++a) {
                    float f = floorf(ext[a] / bin);
                    // `as usize` saturates (NaN, negative -> 0), then +1, clamp [1, 16384].
                    uint32_t d = (f >= (float)kGridDimMax) ? kGridDimMax
                                 : (f > 0.0f ? (uint32_t)f + 1u : 1u);
                    dims[a] = min(max(d, 1u), kGridDimMax);
                }
                if ((uint64_t)dims[0] * dims[1] * dims[2] <= sp.max_bins) break;
                bin *= 2.0f;
            }
            SceneGrid &g = P.grid[s];
            // A box the rule cannot bring under budget (non-finite extent)
            // abandons the substep instead of indexing past the tables.
            if ((uint64_t)dims[0] * dims[1] * dims[2] > sp.max_bins) {
                atomicOr(&bad, 1u);
                dims[0] = dims[1] = dims[2] = 1;
            }
            g.gmin[0] = lo[0]; g.gmin[1] = lo[1]; g.gmin[2] = lo[2];
            g.bin = bin;
            g.inv = 1.0f / bin;
            g.dims[0] = dims[0]; g.dims[1] = dims[1]; g.dims[2] = dims[2];
            g.total = dims[0] * dims[1] * dims[2];
            need = g.total + 1u;  // + pad cell
        }
        // Block-wide inclusive scan of `need` (Hillis-Steele in LDS).
        scan[threadIdx.x] = need;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            uint32_t v = threadIdx.x >= o ? scan[threadIdx.x - o] : 0u;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (s < P.nscenes) {
            const uint32_t base = carry + scan[threadIdx.x] - need;
            P.grid[s].cbase = base;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += scan[1023];
        __syncthreads();
    }
    // Capacity check, then pad-cell counts (counts are all zero on entry).
    const uint32_t ctot = carry;
    const bool over = ctot > P.ccap || bad != 0u;
    for (uint32_t s = threadIdx.x; s < P.nscenes && !over; s += 1024) {
        const SceneParams &sp = P.params[s];
        P.counts[P.grid[s].cbase + P.grid[s].total] = sp.alloc - sp.n;
    }
    // The step counter only advances for a substep that will complete.
    if (threadIdx.x == 0) {
        P.ctl->ctot = over ? 0u : ctot;
        if (bad)
            P.ctl->err |= kErrGrid;
        else if (over)
            P.ctl->err |= kErrCellCapacity;
        else
            P.ctl->step += 1u;
    }
}

// Cell key per cur slot (fluid grid.rs:80-89, z fastest) and its arrival
// rank inside the cell (atomic, order-free: canonicalised later).
__global__ __launch_bounds__(kBlock) void k_keys(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t m = b * kBlock + threadIdx.x;
        if (m >= sp.pbase + sp.n) continue;
        const SceneGrid &g = P.grid[s];
        const uint32_t ux = axis_bin(g.gmin[0], P.cx[m], g.inv, g.dims[0]);
        const uint32_t uy = axis_bin(g.gmin[1], P.cy[m], g.inv, g.dims[1]);
        const uint32_t uz = axis_bin(g.gmin[2], P.cz[m], g.inv, g.dims[2]);
        uint32_t local = (ux * g.dims[1] + uy) * g.dims[2] + uz;
        local = min(local, g.total - 1u);
        const uint32_t c = g.cbase + local;
        P.key[m] = c;
        P.slot[m] = atomicAdd(&P.counts[c], 1u);
    }
}

// --- exclusive scan of counts[0, ctot) into start[0, ctot]; counts zeroed ---
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_reduce(Pool P, uint32_t *tile_sums) {
    if (P.ctl->err) return;
    const uint32_t ctot = P.ctl->ctot;
    const uint32_t ntiles = (ctot + kScanTile - 1) / kScanTile;
    __shared__ uint32_t sh[kBlock / 64];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t acc = 0;
        const uint32_t base = t * kScanTile;
        for (uint32_t k = 0; k < kScanItems; ++k) {
            const uint32_t c = base + k * kBlock + threadIdx.x;
            if (c < ctot) acc += P.counts[c];
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor((int)acc, o, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t s = 0;
            for (uint32_t w = 0; w < kBlock / 64; ++w) s += sh[w];
            tile_sums[t] = s;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_scan_tiles(Pool P, uint32_t *tile_sums) {
    if (P.ctl->err) return;
    const uint32_t ctot = P.ctl->ctot;
    const uint32_t ntiles = (ctot + kScanTile - 1) / kScanTile;
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < ntiles; t0 += 1024) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < ntiles ? tile_sums[t] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        if (t < ntiles) tile_sums[t] = carry + sh[threadIdx.x] - v;  // exclusive
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) P.start[ctot] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(Pool P, const uint32_t *tile_sums) {
    if (P.ctl->err) return;
    const uint32_t ctot = P.ctl->ctot;
    const uint32_t ntiles = (ctot + kScanTile - 1) / kScanTile;
    __shared__ uint32_t sh[kBlock];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // Each thread owns kScanItems consecutive cells.
        const uint32_t base = t * kScanTile + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t local = 0;
        for (uint32_t k = 0; k < kScanItems; ++k) {
            const uint32_t c = base + k;
            v[k] = c < ctot ? P.counts[c] : 0u;
            local += v[k];
        }
        sh[threadIdx.x] = local;
        __syncthreads();
        for (uint32_t o = 1; o < kBlock; o <<= 1) {
            uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        uint32_t run = tile_sums[t] + sh[threadIdx.x] - local;
        for (uint32_t k = 0; k < kScanItems; ++k) {
            const uint32_t c = base + k;
            if (c < ctot) {
                P.start[c] = run;
                P.counts[c] = 0u;
                run += v[k];
            }
        }
        __syncthreads();
    }
}

// Scatter every cur slot to its (not yet canonical) place inside its cell.
__global__ __launch_bounds__(kBlock) void k_place(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t m = b * kBlock + threadIdx.x;
        if (m >= sp.pbase + sp.n) continue;
        const uint32_t dst = P.start[P.key[m]] + P.slot[m];
        P.items_m[dst] = m;
        P.items_g[dst] = P.cg[m];
    }
}

// Canonical position = cell start + number of cell members with a smaller
// original index (cell order = ascending index, fluid grid.rs:53-75), then
// gather the particle into the grid-ordered arrays.
__global__ __launch_bounds__(kBlock) void k_canon_gather(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t u = b * kBlock + threadIdx.x;
        if (u >= sp.pbase + sp.n) continue;
        const uint32_t m = P.items_m[u];
        const uint32_t g = P.items_g[u];
        const uint32_t c = P.key[m];
        const uint32_t s0 = P.start[c], e0 = P.start[c + 1];
        uint32_t rank = 0;
        for (uint32_t k = s0; k < e0; ++k) rank += P.items_g[k] < g ? 1u : 0u;
        const uint32_t d = s0 + rank;
        P.sx[d] = P.cx[m];
        P.sy[d] = P.cy[m];
        P.sz[d] = P.cz[m];
        P.svx[d] = P.cvx[m];
        P.svy[d] = P.cvy[m];
        P.svz[d] = P.cvz[m];
        P.sm[d] = P.cm[m];
        P.sg[d] = g;
    }
}

// Density (fluid_density.rs:17-70, no providers) + Tait (fluid_force.rs:11-29,
// gamma = 7) + the per-particle pressure term p / (rho * rho) used by the force.
__global__ __launch_bounds__(kBlock) void k_density(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * kBlock + threadIdx.x;
        if (i >= sp.pbase + sp.n) continue;
        const SceneGrid &g = P.grid[s];
        const float px = P.sx[i], py = P.sy[i], pz = P.sz[i];
        const bool fluid_i = P.sg[i] < sp.nf;
        const int cx = (int)axis_bin(g.gmin[0], px, g.inv, g.dims[0]);
        const int cy = (int)axis_bin(g.gmin[1], py, g.inv, g.dims[1]);
        const int cz = (int)axis_bin(g.gmin[2], pz, g.inv, g.dims[2]);
        const int nx = (int)g.dims[0], ny = (int)g.dims[1], nz = (int)g.dims[2];
        const float h2 = sp.h2, k6 = sp.k6;
        float sum = fluid_i ? sp.w0 : 0.0f;
        float sum_b = 0.0f;
        uint32_t cnt = 0;
        for (int dz = -1; dz <= 1; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= nz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= ny) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = cx + dx;
                    if (x < 0 || x >= nx) continue;
                    const uint32_t c = g.cbase + ((uint32_t)x * (uint32_t)ny + (uint32_t)y) * (uint32_t)nz + (uint32_t)z;
                    const uint32_t j0 = P.start[c], j1 = P.start[c + 1];
                    for (uint32_t j = j0; j < j1; ++j) {
                        if (j == i) continue;
                        const float ddx = px - P.sx[j], ddy = py - P.sy[j], ddz = pz - P.sz[j];
                        const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (!(r2 <= h2)) continue;
                        ++cnt;
                        const bool fluid_j = P.sg[j] < sp.nf;
                        const float t = h2 - r2;
                        if (fluid_i) {
                            const float w = ((k6 * t) * t) * t;
                            if (fluid_j) sum += w;
                            else sum_b += P.sm[j] * w;
                        } else if (fluid_j) {
                            sum += P.sm[j] * (((k6 * t) * t) * t);
                        }
                    }
                }
            }
        }
        const float rho = fluid_i ? sp.mass * sum + sum_b : P.sm[i] * sp.w0 + sum;
        const float q = rho / sp.rho0;
        const float q2 = q * q;
        const float q4 = q2 * q2;
        float p = sp.b_tait * ((q * q2) * q4 - 1.0f);
        if (sp.clamp_neg && p < 0.0f) p = 0.0f;
        P.dens[i] = rho;
        P.press[i] = p;
        P.c2[i] = p / (rho * rho);
        if (P.ncount) P.ncount[i] = cnt;
    }
}

// Force (fluid_force.rs:165-217: gravity + spiky pressure + Monaghan AV, XSPH)
// fused with the semi-implicit Euler + CFL clamp of fluid_step.rs:110-123.
// Writes the next state into the cur arrays (this substep's grid order) and
// accumulates the next substep's AABB. Boundary slots are carried over.
__global__ __launch_bounds__(kBlock) void k_force(Pool P) {
    if (P.ctl->err) return;
    const uint32_t par = P.ctl->step & 1u;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * kBlock + threadIdx.x;
        const bool live = i < sp.pbase + sp.n;
        float ox = 0, oy = 0, oz = 0;
        if (live) {
            const uint32_t gi = P.sg[i];
            const float px = P.sx[i], py = P.sy[i], pz = P.sz[i];
            const float vix = P.svx[i], viy = P.svy[i], viz = P.svz[i];
            if (gi < sp.nf) {
                const SceneGrid &g = P.grid[s];
                const int cx = (int)axis_bin(g.gmin[0], px, g.inv, g.dims[0]);
                const int cy = (int)axis_bin(g.gmin[1], py, g.inv, g.dims[1]);
                const int cz = (int)axis_bin(g.gmin[2], pz, g.inv, g.dims[2]);
                const int nx = (int)g.dims[0], ny = (int)g.dims[1], nz = (int)g.dims[2];
                const float h = sp.h, h2 = sp.h2, k6 = sp.k6, ks = sp.ks, alpha_c = sp.alpha_c;
                const float rho_i = P.dens[i];
                const float ci2 = P.c2[i];
                float ax = sp.gx, ay = sp.gy, az = sp.gz;
                float xx = 0.0f, xy = 0.0f, xz = 0.0f;
                for (int dz = -1; dz <= 1; ++dz) {
                    const int z = cz + dz;
                    if (z < 0 || z >= nz) continue;
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int y = cy + dy;
                        if (y < 0 || y >= ny) continue;
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int x = cx + dx;
                            if (x < 0 || x >= nx) continue;
                            const uint32_t c = g.cbase + ((uint32_t)x * (uint32_t)ny + (uint32_t)y) * (uint32_t)nz + (uint32_t)z;
                            const uint32_t j0 = P.start[c], j1 = P.start[c + 1];
                            for (uint32_t j = j0; j < j1; ++j) {
                                if (j == i) continue;
                                const float ddx = px - P.sx[j], ddy = py - P.sy[j], ddz = pz - P.sz[j];
                                const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                                if (!(r2 <= h2)) continue;
                                const float rho_j = P.dens[j];
                                const float cj2 = P.c2[j];
                                const float r = sqrtf(r2);
                                const float t = h - r;
                                const float mj = P.sm[j];
                                const float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
                                const float denom = fmaxf(r, 1e-9f);
                                const float qp = coef / denom;
                                ax = ax + ddx * qp;
                                ay = ay + ddy * qp;
                                az = az + ddz * qp;
                                const float vjx = P.svx[j], vjy = P.svy[j], vjz = P.svz[j];
                                const float wx = vix - vjx, wy = viy - vjy, wz = viz - vjz;
                                const float vdn = -(wx * ddx + wy * ddy + wz * ddz);
                                if (vdn > 0.0f) {
                                    const float mu = (vdn * h) / (r2 + 0.01f * h2);
                                    const float cv = (mj * ((alpha_c * mu) / (0.5f * (rho_i + rho_j)))) * ((ks * t) * t);
                                    const float qv = cv / denom;
                                    ax = ax + ddx * qv;
                                    ay = ay + ddy * qv;
                                    az = az + ddz * qv;
                                }
                                const float tt = h2 - r2;
                                const float w = ((k6 * tt) * tt) * tt;
                                const float fx = ((mj * 2.0f) / (rho_i + rho_j)) * w;
                                xx = xx + (vjx - vix) * fx;
                                xy = xy + (vjy - viy) * fx;
                                xz = xz + (vjz - viz) * fx;
                            }
                        }
                    }
                }
                if (P.ax) {
                    P.ax[i] = ax; P.ay[i] = ay; P.az[i] = az;
                    P.xx[i] = xx; P.xy[i] = xy; P.xz[i] = xz;
                }
                // fluid_step.rs:110-123
                const float dt = sp.dt, eps = sp.eps;
                float vx = (vix + ax * dt) + xx * eps;
                float vy = (viy + ay * dt) + xy * eps;
                float vz = (viz + az * dt) + xz * eps;
                const float s2 = vx * vx + vy * vy + vz * vz;
                if (s2 > sp.vmax2) {
                    const float k = sp.vmax / sqrtf(s2);
                    vx = vx * k; vy = vy * k; vz = vz * k;
                }
                ox = px + vx * dt;
                oy = py + vy * dt;
                oz = pz + vz * dt;
                P.cvx[i] = vx; P.cvy[i] = vy; P.cvz[i] = vz;
            } else {
                ox = px; oy = py; oz = pz;
                P.cvx[i] = vix; P.cvy[i] = viy; P.cvz[i] = viz;
            }
            P.cx[i] = ox; P.cy[i] = oy; P.cz[i] = oz;
            P.cm[i] = P.sm[i];
            P.cg[i] = gi;
        }
        block_box(live, ox, oy, oz, &P.box[par * P.nscenes + s]);
    }
}

// Reaction on every boundary particle of a scene with reactions:
// bforce[j] = 0 - X_1 - X_2 - ... over its fluid neighbours i in ascending
// index, X = d * ((m_i * (coef + cv)) / denom) with d = p_i - p_j (the
// reference accumulates it from the fluid side, i ascending:
// fluid_force.rs:196-206). Fluid neighbours are collected with their index
// and subtracted in index order, so the sum is the reference's bit for bit.
__global__ __launch_bounds__(kBlock) void k_bforce(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        if (!sp.react) continue;
        const uint32_t j = b * kBlock + threadIdx.x;
        if (j >= sp.pbase + sp.n) continue;
        const uint32_t gj = P.sg[j];
        if (gj < sp.nf) continue;
        const SceneGrid &g = P.grid[s];
        const float px = P.sx[j], py = P.sy[j], pz = P.sz[j];
        const float vjx = P.svx[j], vjy = P.svy[j], vjz = P.svz[j];
        const float rho_j = P.dens[j], cj2 = P.c2[j], mj = P.sm[j];
        const int cx = (int)axis_bin(g.gmin[0], px, g.inv, g.dims[0]);
        const int cy = (int)axis_bin(g.gmin[1], py, g.inv, g.dims[1]);
        const int cz = (int)axis_bin(g.gmin[2], pz, g.inv, g.dims[2]);
        const int nx = (int)g.dims[0], ny = (int)g.dims[1], nz = (int)g.dims[2];
        const float h = sp.h, h2 = sp.h2, ks = sp.ks, alpha_c = sp.alpha_c;
        uint32_t kid[kReactMaxNeighbours];
        float kx[kReactMaxNeighbours], ky[kReactMaxNeighbours], kz[kReactMaxNeighbours];
        uint32_t cnt = 0;
        bool over = false;
        for (int dz = -1; dz <= 1; ++dz) {
            const int z = cz + dz;
            if (z < 0 || z >= nz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int y = cy + dy;
                if (y < 0 || y >= ny) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = cx + dx;
                    if (x < 0 || x >= nx) continue;
                    const uint32_t c = g.cbase + ((uint32_t)x * (uint32_t)ny + (uint32_t)y) * (uint32_t)nz + (uint32_t)z;
                    const uint32_t i0 = P.start[c], i1 = P.start[c + 1];
                    for (uint32_t i = i0; i < i1; ++i) {
                        const uint32_t gi = P.sg[i];
                        if (gi >= sp.nf) continue;  // fluid neighbours only (cells list fluid first)
                        // d = p_i - p_j as seen from the fluid particle i.
                        const float ddx = P.sx[i] - px, ddy = P.sy[i] - py, ddz = P.sz[i] - pz;
                        const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (!(r2 <= h2)) continue;
                        const float rho_i = P.dens[i], ci2 = P.c2[i], mi = P.sm[i];
                        const float r = sqrtf(r2);
                        const float t = h - r;
                        const float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
                        const float denom = fmaxf(r, 1e-9f);
                        const float wx = P.svx[i] - vjx, wy = P.svy[i] - vjy, wz = P.svz[i] - vjz;
                        const float vdn = -(wx * ddx + wy * ddy + wz * ddz);
                        float cv = 0.0f;
                        if (vdn > 0.0f) {
                            const float mu = (vdn * h) / (r2 + 0.01f * h2);
                            cv = (mj * ((alpha_c * mu) / (0.5f * (rho_i + rho_j)))) * ((ks * t) * t);
                        }
                        const float q = (mi * (coef + cv)) / denom;
                        if (cnt == kReactMaxNeighbours) {
                            over = true;
                            continue;
                        }
                        // insertion by original index
                        uint32_t k = cnt++;
                        while (k > 0 && kid[k - 1] > gi) {
                            kid[k] = kid[k - 1];
                            kx[k] = kx[k - 1]; ky[k] = ky[k - 1]; kz[k] = kz[k - 1];
                            --k;
                        }
                        kid[k] = gi;
                        kx[k] = ddx * q; ky[k] = ddy * q; kz[k] = ddz * q;
                    }
                }
            }
        }
        if (over) atomicOr(&P.ctl->overflow, 1u);
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        for (uint32_t k = 0; k < cnt; ++k) {
            fx = fx - kx[k];
            fy = fy - ky[k];
            fz = fz - kz[k];
        }
        const uint32_t d = sp.pbase + gj;
        P.bfx[d] = fx; P.bfy[d] = fy; P.bfz[d] = fz;
        P.bpx[d] = px; P.bpy[d] = py; P.bpz[d] = pz;
    }
}

// Per span: F = sum bforce[k], tau += (pos[k] - origin) x f_k in index order
// (fluid_boundary.rs:10-31), accumulated over the substeps of a tick and
// scaled by 1/substeps at its end (fluid_step.rs:75-89).
__global__ __launch_bounds__(64) void k_reactions(Pool P) {
    if (P.ctl->err) return;
    const uint32_t k = blockIdx.x * 64 + threadIdx.x;
    if (k >= P.nspans) return;
    const SpanDev sd = P.spans[k];
    float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f, Tx = 0.0f, Ty = 0.0f, Tz = 0.0f;
    for (uint32_t q = sd.start; q < sd.end; ++q) {
        const float fx = P.bfx[q], fy = P.bfy[q], fz = P.bfz[q];
        Fx = Fx + fx; Fy = Fy + fy; Fz = Fz + fz;
        const float rx = P.bpx[q] - sd.origin[0], ry = P.bpy[q] - sd.origin[1], rz = P.bpz[q] - sd.origin[2];
        Tx = Tx + (ry * fz - rz * fy);
        Ty = Ty + (rz * fx - rx * fz);
        Tz = Tz + (rx * fy - ry * fx);
    }
    const uint32_t sub = sd.substeps;
    const uint32_t phase = (P.ctl->step - 1u) % sub;  // substep index inside the tick
    float *o = P.react + 6ull * k;
    if (phase == 0) {
        o[0] = Fx; o[1] = Fy; o[2] = Fz; o[3] = Tx; o[4] = Ty; o[5] = Tz;
    } else {
        o[0] = o[0] + Fx; o[1] = o[1] + Fy; o[2] = o[2] + Fz;
        o[3] = o[3] + Tx; o[4] = o[4] + Ty; o[5] = o[5] + Tz;
    }
    if (sub > 1 && phase == sub - 1) {
        const float inv = 1.0f / (float)sub;
        for (int a = 0; a < 6; ++a) o[a] = o[a] * inv;
    }
}

// Scatter cur state to index order (pool slot pbase + gid).
__global__ __launch_bounds__(kBlock) void k_unsort(Pool P, float *ox, float *oy, float *oz,
                                                   float *ovx, float *ovy, float *ovz) {
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t m = b * kBlock + threadIdx.x;
        if (m >= sp.pbase + sp.n) continue;
        const uint32_t d = sp.pbase + P.cg[m];
        ox[d] = P.cx[m]; oy[d] = P.cy[m]; oz[d] = P.cz[m];
        ovx[d] = P.cvx[m]; ovy[d] = P.cvy[m]; ovz[d] = P.cvz[m];
    }
}

// Scatter sorted per-particle fields to index order (pool slot pbase + gid).
__global__ __launch_bounds__(kBlock) void k_unsort_fields(Pool P, float *odens, float *opress,
                                                          uint32_t *oncount, float *oacc, float *oxs) {
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * kBlock + threadIdx.x;
        if (i >= sp.pbase + sp.n) continue;
        const uint32_t gi = P.sg[i];
        const uint32_t d = sp.pbase + gi;
        odens[d] = P.dens[i];
        opress[d] = P.press[i];
        if (oncount && P.ncount) oncount[d] = P.ncount[i];
        if (oacc && P.ax && gi < sp.nf) {
            oacc[3 * d] = P.ax[i]; oacc[3 * d + 1] = P.ay[i]; oacc[3 * d + 2] = P.az[i];
            oxs[3 * d] = P.xx[i]; oxs[3 * d + 1] = P.xy[i]; oxs[3 * d + 2] = P.xz[i];
        }
    }
}

}  // namespace vxl
