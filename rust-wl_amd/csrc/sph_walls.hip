// gfx950 kernels of provider walls (crates/vxl-phys-core/src/interop.rs:166-224):
// mirror-ghost density, projection of penetrating particles, coincident
// separation. Two sources of wall planes per scene (sph_layout.h):
//
//   half-space providers, queried here every substep with the CPU engine's
//   semantics (fluid_density.rs:17-51, :76-121; fluid_boundary.rs:45-131);
//   plane tables gathered by the host once per tick, with the GPU pipeline's
//   semantics (wall_ghost.wgsl:61-178; walls.rs:104-133).
//
// Same arithmetic contract as the substep kernels: strict f32, every Vec3
// operation written out in the reference's component order (math.rs:54-70),
// neighbours in the reference enumeration order.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_kernels.h"
#include "sph_walls_dev.h"

namespace vxl {

namespace {

using dev::at;
using dev::wave_box;
using dev::Cursor;
using dev::for_rows;
using dev::make_cursor;
using dev::tait;
using namespace wall;

__device__ __forceinline__ uint32_t entry_of(const Pool &P, uint32_t side, uint32_t pbase, uint32_t gid) {
    return P.went[(size_t)side * P.nblocks * kBlock + pbase + gid];
}

}  // namespace

// The fluid particles of table scenes with a non-empty mirror entry, listed
// (wave-aggregated appends) for k_wall_density. (Half-space scenes take their
// ghosts inside the density pass.)
__global__ __launch_bounds__(kBlock) void k_wall_band(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const WallSceneDev &W = P.walls[s];
        if (W.mode != kWallsTable) continue;
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * kBlock + threadIdx.x;
        bool band = false;
        if (i < sp.pbase + sp.n && P.sA[i].w > 0.0f) {
            const uint32_t e = entry_of(P, 0, sp.pbase, P.sg[i]);
            band = e != kWallNone && W.tstart[0][e] < W.tstart[0][e + 1];
        }
        const uint64_t m = __ballot(band);
        if (m == 0) continue;
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t lead = (uint32_t)__builtin_ctzll(m);
        uint32_t base = 0;
        if (lane == lead) base = atomicAdd(P.wband_n, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, (int)lead, 64);
        if (band) P.wband[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    }
}

// Mirror ghosts of the listed table-scene particles (wall_ghost.wgsl:61-136):
// the entry's planes whose band holds pi now (0 < sdf < h); the ghosts of all
// neighbours summed apart in neighbour order, then dens += mass * add and the
// Tait pressure and pressure term again.
__global__ __launch_bounds__(kBlock) void k_wall_density(Pool P) {
    if (P.ctl->err) return;
    const uint32_t count = *P.wband_n;
    for (uint32_t u = blockIdx.x * kBlock + threadIdx.x; u < count; u += gridDim.x * kBlock) {
        const uint32_t i = P.wband[u];
        const uint32_t s = P.block_scene[i / kBlock];
        const WallSceneDev &W = P.walls[s];
        const SceneParams &sp = P.params[s];
        const float4 ai = P.sA[i];
        const V3 pi = xyz(ai);
        const SceneGrid &g = P.grid[s];
        const float h = sp.h, h2 = sp.h2, k6 = sp.k6;
        const uint32_t e = entry_of(P, 0, sp.pbase, P.sg[i]);
        const uint32_t st = W.tstart[0][e], en = W.tstart[0][e + 1];
        const float4 *__restrict__ tp = (const float4 *)W.tplane[0];
        const Cursor cur = make_cursor(g, h, ai.x, ai.y, ai.z);
        float acc = 0.0f;
        for_rows(g, P.start, cur, [&](uint32_t, bool, uint32_t j0, uint32_t j1) {
            for (uint32_t j = j0; j < j1; ++j) {
                if (j == i) continue;
                const float4 aj = *at(P.sA, j);
                const V3 d{ai.x - aj.x, ai.y - aj.y, ai.z - aj.z};
                const float r2 = dot(d, d);
                if (!(r2 <= h2)) continue;
                for (uint32_t q = st; q < en; ++q) {
                    const V3 pt = xyz(tp[2 * q]), n = xyz(tp[2 * q + 1]);
                    const float sdf_i = dot(sub(pi, pt), n);
                    if (sdf_i <= 0.0f || sdf_i >= h) continue;
                    float wg;
                    if (ghost_w(pi, d, pt, n, h2, k6, wg)) acc = acc + wg;
                }
            }
        });
        if (acc == 0.0f) continue;
        const float rho = P.fr[2u * i + 1u].w + sp.mass * acc;  // a fluid particle: rho positive
        float p, c2;
        tait(sp, rho, p, c2);
        P.fr[2u * i].w = c2;
        P.fr[2u * i + 1u].w = rho;
    }
}

// Projection of the integrated fluid state (cur arrays, this substep's grid
// order).
//
// Half-space walls (boundary_pass, fluid_boundary.rs:45-104): contacts of the
// integrated position pi with probe skin = 0.15 h; each with sdf(pi) < 0
// pushes the particle by min(pen + skin, h) along its normal and removes the
// inward normal velocity. A pushed particle is appended to the scene's list
// for the coincident separation.
// Plane tables (wall_project, wall_ghost.wgsl:137-178): the entry's planes in
// order against the position as pushed so far.
__global__ __launch_bounds__(kBlock) void k_wall_project(Pool P) {
    if (P.ctl->err) return;
    const uint32_t tick_sub = P.ctl->tick_sub, tick_len = P.ctl->tick_len;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const WallSceneDev &W = P.walls[s];
        if (W.mode == kWallsOff) continue;
        const SceneParams &sp = P.params[s];
        if (!substep_active(tick_sub, tick_len, sp.substeps)) continue;  // not integrated in this launch
        const uint32_t i = b * kBlock + threadIdx.x;
        if (i >= sp.pbase + sp.n) continue;
        const float4 a = P.cA[i];
        if (!(a.w > 0.0f)) continue;
        const float4 va = P.cV[i];
        const float h = sp.h, skin = 0.15f * sp.h;
        const V3 pi = xyz(a);
        V3 p = pi, v = xyz(va);
        bool hit = false;
        if (W.mode == kWallsHalfspace) {
            V3 n_sum{0.0f, 0.0f, 0.0f};
            for (uint32_t k = 0; k < W.nprov; ++k) {
                const WallProvDev &pv = W.prov[k];
                if (outside(pv, pi, h)) continue;
                prov_contacts(W, pv, pi, skin, [&](uint32_t, V3 point, V3 n) {
                    const float sdf = dot(sub(pi, point), n);
                    const float pen = -sdf;
                    if (pen <= 0.0f) return;
                    p = add(p, mul(n, fminf(pen + skin, h)));
                    n_sum = add(n_sum, n);
                    hit = true;
                    const float vn = dot(v, n);
                    if (vn < 0.0f) v = sub(v, mul(n, vn));
                });
            }
            // Append the pushed ones to the scene's list (one atomic per wave).
            const uint64_t mk = __ballot(hit);
            if (mk) {
                const uint32_t lane = threadIdx.x & 63u, lead = (uint32_t)__builtin_ctzll(mk);
                uint32_t base = 0;
                if (lane == lead) {
                    base = atomicAdd(&P.sep[s].k, (uint32_t)__popcll(mk));
                    atomicAdd(&P.sep[P.nscenes].k, (uint32_t)__popcll(mk));
                }
                base = (uint32_t)__shfl((int)base, (int)lead, 64);
                if (hit) {
                    P.wn[i] = make_float4(n_sum.x, n_sum.y, n_sum.z, 0.0f);
                    P.wsep[2ull * sp.pbase + base + (uint32_t)__popcll(mk & ((1ull << lane) - 1ull))] =
                        ((uint64_t)P.cg[i] << 32) | i;
                }
            }
        } else {
            const uint32_t e = entry_of(P, 1, sp.pbase, P.cg[i]);
            if (e == kWallNone) continue;
            const uint32_t st = W.tstart[1][e], en = W.tstart[1][e + 1];
            const float4 *__restrict__ pl = (const float4 *)W.tplane[1];
            for (uint32_t q = st; q < en; ++q) {
                const V3 pt = xyz(pl[2 * q]), n = xyz(pl[2 * q + 1]);
                const float sdf = dot(sub(p, pt), n);
                const float pen = -sdf;
                if (pen > 0.0f) {
                    const float push = fminf(pen + skin, h);
                    p = add(p, mul(n, push));
                    const float vn = dot(v, n);
                    if (vn < 0.0f) v = sub(v, mul(n, vn));
                    hit = true;
                }
            }
        }
        if (hit) {
            P.cA[i] = make_float4(p.x, p.y, p.z, a.w);
            P.cV[i] = make_float4(v.x, v.y, v.z, va.w);
        }
    }
}

namespace {

// ---- coincident separation ----------------------------------------------
//
// separate_coincident (fluid_boundary.rs:110-131) takes the pushed particles
// in ascending fluid index; particle a starts at its projected position q_a,
// tests the earlier particles b in ascending index against its current
// position and, at each b closer than r_min = 0.2 h (b at its final
// position), moves by (r_min - r) along its normal sum, then goes on with the
// next b. Nothing but a itself moves a.
//
// Between two moves a's position is fixed, so its next move is at the
// smallest index b above the last one whose final position lies within r_min:
// a's result depends only on the final positions of nearby earlier particles.
// The device finds it with a fixed-point iteration: round m recomputes every
// pushed particle's result from the positions of round m - 1 (round 0 from
// the projected ones), finding the candidates through a spatial hash of those
// positions. The sequential result is the unique fixed point (by induction
// over the index order), so a round that changes nothing has reached it.
// Scenes that have not converged after kSepIters rounds take the serial path.

constexpr uint32_t kSepLds = 2048;  // serial path: pushed particles sorted and separated in LDS

// Hash buckets of a scene with k pushed particles: a power of two, at most
// 2 alloc (the scene's share of the bucket arrays), at least the smaller of
// 2k and that.
__device__ __forceinline__ uint32_t sep_buckets(uint32_t k, uint32_t alloc) {
    uint32_t cap = 1;
    while (cap * 2u <= 2u * alloc) cap *= 2u;
    uint32_t hb = 64;
    while (hb < 2u * k) hb *= 2u;
    return min(hb, cap);
}

// Hash cells of side 2.0625 r_min, computed in f64 (exact enough at any
// coordinate): every point closer than r_min to p lies in the 2 x 2 x 2
// block of cells nearest p.
constexpr double kSepCell = 2.0625;
__device__ __forceinline__ int sep_coord(double t) {
    return (int)fmin(fmax(floor(t), -1073741824.0), 1073741824.0);  // NaN -> lower bound
}
struct SepCell {
    int x, y, z;
};
__device__ __forceinline__ SepCell sep_cell(float4 p, double inv) {
    return SepCell{sep_coord((double)p.x * inv), sep_coord((double)p.y * inv), sep_coord((double)p.z * inv)};
}
// Lower corner of the 2 x 2 x 2 block around p.
__device__ __forceinline__ SepCell sep_block(V3 p, double inv) {
    auto lo = [&](float v) {
        const double t = (double)v * inv;
        const int c = sep_coord(t);
        return (t - floor(t)) < 0.5 ? c - 1 : c;
    };
    return SepCell{lo(p.x), lo(p.y), lo(p.z)};
}
__device__ __forceinline__ uint32_t sep_bucket(int x, int y, int z, uint32_t nb) {
    return (((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u)) & (nb - 1u);
}

// The pushed particle entry e of scene s handled by pool slot i, if any.
struct SepEntry {
    bool live;
    uint32_t s, e, k, nb;
    const SceneParams *sp;
};
__device__ __forceinline__ SepEntry sep_entry(const Pool &P, uint32_t b) {
    SepEntry t{false, 0, 0, 0, 0, nullptr};
    const uint32_t s = P.block_scene[b];
    if (P.walls[s].mode != kWallsHalfspace) return t;
    const SceneParams &sp = P.params[s];
    t.s = s;
    t.sp = &sp;
    t.k = P.sep[s].k;
    t.e = b * kBlock + threadIdx.x - sp.pbase;
    t.nb = sep_buckets(t.k, sp.alloc);
    t.live = t.e < t.k;
    return t;
}

__device__ __forceinline__ void sep_insert(const Pool &P, uint32_t copy, const SceneParams &sp, uint32_t nb, uint32_t e,
                                           float4 p, double inv) {
    const SepCell c = sep_cell(p, inv);
    uint32_t *head = P.hhead + (size_t)copy * 2u * P.wslots + 2ull * sp.pbase;
    P.hnext[(size_t)copy * P.wslots + sp.pbase + e] = atomicExch(&head[sep_bucket(c.x, c.y, c.z, nb)], e);
}

}  // namespace

// Round 0's positions (the projected ones) and their hash (copy 0).
__global__ __launch_bounds__(kBlock) void k_sep_begin(Pool P) {
    if (P.ctl->err || P.sep[P.nscenes].k == 0) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const SepEntry t = sep_entry(P, b);
        if (!t.live) continue;
        const uint32_t slot = (uint32_t)P.wsep[2ull * t.sp->pbase + t.e];
        const float4 q = P.cA[slot];
        P.fpos[t.sp->pbase + t.e] = q;
        sep_insert(P, 0, *t.sp, t.nb, t.e, q, 1.0 / (kSepCell * (double)(0.2f * t.sp->h)));
    }
}

// Round m: every pushed particle's separation from the round-m positions of
// the others (hash copy m % 3); its result goes to position copy (m + 1) % 2
// and hash copy (m + 1) % 3; hash copy (m + 2) % 3 is cleared for round m + 1.
__global__ __launch_bounds__(kBlock) void k_sep_round(Pool P, uint32_t m) {
    if (P.ctl->err || P.sep[P.nscenes].k == 0) return;
    if (m > 0 && P.sep[P.nscenes].changed[m - 1] == 0u) return;  // every scene converged
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const SepEntry t = sep_entry(P, b);
        if (t.k == 0 || !t.sp) continue;
        const SepState &st = P.sep[t.s];
        if (m > 0 && st.changed[m - 1] == 0u) continue;  // converged
        const SceneParams &sp = *t.sp;
        const uint32_t alloc = sp.alloc;
        {
            // clear the scene's buckets of copy (m + 2) % 3 (2 alloc slots per copy)
            uint32_t *clr = P.hhead + (size_t)((m + 2u) % 3u) * 2u * P.wslots + 2ull * sp.pbase;
            const uint32_t u = b * kBlock + threadIdx.x - sp.pbase;
            if (u < alloc) {
                if (u < t.nb) clr[u] = kWallNone;
                if (u + alloc < t.nb) clr[u + alloc] = kWallNone;
            }
        }
        if (!t.live) continue;
        const float r_min = 0.2f * sp.h;
        const double inv = 1.0 / (kSepCell * (double)r_min);
        const float4 *cur = P.fpos + (size_t)(m % 2u) * P.wslots + sp.pbase;
        float4 *nxt = P.fpos + (size_t)((m + 1u) % 2u) * P.wslots + sp.pbase;
        const uint32_t copy = m % 3u;
        const uint32_t *head = P.hhead + (size_t)copy * 2u * P.wslots + 2ull * sp.pbase;
        const uint32_t *next = P.hnext + (size_t)copy * P.wslots + sp.pbase;
        const uint64_t *list = P.wsep + 2ull * sp.pbase;
        const uint64_t me = list[t.e];
        const uint32_t gid_a = (uint32_t)(me >> 32), ia = (uint32_t)me;
        const float4 q = P.cA[ia];  // the projected position (written back only after the rounds)
        float4 out = q;
        const V3 na = xyz(P.wn[ia]);
        const float nl = sqrtf(dot(na, na));
        if (!(nl <= 1e-9f)) {
            const V3 dir = mul(na, 1.0f / nl);
            V3 pa = xyz(q);
            uint32_t last = 0;  // indices at or below `last` are done (first pass: from 0)
            bool first = true;
            for (;;) {
                const SepCell c = sep_block(pa, inv);
                uint32_t hd[8];
#pragma unroll
                for (uint32_t k = 0; k < 8u; ++k)
                    hd[k] = head[sep_bucket(c.x + (int)(k & 1u), c.y + (int)((k >> 1) & 1u), c.z + (int)(k >> 2), t.nb)];
                uint32_t best = 0xffffffffu;
                float best_r2 = 0.0f;
#pragma unroll
                for (uint32_t k = 0; k < 8u; ++k) {
                    for (uint32_t f = hd[k]; f != kWallNone; f = next[f]) {
                        const uint32_t gid_b = (uint32_t)(list[f] >> 32);
                        if (gid_b >= gid_a || (!first && gid_b <= last) || gid_b >= best) continue;
                        const V3 d = sub(pa, xyz(cur[f]));
                        const float r2 = dot(d, d);
                        if (r2 < r_min * r_min) {
                            best = gid_b;
                            best_r2 = r2;
                        }
                    }
                }
                if (best == 0xffffffffu) break;
                const float r = sqrtf(best_r2);
                pa = add(pa, mul(dir, r_min - r));
                last = best;
                first = false;
            }
            out = make_float4(pa.x, pa.y, pa.z, q.w);
        }
        nxt[t.e] = out;
        const float4 was = cur[t.e];
        if (__float_as_uint(out.x) != __float_as_uint(was.x) || __float_as_uint(out.y) != __float_as_uint(was.y) ||
            __float_as_uint(out.z) != __float_as_uint(was.z))
        {
            P.sep[t.s].changed[m] = 1u;
            P.sep[P.nscenes].changed[m] = 1u;
        }
        sep_insert(P, (m + 1u) % 3u, sp, t.nb, t.e, out, inv);
    }
}

// Converged scenes: the separated positions back into the state; every
// scene's buckets cleared for the next substep.
__global__ __launch_bounds__(kBlock) void k_sep_finish(Pool P) {
    if (P.ctl->err || P.sep[P.nscenes].k == 0) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const SepEntry t = sep_entry(P, b);
        if (t.k == 0 || !t.sp) continue;
        const SceneParams &sp = *t.sp;
        const uint32_t u = b * kBlock + threadIdx.x - sp.pbase;
        if (u < sp.alloc) {
            for (uint32_t copy = 0; copy < 3u; ++copy) {
                uint32_t *h = P.hhead + (size_t)copy * 2u * P.wslots + 2ull * sp.pbase;
                if (u < t.nb) h[u] = kWallNone;
                if (u + sp.alloc < t.nb) h[u + sp.alloc] = kWallNone;
            }
        }
        if (!t.live) continue;
        const SepState &st = P.sep[t.s];
        uint32_t m = 0;
        while (m < kSepIters && st.changed[m] != 0u) ++m;
        if (m == kSepIters) continue;  // not converged: the serial path
        const float4 f = P.fpos[(size_t)((m + 1u) % 2u) * P.wslots + sp.pbase + t.e];
        const uint32_t slot = (uint32_t)P.wsep[2ull * sp.pbase + t.e];
        P.cA[slot] = f;
    }
}

namespace {

// Ascending bitonic sort of n (a power of two) keys by one block.
__device__ void block_bitonic(uint64_t *key, uint32_t n) {
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < n / 2; t += blockDim.x) {
                const uint32_t lo = 2u * t - (t & (stride - 1u));
                const uint32_t hi = lo + stride;
                const bool asc = (lo & size) == 0u;
                const uint64_t x = key[lo], y = key[hi];
                if ((x > y) == asc) {
                    key[lo] = y;
                    key[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

// separate_coincident by one wave over the pushed particles in ascending
// fluid index (the serial path): the lanes test 64 earlier particles at a
// time; the first one that triggers is applied and the scan resumes after it.
// pos(t) reads and set_pos(t, v) writes the position of list entry t.
template <class Get, class Set>
__device__ void separate_wave(const Pool &P, const uint64_t *key, uint32_t k, float h, Get &&pos, Set &&set_pos) {
    const uint32_t lane = threadIdx.x & 63u;
    const float r_min = 0.2f * h;
    for (uint32_t a = 0; a < k; ++a) {
        const uint32_t ia = (uint32_t)key[a];
        const V3 na = xyz(P.wn[ia]);
        const float nl = sqrtf(dot(na, na));
        if (nl <= 1e-9f) continue;
        const V3 dir = mul(na, 1.0f / nl);
        V3 pa = pos(a);
        bool moved = false;
        for (uint32_t b0 = 0; b0 < a;) {
            const uint32_t bb = b0 + lane;
            float r2 = 0.0f;
            bool trig = false;
            if (bb < a) {
                const V3 d = sub(pa, pos(bb));
                r2 = dot(d, d);
                trig = r2 < r_min * r_min;
            }
            const uint64_t msk = __ballot(trig);
            if (msk == 0) {
                b0 += 64u;
                continue;
            }
            const uint32_t f = (uint32_t)__builtin_ctzll(msk);
            const float rf = __shfl(r2, (int)f, 64);
            const float r = sqrtf(rf);
            pa = add(pa, mul(dir, r_min - r));
            moved = true;
            b0 += f + 1u;
        }
        if (moved) set_pos(a, pa);
    }
}

}  // namespace

// Serial coincident separation of the half-space scenes whose parallel rounds
// did not converge, one block per scene: the pushed particles (appended in
// any order by k_wall_project) sorted by fluid index, then separated by one
// wave. Up to kSepLds of them are sorted and moved in LDS; more are sorted in
// place in the scene's list (2 alloc entries, padded to a power of two) and
// moved in global memory. Every scene's state is reset for the next substep.
__global__ __launch_bounds__(kBlock) void k_wall_separate(Pool P) {
    __shared__ uint64_t skey[kSepLds];
    __shared__ float4 spos[kSepLds];
    if (P.ctl->err || P.sep[P.nscenes].k == 0) return;
    const uint32_t s = blockIdx.x;
    if (s >= P.nscenes || P.walls[s].mode != kWallsHalfspace) return;
    SepState &st = P.sep[s];
    const uint32_t k = st.k;
    if (k == 0) return;
    bool converged = false;
    for (uint32_t m = 0; m < kSepIters; ++m) converged = converged || st.changed[m] == 0u;
    if (!converged) {
        const SceneParams &sp = P.params[s];
        uint64_t *list = P.wsep + 2ull * sp.pbase;
        uint32_t n2 = 1;
        while (n2 < k) n2 <<= 1;
        const bool lds = n2 <= kSepLds;
        uint64_t *key = lds ? skey : list;
        for (uint32_t t = threadIdx.x; t < n2; t += blockDim.x) key[t] = t < k ? list[t] : ~0ull;
        __syncthreads();
        block_bitonic(key, n2);
        if (lds) {
            for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) spos[t] = P.cA[(uint32_t)skey[t]];
            __syncthreads();
            if (threadIdx.x < 64) {
                separate_wave(
                    P, skey, k, sp.h, [&](uint32_t t) { return xyz(spos[t]); },
                    [&](uint32_t t, V3 v) {
                        if ((threadIdx.x & 63u) == 0) spos[t] = make_float4(v.x, v.y, v.z, spos[t].w);
                    });
            }
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) P.cA[(uint32_t)skey[t]] = spos[t];
        } else if (threadIdx.x < 64) {
            separate_wave(
                P, list, k, sp.h, [&](uint32_t t) { return xyz(P.cA[(uint32_t)list[t]]); },
                [&](uint32_t t, V3 v) {
                    const uint32_t ia = (uint32_t)list[t];
                    if ((threadIdx.x & 63u) == 0) P.cA[ia] = make_float4(v.x, v.y, v.z, P.cA[ia].w);
                    __threadfence();
                });
        }
    }
    __syncthreads();
    if (threadIdx.x < kSepIters) st.changed[threadIdx.x] = 0u;
    if (threadIdx.x == 0) st.k = 0;
}

// Partial boxes of the wall scenes' blocks over the projected positions
// (replacing the ones the force pass wrote).
__global__ __launch_bounds__(kBlock) void k_wall_box(Pool P) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *P.wband_n = 0;  // band list of the next substep
    if (blockIdx.x == 0 && threadIdx.x <= kSepIters && P.sep) {
        // pool totals of the separation, for the next substep
        if (threadIdx.x == kSepIters)
            P.sep[P.nscenes].k = 0;
        else
            P.sep[P.nscenes].changed[threadIdx.x] = 0u;
    }
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        if (P.walls[s].mode == kWallsOff) continue;
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * kBlock + threadIdx.x;
        const bool live = i < sp.pbase + sp.n;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (live) a = P.cA[i];
        wave_box(live, a.x, a.y, a.z, &P.part[i / kPartSlots]);
    }
}

__global__ __launch_bounds__(kBlock) void k_wall_index_clear(Pool P, uint32_t s, uint32_t side) {
    const SceneParams &sp = P.params[s];
    uint32_t *w = P.went + (size_t)side * P.nblocks * kBlock + sp.pbase;
    for (uint32_t u = blockIdx.x * kBlock + threadIdx.x; u < sp.nf; u += gridDim.x * kBlock) w[u] = kWallNone;
}

__global__ __launch_bounds__(kBlock) void k_wall_index_set(Pool P, uint32_t s, uint32_t side, const uint32_t *ids,
                                                           uint32_t ne) {
    const SceneParams &sp = P.params[s];
    uint32_t *w = P.went + (size_t)side * P.nblocks * kBlock + sp.pbase;
    for (uint32_t e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock) w[ids[e]] = e;
}

}  // namespace vxl
