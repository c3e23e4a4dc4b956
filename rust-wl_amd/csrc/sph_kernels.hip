// gfx950 kernels of the WCSPH substep (one launch per stage for every scene
// of the pool; a single scene is a pool of one).
//
// Substep = grid_box -> keys -> scan (3 passes) -> place -> canon_gather ->
// density(+Tait) -> force(+AV, XSPH, integrate, next AABB) [-> bforce ->
// reactions].
//
// Arithmetic contract: strict IEEE f32 (built with -ffp-contract=off,
// correctly rounded / and sqrt, denormals kept), every expression in the
// operation order of the reference CPU step, neighbours enumerated cell by
// cell (dz, dy, dx; within a cell ascending original index), so every sum is
// the reference's sum bit for bit. Cells that cannot hold a neighbour within
// h are skipped (a conservative test with a 2 % margin), which removes only
// candidates the r^2 <= h^2 test would reject and keeps the order of the rest.
#include <hip/hip_runtime.h>

#include "sph_device.h"
#include "sph_kernels.h"
#include "sph_walls_dev.h"

namespace vxl {

namespace {

using dev::at;
using dev::axis_bin;
using dev::wave_box;
using dev::Cursor;
using dev::for_row_windows;
using dev::for_rows;
using dev::make_cursor;
using dev::tait;
using dev::wave_max_u32;
using dev::wave_min_u32;
using dev::xcd_block;

// Last cell (dx = 0 or +1) of row r as for_rows chooses it.
__device__ __forceinline__ uint32_t row_last_cell(const SceneGrid &g, const Cursor &c, uint32_t r) {
    const uint32_t q = (r * 11u) >> 5;  // r / 3 for r < 9
    const int dz = (int)q - 1, dy = (int)(r - 3u * q) - 1;
    const float ez = dz < 0 ? c.fz : (dz > 0 ? 1.0f - c.fz : 0.0f);
    const float ey = dy < 0 ? c.fy : (dy > 0 ? 1.0f - c.fy : 0.0f);
    const float rd = ez * ez + ey * ey;
    const float rx = (1.0f - c.fx) * (1.0f - c.fx);
    return (c.cx + 1 < (int)g.dims[0] && rd + rx <= c.cut) ? (uint32_t)c.cx + 1u : (uint32_t)c.cx;
}

// Accepted-neighbour record of one row, as the density pass writes it: bit k
// of the mask = candidate j0 + k passed r^2 <= h^2 (and is not i itself);
// candidates past j0 + 32 are left to the reader.
constexpr uint32_t kRowsPresent = 0x1ffu;
__device__ __forceinline__ uint32_t row_summary(uint32_t r, bool left, bool tail) {
    return (1u << r) | (left ? 1u << (9u + r) : 0u) | (tail ? 1u << (18u + r) : 0u);
}

// The nine neighbour rows of the scene a block works on, kept in LDS: cell =
// cell offset of row r from the cell (cx, cy - 1, cz - 1), mask = r * stride
// (the row's offset in the mask array). Opening a row then costs one LDS read
// instead of the row index arithmetic.
struct RowTab {
    uint32_t cell, mask;
};

__device__ __forceinline__ void row_table(RowTab *t, const SceneGrid &g, uint32_t stride) {
    if (threadIdx.x < 9u) {
        const uint32_t r = threadIdx.x, q = r / 3u;
        t[r] = RowTab{q * g.dims[0] * g.dims[1] + (r - 3u * q) * g.dims[0], r * stride};
    }
}

// Resumable walk over the neighbours the density pass accepted, in the
// reference enumeration order: rows in ascending r, inside a row ascending
// slot (= cell by cell, ascending original index).
struct NbWalk {
    uint32_t summ, pend;  // summary word; rows not yet opened
    uint32_t mk, j0;      // accepted candidates j0 + k (bit k) of the open row
    uint32_t jt, je;      // tail candidates [jt, je) of the open row, tested here
    uint32_t rowb;        // internal cell of (cx, cy - 1, cz - 1) (mod 2^32)
    const RowTab *tab;
};

__device__ __forceinline__ NbWalk nb_begin(uint32_t summ, const SceneGrid &g, const Cursor &c, const RowTab *tab) {
    NbWalk w;
    w.summ = summ;
    w.pend = summ & kRowsPresent;
    w.mk = w.j0 = w.jt = w.je = 0u;
    w.rowb = g.cbase + ((uint32_t)(c.cz - 1) * g.dims[1] + (uint32_t)(c.cy - 1)) * g.dims[0] + (uint32_t)c.cx;
    w.tab = tab;
    return w;
}

enum : int { kWalkDone = 0, kWalkPair = 1, kWalkIdle = 2 };

// One step of the walk: kWalkPair with the next neighbour in jj, kWalkIdle
// when this step opened a row without mask bits or tested a rejected tail
// candidate (more may follow), kWalkDone after the last neighbour.
__device__ __forceinline__ int nb_step(NbWalk &w, const Pool &P, const SceneGrid &g, const Cursor &c, uint32_t i,
                                       const float4 ai, float h2, uint32_t stride, uint32_t &jj) {
    if (w.mk == 0u) {
        if (w.jt < w.je) {
            const uint32_t j = w.jt++;
            const float4 aj = *at(P.sA, j);
            const float ddx = ai.x - aj.x, ddy = ai.y - aj.y, ddz = ai.z - aj.z;
            const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            jj = j;
            return (r2 <= h2 && j != i) ? kWalkPair : kWalkIdle;
        }
        if (w.pend == 0u) return kWalkDone;
        const uint32_t r = (uint32_t)__builtin_ctz(w.pend);
        w.pend &= w.pend - 1u;
        const RowTab t = w.tab[r];
        const uint32_t row = w.rowb + t.cell;
        const uint32_t first = row - ((w.summ >> (9u + r)) & 1u);
        w.j0 = *at(P.start, first);
        w.mk = P.nbm[t.mask + i];  // 36 bytes per slot: a 64-bit offset
        if ((w.summ >> (18u + r)) & 1u) {
            w.jt = w.j0 + 32u;
            w.je = P.start[row - (uint32_t)c.cx + row_last_cell(g, c, r) + 1u];
        }
        if (w.mk == 0u) return kWalkIdle;
    }
    jj = w.j0 + (uint32_t)__builtin_ctz(w.mk);
    w.mk &= w.mk - 1u;
    return kWalkPair;
}

}  // namespace

// Partial boxes of the cur positions (at upload; k_box_reduce folds them into
// the scenes' boxes).
__global__ __launch_bounds__(kBlock) void k_box(Pool P) {
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t m = b * kBlock + threadIdx.x;
        const bool live = m < sp.pbase + sp.n;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (live) a = P.cA[m];
        wave_box(live, a.x, a.y, a.z, &P.part[m / kPartSlots]);
    }
}

// A scene's boundary segment replaced from index-order device arrays (three
// floats per boundary particle, original index nf + q at 3 q): every slot of
// scene s whose original index lies in the segment takes its new position
// and velocity (the state stays in grid order), the scene's blocks rebuild
// their partial boxes, and the scene's current AABB is emptied so that the
// following k_box_reduce folds exactly the new positions into it (the
// FluidStepper's per-tick boundary upload, pipeline/coupling.rs:20-50, and
// the box refresh of pipeline.rs:557-566).
__global__ __launch_bounds__(kBlock) void k_boundary_scatter(Pool P, uint32_t s, const float *__restrict__ bpos,
                                                             const float *__restrict__ bvel) {
    const SceneParams &sp = P.params[s];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SceneBox &bx = P.box[(P.ctl->step & 1u) * P.nscenes + s];
        for (int a = 0; a < 3; ++a) {
            bx.lo[a] = 0xffffffffu;
            bx.hi[a] = 0u;
        }
    }
    const uint32_t b0 = sp.pbase / kBlock, nblk = sp.alloc / kBlock;
    for (uint32_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint32_t b = b0 + k;
        const uint32_t m = b * kBlock + threadIdx.x;
        const bool live = m < sp.pbase + sp.n;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (live) {
            a = P.cA[m];
            const uint32_t g = P.cg[m];
            if (g >= sp.nf) {
                const size_t q = 3ull * (g - sp.nf);
                a.x = bpos[q];
                a.y = bpos[q + 1];
                a.z = bpos[q + 2];
                P.cA[m] = a;
                float4 v = P.cV[m];
                v.x = bvel[q];
                v.y = bvel[q + 1];
                v.z = bvel[q + 2];
                P.cV[m] = v;
            }
        }
        wave_box(live, a.x, a.y, a.z, &P.part[m / kPartSlots]);
    }
}

// Per-scene AABB = reduction of its waves' partial boxes into box[parity]
// (reset to the empty box by k_grid_box / the layout). Each thread folds
// kBoxChunk consecutive partial boxes (scenes own contiguous slot ranges, so
// a thread meets few scene boundaries), each wave folds its lanes when they
// share one scene, and one lane per wave merges the result with atomics.
constexpr uint32_t kBoxChunk = kBoxReduceSpan / kBlock;

__device__ __forceinline__ void box_merge(SceneBox *o, const uint32_t lo[3], const uint32_t hi[3]) {
    for (int a = 0; a < 3; ++a) {
        atomicMin(&o->lo[a], lo[a]);
        atomicMax(&o->hi[a], hi[a]);
    }
}

__global__ __launch_bounds__(kBlock) void k_box_reduce(Pool P) {
    __shared__ uint32_t wsc[kBlock / 64];  // scene each wave folded (~0u none, ~1u several)
    __shared__ uint32_t wlo[kBlock / 64][3], whi[kBlock / 64][3];
    if (P.ctl->err) return;
    const uint32_t par = P.ctl->step & 1u;
    SceneBox *out = P.box + par * P.nscenes;
    const uint32_t npart = P.nblocks * (kBlock / kPartSlots);
    for (uint32_t base = blockIdx.x * kBlock * kBoxChunk; base < npart; base += gridDim.x * kBlock * kBoxChunk) {
        uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
        uint32_t cur = 0xffffffffu;  // scene of the running fold (none yet)
        const uint32_t w0 = base + threadIdx.x * kBoxChunk;
        // The thread's partial boxes and their scenes, all loads in flight at
        // once (past the end: no scene, nothing folded).
        SceneBox qs[kBoxChunk];
        uint32_t scs[kBoxChunk];
#pragma unroll
        for (uint32_t k = 0; k < kBoxChunk; ++k) {
            const uint32_t w = min(w0 + k, npart - 1u);
            qs[k] = P.part[w];
            scs[k] = w0 + k < npart ? P.block_scene[w / (kBlock / kPartSlots)] : 0xffffffffu;
        }
#pragma unroll
        for (uint32_t k = 0; k < kBoxChunk; ++k) {
            const uint32_t sc = scs[k];
            if (sc == 0xffffffffu) break;
            if (sc != cur) {
                if (cur != 0xffffffffu) box_merge(&out[cur], lo, hi);
                cur = sc;
                for (int a = 0; a < 3; ++a) {
                    lo[a] = 0xffffffffu;
                    hi[a] = 0u;
                }
            }
            const SceneBox &q = qs[k];
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], q.lo[a]);
                hi[a] = max(hi[a], q.hi[a]);
            }
        }
        const uint32_t s0 = (uint32_t)__shfl((int)cur, 0, 64);
        const bool one = __all(cur == s0 || cur == 0xffffffffu);  // the wave folds one scene
        if (one) {
            for (int a = 0; a < 3; ++a) {
                lo[a] = wave_min_u32(lo[a]);
                hi[a] = wave_max_u32(hi[a]);
            }
        } else {
            // A wave across scene boundaries: one reduction and one set of
            // atomics per scene it holds (lanes of other scenes contribute
            // the empty box).
            uint64_t left = __ballot(cur != 0xffffffffu);
            while (left) {
                const uint32_t sc = (uint32_t)__builtin_amdgcn_readlane((int)cur, __builtin_ctzll(left));
                const bool mine = cur == sc;
                uint32_t l2[3], h2[3];
                for (int a = 0; a < 3; ++a) {
                    l2[a] = wave_min_u32(mine ? lo[a] : 0xffffffffu);
                    h2[a] = wave_max_u32(mine ? hi[a] : 0u);
                }
                if ((threadIdx.x & 63u) == 0u) box_merge(&out[sc], l2, h2);
                left &= ~__ballot(mine);
            }
        }
        // The workgroup's waves that each folded one scene merge through LDS
        // when it is the same scene: one set of atomics per workgroup instead
        // of per wave (every workgroup's atomics go to the same few words).
        const uint32_t wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63u) == 0u) {
            wsc[wave] = one ? s0 : 0xfffffffeu;
            for (int a = 0; a < 3; ++a) {
                wlo[wave][a] = lo[a];
                whi[wave][a] = hi[a];
            }
        }
        __syncthreads();
        if ((threadIdx.x & 63u) == 0u && one && s0 != 0xffffffffu) {
            uint32_t lead = 0;  // the first wave of this scene merges for all of them
            while (wsc[lead] != s0) ++lead;
            if (lead == wave) {
                for (uint32_t v = wave + 1u; v < kBlock / 64u; ++v)
                    if (wsc[v] == s0)
                        for (int a = 0; a < 3; ++a) {
                            lo[a] = min(lo[a], wlo[v][a]);
                            hi[a] = max(hi[a], whi[v][a]);
                        }
                box_merge(&out[s0], lo, hi);
            }
        }
        __syncthreads();  // the LDS slots are read before the next chunk's
    }
}

// grid_box rule per scene (core/grid.rs:34-51), cell bases, pad-cell counts,
// AABB ping-pong reset, capacity check, and the launch's place in the pool
// tick (tick_sub of tick_len launches). One block.
__global__ __launch_bounds__(1024) void k_grid_box(Pool P, uint32_t tick_sub, uint32_t tick_len) {
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
            // A retired scene (no particles left to step) keeps one cell.
            if (sp.n == 0u)
                for (int a = 0; a < 3; ++a) lo[a] = hi[a] = 0.0f;
            float bin = fmaxf(sp.h, 1e-6f);
            const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
            uint32_t dims[3] = {1, 1, 1};
            for (int it = 0; it < 256; ++it) {
                for (int a = 0; a < 3; ++a) {
                    const float f = floorf(ext[a] / bin);
                    // `as usize` saturates (NaN, negative -> 0), then +1, clamp [1, 16384].
                    const uint32_t d = (f >= (float)kGridDimMax) ? kGridDimMax
                                       : (f > 0.0f ? (uint32_t)f + 1u : 1u);
                    dims[a] = min(max(d, 1u), kGridDimMax);
                }
                if ((uint64_t)dims[0] * dims[1] * dims[2] <= sp.max_bins) break;
                bin *= 2.0f;
            }
            SceneGrid &g = P.grid[s];
            // A box with a non-finite corner (NaN or infinite positions; the
            // rule would give it a 1 x 1 x 1 grid and every particle one cell),
            // or one the rule cannot bring under budget, abandons the substep
            // instead of indexing past the tables.
            bool finite = true;
            for (int a = 0; a < 3; ++a) finite = finite && isfinite(lo[a]) && isfinite(hi[a]);
            if (!finite || (uint64_t)dims[0] * dims[1] * dims[2] > sp.max_bins) {
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
            const uint32_t v = threadIdx.x >= o ? scan[threadIdx.x - o] : 0u;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (s < P.nscenes) P.grid[s].cbase = carry + scan[threadIdx.x] - need;
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
        P.ctl->tick_sub = tick_sub;
        P.ctl->tick_len = tick_len;
        P.ctl->nreact = 0u;
    }
}

// Cell of every cur slot (fluid grid.rs:80-89; stored x fastest) and its
// arrival rank inside the cell. Lanes of a wave that hold the same cell in a
// run (the cur order is the previous grid order) share one atomic.
__global__ __launch_bounds__(kBlock) void k_keys(Pool P) {
    if (P.ctl->err) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t upto = lane == 63u ? ~0ull : ((1ull << (lane + 1u)) - 1ull);
    // kSortBlocks pool blocks per workgroup iteration: every block's position
    // loads, then every block's atomics, are in flight together (one block per
    // iteration waited on a load and then on an atomic, back to back).
    for (uint32_t bb = blockIdx.x; bb * kSortBlocks < P.nblocks; bb += gridDim.x) {
        float4 a[kSortBlocks];
        uint32_t sc[kSortBlocks];
        bool live[kSortBlocks];
#pragma unroll
        for (uint32_t k = 0; k < kSortBlocks; ++k) {
            const uint32_t b = bb * kSortBlocks + k;
            live[k] = false;
            sc[k] = 0u;
            a[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (b < P.nblocks) {
                sc[k] = P.block_scene[b];
                const SceneParams &sp = P.params[sc[k]];
                const uint32_t m = b * kBlock + threadIdx.x;
                live[k] = m < sp.pbase + sp.n;
                if (live[k]) a[k] = P.cA[m];
            }
        }
        uint32_t c[kSortBlocks], base[kSortBlocks];
        int hl[kSortBlocks];
#pragma unroll
        for (uint32_t k = 0; k < kSortBlocks; ++k) {
            c[k] = 0xffffffffu;
            if (live[k]) {
                const SceneGrid &g = P.grid[sc[k]];
                float f;
                const uint32_t ux = axis_bin(g.gmin[0], a[k].x, g.inv, g.dims[0], &f);
                const uint32_t uy = axis_bin(g.gmin[1], a[k].y, g.inv, g.dims[1], &f);
                const uint32_t uz = axis_bin(g.gmin[2], a[k].z, g.inv, g.dims[2], &f);
                c[k] = g.cbase + (uz * g.dims[1] + uy) * g.dims[0] + ux;
            }
            // one atomic per run of equal keys in the wave (slot order is the
            // previous grid order, so runs are long)
            const uint32_t prev = (uint32_t)__shfl_up((int)c[k], 1, 64);
            const bool head = lane == 0u || prev != c[k];
            const uint64_t heads = __ballot(head);
            hl[k] = 63 - __clzll((long long)(heads & upto));
            const uint64_t after = heads & ~upto;
            const int nl = after ? __ffsll((long long)after) - 1 : 64;
            base[k] = 0;
            if (head && live[k]) base[k] = atomicAdd(&P.counts[c[k]], (uint32_t)(nl - (int)lane));
        }
#pragma unroll
        for (uint32_t k = 0; k < kSortBlocks; ++k) {
            const uint32_t bs = (uint32_t)__shfl((int)base[k], hl[k], 64);
            if (live[k]) {
                const uint32_t m = (bb * kSortBlocks + k) * kBlock + threadIdx.x;
                P.key[m] = c[k];
                P.slot[m] = bs + (lane - (uint32_t)hl[k]);
            }
        }
    }
}

// --- exclusive scan of counts[0, ctot) into start[0, ctot]; counts zeroed ---
// A tile of kBlock * ITEMS consecutive cells per block (ITEMS = 4, 8, 16:
// the cells-per-tile knob of the autotuner); tile sums, their scan, then the
// tiles again with their offsets.

template <uint32_t ITEMS>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(Pool P, uint32_t *tile_sums) {
    constexpr uint32_t kTile = kBlock * ITEMS;
    if (P.ctl->err) return;
    const uint32_t ctot = P.ctl->ctot;
    const uint32_t ntiles = (ctot + kTile - 1) / kTile;
    __shared__ uint32_t sh[kBlock / 64];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t acc = 0;
        const uint32_t base = t * kTile;
        for (uint32_t k = 0; k < ITEMS; ++k) {
            const uint32_t c = base + k * kBlock + threadIdx.x;
            if (c < ctot) acc += P.counts[c];
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor((int)acc, o, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t sum = 0;
            for (uint32_t w = 0; w < kBlock / 64; ++w) sum += sh[w];
            tile_sums[t] = sum;
        }
        __syncthreads();
    }
}

template <uint32_t ITEMS>
__global__ __launch_bounds__(1024) void k_scan_tiles(Pool P, uint32_t *tile_sums) {
    constexpr uint32_t kTile = kBlock * ITEMS;
    if (P.ctl->err) return;
    const uint32_t ctot = P.ctl->ctot;
    const uint32_t ntiles = (ctot + kTile - 1) / kTile;
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
            const uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
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

template <uint32_t ITEMS>
__global__ __launch_bounds__(kBlock) void k_scan_apply(Pool P, const uint32_t *tile_sums) {
    static_assert(ITEMS % 4 == 0, "cells move in 16-byte groups");
    constexpr uint32_t kTile = kBlock * ITEMS;
    if (P.ctl->err) return;
    const uint32_t ctot = P.ctl->ctot;
    const uint32_t ntiles = (ctot + kTile - 1) / kTile;
    __shared__ uint32_t sh[kBlock];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        // Each thread owns ITEMS consecutive cells (16-byte loads when all of
        // them are in range: base is a multiple of 4 and the tables are
        // 16-byte aligned).
        const uint32_t base = t * kTile + threadIdx.x * ITEMS;
        uint32_t v[ITEMS];
        uint32_t local = 0;
        const bool whole = base + ITEMS <= ctot;
        if (whole) {
#pragma unroll
            for (uint32_t q = 0; q < ITEMS / 4; ++q) {
                const uint4 a = *(const uint4 *)(P.counts + base + 4u * q);
                v[4 * q] = a.x; v[4 * q + 1] = a.y; v[4 * q + 2] = a.z; v[4 * q + 3] = a.w;
            }
        } else {
            for (uint32_t k = 0; k < ITEMS; ++k) v[k] = base + k < ctot ? P.counts[base + k] : 0u;
        }
        for (uint32_t k = 0; k < ITEMS; ++k) local += v[k];
        sh[threadIdx.x] = local;
        __syncthreads();
        for (uint32_t o = 1; o < kBlock; o <<= 1) {
            const uint32_t a = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += a;
            __syncthreads();
        }
        uint32_t run = tile_sums[t] + sh[threadIdx.x] - local;
        if (whole) {
            uint32_t st[ITEMS];
            for (uint32_t k = 0; k < ITEMS; ++k) {
                st[k] = run;
                run += v[k];
            }
#pragma unroll
            for (uint32_t q = 0; q < ITEMS / 4; ++q) {
                *(uint4 *)(P.start + base + 4u * q) = make_uint4(st[4 * q], st[4 * q + 1], st[4 * q + 2], st[4 * q + 3]);
                *(uint4 *)(P.counts + base + 4u * q) = make_uint4(0u, 0u, 0u, 0u);
            }
        } else {
            for (uint32_t k = 0; k < ITEMS; ++k) {
                const uint32_t c = base + k;
                if (c < ctot) {
                    P.start[c] = run;
                    P.counts[c] = 0u;
                    run += v[k];
                }
            }
        }
        __syncthreads();
    }
}

template __global__ void k_scan_reduce<4>(Pool, uint32_t *);
template __global__ void k_scan_reduce<8>(Pool, uint32_t *);
template __global__ void k_scan_reduce<16>(Pool, uint32_t *);
template __global__ void k_scan_tiles<4>(Pool, uint32_t *);
template __global__ void k_scan_tiles<8>(Pool, uint32_t *);
template __global__ void k_scan_tiles<16>(Pool, uint32_t *);
template __global__ void k_scan_apply<4>(Pool, const uint32_t *);
template __global__ void k_scan_apply<8>(Pool, const uint32_t *);
template __global__ void k_scan_apply<16>(Pool, const uint32_t *);

// Scatter every cur slot to its (not yet canonical) place inside its cell.
__global__ __launch_bounds__(kBlock) void k_place(Pool P) {
    if (P.ctl->err) return;
    // kSortBlocks pool blocks per workgroup iteration, loads of all of them
    // in flight together (key -> cell start -> store is a chain per slot).
    for (uint32_t bb = blockIdx.x; bb * kSortBlocks < P.nblocks; bb += gridDim.x) {
        uint32_t key[kSortBlocks], slot[kSortBlocks], gid[kSortBlocks];
        bool live[kSortBlocks];
#pragma unroll
        for (uint32_t k = 0; k < kSortBlocks; ++k) {
            const uint32_t b = bb * kSortBlocks + k;
            live[k] = false;
            key[k] = slot[k] = gid[k] = 0u;
            if (b < P.nblocks) {
                const SceneParams &sp = P.params[P.block_scene[b]];
                const uint32_t m = b * kBlock + threadIdx.x;
                live[k] = m < sp.pbase + sp.n;
                if (live[k]) {
                    key[k] = P.key[m];
                    slot[k] = P.slot[m];
                    gid[k] = P.cg[m];
                }
            }
        }
        uint32_t st[kSortBlocks];
#pragma unroll
        for (uint32_t k = 0; k < kSortBlocks; ++k) st[k] = live[k] ? P.start[key[k]] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kSortBlocks; ++k)
            if (live[k])
                P.items[st[k] + slot[k]] = make_uint2((bb * kSortBlocks + k) * kBlock + threadIdx.x, gid[k]);
    }
}

// Canonical position = cell start + number of cell members with a smaller
// original index (cell order = ascending index, fluid grid.rs:53-75), then
// gather the particle into the grid-ordered arrays; also the count of fluid
// items per cell (they come first: fluid indices precede boundary ones).
__global__ __launch_bounds__(kBlock) void k_canon_gather(Pool P) {
    if (P.ctl->err) return;
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t u = b * kBlock + threadIdx.x;
        if (u >= sp.pbase + sp.n) continue;
        const uint2 it = P.items[u];
        const uint32_t m = it.x, g = it.y;
        // The particle's state loads first: their latency overlaps the chain
        // key -> cell start -> the cell's members.
        const float4 a = P.cA[m], v = P.cV[m];
        const uint32_t c = P.key[m];
        const uint32_t s0 = P.start[c], e0 = P.start[c + 1];  // e0 > s0: the cell holds u
        uint32_t rank = 0, nfl = 0;
        for (uint32_t k = s0; k < e0; k += 4u) {
            // four members' indices in flight at once (the last index repeats
            // past the cell's end and is not counted)
            uint32_t gk[4];
#pragma unroll
            for (uint32_t q = 0; q < 4u; ++q) gk[q] = P.items[min(k + q, e0 - 1u)].y;
#pragma unroll
            for (uint32_t q = 0; q < 4u; ++q) {
                const bool in = k + q < e0;
                rank += in && gk[q] < g ? 1u : 0u;
                nfl += in && gk[q] < sp.nf ? 1u : 0u;
            }
        }
        const uint32_t d = s0 + rank;
        P.sA[d] = a;
        P.sg[d] = g;
        // the position and velocity halves of the density record (the density
        // pass adds c2 and rho)
        P.fr[2u * d] = make_float4(a.x, a.y, a.z, 0.0f);
        P.fr[2u * d + 1u] = make_float4(v.x, v.y, v.z, 0.0f);
        // Fluid items (smaller indices) head the cell; the cell's first item
        // records how many there are (stale for empty cells: readers clamp).
        if (rank == 0) P.fcount[c] = nfl;
    }
}

// Density sums of one particle (fluid_density.rs:17-70): fluid i sums W over
// fluid neighbours (times its mass at the end) and m_b W over boundary ones;
// boundary i sums m_j W over fluid neighbours only. A candidate with
// acc = false leaves both sums untouched (a select, not a branch).
//
// The class of j is the sign of its packed mass (masses are validated positive
// and finite, so the sign bit alone decides). ALL_FLUID: every lane of the
// wave is a fluid particle, which drops the per-candidate class select of i.
template <bool ALL_FLUID>
struct DensAcc {
    float sum, sum_b;
    bool fluid_i;
    __device__ __forceinline__ void add(bool acc, float r2, float aw, float h2, float k6) {
        const float t = h2 - r2;
        const float w = ((k6 * t) * t) * t;
        const bool fj = __float_as_int(aw) >= 0;
        const float mw = fabsf(aw) * w;
        if (ALL_FLUID) {
            sum = (acc && fj) ? sum + w : sum;
            sum_b = (acc && !fj) ? sum_b + mw : sum_b;
        } else {
            const float add_sum = fluid_i ? w : mw;
            sum = (acc && fj) ? sum + add_sum : sum;
            sum_b = (acc && !fj && fluid_i) ? sum_b + mw : sum_b;
        }
    }
    // All of this candidate's group are fluid particles: the class selects drop.
    __device__ __forceinline__ void add_fluid(bool acc, float r2, float h2, float k6) {
        const float t = h2 - r2;
        const float w = ((k6 * t) * t) * t;
        sum = acc ? sum + w : sum;
    }
    static constexpr bool kFluidGroups = ALL_FLUID;
    __device__ __forceinline__ void ghost(bool, float, float, float) {}
};

// Density sums of a wave holding fluid particles in the band of half-space
// walls: DensAcc<false> plus, after each accepted neighbour of a band
// particle, the neighbour's mirror images in the particle's band faces added
// to `sum` (fluid_density.rs:24-51, the reference's summation order).
struct DensAccWalls {
    static constexpr bool kFluidGroups = false;
    __device__ __forceinline__ void add_fluid(bool, float, float, float) {}
    float sum, sum_b;
    bool fluid_i;
    wall::V3 pi;
    uint32_t fm;  // the lane's band faces (bit k = face k of the scene)
    uint32_t wm;  // their union over the wave
    const WallFaceDev *face;
    float h2, k6;
    __device__ __forceinline__ void add(bool acc, float r2, float aw, float h2_, float k6_) {
        const float t = h2_ - r2;
        const float w = ((k6_ * t) * t) * t;
        const bool fj = __float_as_int(aw) >= 0;
        const float mw = fabsf(aw) * w;
        const float add_sum = fluid_i ? w : mw;
        sum = (acc && fj) ? sum + add_sum : sum;
        sum_b = (acc && !fj && fluid_i) ? sum_b + mw : sum_b;
    }
    __device__ __forceinline__ void ghost(bool acc, float ddx, float ddy, float ddz) {
        for (uint32_t m = wm; m; m &= m - 1u) {
            const uint32_t k = (uint32_t)__builtin_ctz(m);
            float w;
            if (acc && ((fm >> k) & 1u) && wall::face_ghost(face[k], pi, wall::V3{ddx, ddy, ddz}, h2, k6, w))
                sum = sum + w;
        }
    }
};

// Four consecutive candidates jb .. jb + 3 of a row [j0, j1) (loaded without
// clamping: the arrays are padded). SELF: the range may hold i; PART: the
// group may run past j1; MASK: set the mask bits (candidate k = j - j0 < 32),
// else count the accepted ones.
// Where a row's candidate records come from: the grid-ordered array in global
// memory (padded: groups of four load past a range end), or the wave's LDS
// copy of the rows' union (staged from slot `base` on, padded the same way).
struct GlobalSrc {
    const float4 *__restrict__ sA;
    __device__ __forceinline__ const float4 *group(uint32_t jb) const { return at(sA, jb); }
};
struct LdsSrc {
    const float4 *lds;
    uint32_t base;
    __device__ __forceinline__ const float4 *group(uint32_t jb) const { return lds + (jb - base); }
};

template <bool SELF, bool PART, bool MASK, class Acc, class Src>
__device__ __forceinline__ void dens_group(const Src &src, const float4 ai, uint32_t i, uint32_t jb, uint32_t j0,
                                           uint32_t j1, float h2, float k6, Acc &da, uint32_t &mk, uint32_t &cnt) {
    constexpr uint32_t U = 4;
    const float4 *p = src.group(jb);
    float4 aj[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) aj[u] = p[u];
    uint32_t gb = 0;
    auto each = [&](auto &&sum_one) {
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t j = jb + u;
            const float ddx = ai.x - aj[u].x, ddy = ai.y - aj[u].y, ddz = ai.z - aj[u].z;
            const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            bool acc = r2 <= h2;
            if (SELF) acc = acc && j != i;
            if (PART) acc = acc && j < j1;
            if (MASK)
                gb |= acc ? 1u << u : 0u;
            else
                cnt += acc ? 1u : 0u;
            sum_one(acc, r2, aj[u].w, ddx, ddy, ddz);
        }
    };
    // Waves of fluid particles: a group whose four candidates are all fluid
    // (the fluid interior) takes the sum without the class selects.
    bool fluid_group = false;
    if constexpr (Acc::kFluidGroups) {
        const int cls = __float_as_int(aj[0].w) | __float_as_int(aj[1].w) | __float_as_int(aj[2].w) |
                        __float_as_int(aj[3].w);
        fluid_group = !__any(cls < 0);
    }
    if (fluid_group)
        each([&](bool acc, float r2, float, float, float, float) { da.add_fluid(acc, r2, h2, k6); });
    else
        each([&](bool acc, float r2, float aw, float ddx, float ddy, float ddz) {
            da.add(acc, r2, aw, h2, k6);
            da.ghost(acc, ddx, ddy, ddz);
        });
    if (MASK) mk |= gb << (jb - j0);
}

// One row of the fused density loop; returns the mask of its first 32
// candidates (those past them are counted into cnt).
template <bool SELF, class Acc, class Src>
__device__ __forceinline__ uint32_t dens_row(const Src &src, const float4 ai, uint32_t i, uint32_t j0, uint32_t j1,
                                             float h2, float k6, Acc &da, uint32_t &cnt) {
    uint32_t mk = 0;
    const uint32_t jm = min(j1, j0 + 32u);
    uint32_t jb = j0;
    for (; jb + 4u <= jm; jb += 4u) dens_group<SELF, false, true>(src, ai, i, jb, j0, j1, h2, k6, da, mk, cnt);
    if (jb < jm) {
        dens_group<SELF, true, true>(src, ai, i, jb, j0, j1, h2, k6, da, mk, cnt);
        jb += 4u;
    }
    for (; jb + 4u <= j1; jb += 4u) dens_group<SELF, false, false>(src, ai, i, jb, j0, j1, h2, k6, da, mk, cnt);
    if (jb < j1) dens_group<SELF, true, false>(src, ai, i, jb, j0, j1, h2, k6, da, mk, cnt);
    return mk;
}

// Records one wave may stage per neighbour row (the union of its lanes' row
// windows; STAGE = 128, 256 or 512, the LDS staging depth the autotuner
// picks); larger unions read global memory.

// Density (fluid_density.rs:17-70) + Tait (fluid_force.rs:11-29, gamma = 7)
// + the per-particle pressure term c2 = p / (rho * rho) used by the force,
// written with the position and velocity as the particle's density record
// (fr). The sums are taken while the candidates are tested (selects keep the
// reference's accumulation order and bits). Records every fluid particle's
// accepted neighbours (nbm, nbs) for the force pass. WALLS: the pool has
// scenes with half-space walls, whose band particles add the mirror ghosts of
// their neighbours (fluid_density.rs:24-51, :76-121).
// BS = threads per workgroup (64, 128 or 256: a workgroup stays inside one
// kBlock-slot pool block, so inside one scene).
template <bool WALLS, uint32_t BS, uint32_t STAGE>
__device__ __forceinline__ void density_pass(const Pool &P) {
    __shared__ float4 dens_stage[BS / 64][STAGE];
    if (P.ctl->err) return;
    const uint32_t stride = P.nblocks * kBlock;
    constexpr uint32_t kSub = kBlock / BS;
    const uint32_t nsub = P.nblocks * kSub;
    for (uint32_t hb = blockIdx.x; hb < nsub; hb += gridDim.x) {
        const uint32_t b = gridDim.x == nsub ? xcd_block(hb, nsub) : hb;
        const uint32_t s = P.block_scene[b / kSub];
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * BS + threadIdx.x;
        if (i >= sp.pbase + sp.n) continue;
        const float4 ai = P.sA[i];
        if (sp.nf == 0u) {
            // A scene without fluid is not stepped (fluid_step.rs:92-95): its
            // densities keep their initial zeros (the gather left the record's
            // rho and c2 at zero).
            if (P.ncount) P.ncount[i] = 0u;
            continue;
        }
        const SceneGrid &g = P.grid[s];
        const bool fluid_i = ai.w > 0.0f;
        const Cursor cur = make_cursor(g, sp.h, ai.x, ai.y, ai.z);
        const float h2 = sp.h2, k6 = sp.k6;
        float sum = fluid_i ? sp.w0 : 0.0f, sum_b = 0.0f;
        uint32_t cnt = 0, summ = 0;
        const float4 *__restrict__ sA = P.sA;
        // Row r finished with mask mk: count, and record it for a fluid i.
        auto close_row = [&](uint32_t r, bool left, uint32_t j0, uint32_t j1, uint32_t mk) {
            cnt += (uint32_t)__builtin_popcount(mk);
            const bool tail = j1 - j0 > 32u;
            if (fluid_i && (mk != 0u || tail)) {
                P.nbm[r * stride + i] = mk;
                summ |= row_summary(r, left, tail);
            }
        };
        const GlobalSrc gsrc{sA};
        // One row of the full scan from src; only the particle's own row (r = 4,
        // uniform across the wave) can hold i itself.
        auto scan_row = [&](auto &da, const auto &src, uint32_t r, bool left, uint32_t j0, uint32_t j1) {
            const uint32_t mk = r == 4u ? dens_row<true>(src, ai, i, j0, j1, h2, k6, da, cnt)
                                        : dens_row<false>(src, ai, i, j0, j1, h2, k6, da, cnt);
            close_row(r, left, j0, j1, mk);
        };
        // A full wave whose lanes share one (y, z) cell row has the same nine
        // neighbour rows in every lane; only the x windows differ. Each row's
        // union of windows is one contiguous slot range, loaded into LDS once
        // for the wave; every lane then scans its own window from there in
        // the same order. A wave that runs from the end of cell row (y, z)
        // into the start of (y + 1, z) stages too: a lane's row (dz, dy) is
        // then row y + dy or y + 1 + dy of plane z + dz, and the two are
        // consecutive in slot order, so the union is still one range (the
        // tail of the one and the head of the other). (Vector memory data
        // return bounds this kernel.)
        const uint32_t lane = threadIdx.x & 63u;
        const bool full_wave = __builtin_amdgcn_read_exec() == ~0ull;
        const int cy0 = __builtin_amdgcn_readfirstlane(cur.cy), cz0 = __builtin_amdgcn_readfirstlane(cur.cz);
        const bool staged = full_wave && __all(cur.cz == cz0 && (cur.cy == cy0 || cur.cy == cy0 + 1));
        float4 *stage = dens_stage[threadIdx.x >> 6];
        // The full scan: every candidate tested, sums taken through selects.
        auto full_scan = [&](auto &da) {
            if (staged) {
                const int nx = (int)g.dims[0], ny = (int)g.dims[1], nz = (int)g.dims[2];
                const float lx = cur.fx * cur.fx, rx = (1.0f - cur.fx) * (1.0f - cur.fx);
                // One neighbour row (dz, dy) of the wave: the lane's window and
                // the union's slot range. Its cell-start loads are issued one
                // row ahead of the row's scan.
                struct RowPlan {
                    bool live;   // the wave has the row
                    bool has;    // the lane has it
                    bool left;   // the lane's window starts at dx = -1
                    uint32_t r, base, n, j0, j1;
                };
                auto plan = [&](int q) -> RowPlan {
                    RowPlan rp{false, false, false, (uint32_t)q, 0u, 0u, 0u, 0u};
                    const int dz = q / 3 - 1, dy = q % 3 - 1;
                    const int z = cz0 + dz;
                    if (z < 0 || z >= nz) return rp;
                    const float ez = dz < 0 ? cur.fz : (dz > 0 ? 1.0f - cur.fz : 0.0f);
                    const int y = cur.cy + dy;  // the lane's own row
                    // the lane's window, as for_row_windows forms it
                    const float ey = dy < 0 ? cur.fy : (dy > 0 ? 1.0f - cur.fy : 0.0f);
                    const float rd = ez * ez + ey * ey;
                    rp.has = y >= 0 && y < ny && !(rd > cur.cut);  // the skips of for_row_windows
                    const int xl = (cur.cx > 0 && rd + lx <= cur.cut) ? cur.cx - 1 : cur.cx;
                    const int xr = (cur.cx + 1 < nx && rd + rx <= cur.cut) ? cur.cx + 1 : cur.cx;
                    rp.left = xl < cur.cx;
                    const uint32_t row = g.cbase + ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)nx;
                    const uint32_t cf = row + (uint32_t)xl, cl = row + (uint32_t)xr;
                    const uint64_t hm = __ballot(rp.has);
                    if (hm == 0u) return rp;
                    rp.live = true;
                    // The lanes' own cells (row + cx) rise with the lane index
                    // (grid order), so the union runs from the first lane's
                    // cell (or the one left of it) to the last lane's cell (or
                    // the one right of it): at most a cell wider on each side
                    // than the exact union, which costs a few staged records.
                    const int fl = __builtin_ctzll(hm), ll = 63 - __builtin_clzll(hm);
                    const uint32_t own = row + (uint32_t)cur.cx;
                    const uint32_t ofl = (uint32_t)__builtin_amdgcn_readlane((int)own, fl);
                    const int cxf = __builtin_amdgcn_readlane(cur.cx, fl);
                    const uint32_t oll = (uint32_t)__builtin_amdgcn_readlane((int)own, ll);
                    const int cxl = __builtin_amdgcn_readlane(cur.cx, ll);
                    const uint32_t ulo = ofl - (cxf > 0 ? 1u : 0u);
                    const uint32_t uhi = oll + (cxl + 1 < nx ? 1u : 0u);
                    rp.base = P.start[ulo];
                    rp.n = P.start[uhi + 1u];  // the range's end until the row is used
                    if (rp.has) {
                        rp.j0 = P.start[cf];
                        rp.j1 = P.start[cl + 1u];
                    }
                    return rp;
                };
                RowPlan nxt = plan(0);
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    RowPlan rp = nxt;
                    if (q + 1 < 9) nxt = plan(q + 1);
                    if (!rp.live) continue;
                    const uint32_t base = rp.base, n = rp.n - rp.base;
                    if (n + 4u > STAGE) {
                        if (rp.has) scan_row(da, gsrc, rp.r, rp.left, rp.j0, rp.j1);
                        continue;
                    }
                    // The wave's LDS operations complete in order; the wave
                    // barriers keep the compiler from moving the copy and the
                    // reads of other lanes' records across each other.
                    __builtin_amdgcn_wave_barrier();
                    for (uint32_t k = lane; k < n + 4u; k += 64u) stage[k] = *at(sA, base + k);
                    __builtin_amdgcn_wave_barrier();
                    const LdsSrc lsrc{stage, base};
                    if (rp.has) scan_row(da, lsrc, rp.r, rp.left, rp.j0, rp.j1);
                }
            } else {
                for_rows(g, P.start, cur, [&](uint32_t r, bool left, uint32_t j0, uint32_t j1) {
                    scan_row(da, gsrc, r, left, j0, j1);
                });
            }
            sum = da.sum;
            sum_b = da.sum_b;
        };
        uint32_t fm = 0;
        const WallFaceDev *faces = nullptr;
        if constexpr (WALLS) {
            const WallSceneDev &W = P.walls[s];
            faces = W.face;
            if (W.mode == kWallsHalfspace && fluid_i) fm = wall::band_faces(W, wall::V3{ai.x, ai.y, ai.z}, sp.h);
        }
        if (WALLS && __any(fm != 0u)) {
            uint32_t wm = 0;
            for (uint32_t k = 0; k < kWallMaxFaces; ++k)
                if (__any((fm >> k) & 1u)) wm |= 1u << k;
            DensAccWalls da{sum, sum_b, fluid_i, wall::V3{ai.x, ai.y, ai.z}, fm, wm, faces, h2, k6};
            full_scan(da);
        } else if (__all(fluid_i)) {
            DensAcc<true> da{sum, sum_b, true};
            full_scan(da);
        } else if (__any(fluid_i) || P.ncount) {
            // Waves holding a fluid particle take the full scan (a wave running
            // both scans would pay for both).
            DensAcc<false> da{sum, sum_b, fluid_i};
            full_scan(da);
        } else {
            // Boundary particles sum fluid neighbours only
            // (fluid_density.rs:59-69): scan just the fluid head of each cell,
            // in cell order. (With neighbour counts requested, the full scan
            // above counts boundary neighbours too.)
            DensAcc<false> da{sum, sum_b, false};
            for_row_windows(g, cur, [&](uint32_t, uint32_t cf, uint32_t cl, bool) {
                // The window's (at most three) cell starts and fluid counts
                // load together; most tank particles have no fluid near them
                // and add nothing.
                const uint32_t nc = cl - cf;  // 0..2
                uint32_t st[4], fc[3];
#pragma unroll
                for (uint32_t k = 0; k < 4u; ++k) st[k] = P.start[cf + min(k, nc + 1u)];
#pragma unroll
                for (uint32_t k = 0; k < 3u; ++k) fc[k] = P.fcount[cf + min(k, nc)];
#pragma unroll
                for (uint32_t k = 0; k < 3u; ++k) {
                    if (k > nc) break;
                    const uint32_t lo = st[k], jn = lo + min(fc[k], st[k + 1u] - lo);
                    uint32_t unused = 0;
                    for (uint32_t jb = lo; jb < jn; jb += 4u)
                        dens_group<false, true, false>(gsrc, ai, i, jb, lo, jn, h2, k6, da, unused, cnt);
                }
            });
            sum = da.sum;
        }
        if (fluid_i) P.nbs[i] = summ;
        const float rho = fluid_i ? sp.mass * sum + sum_b : (-ai.w) * sp.w0 + sum;
        float p, c2;
        tait(sp, rho, p, c2);
        P.fr[2u * i].w = c2;  // beside the position and velocity the gather wrote
        P.fr[2u * i + 1u].w = fluid_i ? rho : -rho;
        if (P.ncount) P.ncount[i] = cnt;
    }
}

// Six waves per SIMD (at most 80 VGPRs): the kernel is bound by vector
// memory and LDS traffic, so occupancy pays more than the registers the
// compiler would otherwise spend on scheduling (it fits without spills). A
// 512-record stage holds 8 KB of LDS per wave, which allows fewer waves.
template <uint32_t BS, uint32_t STAGE>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(STAGE > 256 ? 4 : 6, 6))) void k_density(Pool P) {
    density_pass<false, BS, STAGE>(P);
}
template <uint32_t BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_density_walls(Pool P) {
    density_pass<true, BS, 256>(P);
}
template __global__ void k_density<64, 128>(Pool);
template __global__ void k_density<64, 256>(Pool);
template __global__ void k_density<64, 512>(Pool);
template __global__ void k_density<128, 128>(Pool);
template __global__ void k_density<128, 256>(Pool);
template __global__ void k_density<128, 512>(Pool);
template __global__ void k_density<256, 128>(Pool);
template __global__ void k_density<256, 256>(Pool);
template __global__ void k_density<256, 512>(Pool);
template __global__ void k_density_walls<64>(Pool);
template __global__ void k_density_walls<128>(Pool);
template __global__ void k_density_walls<256>(Pool);



// One neighbour pair of the force sum, fluid_force.rs:165-217 operation order.
struct ForceAcc {
    float ax, ay, az, xx, xy, xz;
};
// Scene constants of the pair terms, held in registers for the whole walk.
struct ForceConst {
    float h, h2, k6, ks, alpha_c, h2c;  // h2c = 0.01 h^2 (the AV softening term)
};
__device__ __forceinline__ ForceConst force_const(const SceneParams &sp) {
    return ForceConst{sp.h, sp.h2, sp.k6, sp.ks, sp.alpha_c, 0.01f * sp.h2};
}
__device__ __forceinline__ void force_pair(ForceAcc &f, const float4 ai, const float4 vi, const float ci2,
                                           const float4 aj, const float4 vj, const float cj2, const ForceConst &fc) {
    const float h = fc.h, h2 = fc.h2, k6 = fc.k6, ks = fc.ks, alpha_c = fc.alpha_c;
    const float ddx = ai.x - aj.x, ddy = ai.y - aj.y, ddz = ai.z - aj.z;
    const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
    const float rho_i = vi.w, rho_j = vj.w;
    const float r = sqrtf(r2);
    const float t = h - r;
    const float mj = fabsf(aj.w);
    const float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
    const float denom = fmaxf(r, 1e-9f);
    const float qp = coef / denom;
    f.ax = f.ax + ddx * qp;
    f.ay = f.ay + ddy * qp;
    f.az = f.az + ddz * qp;
    const float wx = vi.x - vj.x, wy = vi.y - vj.y, wz = vi.z - vj.z;
    const float vdn = -(wx * ddx + wy * ddy + wz * ddz);
    if (vdn > 0.0f) {
        const float mu = (vdn * h) / (r2 + fc.h2c);
        const float cv = (mj * ((alpha_c * mu) / (0.5f * (rho_i + rho_j)))) * ((ks * t) * t);
        const float qv = cv / denom;
        f.ax = f.ax + ddx * qv;
        f.ay = f.ay + ddy * qv;
        f.az = f.az + ddz * qv;
    }
    const float tt = h2 - r2;
    const float w = ((k6 * tt) * tt) * tt;
    const float fx = ((mj * 2.0f) / (rho_i + rho_j)) * w;
    f.xx = f.xx + (vj.x - vi.x) * fx;
    f.xy = f.xy + (vj.y - vi.y) * fx;
    f.xz = f.xz + (vj.z - vi.z) * fx;
}

// Neighbour j of the force sum from its density record (r0, r1): A with the
// mass (the scene's for a fluid neighbour when every fluid mass is the
// scene's, else read from sA), V with rho, and c2.
__device__ __forceinline__ void nb_record(const float4 r0, const float4 r1, uint32_t j, const float4 *__restrict__ sA,
                                          float smass, bool own_mass, float4 &aj, float4 &vj, float &cj) {
    float mj = smass;
    if (own_mass || r1.w < 0.0f) mj = fabsf(at(sA, j)->w);
    aj = make_float4(r0.x, r0.y, r0.z, mj);
    vj = make_float4(r1.x, r1.y, r1.z, fabsf(r1.w));
    cj = r0.w;
}

// The walk of the recorded neighbours with each pair's record loaded one
// pair ahead of its evaluation (same pairs, same order as a plain walk).
__device__ __forceinline__ void walk_ahead(ForceAcc &f, const Pool &P, const SceneGrid &g, const SceneParams &sp,
                                           uint32_t i, bool fluid, const float4 ai, const float4 vi, float ci2,
                                           const RowTab *rowtab, uint32_t stride) {
    const float4 *__restrict__ sA = P.sA;
    const float4 *__restrict__ fr = P.fr;
    const bool own_mass = sp.fmass_uniform == 0u;
    const float h2 = sp.h2;
    Cursor cur{};
    if (fluid) cur = make_cursor(g, sp.h, ai.x, ai.y, ai.z);
    NbWalk w = nb_begin(fluid ? P.nbs[i] : 0u, g, cur, rowtab);
    const ForceConst fc = force_const(sp);
    uint32_t jn = 0;
    int stn = nb_step(w, P, g, cur, i, ai, h2, stride, jn);
    float4 rn0 = ai, rn1 = vi;
    if (stn == kWalkPair) {
        rn0 = *at(fr, 2u * jn);
        rn1 = *at(fr, 2u * jn + 1u);
    }
    for (;;) {
        const int st = stn;
        if (!__any(st != kWalkDone)) break;
        const float4 r0 = rn0, r1 = rn1;
        const uint32_t j = jn;
        if (st != kWalkDone) {
            stn = nb_step(w, P, g, cur, i, ai, h2, stride, jn);
            if (stn == kWalkPair) {
                rn0 = *at(fr, 2u * jn);
                rn1 = *at(fr, 2u * jn + 1u);
            }
        }
        if (st == kWalkPair) {
            float4 aj, vj;
            float cj;
            nb_record(r0, r1, j, sA, sp.mass, own_mass, aj, vj, cj);
            force_pair(f, ai, vi, ci2, aj, vj, cj, fc);
        }
    }
}

// The walk of the recorded neighbours from a per-lane queue of row chunks.
// Before the pair loop every lane opens all its rows at once (the masks and
// row starts load together) and lists, in row order, each chunk that holds a
// neighbour as (mask, first slot): the density pass's mask of the row's first
// 32 candidates, then the row's further candidates tested here 32 at a time.
// The pair loop then only takes the lowest bit of the current chunk and, when
// the chunk runs out, reads the lane's next chunk from LDS: no row opening and
// no global load but the neighbour's record. Same pairs, same order as
// nb_step's walk. A lane that needs more than CAP chunks makes its wave walk
// with walk_ahead instead.
template <uint32_t CAP>
struct QueueLds {
    uint2 e[CAP][64];  // chunk k of lane l at e[k][l]
};
constexpr uint32_t kQueueCap = 12;

// Lists the lane's chunks (returns their count; above CAP the queue is
// incomplete) and the slot range [lo, hi) of its accepted neighbours in its
// own cell row (row 4: dz = dy = 0), empty (lo >= hi) when it has none there.
template <uint32_t CAP>
__device__ __forceinline__ uint32_t queue_build(const Pool &P, const SceneGrid &g, const SceneParams &sp, uint32_t i,
                                                bool fluid, const float4 ai, const RowTab *rowtab, QueueLds<CAP> &q,
                                                uint32_t &lo, uint32_t &hi) {
    const float4 *__restrict__ sA = P.sA;
    const float h2 = sp.h2;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t n = 0;
    lo = 0xffffffffu;
    hi = 0u;
    if (fluid) {
        const Cursor c = make_cursor(g, sp.h, ai.x, ai.y, ai.z);
        const uint32_t summ = P.nbs[i];
        const uint32_t rowb =
            g.cbase + ((uint32_t)(c.cz - 1) * g.dims[1] + (uint32_t)(c.cy - 1)) * g.dims[0] + (uint32_t)c.cx;
        uint32_t j0s[9], mks[9];
#pragma unroll
        for (uint32_t r = 0; r < 9u; ++r) {
            j0s[r] = 0u;
            mks[r] = 0u;
            if ((summ >> r) & 1u) {
                const RowTab t = rowtab[r];
                const uint32_t row = rowb + t.cell;
                j0s[r] = *at(P.start, row - ((summ >> (9u + r)) & 1u));
                mks[r] = P.nbm[t.mask + i];
            }
        }
        auto list = [&](uint32_t r, uint32_t m, uint32_t base) {
            if (n < CAP) q.e[n][lane] = make_uint2(m, base);
            ++n;
            if (r == 4u) {
                lo = min(lo, base + (uint32_t)__builtin_ctz(m));
                hi = max(hi, base + 32u - (uint32_t)__builtin_clz(m));
            }
        };
#pragma unroll
        for (uint32_t r = 0; r < 9u; ++r) {
            if (mks[r] != 0u) list(r, mks[r], j0s[r]);
            if ((summ >> (18u + r)) & 1u) {
                // Candidates past the first 32 of the row, tested 32 at a time.
                const uint32_t row = rowb + rowtab[r].cell;
                const uint32_t je = P.start[row - (uint32_t)c.cx + row_last_cell(g, c, r) + 1u];
                for (uint32_t jb = j0s[r] + 32u; jb < je; jb += 32u) {
                    uint32_t m = 0;
                    const uint32_t kn = min(32u, je - jb);
                    for (uint32_t k = 0; k < kn; ++k) {
                        const uint32_t j = jb + k;
                        const float4 aj = *at(sA, j);
                        const float ddx = ai.x - aj.x, ddy = ai.y - aj.y, ddz = ai.z - aj.z;
                        const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (r2 <= h2 && j != i) m |= 1u << k;
                    }
                    if (m != 0u) list(r, m, jb);
                }
            }
        }
    }
    return n;
}

// One wave's copy of the neighbour records of its lanes' own cell rows (row
// 4): the union of the lanes' row-4 ranges, slots [slo, slo + sn), its first
// CAP records copied with coalesced loads. The whole wave calls it.
template <uint32_t CAP>
__device__ __forceinline__ void stage_own_rows(float4 (*st)[2], const float4 *__restrict__ fr, uint32_t lo,
                                               uint32_t hi, uint32_t &slo, uint32_t &sn) {
    slo = __builtin_amdgcn_readfirstlane(wave_min_u32(lo));
    const uint32_t shi = __builtin_amdgcn_readfirstlane(wave_max_u32(hi));
    sn = slo < shi ? min(shi - slo, CAP) : 0u;
    for (uint32_t k = threadIdx.x & 63u; k < 2u * sn; k += 64u) (&st[0][0])[k] = *at(fr, 2u * slo + k);
    __builtin_amdgcn_wave_barrier();
}

// Exchange with the other lane of the lane pair (2k, 2k + 1): DPP quad_perm
// [1, 0, 3, 2]; the whole wave must be active.
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ float4 pair_swap(const float4 v) {
    return make_float4(__uint_as_float(pair_swap(__float_as_uint(v.x))), __uint_as_float(pair_swap(__float_as_uint(v.y))),
                       __uint_as_float(pair_swap(__float_as_uint(v.z))), __uint_as_float(pair_swap(__float_as_uint(v.w))));
}

// How the pair loop of the queue walk reads a neighbour's density record.
enum : uint32_t {
    kLoadPlain = 0,   // each lane loads its record's two 16-byte halves
    kLoadPaired = 1,  // lane pairs load both halves of one record per instruction
    kLoadStaged = 2,  // records of the wave's stage from LDS, the others as kLoadPlain
};

// The pair loop over the lane's n listed chunks. kLoadPaired: the lanes of a
// pair (2k, 2k + 1) load the even lane's record (half `odd` each), then the
// odd lane's, and exchange halves through DPP, so a wave-wide load touches
// the cache lines of 32 whole records instead of 64 half records (a lane
// without a pair loads its own record for its partner). kLoadStaged: a
// neighbour in the wave's stage (slots [slo, slo + sn), stage_own_rows) is
// read from LDS. Same pairs, same order, same values in every mode.
template <uint32_t LOAD, uint32_t CAP>
__device__ __forceinline__ void queue_pairs(ForceAcc &f, const Pool &P, const SceneParams &sp, uint32_t i,
                                            const float4 ai, const float4 vi, float ci2, uint32_t n,
                                            const QueueLds<CAP> &q, const float4 (*st)[2], uint32_t slo,
                                            uint32_t sn) {
    const float4 *__restrict__ sA = P.sA;
    const float4 *__restrict__ fr = P.fr;
    const bool own_mass = sp.fmass_uniform == 0u;
    const uint32_t lane = threadIdx.x & 63u;
    const ForceConst fc = force_const(sp);
    uint32_t e = 0, mk = 0, j0 = 0;
    if (n != 0u) {
        const uint2 v = q.e[0][lane];
        mk = v.x;
        j0 = v.y;
    }
    // The lane's next pair, if any (chunks in order, bits in order).
    auto pop = [&](uint32_t &j) -> bool {
        if (mk == 0u) return false;
        j = j0 + (uint32_t)__builtin_ctz(mk);
        mk &= mk - 1u;
        if (mk == 0u && ++e < n) {
            const uint2 v = q.e[e][lane];
            mk = v.x;
            j0 = v.y;
        }
        return true;
    };
    auto eval = [&](float4 r0, float4 r1, uint32_t j) {
        if constexpr (LOAD == kLoadStaged) {
            // A staged neighbour (its global load read a placeholder).
            const uint32_t k = j - slo;
            if (k < sn) {
                r0 = st[k][0];
                r1 = st[k][1];
            }
        }
        float4 aj, vj;
        float cj;
        nb_record(r0, r1, j, sA, sp.mass, own_mass, aj, vj, cj);
        force_pair(f, ai, vi, ci2, aj, vj, cj, fc);
    };
    const uint32_t odd = lane & 1u;
    const uint32_t i0 = __builtin_amdgcn_readfirstlane(i);
    auto load = [&](bool h, uint32_t j, float4 &r0, float4 &r1) {
        if constexpr (LOAD == kLoadPaired) {
            const uint32_t js = h ? j : i;
            const uint32_t jp = pair_swap(js);
            const float4 la = *at(fr, 2u * (odd ? jp : js) + odd);  // half `odd` of the even lane's record
            const float4 lb = *at(fr, 2u * (odd ? js : jp) + odd);  // half `odd` of the odd lane's record
            const float4 rv = pair_swap(odd ? la : lb);
            r0 = odd ? rv : la;
            r1 = odd ? lb : rv;
        } else if constexpr (LOAD == kLoadStaged) {
            // Every lane issues the global loads in every iteration (a lane
            // whose neighbour is staged, or that has none, reads the wave's
            // first slot); the staged record is read from LDS when the pair
            // is evaluated.
            const uint32_t jg = h && j - slo >= sn ? j : i0;
            r0 = *at(fr, 2u * jg);
            r1 = *at(fr, 2u * jg + 1u);
        } else {
            if (h) {
                r0 = *at(fr, 2u * j);
                r1 = *at(fr, 2u * j + 1u);
            }
        }
    };
    // Two pairs in flight: each pair's record loads are issued one pair
    // before it is evaluated, alternating between two register sets.
    uint32_t ja = 0, jb = 0;
    float4 a0 = ai, a1 = vi, b0 = ai, b1 = vi;
    bool ha = pop(ja);
    load(ha, ja, a0, a1);
    for (;;) {
        if (!__any(ha)) break;
        const bool hb = pop(jb);
        load(hb, jb, b0, b1);
        if (ha) eval(a0, a1, ja);
        if (!__any(hb)) break;
        ha = pop(ja);
        load(ha, ja, a0, a1);
        if (hb) eval(b0, b1, jb);
    }
}

// Returns false (nothing evaluated) when a lane of the wave needs more than
// kQueueCap chunks; the caller then walks the wave with walk_ahead.
template <uint32_t LOAD>
__device__ __forceinline__ bool queue_walk(ForceAcc &f, const Pool &P, const SceneGrid &g, const SceneParams &sp,
                                           uint32_t i, bool fluid, const float4 ai, const float4 vi, float ci2,
                                           const RowTab *rowtab, QueueLds<kQueueCap> &q) {
    uint32_t lo, hi;
    const uint32_t n = queue_build(P, g, sp, i, fluid, ai, rowtab, q, lo, hi);
    if (__any(n > kQueueCap)) return false;
    queue_pairs<LOAD, kQueueCap>(f, P, sp, i, ai, vi, ci2, n, q, nullptr, 0u, 0u);
    return true;
}

// The three terms one neighbour adds to a particle's force sums, as
// force_pair forms them (fluid_force.rs:165-217), kept apart so that another
// lane can add them in the particle's order: pressure d * (coef / denom),
// viscosity d * (cv / denom) when approaching (av), XSPH (v_j - v_i) * fx.
struct PairTerms {
    float px, py, pz, vx, vy, vz, xx, xy, xz;
    uint32_t av;
};

__device__ __forceinline__ PairTerms pair_terms(const float4 ai, const float4 vi, const float ci2, const float4 aj,
                                                const float4 vj, const float cj2, const ForceConst &fc) {
    PairTerms o;
    const float h = fc.h, h2 = fc.h2, k6 = fc.k6, ks = fc.ks, alpha_c = fc.alpha_c;
    const float ddx = ai.x - aj.x, ddy = ai.y - aj.y, ddz = ai.z - aj.z;
    const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
    const float rho_i = vi.w, rho_j = vj.w;
    const float r = sqrtf(r2);
    const float t = h - r;
    const float mj = fabsf(aj.w);
    const float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
    const float denom = fmaxf(r, 1e-9f);
    const float qp = coef / denom;
    o.px = ddx * qp;
    o.py = ddy * qp;
    o.pz = ddz * qp;
    const float wx = vi.x - vj.x, wy = vi.y - vj.y, wz = vi.z - vj.z;
    const float vdn = -(wx * ddx + wy * ddy + wz * ddz);
    o.av = vdn > 0.0f ? 1u : 0u;
    o.vx = o.vy = o.vz = 0.0f;
    if (vdn > 0.0f) {
        const float mu = (vdn * h) / (r2 + fc.h2c);
        const float cv = (mj * ((alpha_c * mu) / (0.5f * (rho_i + rho_j)))) * ((ks * t) * t);
        const float qv = cv / denom;
        o.vx = ddx * qv;
        o.vy = ddy * qv;
        o.vz = ddz * qv;
    }
    const float tt = h2 - r2;
    const float w = ((k6 * tt) * tt) * tt;
    const float fx = ((mj * 2.0f) / (rho_i + rho_j)) * w;
    o.xx = (vj.x - vi.x) * fx;
    o.xy = (vj.y - vi.y) * fx;
    o.xz = (vj.z - vi.z) * fx;
    return o;
}

// LDS of one wave in the pair-parallel force pass.
constexpr uint32_t kPairCap = 256;    // pairs listed at once (groups of ranks)
constexpr uint32_t kPairRanks = 33;   // ranks of one row (32 mask bits) + 1
constexpr uint32_t kPairStage = 128;  // neighbour records of one row staged per wave
struct PairLds {
    float4 stage[kPairStage][2];    // the row's neighbour records (density record halves)
    float4 term[64][3];             // a chunk's terms: pressure (w: viscosity flag), viscosity, XSPH
    float4 own[64][2];              // the wave's particles: (x, y, z, c2), (v, rho)
    uint16_t list[kPairCap];        // (staged index << 6) | owner lane
    uint16_t link[kPairCap];        // position of the owner's next pair in the group (0xffff: none)
    uint64_t act[kPairRanks];       // lanes holding a pair of rank k
};

// Pair-parallel force sums of one wave (every lane without a row longer than
// 32 candidates). Row by row in the reference order: the records of the
// row's accepted neighbours are copied to LDS once for the wave with
// coalesced loads, the lanes' neighbours (the density pass's masks) are
// listed rank by rank (rank k = every lane's k-th neighbour in the row), all
// 64 lanes evaluate the listed pairs' terms evenly, a chunk of 64 at a time,
// and each particle's lane then adds its own terms in rank order, which is
// its neighbour order: the reference's sums, bit for bit. A row whose
// neighbours span more than the stage is walked lane by lane from global
// memory.
__device__ __forceinline__ void pair_parallel(ForceAcc &f, const Pool &P, const SceneGrid &g, const SceneParams &sp,
                                              uint32_t i, bool fluid, const float4 ai, const float4 vi, float ci2,
                                              const RowTab *rowtab, PairLds &L) {
    const float4 *__restrict__ sA = P.sA;
    const float4 *__restrict__ fr = P.fr;
    const bool own_mass = sp.fmass_uniform == 0u;
    const ForceConst fc = force_const(sp);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    L.own[lane][0] = make_float4(ai.x, ai.y, ai.z, ci2);
    L.own[lane][1] = vi;
    Cursor cur{};
    if (fluid) cur = make_cursor(g, sp.h, ai.x, ai.y, ai.z);
    const uint32_t summ = fluid ? P.nbs[i] : 0u;
    const uint32_t rowb =
        g.cbase + ((uint32_t)(cur.cz - 1) * g.dims[1] + (uint32_t)(cur.cy - 1)) * g.dims[0] + (uint32_t)cur.cx;
    for (uint32_t r = 0; r < 9u; ++r) {
        const bool has = ((summ >> r) & 1u) != 0u;
        if (!__any(has)) continue;
        uint32_t m = 0, j0 = 0;
        if (has) {
            const RowTab t = rowtab[r];
            const uint32_t row = rowb + t.cell;
            j0 = *at(P.start, row - ((summ >> (9u + r)) & 1u));
            m = P.nbm[t.mask + i];
        }
        // the wave's neighbours of this row lie in [jb, je)
        const uint32_t jb = dev::wave_min_u32(m ? j0 : 0xffffffffu);
        const uint32_t je = dev::wave_max_u32(m ? j0 + 32u - (uint32_t)__builtin_clz(m) : 0u);
        if (jb == 0xffffffffu) continue;
        if (je - jb > kPairStage) {
            for (uint32_t mm = m; mm != 0u; mm &= mm - 1u) {
                const uint32_t j = j0 + (uint32_t)__builtin_ctz(mm);
                float4 aj, vj;
                float cj;
                nb_record(*at(fr, 2u * j), *at(fr, 2u * j + 1u), j, sA, sp.mass, own_mass, aj, vj, cj);
                force_pair(f, ai, vi, ci2, aj, vj, cj, fc);
            }
            continue;
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k = lane; k < 2u * (je - jb); k += 64u) (&L.stage[0][0])[k] = *at(fr, 2u * jb + k);
        const uint32_t k_mine = (uint32_t)__builtin_popcount(m);
        uint32_t k_fold = 0;  // this lane's next rank to add
        uint32_t rank = 0;    // ranks listed so far in this row
        for (;;) {
            // List ranks while a whole rank still fits; link each lane's pairs.
            const uint32_t rank0 = rank;
            uint32_t n = 0, prev = 0xffffu, first = 0xffffu;
            while (n + 64u <= kPairCap) {
                const uint64_t act = __ballot(m != 0u);
                if (act == 0ull) break;
                if (m != 0u) {
                    const uint32_t pos = n + (uint32_t)__popcll(act & below);
                    L.list[pos] = (uint16_t)(((j0 + (uint32_t)__builtin_ctz(m) - jb) << 6) | lane);
                    L.link[pos] = 0xffffu;
                    if (prev != 0xffffu) L.link[prev] = (uint16_t)pos;
                    if (first == 0xffffu) first = pos;
                    prev = pos;
                    m &= m - 1u;
                }
                if (lane == 0u) L.act[rank - rank0] = act;
                n += (uint32_t)__popcll(act);
                ++rank;
            }
            if (n == 0u) break;
            __builtin_amdgcn_wave_barrier();
            uint32_t pos = first;
            for (uint32_t c0 = 0; c0 < n; c0 += 64u) {
                const uint32_t p = c0 + lane;
                if (p < n) {
                    const uint32_t e = L.list[p];
                    const uint32_t il = e & 63u, k = e >> 6, j = jb + k;
                    const float4 r0 = L.stage[k][0], r1 = L.stage[k][1];
                    float4 aj, vj;
                    float cj;
                    nb_record(r0, r1, j, sA, sp.mass, own_mass, aj, vj, cj);
                    const float4 oi = L.own[il][0];
                    const PairTerms t =
                        pair_terms(make_float4(oi.x, oi.y, oi.z, 0.0f), L.own[il][1], oi.w, aj, vj, cj, fc);
                    L.term[lane][0] = make_float4(t.px, t.py, t.pz, __uint_as_float(t.av));
                    L.term[lane][1] = make_float4(t.vx, t.vy, t.vz, 0.0f);
                    L.term[lane][2] = make_float4(t.xx, t.xy, t.xz, 0.0f);
                }
                __builtin_amdgcn_wave_barrier();
                while (pos < c0 + 64u) {
                    const uint32_t q = pos - c0;
                    const float4 tp = L.term[q][0], tx = L.term[q][2];
                    f.ax = f.ax + tp.x;
                    f.ay = f.ay + tp.y;
                    f.az = f.az + tp.z;
                    if (__float_as_uint(tp.w) != 0u) {
                        const float4 tv = L.term[q][1];
                        f.ax = f.ax + tv.x;
                        f.ay = f.ay + tv.y;
                        f.az = f.az + tv.z;
                    }
                    f.xx = f.xx + tx.x;
                    f.xy = f.xy + tx.y;
                    f.xz = f.xz + tx.z;
                    ++k_fold;
                    pos = L.link[pos];
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        (void)k_mine;
        (void)k_fold;
    }
}

// Force (fluid_force.rs:165-217: gravity + spiky pressure + Monaghan AV, XSPH)
// fused with the semi-implicit Euler + CFL clamp of fluid_step.rs:110-123.
// Writes the next state into the cur arrays (this substep's grid order) and
// the block's partial AABB of the next positions. Boundary slots are carried
// over.
//
// WALK = 0: every lane tests its candidates and evaluates the accepted pairs
// in its candidate loop.
// WALK = 1: every lane walks the neighbours the density pass recorded (same
// positions, same test), so the candidate scan is not repeated; the wave
// evaluates one pair per lane per iteration.
// WALK = 2: the same walk with the next pair's loads issued one pair ahead.
// WALK = 3: pair-parallel rows (pair_parallel).
// WALK = 4: the walk from a per-lane queue of row chunks (queue_walk).
// WALK = 5: the queue walk with each record loaded by a lane pair
// (kLoadPaired).
// WALK = 6: the queue walk reading each wave's own cell row from LDS
// (kLoadStaged).
//
// BS = threads per workgroup (64, 128 or 256, inside one pool block).
template <uint32_t WALK, uint32_t BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WALK == 6 ? 6 : 1, 8))) void k_force_t(Pool P) {
    __shared__ RowTab rowtab[9];
    if (P.ctl->err) return;
    const uint32_t tick_sub = P.ctl->tick_sub, tick_len = P.ctl->tick_len;
    const uint32_t stride = P.nblocks * kBlock;
    constexpr uint32_t kSub = kBlock / BS;
    const uint32_t nsub = P.nblocks * kSub;
    for (uint32_t hb = blockIdx.x; hb < nsub; hb += gridDim.x) {
        const uint32_t b = gridDim.x == nsub ? xcd_block(hb, nsub) : hb;
        const uint32_t s = P.block_scene[b / kSub];
        const SceneParams &sp = P.params[s];
        // A scene that does not step in this launch keeps its state (every
        // particle is carried over like a boundary one).
        const bool active = substep_active(tick_sub, tick_len, sp.substeps);
        if constexpr (WALK != 0) {
            __syncthreads();  // the previous block's walks are done with the table
            row_table(rowtab, P.grid[s], stride);
            __syncthreads();
        }
        const uint32_t i = b * BS + threadIdx.x;
        const bool live = i < sp.pbase + sp.n;
        // Own state: A (position, signed mass) and the density record
        // (velocity, rho, c2).
        float4 ai = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = ai;
        float ci2 = 0.0f;
        if (live) {
            ai = P.sA[i];
            vi = P.fr[2u * i + 1u];
            ci2 = P.fr[2u * i].w;
        }
        const bool fluid = live && ai.w > 0.0f && active;
        const SceneGrid &g = P.grid[s];
        const float4 *__restrict__ sA = P.sA;
        const float4 *__restrict__ fr = P.fr;
        // A boundary neighbour, or any neighbour in a scene of unequal fluid
        // masses, reads its mass from sA.
        const bool own_mass = sp.fmass_uniform == 0u;
        ForceAcc f{sp.gx, sp.gy, sp.gz, 0.0f, 0.0f, 0.0f};
        const float h2 = sp.h2;
        if constexpr (WALK == 0) {
            if (fluid) {
                const Cursor cur = make_cursor(g, sp.h, ai.x, ai.y, ai.z);
                const ForceConst fc = force_const(sp);
                for_rows(g, P.start, cur, [&](uint32_t, bool, uint32_t j0, uint32_t j1) {
                    for (uint32_t j = j0; j < j1; ++j) {
                        if (j == i) continue;
                        const float4 r0 = fr[2u * j];
                        const float ddx = ai.x - r0.x, ddy = ai.y - r0.y, ddz = ai.z - r0.z;
                        const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (!(r2 <= h2)) continue;
                        float4 aj, vj;
                        float cj;
                        nb_record(r0, fr[2u * j + 1u], j, sA, sp.mass, own_mass, aj, vj, cj);
                        force_pair(f, ai, vi, ci2, aj, vj, cj, fc);
                    }
                });
            }
        } else if constexpr (WALK == 1) {
            Cursor cur{};
            if (fluid) cur = make_cursor(g, sp.h, ai.x, ai.y, ai.z);
            NbWalk w = nb_begin(fluid ? P.nbs[i] : 0u, g, cur, rowtab);
            const ForceConst fc = force_const(sp);
            for (;;) {
                uint32_t j = 0;
                const int st = nb_step(w, P, g, cur, i, ai, h2, stride, j);
                if (!__any(st != kWalkDone)) break;
                if (st == kWalkPair) {
                    float4 aj, vj;
                    float cj;
                    nb_record(*at(fr, 2u * j), *at(fr, 2u * j + 1u), j, sA, sp.mass, own_mass, aj, vj, cj);
                    force_pair(f, ai, vi, ci2, aj, vj, cj, fc);
                }
            }
        } else if constexpr (WALK == 2) {
            walk_ahead(f, P, g, sp, i, fluid, ai, vi, ci2, rowtab, stride);
        } else if constexpr (WALK == 4 || WALK == 5) {
            __shared__ QueueLds<kQueueCap> queue_lds[BS / 64];
            if (!queue_walk<WALK == 5 ? kLoadPaired : kLoadPlain>(f, P, g, sp, i, fluid, ai, vi, ci2, rowtab,
                                                                  queue_lds[threadIdx.x >> 6]))
                walk_ahead(f, P, g, sp, i, fluid, ai, vi, ci2, rowtab, stride);
        } else if constexpr (WALK == 6) {
            // The queue walk with each wave's own cell rows staged in LDS.
            // Nine chunks per lane and 62 staged records per wave keep a
            // 256-thread workgroup under 26.6 KB of LDS: six per CU, six waves
            // per SIMD.
            constexpr uint32_t kCap = 9, kStage = 62;
            __shared__ QueueLds<kCap> queue_lds[BS / 64];
            __shared__ float4 stage_lds[BS / 64][kStage][2];
            QueueLds<kCap> &q = queue_lds[threadIdx.x >> 6];
            float4(*st)[2] = stage_lds[threadIdx.x >> 6];
            uint32_t lo, hi, slo, sn;
            const uint32_t n = queue_build(P, g, sp, i, fluid, ai, rowtab, q, lo, hi);
            if (__any(n > kCap)) {
                walk_ahead(f, P, g, sp, i, fluid, ai, vi, ci2, rowtab, stride);
            } else {
                stage_own_rows<kStage>(st, fr, lo, hi, slo, sn);
                queue_pairs<kLoadStaged, kCap>(f, P, sp, i, ai, vi, ci2, n, q, st, slo, sn);
            }
        } else {
            // Pair-parallel rows (pair_parallel) for waves whose rows all fit a
            // 32-bit mask and whose neighbours lie within 2^26 slots of the
            // row's first candidate; the walk otherwise.
            const uint32_t summ = fluid ? P.nbs[i] : 0u;
            const bool tail = ((summ >> 18) & kRowsPresent) != 0u;
            __shared__ PairLds pair_lds[BS / 64];
            if (__any(tail)) {
                walk_ahead(f, P, g, sp, i, fluid, ai, vi, ci2, rowtab, stride);
            } else {
                pair_parallel(f, P, g, sp, i, fluid, ai, vi, ci2, rowtab, pair_lds[threadIdx.x >> 6]);
            }
        }
        float ox = ai.x, oy = ai.y, oz = ai.z;
        if (live) {
            float4 vo = make_float4(vi.x, vi.y, vi.z, 0.0f);
            if (fluid) {
                if (P.acc) {
                    P.acc[i] = make_float4(f.ax, f.ay, f.az, 0.0f);
                    P.xsph[i] = make_float4(f.xx, f.xy, f.xz, 0.0f);
                }
                // fluid_step.rs:110-123
                const float dt = sp.dt, eps = sp.eps;
                float vx = (vi.x + f.ax * dt) + f.xx * eps;
                float vy = (vi.y + f.ay * dt) + f.xy * eps;
                float vz = (vi.z + f.az * dt) + f.xz * eps;
                const float s2 = vx * vx + vy * vy + vz * vz;
                if (s2 > sp.vmax2) {
                    const float k = sp.vmax / sqrtf(s2);
                    vx = vx * k; vy = vy * k; vz = vz * k;
                }
                ox = ai.x + vx * dt;
                oy = ai.y + vy * dt;
                oz = ai.z + vz * dt;
                vo = make_float4(vx, vy, vz, 0.0f);
            }
            P.cA[i] = make_float4(ox, oy, oz, ai.w);
            P.cV[i] = vo;
            P.cg[i] = P.sg[i];
        }
        wave_box(live, ox, oy, oz, &P.part[i / kPartSlots]);
    }
}

template __global__ void k_force_t<0, 64>(Pool P);
template __global__ void k_force_t<1, 64>(Pool P);
template __global__ void k_force_t<2, 64>(Pool P);
template __global__ void k_force_t<0, 128>(Pool P);
template __global__ void k_force_t<1, 128>(Pool P);
template __global__ void k_force_t<2, 128>(Pool P);
template __global__ void k_force_t<0, 256>(Pool P);
template __global__ void k_force_t<1, 256>(Pool P);
template __global__ void k_force_t<2, 256>(Pool P);
template __global__ void k_force_t<4, 64>(Pool P);
template __global__ void k_force_t<4, 128>(Pool P);
template __global__ void k_force_t<4, 256>(Pool P);
template __global__ void k_force_t<3, 64>(Pool P);
template __global__ void k_force_t<3, 128>(Pool P);
template __global__ void k_force_t<3, 256>(Pool P);
template __global__ void k_force_t<5, 64>(Pool P);
template __global__ void k_force_t<5, 128>(Pool P);
template __global__ void k_force_t<5, 256>(Pool P);
template __global__ void k_force_t<6, 64>(Pool P);
template __global__ void k_force_t<6, 128>(Pool P);
template __global__ void k_force_t<6, 256>(Pool P);

// The boundary particles whose reaction this substep needs (every boundary
// particle of a scene with reactions on all spans, the dynamic bodies'
// particles of a coupled scene), listed by grid slot so the reaction pass
// below runs on dense waves: a dynamic body's particles, and the wall
// particles of the tank's x walls, are scattered over the fluid's slots. The
// list's order is free (each entry writes its own reaction). Each workgroup
// takes kListBlocks consecutive pool blocks: a first sweep keeps every wave's
// ballot per block in LDS, one atomic per workgroup reserves the workgroup's
// range of the list, and each wave writes its entries into its own part of
// that range. (One atomic per wave on the single counter serialised about
// 50,000 atomics per substep on the 10M tank.)
__global__ __launch_bounds__(kBlock) void k_bforce_list(Pool P) {
    constexpr uint32_t kWaves = kBlock / 64u;
    __shared__ uint64_t mk[kListBlocks][kWaves];
    __shared__ uint32_t wcnt[kWaves];
    __shared__ uint32_t wbase;
    if (P.ctl->err) return;
    const uint32_t tick_sub = P.ctl->tick_sub, tick_len = P.ctl->tick_len;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    const uint32_t b0 = blockIdx.x * kListBlocks, nb = min(kListBlocks, P.nblocks - min(b0, P.nblocks));
    // The workgroup's original indices, all loads in flight at once (every
    // slot of a pool block is inside the gathered arrays).
    uint32_t gall[kListBlocks];
#pragma unroll
    for (uint32_t k = 0; k < kListBlocks; ++k) gall[k] = k < nb ? P.sg[(b0 + k) * kBlock + threadIdx.x] : 0u;
    uint32_t cnt = 0;  // this wave's entries (wave-uniform)
#pragma unroll
    for (uint32_t k = 0; k < kListBlocks; ++k) {
        if (k >= nb) break;
        const uint32_t b = b0 + k;
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        bool want = false;
        if (sp.react && substep_active(tick_sub, tick_len, sp.substeps)) {  // uniform per block
            const uint32_t j = b * kBlock + threadIdx.x;
            if (j < sp.pbase + sp.n) {
                const uint32_t gj = gall[k];
                want = gj >= sp.nf;  // boundary particles follow the fluid prefix
                if (want && sp.react == kReactDynamic) {  // only dynamic bodies consume a reaction
                    bool dyn = false;
                    for (uint32_t r = sp.rbody0; r < sp.rbody0 + sp.nrbody; ++r)
                        dyn = dyn || (gj >= P.rigid[r].gstart && gj < P.rigid[r].gend);
                    want = dyn;
                }
            }
        }
        const uint64_t m = __ballot(want);
        if (lane == 0u) mk[k][wave] = m;
        cnt += (uint32_t)__popcll(m);
    }
    if (lane == 0u) wcnt[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0u) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < kWaves; ++w) tot += wcnt[w];
        wbase = tot ? atomicAdd(&P.ctl->nreact, tot) : 0u;
    }
    __syncthreads();
    if (cnt == 0u) return;
    uint32_t base = wbase;
    for (uint32_t w = 0; w < wave; ++w) base += wcnt[w];
    for (uint32_t k = 0; k < nb; ++k) {
        const uint64_t m = mk[k][wave];
        if (m == 0ull) continue;
        if ((m >> lane) & 1ull) P.rlist[base + (uint32_t)__popcll(m & below)] = (b0 + k) * kBlock + threadIdx.x;
        base += (uint32_t)__popcll(m);
    }
}

// Reaction on every listed boundary particle:
// bforce[j] = 0 - X_1 - X_2 - ... over its fluid neighbours i in ascending
// index, X = d * ((m_i * (coef + cv)) / denom) with d = p_i - p_j (the
// reference accumulates it from the fluid side, i ascending:
// fluid_force.rs:196-206). The neighbours arrive in cell order, so they are
// subtracted in batches: each pass over the candidates keeps, sorted by
// original index, the kReactBatch smallest indices above the last one already
// subtracted (insertion into the lane's column of an LDS table of index and
// slot), then evaluates and subtracts them in that order, and passes again
// while any were left out. Any neighbour count gives the reference's sum bit
// for bit; one or two passes cover the usual case (a wall particle sees about
// half a neighbourhood, ~16 fluid particles), and the small table keeps the
// workgroup's LDS low enough for five waves per SIMD. Only each cell's fluid
// head is scanned, so a dry tank
// particle tests no candidate. A grid-stride loop over the list
// (kBforceThreads entries per workgroup and round).
__global__ __launch_bounds__(kBforceThreads) void k_bforce(Pool P) {
    __shared__ uint32_t skey[kReactBatch][kBforceThreads];   // original index, ascending
    __shared__ uint32_t sslot[kReactBatch][kBforceThreads];  // its slot
    if (P.ctl->err) return;
    const uint32_t nlist = P.ctl->nreact;
    const uint32_t lane = threadIdx.x;
    for (uint32_t q0 = blockIdx.x * kBforceThreads; q0 < nlist; q0 += gridDim.x * kBforceThreads) {
        const uint32_t qi = q0 + lane;
        if (qi >= nlist) continue;
        const uint32_t j = P.rlist[qi];
        const uint32_t s = P.block_scene[j / kBlock];
        const SceneParams &sp = P.params[s];
        const uint32_t gj = P.sg[j];
        const float4 aj = P.sA[j];
        const SceneGrid &g = P.grid[s];
        const float4 vj = P.fr[2u * j + 1u];
        const float rho_j = -vj.w, cj2 = P.fr[2u * j].w, mj = -aj.w;  // a boundary record: rho negated
        const Cursor cur = make_cursor(g, sp.h, aj.x, aj.y, aj.z);
        const float h = sp.h, h2 = sp.h2, ks = sp.ks, alpha_c = sp.alpha_c;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        bool first = true;
        uint32_t last = 0;  // original index of the last neighbour subtracted
        for (;;) {
            uint32_t cnt = 0;
            bool more = false;
            // Fluid neighbours only: each cell's fluid head (its members in
            // ascending original index put the fluid prefix first).
            for_row_windows(g, cur, [&](uint32_t, uint32_t cf, uint32_t cl, bool) {
                // The window's (at most three) cell starts and fluid counts
                // load together.
                const uint32_t nc = cl - cf;  // 0..2
                uint32_t st[4], fc[3];
#pragma unroll
                for (uint32_t k = 0; k < 4u; ++k) st[k] = P.start[cf + min(k, nc + 1u)];
#pragma unroll
                for (uint32_t k = 0; k < 3u; ++k) fc[k] = P.fcount[cf + min(k, nc)];
#pragma unroll
                for (uint32_t w = 0; w < 3u; ++w) {
                    if (w > nc) break;
                    const uint32_t i0 = st[w], i1 = i0 + min(fc[w], st[w + 1u] - i0);
                    for (uint32_t ib = i0; ib < i1; ib += 4u) {
                        // four candidates' records in flight at once (the grid-ordered
                        // array is padded past its end)
                        float4 a4[4];
#pragma unroll
                        for (uint32_t u = 0; u < 4u; ++u) a4[u] = *at(P.sA, ib + u);
#pragma unroll
                        for (uint32_t u = 0; u < 4u; ++u) {
                            const uint32_t i = ib + u;
                            const float4 ai = a4[u];
                            if (i >= i1 || !(ai.w > 0.0f)) continue;  // fluid neighbours only
                            const float ddx = ai.x - aj.x, ddy = ai.y - aj.y, ddz = ai.z - aj.z;
                            const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                            if (!(r2 <= h2)) continue;
                            const uint32_t gi = P.sg[i];
                            if (!first && gi <= last) continue;  // subtracted by an earlier pass
                            if (cnt == kReactBatch) {
                                more = true;
                                if (gi > skey[kReactBatch - 1][lane]) continue;  // a later pass takes it
                            }
                            // insertion by original index (a full batch drops its largest)
                            uint32_t k = cnt < kReactBatch ? cnt++ : kReactBatch - 1u;
                            while (k > 0 && skey[k - 1][lane] > gi) {
                                skey[k][lane] = skey[k - 1][lane];
                                sslot[k][lane] = sslot[k - 1][lane];
                                --k;
                            }
                            skey[k][lane] = gi;
                            sslot[k][lane] = i;
                        }
                    }
                }
            });
            // The batch in ascending original index: d = p_i - p_j as seen
            // from the fluid particle i. Each neighbour's records load one
            // neighbour ahead of its evaluation.
            float4 an = aj, vn = vj;
            float cn = cj2;
            if (cnt != 0u) {
                const uint32_t i = sslot[0][lane];
                an = *at(P.sA, i);
                vn = *at(P.fr, 2u * i + 1u);
                cn = at(P.fr, 2u * i)->w;
            }
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4 ai = an, vi = vn;
                const float ci2 = cn;
                if (k + 1u < cnt) {
                    const uint32_t i = sslot[k + 1u][lane];
                    an = *at(P.sA, i);
                    vn = *at(P.fr, 2u * i + 1u);
                    cn = at(P.fr, 2u * i)->w;
                }
                const float ddx = ai.x - aj.x, ddy = ai.y - aj.y, ddz = ai.z - aj.z;
                const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                const float rho_i = vi.w, mi = ai.w;
                const float r = sqrtf(r2);
                const float t = h - r;
                const float coef = (mj * ((ks * t) * t)) * (ci2 + cj2);
                const float denom = fmaxf(r, 1e-9f);
                const float wx = vi.x - vj.x, wy = vi.y - vj.y, wz = vi.z - vj.z;
                const float vdn = -(wx * ddx + wy * ddy + wz * ddz);
                float cv = 0.0f;
                if (vdn > 0.0f) {
                    const float mu = (vdn * h) / (r2 + 0.01f * h2);
                    cv = (mj * ((alpha_c * mu) / (0.5f * (rho_i + rho_j)))) * ((ks * t) * t);
                }
                const float q = (mi * (coef + cv)) / denom;
                fx = fx - ddx * q;
                fy = fy - ddy * q;
                fz = fz - ddz * q;
            }
            if (!more) break;
            first = false;
            last = skey[kReactBatch - 1][lane];
        }
        const uint32_t d = sp.pbase + gj;
        P.bf[d] = make_float4(fx, fy, fz, 0.0f);
        P.bp[d] = make_float4(aj.x, aj.y, aj.z, 0.0f);
    }
}

// Per span: F = sum bforce[k], tau += (pos[k] - origin) x f_k in index order
// (fluid_boundary.rs:10-31), accumulated over the substeps of a tick and
// scaled by 1/substeps at its end (fluid_step.rs:75-89).
//
// One workgroup per span. The terms (f_k and the cross products) do not
// depend on the running sums, so waves 1..3 compute them, kFold particles per
// chunk (each wave a contiguous segment of kSeg), into one of two LDS buffers
// (component-major rows, padded by one float4 so the six rows start in
// different banks), while lane c (c < 6) of wave 0 adds component c of the
// other buffer's terms in index order: six independent chains of dependent
// additions, fed from LDS 64 (then 16) terms at a time. A particle whose six
// terms are all zero (a tank particle without fluid neighbours: bforce = +0,
// so the cross products are +-0) is left out of its segment, order kept:
// adding +-0 leaves a sum unchanged, and no sum is ever -0 (it starts at +0,
// and x + y is -0 in round-to-nearest only when both are -0); for the same
// reason each segment is padded with +0 terms to a multiple of 16. Same
// values added in the same order as a serial fold over every particle, same
// bits; the chain is the kernel's time (about 2.9 ns per addition on MI355X,
// against 2.6 ns for operands already in registers), and the terms' loads
// and products run beside it. A chunk (1536 particles, ~4.5 us of additions
// when all are wet) outlasts the latency of the producers' loads, issued one
// chunk ahead.
__global__ __launch_bounds__(kReactThreads) void k_reactions(Pool P) {
    constexpr uint32_t kProd = kReactThreads - 64u;           // producer lanes
    constexpr uint32_t kPer = 8, kFold = kPer * kProd;         // particles per chunk
    constexpr uint32_t kWaves = kProd / 64u, kSeg = kFold / kWaves;  // one segment per producer wave
    constexpr uint32_t kRow = kFold / 4u + 1u;                 // float4 per component row
    __shared__ float4 term[2][6][kRow];
    __shared__ uint32_t segn[2][kWaves];  // terms kept in each segment
    if (P.ctl->err) return;
    const uint32_t k = blockIdx.x;
    if (k >= P.nspans) return;
    const SpanDev sd = P.spans[k];
    if (!sd.active) return;
    const uint32_t sub = sd.substeps;
    const uint32_t tick_sub = P.ctl->tick_sub, tick_len = max(P.ctl->tick_len, sub);
    if (!substep_active(tick_sub, tick_len, sub)) return;  // the scene does not step in this launch
    const uint32_t phase = tick_sub + sub - tick_len;      // the scene's substep index inside its tick
    const bool folder = threadIdx.x < 64u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t pw = (threadIdx.x >> 6) - 1u;  // producer wave (waves 1..3)
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    float acc = 0.0f;  // component `lane` of (F, tau) for lanes 0..5 of wave 0
    float4 f[kPer], p[kPer];
    auto fetch = [&](uint32_t q) {
#pragma unroll
        for (uint32_t m = 0; m < kPer; ++m) {
            const uint32_t j = q + pw * kSeg + m * 64u + lane;
            f[m] = p[m] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (j < sd.end) {
                f[m] = P.bf[j];
                p[m] = P.bp[j];
            }
        }
    };
    // The segment's nonzero terms, in index order (past the span's end the
    // terms are zero too: nothing is written for them).
    auto put = [&](uint32_t buf) {
        float *tc = &term[buf][0][0].x;
        uint32_t n = 0;  // terms kept so far (wave-uniform)
#pragma unroll
        for (uint32_t m = 0; m < kPer; ++m) {
            const float rx = p[m].x - sd.origin[0], ry = p[m].y - sd.origin[1], rz = p[m].z - sd.origin[2];
            const float t3 = ry * f[m].z - rz * f[m].y;
            const float t4 = rz * f[m].x - rx * f[m].z;
            const float t5 = rx * f[m].y - ry * f[m].x;
            const bool keep = f[m].x != 0.0f || f[m].y != 0.0f || f[m].z != 0.0f || t3 != 0.0f || t4 != 0.0f ||
                              t5 != 0.0f;
            const uint64_t kb = __ballot(keep);
            if (keep) {
                const uint32_t u = pw * kSeg + n + (uint32_t)__popcll(kb & below);
                tc[0 * 4u * kRow + u] = f[m].x;
                tc[1 * 4u * kRow + u] = f[m].y;
                tc[2 * 4u * kRow + u] = f[m].z;
                tc[3 * 4u * kRow + u] = t3;
                tc[4 * 4u * kRow + u] = t4;
                tc[5 * 4u * kRow + u] = t5;
            }
            n += (uint32_t)__popcll(kb);
        }
        // Pad the segment with +0 terms to a multiple of 16 (a no-op for the
        // sums, as the left-out ones), so the fold reads whole float4 groups.
        const uint32_t npad = (n + 15u) & ~15u;
        if (n + lane < npad) {
            const uint32_t u = pw * kSeg + n + lane;
#pragma unroll
            for (uint32_t c = 0; c < 6u; ++c) tc[c * 4u * kRow + u] = 0.0f;
        }
        if (lane == 0u) segn[buf][pw] = npad;
    };
    if (!folder) {
        fetch(sd.start);
        put(0);
        fetch(sd.start + kFold);
    }
    __syncthreads();
    uint32_t buf = 0;
    for (uint32_t q = sd.start; q < sd.end; q += kFold, buf ^= 1u) {
        if (folder) {
            if (lane < 6u) {
#pragma unroll 1
                for (uint32_t w = 0; w < kWaves; ++w) {
                    const uint32_t nq = segn[buf][w];
                    const float4 *row = &term[buf][lane][w * (kSeg / 4u)];
                    uint32_t u = 0;
                    for (; u + 64u <= nq; u += 64u) {
                        float4 tv[16];
#pragma unroll
                        for (uint32_t v = 0; v < 16u; ++v) tv[v] = row[u / 4u + v];
#pragma unroll
                        for (uint32_t v = 0; v < 16u; ++v) {
                            acc = acc + tv[v].x;
                            acc = acc + tv[v].y;
                            acc = acc + tv[v].z;
                            acc = acc + tv[v].w;
                        }
                    }
                    for (; u < nq; u += 16u) {  // nq is a multiple of 16
                        float4 tv[4];
#pragma unroll
                        for (uint32_t v = 0; v < 4u; ++v) tv[v] = row[u / 4u + v];
#pragma unroll
                        for (uint32_t v = 0; v < 4u; ++v) {
                            acc = acc + tv[v].x;
                            acc = acc + tv[v].y;
                            acc = acc + tv[v].z;
                            acc = acc + tv[v].w;
                        }
                    }
                }
            }
        } else if (q + kFold < sd.end) {
            put(buf ^ 1u);  // the next chunk, into the buffer folded in the previous iteration
            fetch(q + 2u * kFold);
        }
        __syncthreads();  // this chunk is folded and the next one written
    }
    if (!folder || lane >= 6u) return;
    float *o = P.react + 6ull * k + lane;
    *o = phase == 0 ? acc : *o + acc;
    if (sub > 1 && phase == sub - 1) *o = *o * (1.0f / (float)sub);
}

// Resident scenes move to their new slots at a layout: move k copies the
// records of slots [from_k, from_k + n_k) to [to_k, to_k + n_k); moves are
// listed as (from, to, first) with first the running count, and a sentinel
// entry whose first is the total.
__global__ __launch_bounds__(kBlock) void k_move_records(const float4 *__restrict__ sA, const float4 *__restrict__ sV,
                                                         const uint32_t *__restrict__ sG, float4 *dA, float4 *dV,
                                                         uint32_t *dG, const uint32_t *__restrict__ moves,
                                                         uint32_t nmoves, uint32_t total) {
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
        uint32_t lo = 0, hi = nmoves;  // the move whose range holds t
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (moves[3u * mid + 2u] <= t)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t k = t - moves[3u * lo + 2u];
        const uint32_t from = moves[3u * lo] + k, to = moves[3u * lo + 1u] + k;
        dA[to] = sA[from];
        dV[to] = sV[from];
        dG[to] = sG[from];
    }
}

// p[k] = k for k < n (original indices of a new scene).
__global__ __launch_bounds__(kBlock) void k_iota(uint32_t *p, uint32_t n) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) p[k] = k;
}

// Retire scene s in place: it keeps its slots and cells until the next
// layout, but every kernel now treats it as empty (no particle, no fluid, no
// reactions, no bodies), so it costs nothing and changes nothing.
__global__ void k_freeze(SceneParams *params, uint32_t s) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        params[s].n = 0u;
        params[s].nf = 0u;
        params[s].react = 0u;
        params[s].nrbody = 0u;
    }
}

// Scatter scene s's cur state to index order (AoS, 3 floats per particle at
// its original index).
__global__ __launch_bounds__(kBlock) void k_unsort(Pool P, uint32_t s, float *opos, float *ovel) {
    const SceneParams &sp = P.params[s];
    const uint32_t b0 = sp.pbase / kBlock, nblk = sp.alloc / kBlock;
    for (uint32_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint32_t m = (b0 + k) * kBlock + threadIdx.x;
        if (m >= sp.pbase + sp.n) continue;
        const size_t d = 3ull * P.cg[m];
        const float4 a = P.cA[m], v = P.cV[m];
        opos[d] = a.x; opos[d + 1] = a.y; opos[d + 2] = a.z;
        ovel[d] = v.x; ovel[d + 1] = v.y; ovel[d + 2] = v.z;
    }
}

// Scatter sorted per-particle fields to index order (pool slot pbase + gid).
__global__ __launch_bounds__(kBlock) void k_unsort_fields(Pool P, float *odens, float *opress, uint32_t *oncount,
                                                          float *oacc, float *oxs) {
    for (uint32_t b = blockIdx.x; b < P.nblocks; b += gridDim.x) {
        const uint32_t s = P.block_scene[b];
        const SceneParams &sp = P.params[s];
        const uint32_t i = b * kBlock + threadIdx.x;
        if (i >= sp.pbase + sp.n) continue;
        const uint32_t gi = P.sg[i];
        const uint32_t d = sp.pbase + gi;
        // rho from the density record, p from it by the same Tait step (a
        // scene without fluid is never stepped: both stay zero)
        const float rho = fabsf(P.fr[2u * i + 1u].w);
        float p = 0.0f, c2;
        if (sp.nf != 0u) tait(sp, rho, p, c2);
        odens[d] = rho;
        opress[d] = p;
        if (oncount && P.ncount) oncount[d] = P.ncount[i];
        if (oacc && P.acc && gi < sp.nf) {
            const float4 a = P.acc[i], x = P.xsph[i];
            oacc[3ull * d] = a.x; oacc[3ull * d + 1] = a.y; oacc[3ull * d + 2] = a.z;
            oxs[3ull * d] = x.x; oxs[3ull * d + 1] = x.y; oxs[3ull * d + 2] = x.z;
        }
    }
}

}  // namespace vxl
