// Kernel argument bundle and kernel declarations of the WCSPH substep.
#ifndef VXL_SPH_KERNELS_H
#define VXL_SPH_KERNELS_H
#include <hip/hip_runtime.h>

#include "sph_layout.h"

namespace vxl {

// Device pointers of the particle pool and cell tables (see sph_layout.h).
//
// Particle records are packed for one 16-byte load per neighbour candidate:
//   A = (x, y, z, m)   m = +pmass for fluid, -pmass for boundary particles
//   V = (vx, vy, vz, 0)
// In grid order the gather writes A and the position and velocity halves of
// each particle's 32-byte density record fr (below), to which the density
// pass adds c2 and rho.
// Cell tables are stored with x fastest (internal cell = (z*ny + y)*nx + x),
// so the three cells dx = -1, 0, +1 of one (dz, dy) row of the reference's
// dz -> dy -> dx enumeration are one contiguous slot range.
struct Pool {
    // state in the previous substep's grid order
    float4 *cA, *cV;
    uint32_t *cg;
    // state gathered into this substep's grid order
    float4 *sA;
    uint32_t *sg;
    // [2 * slots] the density pass's record of each particle, one 32-byte
    // pair (x, y, z, c2), (v, rho) with c2 = p / (rho * rho) and rho negated
    // for boundary particles (the sign is the class): the force pass loads a
    // neighbour in one line. The only store of rho and c2.
    float4 *fr;
    // accepted neighbours, written by the density pass, read by the force pass:
    //   nbm[r * nblocks * kBlock + i]  bit k = candidate j0 + k of (dz, dy) row r
    //                                  (r = 3 (dz + 1) + dy + 1) is a neighbour
    //   nbs[i]  bits 0-8: row r holds a neighbour or a tail; 9-17: row r starts
    //           at cell dx = -1; 18-26: row r has more than 32 candidates (the
    //           tail j0 + 32 .. end is tested again by the reader)
    uint32_t *nbm, *nbs;
    // sort
    uint32_t *key, *slot;
    uint2 *items;              // [slots] (cur slot, gid) of each cell member in arrival order
    uint32_t *counts, *start;  // counts[ccap], start[ccap + 1]
    uint32_t *fcount;          // [ccap] fluid items heading each non-empty cell (this substep)
    // optional diagnostics (grid order; null = off)
    uint32_t *ncount;
    float4 *acc, *xsph;
    // scene tables
    const SceneParams *params;
    SceneGrid *grid;
    SceneBox *box;  // [2][nscenes] ping-pong
    SceneBox *part; // [nblocks] per-block partial boxes
    Control *ctl;
    const uint32_t *block_scene;
    uint32_t nblocks;  // pool slots / kBlock
    uint32_t nscenes;
    uint32_t ccap;     // cell table capacity
    uint32_t nspans;   // reaction spans (0 = no reaction pass)
    // reactions (index order, pool slots): boundary force and position
    float4 *bf, *bp;
    uint32_t *rlist;   // [slots] the grid slots of this substep's reacting boundary particles (ctl->nreact)
    SpanDev *spans;    // torque origins of dynamic bodies follow their pose
    float *react;      // [nspans][6]: F, tau (tick accumulator)
    // dynamic rigid bodies (all scenes) and their body-frame Akinci lattices
    RigidDev *rigid;
    uint32_t nrigid;
    const float4 *rlat;
    // provider walls (walls = null: no scene of the pool has walls)
    const WallSceneDev *walls;  // [nscenes]
    uint32_t *went;   // [2][nblocks * kBlock] table entry of slot pbase + fluid index (kWallNone: none)
    SepState *sep;    // [nscenes + 1] coincident separation of this substep; [nscenes]: pool totals
                      // (k = pushed particles of all scenes, changed[m] = some scene moved in round m)
    uint64_t *wsep;   // [2 * slots] (fluid index << 32 | slot) of the pushed ones, from 2 * pbase
    float4 *wn;       // [slots] contact normal sum of a pushed particle (grid order)
    float4 *fpos;     // [2][slots] positions of the pushed ones (fixed-point rounds), from pbase
    uint32_t *hhead;  // [3][2 * slots] spatial hash buckets of the pushed ones, from 2 * pbase
    uint32_t *hnext;  // [3][slots] bucket chains, from pbase
    uint32_t wslots;  // slots of the wall arrays (their per-copy stride)
    uint32_t *wband;  // [slots] fluid slots with mirror ghosts this substep
    uint32_t *wband_n;
};

__global__ void k_box(Pool P);
// Boundary segment of scene s from index-order device arrays (3 floats per
// boundary particle), the scene's partial boxes, its AABB emptied for
// k_box_reduce.
__global__ void k_boundary_scatter(Pool P, uint32_t s, const float *bpos, const float *bvel);
constexpr uint32_t kBoxReduceSpan = kBlock * 8;  // partial boxes per k_box_reduce block
__global__ void k_box_reduce(Pool P);
__global__ void k_grid_box(Pool P, uint32_t tick_sub, uint32_t tick_len);
__global__ void k_keys(Pool P);
// ITEMS cells per thread: a scan tile holds kBlock * ITEMS cells (4, 8, 16).
template <uint32_t ITEMS>
__global__ void k_scan_reduce(Pool P, uint32_t *tile_sums);
template <uint32_t ITEMS>
__global__ void k_scan_tiles(Pool P, uint32_t *tile_sums);
template <uint32_t ITEMS>
__global__ void k_scan_apply(Pool P, const uint32_t *tile_sums);
constexpr uint32_t kScanMinTile = kBlock * 4;  // smallest tile (sizes the tile-sum array)
__global__ void k_place(Pool P);
__global__ void k_canon_gather(Pool P);
// BS threads per workgroup (64, 128, 256); STAGE records of LDS staging per
// wave and neighbour row (128, 256, 512).
template <uint32_t BS, uint32_t STAGE>
__global__ void k_density(Pool P);
template <uint32_t BS>
__global__ void k_density_walls(Pool P);  // pools with half-space walls: mirror ghosts in the sums
template <uint32_t WALK, uint32_t BS>
__global__ void k_force_t(Pool P);    // 0: candidate scan; 1: walk of the density pass's masks;
                                      // 2: the walk, loads one pair ahead; 3: pair-parallel rows;
                                      // 4: queue walk; 5: queue walk, own rows staged in LDS
__global__ void k_bforce_list(Pool P);  // lists the boundary slots the reaction pass works on
__global__ void k_bforce(Pool P);
// Provider walls (scenes with WallSceneDev::mode != kWallsOff):
// k_wall_band + k_wall_density after the density pass (table scenes: the
// particles with mirror ghosts listed, then their density and Tait again),
// k_wall_project after the force pass (projection of the integrated state),
// k_wall_separate (one block per scene: the serial coincident separation
// where the parallel rounds did not converge), k_wall_box (the blocks'
// partial boxes again).
__global__ void k_wall_band(Pool P);
__global__ void k_wall_density(Pool P);
__global__ void k_wall_project(Pool P);
__global__ void k_wall_separate(Pool P);
__global__ void k_wall_box(Pool P);
// Parallel coincident separation (between k_wall_project and k_wall_separate):
// k_sep_begin, k_sep_round(m) for m < kSepIters, k_sep_finish.
__global__ void k_sep_begin(Pool P);
__global__ void k_sep_round(Pool P, uint32_t m);
__global__ void k_sep_finish(Pool P);
// Table entry index of scene s, side: every fluid index of the scene to
// kWallNone, then went[pbase + ids[e]] = e.
__global__ void k_wall_index_clear(Pool P, uint32_t s, uint32_t side);
__global__ void k_wall_index_set(Pool P, uint32_t s, uint32_t side, const uint32_t *ids, uint32_t ne);
constexpr uint32_t kSortBlocks = 4;      // k_keys, k_place: pool blocks per workgroup iteration
constexpr uint32_t kListBlocks = 16;     // k_bforce_list: pool blocks per workgroup
constexpr uint32_t kBforceThreads = 128;  // k_bforce: slots per workgroup (LDS: 8 B per slot and batch entry)
constexpr uint32_t kReactThreads = 256;  // k_reactions: one workgroup per span (1 folding wave, 3 producers)
__global__ void k_reactions(Pool P);
// Rigid leg, once per tick: integrate the bodies (before the first substep),
// rebuild their particles from the pose and the scenes' AABBs, keep the
// tick's reaction for the next tick (after the last substep).
__global__ void k_rigid(Pool P);
__global__ void k_body_refresh(Pool P);
__global__ void k_rigid_collect(Pool P);
__global__ void k_unsort(Pool P, uint32_t s, float *opos, float *ovel);  // scene s, index order
__global__ void k_freeze(SceneParams *params, uint32_t s);               // scene s: no particles from now on
__global__ void k_move_records(const float4 *sA, const float4 *sV, const uint32_t *sG, float4 *dA, float4 *dV,
                               uint32_t *dG, const uint32_t *moves, uint32_t nmoves, uint32_t total);
__global__ void k_iota(uint32_t *p, uint32_t n);
__global__ void k_unsort_fields(Pool P, float *odens, float *opress, uint32_t *oncount, float *oacc, float *oxs);

}  // namespace vxl
#endif
