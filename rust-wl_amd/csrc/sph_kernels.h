// Kernel argument bundle and kernel declarations of the WCSPH substep.
#ifndef VXL_SPH_KERNELS_H
#define VXL_SPH_KERNELS_H
#include <hip/hip_runtime.h>

#include "sph_layout.h"

namespace vxl {

// Device pointers of the particle pool and cell tables (see sph_layout.h).
struct Pool {
    // state in the previous substep's grid order
    float *cx, *cy, *cz, *cvx, *cvy, *cvz, *cm;
    uint32_t *cg;
    // state gathered into this substep's grid order
    float *sx, *sy, *sz, *svx, *svy, *svz, *sm;
    uint32_t *sg;
    // sort
    uint32_t *key, *slot, *items_m, *items_g;
    uint32_t *counts, *start;  // counts[ccap], start[ccap + 1]
    // per-particle fields (grid order)
    float *dens, *press, *c2;
    // optional diagnostics (grid order; null = off)
    uint32_t *ncount;
    float *ax, *ay, *az, *xx, *xy, *xz;
    // scene tables
    const SceneParams *params;
    SceneGrid *grid;
    SceneBox *box;  // [2][nscenes] ping-pong
    Control *ctl;
    const uint32_t *block_scene;
    uint32_t nblocks;  // pool slots / kBlock
    uint32_t nscenes;
    uint32_t ccap;     // cell table capacity
    uint32_t nspans;   // reaction spans (0 = no reaction pass)
    // reactions (index order, pool slots): boundary force and position
    float *bfx, *bfy, *bfz, *bpx, *bpy, *bpz;
    const SpanDev *spans;
    float *react;      // [nspans][6]: F, tau (tick accumulator)
};

__global__ void k_box(Pool P);
__global__ void k_grid_box(Pool P);
__global__ void k_keys(Pool P);
__global__ void k_scan_reduce(Pool P, uint32_t *tile_sums);
__global__ void k_scan_tiles(Pool P, uint32_t *tile_sums);
__global__ void k_scan_apply(Pool P, const uint32_t *tile_sums);
__global__ void k_place(Pool P);
__global__ void k_canon_gather(Pool P);
__global__ void k_density(Pool P);
__global__ void k_force(Pool P);
__global__ void k_bforce(Pool P);
__global__ void k_reactions(Pool P);
__global__ void k_unsort(Pool P, float *ox, float *oy, float *oz, float *ovx, float *ovy, float *ovz);
__global__ void k_unsort_fields(Pool P, float *odens, float *opress, uint32_t *oncount, float *oacc,
                                float *oxs);

}  // namespace vxl
#endif
