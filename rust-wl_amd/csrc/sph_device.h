// Device helpers shared by the kernel translation units.
#ifndef VXL_SPH_DEVICE_H
#define VXL_SPH_DEVICE_H
#include <hip/hip_runtime.h>

#include "sph_layout.h"

namespace vxl {
namespace dev {

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// Block-wide AABB of the positions handed in by live lanes, written as this
// block's partial box (ordered u32; the per-scene reduction runs after the
// launch). No atomics: thousands of blocks merging into six words would
// serialise on one L2 channel.
__device__ __forceinline__ void block_box(bool live, float x, float y, float z, SceneBox *part) {
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
        part->lo[0] = lx; part->lo[1] = ly; part->lo[2] = lz;
        part->hi[0] = hx; part->hi[1] = hy; part->hi[2] = hz;
    }
    __syncthreads();
}

}  // namespace dev
}  // namespace vxl
#endif
