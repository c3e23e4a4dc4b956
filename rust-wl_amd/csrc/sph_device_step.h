// A single scene held in caller-owned device arrays (index order, 3 floats
// per particle position/velocity, 1 per mass), advanced in place by whole
// ticks with the engine's substep kernels. The pool of the one scene (state,
// sort, neighbour and cell arrays, scene tables) is carved from a workspace
// the caller of device_step provides.
#ifndef VXL_SPH_DEVICE_STEP_H
#define VXL_SPH_DEVICE_STEP_H
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "sph_kernels.h"
#include "sph_scene.h"

namespace vxl {

// Bytes of the workspace of a scene of n particles and ccap grid cells (the
// scene's grid plus its pad cell). With base != nullptr, also the pool carved
// from it and the pointers of the tables the host fills.
size_t device_step_workspace(uint32_t n, uint32_t ccap, char *base, Pool *out, uint32_t **tile_sums,
                             SceneParams **params, uint32_t **block_scene);

// Advance the scene `ticks` ticks on `stream` inside workspace ws (ws_bytes
// at least device_step_workspace(n, ccap)); pos/vel are read at the start and
// written at the end, only when every substep completed. Returns after the
// stream has finished. Throws CapacityError (grid beyond ccap - 1 cells),
// std::domain_error (non-finite positions), std::invalid_argument (a mass not
// positive and finite).
void device_step(const SceneConfig &cfg, uint32_t n, uint32_t nf, float *pos, float *vel, const float *pmass,
                 uint32_t ticks, hipStream_t stream, char *ws, size_t ws_bytes, uint32_t ccap);

// Largest scene of device_step_noalloc.
constexpr uint32_t kNoAllocParticles = 1u << 17;  // 131,072 particles (fluid + boundary)
constexpr uint32_t kNoAllocCells = 1u << 19;      // grid cells + the pad cell

// device_step with the workspace taken from a file-scope device array of the
// library (one per device, kept for the life of the process) instead of an
// allocation: nothing is allocated. Throws CapacityError beyond
// kNoAllocParticles particles (before any launch) or kNoAllocCells - 1 grid
// cells. Calls on one device are serialised (they share the array); the
// array keeps the last call's bytes, and every word a call reads first is
// set by that call.
void device_step_noalloc(const SceneConfig &cfg, uint32_t n, uint32_t nf, float *pos, float *vel, const float *pmass,
                         uint32_t ticks, hipStream_t stream);

// Throws unless [p, p + bytes) lies inside one device allocation of this
// process (a host pointer or a short array would make the kernels fault).
void check_device_range(const void *p, size_t bytes, const char *what);

}  // namespace vxl
#endif
