// One scene held in caller-owned device arrays, stepped in place (see
// sph_device_step.h): the pool of a single scene carved from one workspace,
// the substep kernels of the engine, no host copy of the particles.
#include "sph_device_step.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "sph_engine.h"

namespace vxl {

namespace {

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(e, std::string(what) + ": " + hipGetErrorString(e));
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// The scene's tables and constants at the start of a call: every word a
// substep reads before writing it is set here (the cell counts, the control
// block, both AABB slots), since a workspace keeps the previous call's bytes.
__global__ void k_ds_init(Pool P, SceneParams sp, SceneParams *params) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *params = sp;
        *P.ctl = Control{};
        for (int k = 0; k < 2; ++k)
            for (int a = 0; a < 3; ++a) {
                P.box[k].lo[a] = 0xffffffffu;
                P.box[k].hi[a] = 0u;
            }
    }
}

// Particles into the pool records (A = position and signed mass, V =
// velocity, original index), zero cell counts, the block -> scene map.
// A mass that is not positive and finite abandons the call.
__global__ void k_ds_load(Pool P, uint32_t n, uint32_t nf, const float *__restrict__ pos,
                          const float *__restrict__ vel, const float *__restrict__ pmass, uint32_t *block_scene) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (uint32_t c = tid; c < P.ccap; c += stride) P.counts[c] = 0u;
    for (uint32_t b = tid; b < P.nblocks; b += stride) block_scene[b] = 0u;
    for (uint32_t i = tid; i < n; i += stride) {
        const float m = pmass[i];
        if (!(m > 0.0f) || !isfinite(m)) atomicOr(&P.ctl->err, kErrMass);
        const size_t q = 3ull * i;
        P.cA[i] = make_float4(pos[q], pos[q + 1], pos[q + 2], i < nf ? m : -m);
        P.cV[i] = make_float4(vel[q], vel[q + 1], vel[q + 2], 0.0f);
        P.cg[i] = i;
    }
}

}  // namespace

size_t device_step_workspace(uint32_t n, uint32_t ccap, char *base, Pool *out, uint32_t **tile_sums,
                             SceneParams **params, uint32_t **block_scene) {
    const uint64_t slots = ((uint64_t)n + kBlock - 1) / kBlock * kBlock;
    const uint64_t blocks = slots / kBlock;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char * {
        char *p = base ? base + off : nullptr;
        off = align256(off + bytes);
        return p;
    };
    Pool P{};
    P.cA = (float4 *)take(16 * slots);
    P.cV = (float4 *)take(16 * slots);
    P.cg = (uint32_t *)take(4 * slots);
    P.sA = (float4 *)take(16 * (slots + kReadPad));
    P.sg = (uint32_t *)take(4 * slots);
    P.fr = (float4 *)take(32 * slots);
    P.nbm = (uint32_t *)take(36 * slots);
    P.nbs = (uint32_t *)take(4 * slots);
    P.key = (uint32_t *)take(4 * slots);
    P.slot = (uint32_t *)take(4 * slots);
    P.items = (uint2 *)take(8 * slots);
    P.counts = (uint32_t *)take(4ull * ccap);
    P.start = (uint32_t *)take(4ull * ccap + 4);
    P.fcount = (uint32_t *)take(4ull * ccap);
    uint32_t *ts = (uint32_t *)take(4ull * ((ccap + 2047) / 2048 + 1));
    SceneParams *pr = (SceneParams *)take(sizeof(SceneParams));
    P.params = pr;
    P.grid = (SceneGrid *)take(sizeof(SceneGrid));
    P.box = (SceneBox *)take(2 * sizeof(SceneBox));
    P.part = (SceneBox *)take(sizeof(SceneBox) * (slots / 64 + 1));
    P.ctl = (Control *)take(sizeof(Control));
    uint32_t *bs = (uint32_t *)take(4 * blocks);
    P.block_scene = bs;
    P.nblocks = (uint32_t)blocks;
    P.nscenes = 1;
    P.ccap = ccap;
    if (out) *out = P;
    if (tile_sums) *tile_sums = ts;
    if (params) *params = pr;
    if (block_scene) *block_scene = bs;
    return off;
}

void device_step(const SceneConfig &cfg, uint32_t n, uint32_t nf, float *pos, float *vel, const float *pmass,
                 uint32_t ticks, hipStream_t stream, char *ws, size_t ws_bytes, uint32_t ccap) {
    Pool P;
    uint32_t *tile_sums = nullptr, *block_scene = nullptr;
    SceneParams *params = nullptr;
    if (device_step_workspace(n, ccap, ws, &P, &tile_sums, &params, &block_scene) > ws_bytes)
        throw CapacityError("workspace smaller than the scene needs");
    SceneParams sp = scene_params(cfg, n, nf);
    sp.pbase = 0;
    sp.alloc = P.nblocks * kBlock;
    sp.fmass_uniform = 0u;  // every neighbour reads its own mass (no host pass over the masses)
    const uint32_t nb = P.nblocks;
    const uint32_t grid_load = std::max<uint32_t>(1u, std::min<uint32_t>((std::max(n, ccap) + kBlock - 1) / kBlock,
                                                                         8192u));
    const uint32_t grid_scan = std::min<uint32_t>((ccap + 2047) / 2048 + 1, 4096u);  // tiles of 2048 cells
    const uint32_t npart = nb * (kBlock / 64u);  // partial boxes (one per wave of slots)
    const uint32_t grid_reduce =
        std::max<uint32_t>(1u, std::min<uint32_t>((npart + kBoxReduceSpan - 1) / kBoxReduceSpan, 4096u));
    hipLaunchKernelGGL(k_ds_init, dim3(1), dim3(64), 0, stream, P, sp, params);
    hipLaunchKernelGGL(k_ds_load, dim3(grid_load), dim3(kBlock), 0, stream, P, n, nf, (const float *)pos,
                       (const float *)vel, pmass, block_scene);
    hipLaunchKernelGGL(k_box, dim3(nb), dim3(kBlock), 0, stream, P);
    hipLaunchKernelGGL(k_box_reduce, dim3(grid_reduce), dim3(kBlock), 0, stream, P);
    // The engine's substep for a pool without walls or reactions
    // (Engine::enqueue_substep), the force pass walking the recorded
    // neighbours from its per-lane queue of row chunks.
    const uint64_t substeps = (uint64_t)ticks * std::max<uint32_t>(cfg.substeps, 1u);
    for (uint64_t k = 0; k < substeps; ++k) {
        hipLaunchKernelGGL(k_grid_box, dim3(1), dim3(1024), 0, stream, P, 0u, 0u);
        hipLaunchKernelGGL(k_keys, dim3((nb + kSortBlocks - 1) / kSortBlocks), dim3(kBlock), 0, stream, P);
        hipLaunchKernelGGL(k_scan_reduce<8>, dim3(grid_scan), dim3(kBlock), 0, stream, P, tile_sums);
        hipLaunchKernelGGL(k_scan_tiles<8>, dim3(1), dim3(1024), 0, stream, P, tile_sums);
        hipLaunchKernelGGL(k_scan_apply<8>, dim3(grid_scan), dim3(kBlock), 0, stream, P, (const uint32_t *)tile_sums);
        hipLaunchKernelGGL(k_place, dim3((nb + kSortBlocks - 1) / kSortBlocks), dim3(kBlock), 0, stream, P);
        hipLaunchKernelGGL(k_canon_gather, dim3(nb), dim3(kBlock), 0, stream, P);
        hipLaunchKernelGGL((k_density<256, 256>), dim3(nb), dim3(kBlock), 0, stream, P);
        hipLaunchKernelGGL((k_force_t<4, 256>), dim3(nb), dim3(kBlock), 0, stream, P);
        hipLaunchKernelGGL(k_box_reduce, dim3(grid_reduce), dim3(kBlock), 0, stream, P);
    }
    hip_ok(hipGetLastError(), "device step launch");
    Control ctl{};
    hip_ok(hipMemcpyAsync(&ctl, P.ctl, sizeof(Control), hipMemcpyDeviceToHost, stream), "control read");
    hip_ok(hipStreamSynchronize(stream), "device step");
    if (ctl.err & kErrMass) throw std::invalid_argument("particle masses must be positive and finite");
    if (ctl.err & kErrGrid) throw std::domain_error("the particle box is not finite (NaN or infinite positions)");
    if (ctl.err & kErrCellCapacity) throw CapacityError("the scene's grid has more cells than the workspace holds");
    hipLaunchKernelGGL(k_unsort, dim3(std::min<uint32_t>(nb, 8192u)), dim3(kBlock), 0, stream, P, 0u, pos, vel);
    hip_ok(hipGetLastError(), "k_unsort");
    hip_ok(hipStreamSynchronize(stream), "device step");
}

// ---- the allocation-free variant: the workspace is a file-scope device array ----

namespace {

// One scene of up to kNoAllocParticles particles and kNoAllocCells cells
// (device_step_workspace(kNoAllocParticles, kNoAllocCells) bytes, checked
// when the address is taken). Each device holds its own copy for the life of
// the process; calls on one device share it.
constexpr size_t kScratchBytes = 26ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch[kMaxDevices] = {};
std::mutex g_device_mu[kMaxDevices];

// Device address of this device's scratch: looked up once per device, with
// that device current. In host code `scratch` names a host-side stand-in, so
// only this address may reach a kernel or a copy.
char *scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        if (device_step_workspace(kNoAllocParticles, kNoAllocCells, nullptr, nullptr, nullptr, nullptr, nullptr) >
            kScratchBytes)
            throw std::logic_error("allocation-free scratch smaller than its stated maximum scene");
        void *p = nullptr;
        hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw DeviceError(hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch): null address");
        g_scratch[dev] = (char *)p;
    }
    return g_scratch[dev];
}

}  // namespace

void device_step_noalloc(const SceneConfig &cfg, uint32_t n, uint32_t nf, float *pos, float *vel, const float *pmass,
                         uint32_t ticks, hipStream_t stream) {
    if (n > kNoAllocParticles)
        throw CapacityError("the allocation-free step holds at most " + std::to_string(kNoAllocParticles) +
                            " particles (" + std::to_string(n) + " given)");
    int dev = 0;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    char *ws = scratch_of(dev);
    if (ticks == 0) return;
    // Cells beyond the scratch's table end the call with CapacityError (the
    // substep is abandoned on the device, the arrays are left untouched).
    const uint64_t cells = std::min<uint64_t>(cfg.max_bins ? cfg.max_bins : (1ull << 21), (uint64_t)kMaxCells - 1u);
    const uint32_t ccap = (uint32_t)std::min<uint64_t>(cells + 1u, kNoAllocCells);
    std::lock_guard<std::mutex> lk(g_device_mu[dev]);
    device_step(cfg, n, nf, pos, vel, pmass, ticks, stream, ws, kScratchBytes, ccap);
}

void check_device_range(const void *p, size_t bytes, const char *what) {
    void *base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        throw std::invalid_argument(std::string(what) + " is not device memory of this process");
    }
    const char *b = (const char *)base, *q = (const char *)p;
    if (q < b || (size_t)(q - b) + bytes > size)
        throw std::length_error(std::string(what) + " is shorter than the particle count needs");
}

}  // namespace vxl
