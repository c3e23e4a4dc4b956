// Plain-old-data layout shared by the host runtime and the gfx950 kernels.
//
// Pool layout (device, SoA f32/u32):
//   * every scene owns the slot range [pbase, pbase + alloc) of the particle
//     pool, alloc = n rounded up to kBlock, so a block of kBlock slots always
//     belongs to exactly one scene (block_scene[] maps block -> scene);
//   * "cur" arrays hold the state in the grid order of the previous substep;
//     "srt" arrays hold the state gathered into the current grid order;
//   * gid is the scene-local original particle index (fluid prefix first),
//     the key the reference orders every cell by (fluid grid.rs:53-75);
//   * every scene owns the cell range [cbase, cbase + total + 1) of the cell
//     tables: total grid cells plus one pad cell that absorbs the unused
//     slots of the scene, so one global exclusive scan yields per-scene starts
//     that begin exactly at pbase.
#ifndef VXL_SPH_LAYOUT_H
#define VXL_SPH_LAYOUT_H
#include <stdint.h>

namespace vxl {

constexpr uint32_t kBlock = 256;        // slots per block (4 wave64)
constexpr uint32_t kPartSlots = 64;     // slots per partial AABB (one wave)
constexpr uint32_t kGridDimMax = 1u << 14;  // core/grid.rs:8
// Slots of one pool. The neighbour loops address per-slot arrays through
// 32-bit byte offsets from a scalar base (dev::at); the widest array read
// that way is the 32-byte density record, so a pool holds at most 2^27 slots
// (the record's last byte offset stays below 2^32). Arrays wider per slot
// (the 36-byte neighbour masks) are indexed in 64 bits.
constexpr uint32_t kMaxPoolSlots = 1u << 27;
// Cells of the cell tables: `start` is read through 32-bit byte offsets too.
constexpr uint32_t kMaxCells = (1u << 30) - 16u;
// Bytes per slot of the widest array addressed through 32-bit offsets, and
// per cell of the cell tables.
constexpr uint32_t kAtRecordBytes = 32;
constexpr uint32_t kAtCellBytes = 4;
static_assert((uint64_t)kMaxPoolSlots * kAtRecordBytes <= (1ull << 32), "density record offsets must fit 32 bits");
static_assert(((uint64_t)kMaxPoolSlots + 4u) * 16u <= (1ull << 32), "A record offsets must fit 32 bits");
static_assert(((uint64_t)kMaxCells + 1u) * kAtCellBytes <= (1ull << 32), "cell start offsets must fit 32 bits");
// Extra records after the last slot of the grid-ordered arrays: neighbour
// candidates are loaded in groups of 4 without clamping to the range end.
constexpr uint32_t kReadPad = 4;

// Host-computed constants of one scene (FluidSystem::new, fluid_access.rs:8-13,
// plus the substep constants of fluid_step.rs:110-113). Values are computed
// once on the host in strict f32 and passed bit-for-bit.
struct SceneParams {
    uint32_t pbase;      // first pool slot
    uint32_t n;          // particles (fluid + boundary)
    uint32_t nf;         // fluid prefix length
    uint32_t alloc;      // slots owned (multiple of kBlock)
    uint64_t max_bins;   // GRID_MAX_BINS budget of the grid rule
    float h, h2, k6, ks, w0;
    float mass;          // FluidSystem::mass (fluid density sum scale)
    float rho0, b_tait, alpha_c, eps;
    float gx, gy, gz;
    float dt, vmax, vmax2;
    uint32_t clamp_neg;  // tensile_instability_suppression
    uint32_t react;      // boundary reactions: 0 none, kReactAll every span, kReactDynamic dynamic bodies' spans
    uint32_t rbody0;     // first dynamic rigid body of the scene (pool rigid table)
    uint32_t nrbody;     // dynamic rigid bodies of the scene (0: none)
    uint32_t fmass_uniform;  // every fluid particle's mass is `mass`, bit for bit
    uint32_t substeps;   // substeps per tick of this scene (dt = dt_tick / substeps)
    uint32_t pad[3];
};

// A dynamic Box body coupled to its scene's fluid (the rigid leg of
// crates/vxl-phys/src/world_step.rs:13-123): advanced once per tick before
// the fluid substeps by gravity and the fluid's tick-averaged reaction, then
// its Akinci particles are rebuilt from the new pose.
struct RigidDev {
    float pos[3];
    float rot[4];            // quaternion x, y, z, w
    float linvel[3];
    float angvel[3];
    float inv_mass;
    float inv_inertia[3];    // body-frame diagonal (mass.rs:35-48)
    float half[3];
    float react[6];          // (F, tau) of the last fluid tick (zero before the first)
    // rigid configuration of the scene (PhysConfig, config.rs:206-233)
    float dt_tick;
    float gravity[3];
    float max_linear, max_angular;
    float tank[3];           // tank contact planes; tank[0] <= 0: none
    uint32_t substeps;
    uint32_t scene;          // pool scene index
    uint32_t span;           // pool span index (reaction in, torque origin out)
    uint32_t gstart, gend;   // the body's particles, scene index order
    uint32_t lat;            // first body-frame lattice point in the pool lattice table
    uint32_t pad[2];
};

// A boundary span of a scene with reactions (fluid_boundary.rs:10-31):
// pool slots [start, end) = pbase + local [start, end).
struct SpanDev {
    uint32_t scene;
    uint32_t start, end;
    uint32_t substeps;   // tick = substeps substeps; reactions are tick averages
    float origin[3];
    uint32_t active;     // 0: the span's reaction is not computed (left at zero)
};

constexpr uint32_t kReactAll = 1;      // every span (FluidStepper semantics)
constexpr uint32_t kReactDynamic = 2;  // only spans of dynamic rigid bodies (coupled scenes)

constexpr uint32_t kReactBatch = 16;  // fluid neighbours a boundary particle's reaction pass holds at once

// Provider walls of a scene (crates/vxl-phys-core/src/interop.rs:166-224).
//
// kWallsHalfspace: providers made of axis-aligned half-space faces, queried
// on the device every substep with the CPU engine's semantics: mirror ghosts
// summed with each neighbour (fluid_density.rs:17-51, :76-121), projection
// of the integrated position and coincident separation
// (fluid_boundary.rs:45-131).
// kWallsTable: sparse plane tables gathered by the host from any provider
// once per tick and held until replaced, with the GPU pipeline's semantics
// (wall_ghost.wgsl:61-178): ghosts summed apart and added after the density
// pass, projection against the pushed position, no separation.
constexpr uint32_t kWallsOff = 0;
constexpr uint32_t kWallsHalfspace = 1;
constexpr uint32_t kWallsTable = 2;
constexpr uint32_t kWallMaxProviders = 8;
constexpr uint32_t kWallMaxFaces = 16;
constexpr uint32_t kWallMirrorPlanes = 8;  // planes kept per particle (fluid_density.rs:108)
constexpr uint32_t kWallNone = 0xffffffffu;
// Coincident separation of a half-space scene: rounds of the parallel
// fixed-point iteration before the serial fallback.
constexpr uint32_t kSepIters = 6;

// Per-scene state of the coincident separation of one substep.
struct SepState {
    uint32_t k;                  // particles the projection pushed
    uint32_t changed[kSepIters]; // round m moved some particle (0: converged)
    uint32_t pad;
};

struct WallFaceDev {
    uint32_t axis;   // 0, 1, 2
    float plane;
    float sign;      // outward normal sign * e_axis (the fluid side)
    uint32_t pad;
};

struct WallProvDev {
    float lo[3], hi[3];  // bounds() for the AABB prefilter
    uint32_t has_bounds;
    uint32_t face0, nface;  // faces of this provider, in contact order
    uint32_t pad[3];
};

struct WallSceneDev {
    uint32_t mode;       // kWallsOff | kWallsHalfspace | kWallsTable
    uint32_t nprov;
    uint32_t ne[2];      // table entries (0 mirror, 1 project)
    WallProvDev prov[kWallMaxProviders];
    WallFaceDev face[kWallMaxFaces];
    // Plane tables of the scene (kWallsTable): start[ne + 1] plane ranges and
    // planes as 8 floats each (point xyz, 0, outward normal xyz, 0).
    const uint32_t *tstart[2];
    const float *tplane[2];
};

// Device-written per-substep grid of one scene (grid_box rule, core/grid.rs:34-51).
struct SceneGrid {
    float gmin[3];
    float inv;
    float bin;
    uint32_t dims[3];
    uint32_t total;      // dims product (<= max_bins)
    uint32_t cbase;      // first cell of this scene in the global tables
    uint32_t pad[2];
};

// Ordered-u32 AABB accumulator (min over lo, max over hi).
struct SceneBox {
    uint32_t lo[3];
    uint32_t hi[3];
};

// Device-side control block.
struct Control {
    uint32_t step;       // substep counter: AABB ping-pong parity
    uint32_t ctot;       // cells used this substep (sum of total + 1)
    uint32_t err;        // nonzero: the substep was abandoned (see kErr*)
    uint32_t tick_sub;   // launch index inside the pool tick (set by k_grid_box)
    uint32_t tick_len;   // launches per pool tick: the most substeps of any scene (0: every launch steps every scene)
    uint32_t nreact;     // boundary slots listed for this substep's reaction pass (k_bforce_list)
    uint32_t pad[2];
};

// Scenes of one pool may differ in substeps per tick. A pool tick runs
// tick_len launches and a scene with s substeps steps in the last s of them;
// in the others it is carried over unchanged (sorted and its densities taken
// like the others, not integrated, no reactions). Ending every scene's tick
// on the pool tick's last launch leaves each scene's fields, reactions and
// rigid coupling as they are after its own last substep.
__host__ __device__ inline bool substep_active(uint32_t tick_sub, uint32_t tick_len, uint32_t substeps) {
    return tick_sub + substeps >= tick_len;
}

constexpr uint32_t kErrCellCapacity = 1u;
constexpr uint32_t kErrGrid = 2u;  // non-finite particle box
constexpr uint32_t kErrMass = 4u;  // a particle mass not positive and finite (device-array step)

// Order-preserving map f32 -> u32 (total order; -0 < +0).
__host__ __device__ inline uint32_t f2o(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __builtin_bit_cast(float, u);
}

}  // namespace vxl
#endif
