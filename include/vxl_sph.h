/*
 * vxl_sph.h — C ABI of the MI355X (gfx950) WCSPH fluid step.
 *
 * Drop-in boundary for the reference engine's GPU fluid path:
 *   - PhysGpuBackend  (crates/vxl-phys-gpu/src/lib.rs:20-29)      -> vxl_sph_backend_*
 *   - FluidStepper    (crates/vxl-phys-core/src/interop.rs:293-316) -> vxl_sph_stepper_*
 *     semantics of GpuFluidStepper (crates/vxl-phys-gpu/src/pipeline/stepper.rs:70-147)
 *   - FluidSystem step/accessors (crates/vxl-phys-fluid/src/fluid_step.rs:13-147,
 *     fluid_access.rs:78-186) over a pool of scenes               -> vxl_sph_engine_*, vxl_sph_scene_*
 *   - boundary::lattice (crates/vxl-phys-fluid/src/boundary.rs:67-88) -> vxl_sph_box_lattice
 * Conventions follow crates/vxl-phys-ffi/src/lib.rs:1-41: a versioned ABI,
 * every configuration block starts with {size, version}, batch calls only,
 * errors are status codes (exceptions never cross the boundary).
 *
 * Vec3 arrays are 3 floats per particle (the repr(C) layout of &[Vec3]).
 * Handles are not thread-safe: one host thread per handle.
 */
#ifndef VXL_SPH_H
#define VXL_SPH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ffi/src/lib.rs:13 — bump on any incompatible change. */
#define VXL_SPH_ABI_VERSION 1u

/* ffi/src/lib.rs:24-34 (codes 0-5 keep the reference's meaning). */
typedef enum VxlSphStatus {
    VXL_SPH_OK = 0,
    VXL_SPH_NULL_POINTER = 1,
    VXL_SPH_VERSION_MISMATCH = 2,
    VXL_SPH_SIZE_MISMATCH = 3,
    VXL_SPH_NOT_READY = 4,
    VXL_SPH_PANIC_CAUGHT = 5,
    VXL_SPH_INVALID_ARGUMENT = 16,
    VXL_SPH_UNSUPPORTED = 17,
    VXL_SPH_DEVICE_ERROR = 18,
    VXL_SPH_OUT_OF_MEMORY = 19,
    VXL_SPH_CAPACITY = 20,
    VXL_SPH_NOT_FOUND = 21
} VxlSphStatus;

/* ffi/src/lib.rs:17-22 */
typedef struct VxlSphHeader {
    uint32_t size;    /* sizeof the whole block */
    uint32_t version; /* VXL_SPH_ABI_VERSION */
} VxlSphHeader;

/* FluidConfig (fluid/src/config.rs:6-69) + the values a restated scene needs. */
typedef struct VxlSphFluidConfig {
    VxlSphHeader hdr;
    float rest_density;         /* rho0 */
    float smoothing_radius;     /* h */
    float xsph_viscosity;       /* eps */
    float sound_speed;          /* c */
    float gamma_tait;           /* only 7 is supported (fluid_force.rs:17-21) */
    uint32_t substeps;          /* substeps per tick */
    uint32_t tensile_clamp;     /* tensile_instability_suppression: p < 0 -> 0 */
    uint32_t boundary_layers;   /* Akinci layers of vxl_sph_box_lattice users */
    float gravity[3];
    float max_speed_frac;       /* CFL clamp |v| <= frac * h / dt */
    float artificial_viscosity; /* Monaghan alpha */
    float particle_mass;        /* FluidSystem::mass; 0 = lattice calibration (fluid_access.rs:27-42) */
    float spacing;              /* lattice spacing (mass calibration, Akinci sampling) */
    uint64_t max_bins;          /* grid budget GRID_MAX_BINS (core/grid.rs:20); 0 = 1 << 21 */
    float dt_tick;              /* tick length used by vxl_sph_engine_step */
    uint32_t reserved;
} VxlSphFluidConfig;

/* Boundary span (body, origin, [start, end)) — FluidSystem::boundary_spans (fluid_boundary.rs:129-135). */
typedef struct VxlSphSpan {
    uint32_t body;
    float origin[3];
    uint32_t start;
    uint32_t end;
} VxlSphSpan;

/* (body, force, torque about the body origin) — FluidSystem::boundary_reactions (fluid_access.rs:113-118). */
typedef struct VxlSphReaction {
    uint32_t body;
    float force[3];
    float torque[3];
} VxlSphReaction;

/* Grid of the last rebuild — FluidSystem::neighbor_grid (fluid_access.rs:128-140). */
typedef struct VxlSphGridInfo {
    float min[3];
    float inv;
    float bin;
    uint32_t dims[3];
    uint32_t total;
} VxlSphGridInfo;

typedef struct VxlSphEngine VxlSphEngine;
typedef struct VxlSphStepper VxlSphStepper;

/* ---- library / backend (PhysGpuBackend, gpu/src/lib.rs:20-29) ---- */
uint32_t vxl_sph_abi_version(void);
const char *vxl_sph_backend_name(void);                 /* "hip-gfx950" */
int32_t vxl_sph_backend_available(void);                /* 1 when a gfx950 device is visible */
int32_t vxl_sph_backend_submit(const char *tag);        /* 1 for "fluid" when available */
VxlSphStatus vxl_sph_device_count(int32_t *out);
const char *vxl_sph_last_error(void);                   /* thread-local message of the last failure */
VxlSphStatus vxl_sph_default_config(VxlSphFluidConfig *out); /* FluidConfig::default (config.rs:51-69) */

/* ---- engine: a pool of independent scenes on one device ---- */
VxlSphStatus vxl_sph_engine_create(uint32_t abi_version, int32_t device, VxlSphEngine **out);
VxlSphStatus vxl_sph_engine_destroy(VxlSphEngine *e);
/* Enable per-particle diagnostics (neighbour counts, acceleration, XSPH). */
VxlSphStatus vxl_sph_engine_set_diagnostics(VxlSphEngine *e, int32_t on);
/* Launch geometry of the substep kernels (chosen by the autotuner). No
 * counterpart in the reference (its workgroup size is fixed at 64,
 * crates/vxl-phys-gpu/src/density.wgsl:52). 0 keeps a field's default. */
typedef struct VxlSphTuning {
    VxlSphHeader hdr;
    uint32_t persist_blocks;  /* grid size cap of the per-slot kernels (0 = one block per 256 slots) */
    uint32_t force_kernel;    /* VXL_SPH_FORCE_* (0 = VXL_SPH_FORCE_STAGED) */
    uint32_t block_threads;   /* workgroup size of the density and force kernels: 64, 128, 256 (0 = 256) */
    uint32_t stage_records;   /* LDS staging depth of the density pass, records per wave and row: 128, 256, 512 (0 = 256) */
    uint32_t scan_tile_cells; /* cells per tile of the cell-count scan: 1024, 2048, 4096 (0 = 2048) */
    uint32_t reserved;
} VxlSphTuning;
/* Force pass: every particle tests its neighbour candidates again (SCAN), or
 * walks the accepted neighbours the density pass recorded (the others). */
#define VXL_SPH_FORCE_DEFAULT 0u
#define VXL_SPH_FORCE_SCAN 1u
#define VXL_SPH_FORCE_WALK 2u
#define VXL_SPH_FORCE_WALK_AHEAD 3u  /* the walk with each pair's loads issued one pair ahead */
#define VXL_SPH_FORCE_PAIRS 4u       /* row by row, the wave's recorded pairs spread evenly over its lanes */
#define VXL_SPH_FORCE_QUEUE 5u       /* the walk from a per-lane queue of row chunks opened before the pair loop */
#define VXL_SPH_FORCE_PAIRED 6u      /* the queue walk with each neighbour record loaded by a lane pair */
#define VXL_SPH_FORCE_STAGED 7u      /* the queue walk with each wave's own cell row staged in LDS */
VxlSphStatus vxl_sph_engine_set_tuning(VxlSphEngine *e, const VxlSphTuning *t);
VxlSphStatus vxl_sph_engine_get_tuning(VxlSphEngine *e, VxlSphTuning *out);
/* Autotuner: coordinate descent over every field above, each candidate timed
 * for `substeps` real substeps on the resident pool (results are identical
 * for all of them), the fastest kept; cached per (arch, log2 particle count)
 * for the process. `out` may be NULL. */
VxlSphStatus vxl_sph_engine_autotune(VxlSphEngine *e, uint32_t substeps, VxlSphTuning *out);
/* Replay each substep as one hipGraph (default on). */
VxlSphStatus vxl_sph_engine_set_graphs(VxlSphEngine *e, int32_t on);
/* Advance every scene by `ticks` ticks (enqueue only; see vxl_sph_engine_sync).
 * Scenes may differ in substeps per tick: a tick is as many launches as the
 * scene with the most substeps takes, and a scene with s substeps steps in
 * the last s of them (results as if it were stepped alone). */
VxlSphStatus vxl_sph_engine_step(VxlSphEngine *e, uint32_t ticks);
VxlSphStatus vxl_sph_engine_sync(VxlSphEngine *e);
VxlSphStatus vxl_sph_engine_device_bytes(VxlSphEngine *e, uint64_t *out);
/* Tick graphs instantiated and updated in place so far. A layout (scenes
 * added or removed) captures the tick again; when its launch sequence is
 * unchanged the executable graph is updated (hipGraphExecUpdate) rather than
 * instantiated. No counterpart in the reference. */
VxlSphStatus vxl_sph_engine_graph_stats(VxlSphEngine *e, uint64_t *built, uint64_t *updated);

/* Stage timings. No C-ABI counterpart in the reference; they replace its
 * per-tick stage timer TickMs (crates/vxl-phys-gpu/src/pipeline/gpu_types.rs:48-58)
 * with HIP events on the engine's stream. */
typedef struct VxlSphKernelTime {
    char name[32];
    float ms_total;   /* summed over the profiled substeps */
} VxlSphKernelTime;
/* Run `substeps` substeps eagerly with a HIP event between consecutive kernels. */
VxlSphStatus vxl_sph_engine_profile(VxlSphEngine *e, uint32_t substeps, uint32_t cap, VxlSphKernelTime *out,
                                    uint32_t *count);
/* Run `ticks` ticks and return the device time of each (ms_out[ticks]). */
VxlSphStatus vxl_sph_engine_step_timed(VxlSphEngine *e, uint32_t ticks, float *ms_out);

/* Add a scene: all particles (fluid prefix, then boundary) in index order,
 * FluidSystem::raw_particles layout (fluid_access.rs:147-155). vel / pmass
 * may be NULL (zero velocity / cfg mass). */
VxlSphStatus vxl_sph_scene_add(VxlSphEngine *e, const VxlSphFluidConfig *cfg, uint32_t n, uint32_t n_fluid,
                               const float *pos, const float *vel, const float *pmass, const VxlSphSpan *spans,
                               uint32_t nspans, uint32_t *out_scene);
VxlSphStatus vxl_sph_scene_remove(VxlSphEngine *e, uint32_t scene);
/* Replace the boundary segment's positions/velocities (index order, nb = n - n_fluid);
 * GpuFluidStepper's per-tick boundary upload (pipeline/coupling.rs:20-50). */
VxlSphStatus vxl_sph_scene_upload_boundary(VxlSphEngine *e, uint32_t scene, uint32_t nb, const float *pos,
                                           const float *vel);
/* Positions / velocities of all n particles, index order (either may be NULL). */
VxlSphStatus vxl_sph_scene_read_state(VxlSphEngine *e, uint32_t scene, uint32_t n, float *pos, float *vel);
/* dens/press: n; ncount: n (diagnostics); acc/xsph: 3*n_fluid (diagnostics). Any may be NULL. */
VxlSphStatus vxl_sph_scene_read_fields(VxlSphEngine *e, uint32_t scene, uint32_t n, float *dens, float *press,
                                       uint32_t *ncount, float *acc, float *xsph);
/* start: total+1 entries, items: n entries (scene-local indices), keys: n entries (cell per particle). */
VxlSphStatus vxl_sph_scene_read_grid(VxlSphEngine *e, uint32_t scene, VxlSphGridInfo *info, uint32_t start_cap,
                                     uint32_t *start, uint32_t n, uint32_t *items, uint32_t *keys);
/* AABB of all particles at the next grid rebuild (FluidStepper::bounds). */
VxlSphStatus vxl_sph_scene_read_bounds(VxlSphEngine *e, uint32_t scene, float lo[3], float hi[3]);
/* FNV-1a-64 over fluid pos then vel f32 bits per particle, index order
 * (crates/vxl-phys-fluid/examples/sph_scale.rs:52-63). */
VxlSphStatus vxl_sph_scene_state_hash(VxlSphEngine *e, uint32_t scene, uint64_t *out);

/* FNV-1a-64 byte-wise over every particle's position (index order, x y z, f32
 * little-endian bytes) then every velocity: the A/B state hash of
 * crates/vxl-phys-gpu/tests/packet_grid_canary.rs:102-111 (the other
 * convention, vxl_sph_scene_state_hash, follows sph_scale.rs). */
VxlSphStatus vxl_sph_scene_state_hash_all(VxlSphEngine *e, uint32_t scene, uint64_t *out);

/* ---- provider walls (crates/vxl-phys-core/src/interop.rs:166-224) ---- */

/* A half-space wall provider: the AABB of ProviderColliders::bounds
 * (has_bounds = 0: None, no prefilter) and its faces. Face (axis, plane,
 * sign) is the half-space whose fluid side is sign * (x[axis] - plane) > 0;
 * its contact for a query point p is p with p[axis] = plane and the outward
 * normal sign * e_axis, reported while sign * (p[axis] - plane) <= 2 * probe
 * (the reference's test doubles FloorY and Tank, fluid/src/tests.rs:11-110,
 * are such providers). */
typedef struct VxlSphWallProvider {
    uint32_t id;
    uint32_t has_bounds;
    float lo[3];
    float hi[3];
} VxlSphWallProvider;
typedef struct VxlSphWallFace {
    uint32_t axis;      /* 0, 1, 2 */
    float plane;
    float sign;         /* +1 or -1 */
    uint32_t provider;  /* index into the provider array; faces listed provider by provider */
} VxlSphWallFace;

/* Half-space providers queried on the device every substep with the CPU
 * engine's semantics: FluidSystem::set_boundaries (fluid_access.rs:184-186)
 * + FluidSystem::step(dt, providers) (fluid_step.rs:75-127): mirror ghosts in
 * the density pass (fluid_density.rs:17-51, :76-121), projection and
 * coincident separation after the integration (fluid_boundary.rs:45-131).
 * At most 8 providers and 16 faces; nprov = 0 removes the walls. Takes effect
 * at the next step. */
VxlSphStatus vxl_sph_scene_set_halfspace_walls(VxlSphEngine *e, uint32_t scene, uint32_t nprov,
                                               const VxlSphWallProvider *provs, uint32_t nface,
                                               const VxlSphWallFace *faces);
/* Sparse plane table of side 0 (mirror) or 1 (project), gathered by the host
 * from any provider (gather_wall_contacts / gather_wall_project_contacts,
 * pipeline/walls.rs:145-216) and held until replaced, with the GPU pipeline's
 * semantics (WallStage::upload, walls.rs:104-116; wall_ghost.wgsl:61-178).
 * ids: ne strictly ascending fluid indices; start: ne + 1 plane offsets from
 * 0 to np; planes: np x (point xyz, outward normal xyz). */
VxlSphStatus vxl_sph_scene_set_wall_tables(VxlSphEngine *e, uint32_t scene, uint32_t side, uint32_t ne,
                                           const uint32_t *ids, const uint32_t *start, uint32_t np,
                                           const float *planes);
/* Remove a scene's walls (either kind). */
VxlSphStatus vxl_sph_scene_clear_walls(VxlSphEngine *e, uint32_t scene);

/* ---- one scene in caller-owned device arrays ---- */
/* FluidSystem::step (fluid_step.rs:75-127) for `ticks` ticks of a scene held
 * in device memory of this process: pos/vel (3 floats per particle) and pmass
 * (1 per particle) of all n particles, fluid prefix first, index order. pos
 * and vel are read at the start and overwritten with the new state at the
 * end, only when every substep completed. No reactions, no walls. `stream`
 * is a hipStream_t (NULL: the null stream); the call returns after the
 * stream has finished its work. A device workspace for the scene's pool is
 * allocated and freed by the call. No counterpart in the reference (its
 * FluidSystem owns host vectors). */
VxlSphStatus vxl_sph_step_device(const VxlSphFluidConfig *cfg, uint32_t n, uint32_t n_fluid, float *pos, float *vel,
                                 const float *pmass, uint32_t ticks, void *stream);
/* vxl_sph_step_device for callers that cannot allocate (code between HIP
 * calls that must not allocate, tight serving loops, harnesses that forbid
 * hipMalloc): the same kernels, results equal bit for bit, the workspace taken
 * from a device array of the library (about 26 MiB, one per device, held until
 * the process ends) instead of an allocation. Nothing is allocated. Limits:
 * n <= 131072 particles and a grid of at most 524287 cells; beyond them the
 * call returns VXL_SPH_CAPACITY and leaves the arrays untouched. Calls on one
 * device share the array and are serialised; each call sets every word it
 * reads before reading it. No counterpart in the reference: it exists for the
 * callers that cannot allocate. */
VxlSphStatus vxl_sph_step_device_noalloc(const VxlSphFluidConfig *cfg, uint32_t n, uint32_t n_fluid, float *pos,
                                         float *vel, const float *pmass, uint32_t ticks, void *stream);

/* ---- FluidStepper drop-in (single scene; GpuFluidStepper semantics) ---- */
VxlSphStatus vxl_sph_stepper_create(uint32_t abi_version, const VxlSphFluidConfig *cfg, int32_t device,
                                    VxlSphStepper **out);
VxlSphStatus vxl_sph_stepper_destroy(VxlSphStepper *s);
/* FluidStepper::step: pos/vel/pmass are all particles (fluid prefix + boundary
 * segment). The first call uploads the whole scene; later calls upload only the
 * boundary segment (fluid state is device-resident) and length-check pmass. */
VxlSphStatus vxl_sph_stepper_step(VxlSphStepper *s, float dt_tick, uint32_t n_fluid, uint32_t n, const float *pos,
                                  const float *vel, const float *pmass, const VxlSphSpan *spans, uint32_t nspans);
VxlSphStatus vxl_sph_stepper_reactions(VxlSphStepper *s, uint32_t cap, VxlSphReaction *out, uint32_t *count);
VxlSphStatus vxl_sph_stepper_bounds(VxlSphStepper *s, float lo[3], float hi[3]);
VxlSphStatus vxl_sph_stepper_read_state(VxlSphStepper *s, uint32_t n, float *pos, float *vel);
/* FluidStepper::gather_walls (interop.rs:313-315; GpuFluidStepper, pipeline/stepper.rs:118-137):
 * the host gathers both tables from its providers at the stepper's fluid
 * positions (vxl_sph_stepper_read_state; before the first step, the positions
 * it will pass) and hands them over here; they are used from the next step
 * on. Same arguments as vxl_sph_scene_set_wall_tables. */
VxlSphStatus vxl_sph_stepper_set_wall_tables(VxlSphStepper *s, uint32_t side, uint32_t ne, const uint32_t *ids,
                                             const uint32_t *start, uint32_t np, const float *planes);
/* Half-space providers for the stepper's scene (vxl_sph_scene_set_halfspace_walls). */
VxlSphStatus vxl_sph_stepper_set_halfspace_walls(VxlSphStepper *s, uint32_t nprov, const VxlSphWallProvider *provs,
                                                 uint32_t nface, const VxlSphWallFace *faces);

/* ---- scene construction ---- */
/* boundary::lattice for Shape::Box (boundary.rs:67-88, :208-232): writes up to cap
 * points (3 floats each), returns the count, points per layer and V_b. */
VxlSphStatus vxl_sph_box_lattice(const float half[3], float s, float h, uint32_t layers, uint32_t cap,
                                 float *out_pts, uint32_t *out_n, uint32_t *out_n0, float *out_volume);
/* A Box body and its pose (boundary.rs BodyPose, fluid/src/types.rs:6-15). */
typedef struct VxlSphBoxBody {
    uint32_t id;
    float half[3];
    float pos[3];
    float rot[4];     /* quaternion x, y, z, w */
    float linvel[3];
    float angvel[3];
} VxlSphBoxBody;
/* set_boundary_particles for Box bodies (fluid_step.rs:13-53): Akinci particles
 * of cfg->boundary_layers layers at cfg->spacing, world pose p + q*lp, velocity
 * v + w x r, pmass = rho0 * V_b, body order x layer x face order, one span per
 * body with indices starting at `first`. Call with pos == NULL to query n / nspans. */
VxlSphStatus vxl_sph_box_particles(const VxlSphFluidConfig *cfg, uint32_t nbodies, const VxlSphBoxBody *bodies,
                                   uint32_t first, uint32_t *n, uint32_t *nspans, float *pos, float *vel,
                                   float *pmass, VxlSphSpan *spans);
/* FluidSystem::new constants (fluid_access.rs:8-42): h, h2, k6, ks, w0, b_tait, mass, alpha_c. */
VxlSphStatus vxl_sph_fluid_constants(const VxlSphFluidConfig *cfg, float out[8]);

/* ---- coupled scenes: fluid + Box bodies, dynamic bodies advanced on the device ---- */
/* The rigid leg of World::step for dynamic Box bodies
 * (crates/vxl-phys/src/world_step.rs:13-123, crates/vxl-phys-integrate/src/lib.rs:18-64):
 * per tick, `substeps` rigid substeps driven by gravity and the previous fluid
 * tick's averaged reaction, then the bodies' Akinci particles are rebuilt from
 * the new pose and the fluid advances one tick. Defaults are PhysConfig's
 * (crates/vxl-phys-core/src/config.rs:206-233). `tank` holds the interior
 * planes of an open tank [0,L0] x [0,inf) x [0,L2] that dynamic boxes cannot
 * cross: a contact rule of this library (the reference has no contact between
 * a box and a particle-only tank); L0 <= 0 disables it. */
typedef struct VxlSphRigidConfig {
    VxlSphHeader hdr;
    uint32_t substeps;
    float gravity[3];
    float max_linear;
    float max_angular;
    float tank[3];
    uint32_t reserved;
} VxlSphRigidConfig;
VxlSphStatus vxl_sph_default_rigid_config(VxlSphRigidConfig *out);
/* A Box body of a coupled scene: density > 0 makes it dynamic (mass properties
 * of crates/vxl-phys-core/src/mass.rs:35-48), otherwise it is static. */
typedef struct VxlSphRigidBox {
    VxlSphBoxBody box;
    float density;
} VxlSphRigidBox;
/* Fluid particles (n_fluid x 3 floats; fluid_vel may be NULL) followed by the
 * bodies' Akinci particles in body order (set_boundary_particles,
 * fluid_step.rs:13-53). Body ids must be unique; reactions are computed. */
VxlSphStatus vxl_sph_scene_add_coupled(VxlSphEngine *e, const VxlSphFluidConfig *cfg, const VxlSphRigidConfig *rc,
                                       uint32_t n_fluid, const float *fluid_pos, const float *fluid_vel,
                                       uint32_t nbodies, const VxlSphRigidBox *bodies, uint32_t *out_scene);
/* Pose and velocities of a coupled scene's dynamic bodies, in creation order
 * (call with out == NULL to query the count). */
VxlSphStatus vxl_sph_scene_read_bodies(VxlSphEngine *e, uint32_t scene, uint32_t cap, VxlSphRigidBox *out,
                                       uint32_t *count);

/* Seeded dam-break scene (tank [0,L]^3 open at the top, fluid block in the
 * corner, 5 Akinci Box bodies of thickness 2d). Call with pos == NULL to
 * query n / n_fluid / nspans, then again with buffers of those sizes. */
typedef struct VxlSphDamBreak {
    VxlSphHeader hdr;
    float tank[3];
    uint32_t block[3];
    float d;
    float jitter;      /* jitter amplitude as a fraction of d */
    uint64_t seed;
    uint32_t walls;    /* 0 = no tank particles */
    uint32_t reserved;
} VxlSphDamBreak;
VxlSphStatus vxl_sph_dam_break(const VxlSphDamBreak *spec, const VxlSphFluidConfig *cfg, uint32_t *n,
                               uint32_t *n_fluid, uint32_t *nspans, float *pos, float *vel, float *pmass,
                               VxlSphSpan *spans);

/* ---- frame dumps (VXLD v3) ---- */
/* The binary frame format of the reference demos and renderer
 * (crates/vxl-phys/examples/dam_break.rs:14-44, :84-128; body records as
 * showcase.rs:311-360; read by scripts/render_demo.py:39-100): header, then per
 * frame the tick, a time in ms, Box body records (kind, awake, pose, half
 * extents) and `nfluid` fluid systems of particle positions. No voxels, splats
 * or meshes are written. */
typedef struct VxlSphVxld VxlSphVxld;
VxlSphStatus vxl_sph_vxld_open(const char *path, uint32_t ticks, uint32_t ticks_per_frame, float dt,
                               uint32_t nfluid, VxlSphVxld **out);
/* A frame of engine scenes (nscenes == nfluid): each coupled scene's static
 * bodies (asleep) and dynamic bodies (awake, current pose), then each scene's
 * fluid positions (index order) as one fluid system. */
VxlSphStatus vxl_sph_vxld_frame(VxlSphVxld *w, VxlSphEngine *e, uint32_t tick, float ms, uint32_t nscenes,
                                const uint32_t *scenes);
/* A frame from host arrays: bodies (density > 0 written awake), counts[nfluid]
 * and the systems' positions back to back (3 floats per particle). */
VxlSphStatus vxl_sph_vxld_frame_raw(VxlSphVxld *w, uint32_t tick, float ms, uint32_t nbodies,
                                    const VxlSphRigidBox *bodies, uint32_t nfluid, const uint32_t *counts,
                                    const float *pos);
VxlSphStatus vxl_sph_vxld_close(VxlSphVxld *w);

/* ---- batched and served forms ---- */
/* A sampled dam-break request: scene geometry and physics, requested steps,
 * arrival time (served form) and dynamic boxes (served form: a request with
 * boxes is served as a coupled scene, see vxl_sph_scene_add_coupled). */
typedef struct VxlSphRequest {
    VxlSphDamBreak dam;
    float sound_speed;
    float artificial_viscosity;
    uint32_t steps;
    uint32_t nboxes;
    double arrival_s;
    double fill;              /* realised fluid volume / tank volume */
    VxlSphBoxBody boxes[4];
    float box_density[4];
} VxlSphRequest;

/* Batched form: `count` scenes of a seeded set (fluid count log-uniform in
 * [nmin, nmax]; tank x, z ~ U[1, 4] m, y = 1 m; fill ~ U[0.1, 0.5] sets the block
 * height; c0 ~ U[15, 30] m/s; alpha ~ U[0.01, 0.05]; d = 0.02 m). */
VxlSphStatus vxl_sph_sample_batch(uint64_t seed, uint32_t count, uint32_t nmin, uint32_t nmax, VxlSphRequest *out);
/* Served form: a seeded Poisson stream of `rate` requests/s, fluid count
 * log-uniform in [nmin, nmax], steps ~ U[smin, smax], 0..max_boxes boxes. */
VxlSphStatus vxl_sph_sample_requests(uint64_t seed, uint32_t count, double rate, uint32_t nmin, uint32_t nmax,
                                     uint32_t smin, uint32_t smax, uint32_t max_boxes, VxlSphRequest *out);
/* The fluid configuration a request's scene is stepped with
 * (h = 2d, mass = rho0 d^3, gamma 7, eps 0.5, one substep of dt). */
VxlSphStatus vxl_sph_request_config(const VxlSphRequest *req, float dt, uint64_t max_bins, VxlSphFluidConfig *out);
/* The Box bodies a request with boxes is served with (the tank's five static
 * bodies, then the dynamic boxes) and its rigid configuration; count = 0 for a
 * request without boxes. Call with out == NULL to query the count. */
VxlSphStatus vxl_sph_request_bodies(const VxlSphRequest *req, uint32_t cap, VxlSphRigidBox *out, uint32_t *count,
                                    VxlSphRigidConfig *rc);

/* Peak device bytes of serving this request alone: its pool (state, sort,
 * neighbour and cell arrays; see vxl_sph_pool_limits), readback staging and
 * scene tables. No counterpart in the reference; for sizing budget_bytes. */
VxlSphStatus vxl_sph_request_bytes(const VxlSphRequest *req, float dt, uint64_t max_bins, uint64_t *out);
/* Capacity of one engine's pool: particle slots, grid cells, and the bytes per
 * slot of the widest per-slot record the kernels address with 32-bit offsets
 * (max_slots x record_bytes stays within 2^32). No counterpart in the reference
 * (its packet has a fixed size, crates/vxl-phys-gpu/src/pipeline/gpu_types.rs:38-43). */
VxlSphStatus vxl_sph_pool_limits(uint64_t *max_slots, uint64_t *max_cells, uint32_t *record_bytes);

typedef struct VxlSphServeConfig {
    VxlSphHeader hdr;
    uint64_t budget_bytes;   /* device bytes the resident scenes may use */
    double time_scale;       /* arrival schedule speed-up (1 = real time) */
    double max_wall_s;
    uint32_t max_resident;   /* 0 = budget only */
    uint32_t builders;       /* host threads building arrived scenes */
    uint64_t max_bins;
    float dt;
    int32_t device;
    uint64_t max_particles;  /* resident particles (fluid + boundary) per pool step at most; 0 = no limit */
    uint32_t admit_every;    /* pool steps between admissions, batched into one layout (0 = 128) */
    uint32_t reserved;
} VxlSphServeConfig;

typedef struct VxlSphServeStats {
    uint32_t requests, completed;
    uint64_t pool_steps, fluid_updates;
    double wall_s, updates_per_s;
    double step_ms_p50, step_ms_p99, step_ms_mean; /* device time of one pool step (HIP events) */
    double loop_ms_p50, loop_ms_p99;  /* interval between consecutive pool-step completions on the
                                         device clock: the step plus any gap admissions, layouts or
                                         retirements left on the device (idle pools excluded) */
    double latency_s_p50, latency_s_p99;
    double wait_s_p50, wait_s_p99;
    uint32_t peak_resident;
    uint32_t rejected;        /* requests whose scene alone exceeds budget_bytes (not served) */
    uint64_t peak_bytes;      /* highest device bytes held (never above budget_bytes) */
    uint64_t layouts;
    double layout_ms_p50, layout_ms_p99; /* host time of one pool layout (admissions) */
    double layout_phase_ms[5];           /* summed over the run: waiting for the device, moving resident
                                            scenes, uploading new ones, scene tables, building the graph */
} VxlSphServeStats;

typedef struct VxlSphServeResult {
    uint32_t index, n, n_fluid, steps;
    double arrival_s, admit_s, finish_s;
    uint64_t state_hash;      /* same FNV-1a as vxl_sph_scene_state_hash */
} VxlSphServeResult;

/* Continuous batching: requests are built after they arrive, admitted in
 * arrival order as the HBM budget allows (the engine's projected peak device
 * bytes, pool growth included, stay within budget_bytes), co-stepped in one
 * launch per stage with every other resident request, and retired at step
 * boundaries. A request whose scene alone exceeds the budget is rejected:
 * its result has steps = 0, the others are served, and the call returns
 * VXL_SPH_CAPACITY with stats and results filled in. */
VxlSphStatus vxl_sph_serve(const VxlSphServeConfig *cfg, uint32_t n, const VxlSphRequest *reqs,
                           VxlSphServeStats *stats, VxlSphServeResult *results);

#ifdef __cplusplus
}
#endif
#endif
