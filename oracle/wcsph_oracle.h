/*
 * WCSPH oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A strict-f32 CPU restatement of the reference engine's WCSPH fluid step
 * (crates/vxl-phys-fluid). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / the timed CPU
 * baseline. The product library (rust-wl_amd/) never links or calls it.
 *
 * Every function cites the reference file:line whose operation order it
 * follows. All arithmetic is IEEE f32, built with -ffp-contract=off.
 *
 * Pins: every function is checked against the reference's known answers,
 * recorded hashes and readouts, its own tests' invariants or closed forms
 * (tests/test_oracle.py, test_golden.py, test_walls_oracle.py,
 * test_reference_boundary_oracle.py). One branch has no such check and its
 * parity is unpinned: the Tait equation for gamma != 7 (the reference's
 * powf, not bit-reproducible here; the product rejects gamma != 7).
 */
#ifndef WCSPH_ORACLE_H
#define WCSPH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FluidConfig (crates/vxl-phys-fluid/src/config.rs:6-69), plus the grid budget
 * (crates/vxl-phys-core/src/grid.rs:20) as a parameter. */
typedef struct OrConfig {
    float rest_density;
    float smoothing_radius;
    float xsph_viscosity;
    float sound_speed;
    float gamma_tait;
    uint32_t substeps;
    uint32_t threads;
    uint32_t tensile_clamp;      /* tensile_instability_suppression */
    uint32_t boundary_layers;
    float gravity[3];
    float max_speed_frac;
    float artificial_viscosity;
    uint64_t max_bins;           /* GRID_MAX_BINS */
    uint32_t skip_reactions;     /* 1: no boundary-force pass (timing a GPU line that computes none) */
    uint32_t pad;
} OrConfig;

/* A Box body for set_boundary_particles (fluid_step.rs:13-53). */
typedef struct OrBoxBody {
    uint32_t id;
    float half[3];
    float pos[3];
    float rot[4];     /* quaternion x, y, z, w */
    float linvel[3];
    float angvel[3];
} OrBoxBody;

typedef struct OrSystem OrSystem;

void or_default_config(OrConfig *cfg);

/* FluidSystem::new: lattice block + mass calibration (fluid_access.rs:7-76). */
OrSystem *or_new_lattice(const OrConfig *cfg, const float origin[3],
                         const uint32_t dims[3], float spacing);

/* Restatement-only constructor: explicit fluid positions/velocities (3 floats
 * per particle) and an explicit particle mass. Constants are derived exactly
 * as FluidSystem::new does (fluid_access.rs:8-13). */
OrSystem *or_new_particles(const OrConfig *cfg, uint32_t n_fluid,
                           const float *pos, const float *vel, float mass,
                           float spacing);

void or_free(OrSystem *s);

/* set_boundary_particles for Box bodies (fluid_step.rs:13-53). Returns the
 * boundary particle count. */
uint32_t or_set_boundary_boxes(OrSystem *s, uint32_t nbodies, const OrBoxBody *bodies);

/* Raw append of boundary particles (restatement-only): positions, velocities,
 * per-particle masses and span table (body, origin[3], start, end as floats
 * would be lossy, so spans are given as separate arrays). */
void or_set_boundary_raw(OrSystem *s, uint32_t nb, const float *pos, const float *vel,
                         const float *pmass, uint32_t nspans, const uint32_t *span_body,
                         const float *span_origin, const uint32_t *span_start,
                         const uint32_t *span_end);

void or_set_velocities(OrSystem *s, uint32_t n, const float *vel);
void or_set_threads(OrSystem *s, uint32_t threads);
/* Test-only planted mistakes (1 = poly6 reassociated, 2 = XSPH dropped,
 * 3 = neighbour order transposed, 4 = mirror ghost reflected the wrong way,
 * 5 = coincident separation moves the earlier particle); 0 restores the
 * reference algorithm. */
void or_set_fault(OrSystem *s, int fault);

/* Provider walls (crates/vxl-phys-core/src/interop.rs:166-224). A provider is
 * a set of axis-aligned half-space faces, as the reference's test doubles
 * FloorY and Tank (crates/vxl-phys-fluid/src/tests.rs:11-110): for a query
 * point p, face (axis, plane, sign) within 2 * probe of p ((p[axis] - plane) *
 * sign <= 2 probe) yields the contact point p with p[axis] = plane and the
 * outward normal sign * e_axis. Providers are queried in the order given (the
 * boundaries list of set_boundaries, fluid_access.rs:184-186), a provider's
 * faces in the order given. */
typedef struct OrWallProvider {
    uint32_t id;
    uint32_t has_bounds;   /* 0: bounds() = None, the AABB prefilter passes */
    float lo[3], hi[3];
} OrWallProvider;
typedef struct OrWallFace {
    uint32_t axis;         /* 0, 1, 2 */
    float plane;
    float sign;            /* +1 or -1: the fluid side */
    uint32_t provider;     /* index into the provider list */
} OrWallFace;

/* Half-space providers queried every substep, the CPU path: mirror ghosts in
 * the density pass (fluid_density.rs:17-51, :76-121) and the projection with
 * coincident separation after the integration (fluid_boundary.rs:45-131).
 * Returns 0, or -1 for a face naming no provider or an axis > 2. */
int or_set_halfspace_walls(OrSystem *s, uint32_t nprov, const OrWallProvider *provs,
                           uint32_t nface, const OrWallFace *faces);
/* The sparse plane tables of the current fluid positions (gather_in,
 * wall_gather.rs:52-100; side 0 mirror, 1 project): returns the entry count;
 * writes up to cap_e ids and start[0..ne] and up to cap_p planes (point xyz,
 * normal xyz); *out_np = plane count. */
uint32_t or_gather_wall_tables(const OrSystem *s, uint32_t side, uint32_t cap_e, uint32_t *ids,
                               uint32_t *start, uint32_t cap_p, float *planes, uint32_t *out_np);
/* Plane tables held until replaced, the GPU path (WallStage::upload,
 * walls.rs:104-116; kernels wall_ghost.wgsl:61-178): ghosts summed apart after
 * the density pass, projection after the integration without separation.
 * ids strictly ascending fluid indices. Returns 0 or -1 (malformed). */
int or_set_wall_tables(OrSystem *s, uint32_t side, uint32_t ne, const uint32_t *ids,
                       const uint32_t *start, uint32_t np, const float *planes);
void or_clear_walls(OrSystem *s);
/* Particles the last boundary_pass pushed (half-space walls). */
uint32_t or_last_pushed(const OrSystem *s);

/* FluidSystem::step (fluid_step.rs:75-89). */
void or_step(OrSystem *s, float dt_tick);

/* Sizes and scalars. */
uint32_t or_n_fluid(const OrSystem *s);
uint32_t or_n_total(const OrSystem *s);
uint32_t or_n_spans(const OrSystem *s);
float or_mass(const OrSystem *s);
/* h, h2, k6, ks, w0, b_tait, mass, alpha_c  (8 floats) */
void or_constants(const OrSystem *s, float out[8]);

/* Arrays (all particles, index order; Vec3 arrays are 3 floats per particle). */
void or_get_pos(const OrSystem *s, float *out);
void or_get_vel(const OrSystem *s, float *out);
void or_get_pmass(const OrSystem *s, float *out);
void or_get_dens(const OrSystem *s, float *out);
void or_get_press(const OrSystem *s, float *out);
void or_get_acc(const OrSystem *s, float *out);   /* fluid prefix: n_fluid*3 */
void or_get_xsph(const OrSystem *s, float *out);  /* fluid prefix: n_fluid*3 */
void or_get_bforce(const OrSystem *s, float *out); /* boundary suffix: nb*3 */
void or_get_ncount(const OrSystem *s, uint32_t *out); /* neighbours of last density pass */
/* (body, F[3], tau[3]) per span: out_body[nspans], out_ft[nspans*6] */
void or_get_reactions(const OrSystem *s, uint32_t *out_body, float *out_ft);

/* Grid of the last rebuild (fluid grid.rs:29-76): min[3], inv, bin, dims[3]. */
void or_grid_info(const OrSystem *s, float out_min[3], float *out_inv, float *out_bin,
                  uint32_t out_dims[3]);
uint64_t or_grid_total(const OrSystem *s);
void or_grid_tables(const OrSystem *s, uint32_t *start /* total+1 */, uint32_t *items /* n */,
                    uint32_t *keys /* n, per particle cell index */);

/* FNV-1a over start, 0xffffffff, items, 0xffffffff, per-particle bin
 * (crates/vxl-phys-gpu/examples/gpu_grid_probe.rs:50-68). */
uint64_t or_grid_hash(const OrSystem *s);
/* FNV-1a over fluid pos then vel bits per particle (sph_scale.rs:52-63). */
uint64_t or_state_hash(const OrSystem *s);

/* grid_box rule (crates/vxl-phys-core/src/grid.rs:34-51). */
void or_grid_box(const float lo[3], const float hi[3], float h, uint64_t max_bins,
                 float out_min[3], float *out_bin, uint32_t out_dims[3]);

/* boundary::lattice for a Box (boundary.rs:67-88, :208-232). Writes up to cap
 * points; returns the total count. n0 = points per layer. */
uint32_t or_lattice_box(const float half[3], float s, float h, uint32_t layers,
                        float *out_pts, uint32_t cap, uint32_t *out_n0, float *out_volume);
/* boundary_self_sum restated literally, O(n^2) (boundary.rs:106-129). */
float or_boundary_self_sum_bruteforce(const float *pts, uint32_t n, float h);
/* Same sum through a uniform grid with per-particle neighbour lists sorted by
 * index: the identical summation order, O(n k log k). */
float or_boundary_self_sum_grid(const float *pts, uint32_t n, float h);

/* Neighbour set of particle i from the last grid (grid.rs:115-157) in
 * enumeration order; returns count, writes up to cap indices. */
uint32_t or_neighbors(const OrSystem *s, uint32_t i, uint32_t *out, uint32_t cap);

#ifdef __cplusplus
}
#endif
#endif
