/*
 * Rigid-box leg of the coupled fluid step — TEST INFRASTRUCTURE ONLY (same
 * rules as wcsph_oracle.h: only tests/, smoke() and bench.py's CPU baseline
 * may load it; the product never links it).
 *
 * A strict-f32 restatement of the reference's rigid integrator for dynamic
 * Box bodies driven by gravity and the fluid's tick-averaged reaction:
 *   - mass properties of a box (crates/vxl-phys-core/src/mass.rs:35-48);
 *   - gravity as a force g * (1 / inv_mass) (crates/vxl-phys-field/src/lib.rs:22-31),
 *     then the fluid reaction (F, tau) added through the coupling applier
 *     (crates/vxl-phys/src/world_step.rs:235-250, world_step/coupling.rs:52-56);
 *   - Integrator::integrate_velocities / integrate_positions
 *     (crates/vxl-phys-integrate/src/lib.rs:18-64) with the world inverse
 *     inertia R diag(I^-1) R^T (crates/vxl-phys-core/src/body.rs:172-176) and
 *     Quat::integrate_angular (crates/vxl-phys-core/src/math.rs:228-240);
 *   - PhysConfig defaults: 2 substeps per tick, |v| <= 100, |w| <= 50
 *     (crates/vxl-phys-core/src/config.rs:206-233).
 * The reference has no contact between a box and a particle-only tank; this
 * build adds one rule (documented with or_rigid_tick below).
 */
#ifndef RIGID_ORACLE_H
#define RIGID_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrRigidBody {
    float half[3];
    float pos[3];
    float rot[4]; /* quaternion x, y, z, w */
    float linvel[3];
    float angvel[3];
    float inv_mass;
    float inv_inertia[3]; /* body-frame diagonal */
} OrRigidBody;

typedef struct OrRigidConfig {
    float dt_tick;
    uint32_t substeps;
    float gravity[3];
    float max_linear;
    float max_angular;
    float tank[3]; /* contact planes of the open tank [0,L0] x [0,inf) x [0,L2]; L0 <= 0: none */
} OrRigidConfig;

/* mass_props for Shape::Box (mass.rs:35-48); density <= 0 falls back to 1. */
void or_box_mass_props(const float half[3], float density, float *inv_mass, float inv_inertia[3]);

/* One tick: `substeps` substeps of dt = dt_tick / substeps. Per substep and
 * body (in order): force = g * (1 / inv_mass) + F, torque = tau, with
 * react[6k..6k+6) = (F, tau) of body k (the previous fluid tick's average);
 * integrate velocities (clamped), integrate positions, then the tank contact
 * rule: for each plane in the order floor (y >= 0), x >= 0, x <= L0, z >= 0,
 * z <= L2, the deepest box corner (first maximum over corners
 * k = 0..7, corner = (+-hx, +-hy, +-hz) with bit 0/1/2 of k selecting + on
 * x/y/z) that penetrates by d > 0 moves the box by d along the plane normal;
 * if that corner's velocity v + w x r approaches the plane (v.n < 0), a
 * frictionless, fully inelastic impulse j = -v.n / (1/m + n.((I^-1 (r x n)) x r))
 * is applied at it. */
void or_rigid_tick(const OrRigidConfig *cfg, uint32_t n, OrRigidBody *bodies, const float *react);

#ifdef __cplusplus
}
#endif
#endif
