/*
 * A plain C consumer of include/vxl_sph.h: the call sequences a host binding
 * (the vxl-phys-hip crate's HipFluidStepper / HipBackend, or a cgo/ctypes
 * wrapper) makes, with no Python in the process. The tests run it as is and
 * under host AddressSanitizer + UndefinedBehaviorSanitizer (the stepper and
 * the CPU-side entry points) and ThreadSanitizer (the threaded serve loop).
 *
 *   abi_harness cpu     OUT   CPU-side entry points and their error codes
 *   abi_harness stepper OUT   FluidStepper drop-in: the sequence
 *                             World::fluid_stepper_pass drives
 *                             (crates/vxl-phys/src/world_step/fluid_stepper.rs:66-79):
 *                             step, reactions, bounds every tick
 *   abi_harness serve   OUT   the served form over a small seeded stream
 *
 * OUT receives little-endian records the tests compare with the oracle (or
 * with another build of the library). Exit status 0 means every status code
 * was the expected one; a message names the first one that was not.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "vxl_sph.h"

static int fail(const char *what, VxlSphStatus got, VxlSphStatus want) {
    fprintf(stderr, "abi_harness: %s returned %d, expected %d (%s)\n", what, (int)got, (int)want,
            vxl_sph_last_error());
    return 1;
}

#define EXPECT(call, want)                                   \
    do {                                                     \
        VxlSphStatus st_ = (call);                           \
        if (st_ != (want)) return fail(#call, st_, (want));  \
    } while (0)
#define OK(call) EXPECT(call, VXL_SPH_OK)

static void put(FILE *f, const void *p, size_t bytes) {
    if (bytes && fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "abi_harness: short write\n");
        exit(2);
    }
}

/* The dam-break physics of the test scenes: h = 2d, mass = rho0 d^3 rounded
 * as three f32 products, one substep of dt (substeps may be raised). */
static VxlSphStatus dam_config(float d, float dt, uint32_t substeps, VxlSphFluidConfig *c) {
    VxlSphStatus st = vxl_sph_default_config(c);
    if (st != VXL_SPH_OK) return st;
    c->smoothing_radius = 2.0f * d;
    c->sound_speed = 20.0f;
    c->artificial_viscosity = 0.02f;
    c->xsph_viscosity = 0.5f;
    c->substeps = substeps;
    c->boundary_layers = 1;
    c->particle_mass = ((1000.0f * d) * d) * d;
    c->spacing = d;
    c->max_bins = 1ull << 21;
    c->dt_tick = dt;
    return VXL_SPH_OK;
}

typedef struct Scene {
    uint32_t n, nf, ns;
    float *pos, *vel, *pmass;
    VxlSphSpan *spans;
} Scene;

static void scene_free(Scene *s) {
    free(s->pos);
    free(s->vel);
    free(s->pmass);
    free(s->spans);
}

static VxlSphStatus dam_scene(const VxlSphFluidConfig *c, const float tank[3], const uint32_t block[3], float d,
                              uint64_t seed, Scene *s) {
    VxlSphDamBreak spec;
    memset(&spec, 0, sizeof spec);
    spec.hdr.size = sizeof spec;
    spec.hdr.version = VXL_SPH_ABI_VERSION;
    memcpy(spec.tank, tank, sizeof spec.tank);
    memcpy(spec.block, block, sizeof spec.block);
    spec.d = d;
    spec.jitter = 0.05f;
    spec.seed = seed;
    spec.walls = 1;
    memset(s, 0, sizeof *s);
    VxlSphStatus st = vxl_sph_dam_break(&spec, c, &s->n, &s->nf, &s->ns, NULL, NULL, NULL, NULL);
    if (st != VXL_SPH_OK) return st;
    s->pos = calloc(3 * (size_t)s->n, sizeof(float));
    s->vel = calloc(3 * (size_t)s->n, sizeof(float));
    s->pmass = calloc(s->n, sizeof(float));
    s->spans = calloc(s->ns ? s->ns : 1, sizeof(VxlSphSpan));
    if (!s->pos || !s->vel || !s->pmass || !s->spans) return VXL_SPH_OUT_OF_MEMORY;
    return vxl_sph_dam_break(&spec, c, &s->n, &s->nf, &s->ns, s->pos, s->vel, s->pmass, s->spans);
}

/* ---- cpu: entry points that need no device ---- */
static int run_cpu(FILE *out) {
    if (vxl_sph_abi_version() != VXL_SPH_ABI_VERSION) {
        fprintf(stderr, "abi_harness: abi version %u\n", vxl_sph_abi_version());
        return 1;
    }
    VxlSphFluidConfig c;
    EXPECT(vxl_sph_default_config(NULL), VXL_SPH_NULL_POINTER);
    OK(dam_config(0.02f, 5e-4f, 1, &c));

    float k[8];
    OK(vxl_sph_fluid_constants(&c, k));
    VxlSphFluidConfig bad = c;
    bad.hdr.version = VXL_SPH_ABI_VERSION + 1;
    EXPECT(vxl_sph_fluid_constants(&bad, k), VXL_SPH_VERSION_MISMATCH);
    bad = c;
    bad.hdr.size = 8;
    EXPECT(vxl_sph_fluid_constants(&bad, k), VXL_SPH_SIZE_MISMATCH);

    /* boundary::lattice of a box: query, then fill; a short buffer receives its
     * first points and the call still reports the whole count. */
    const float half[3] = {0.2f, 0.05f, 0.15f};
    uint32_t n = 0, n0 = 0;
    float vol = 0.0f;
    OK(vxl_sph_box_lattice(half, 0.02f, 0.04f, 2, 0, NULL, &n, &n0, &vol));
    float *pts = calloc(3 * (size_t)n, sizeof(float));
    if (!pts) return 2;
    OK(vxl_sph_box_lattice(half, 0.02f, 0.04f, 2, n, pts, &n, &n0, &vol));
    uint32_t n_short = 0;
    OK(vxl_sph_box_lattice(half, 0.02f, 0.04f, 2, n / 2, pts, &n_short, NULL, NULL));
    if (n_short != n) {
        fprintf(stderr, "abi_harness: lattice count %u, then %u\n", n, n_short);
        return 1;
    }
    uint32_t hdr[4] = {0x43505531u /* "1UPC" */, n, n0, 0};
    put(out, hdr, sizeof hdr);
    put(out, &vol, sizeof vol);
    put(out, pts, 3 * (size_t)n * sizeof(float));
    free(pts);

    /* A lattice large enough for the threaded self-sum (64Ki points and up). */
    const float plate[3] = {1.0f, 0.02f, 1.0f};
    uint32_t np = 0, np0 = 0;
    float pvol = 0.0f;
    OK(vxl_sph_box_lattice(plate, 0.01f, 0.02f, 1, 0, NULL, &np, &np0, &pvol));
    if (np < 65536) {
        fprintf(stderr, "abi_harness: plate lattice of %u points\n", np);
        return 1;
    }
    uint32_t ph[2] = {np, np0};
    put(out, ph, sizeof ph);
    put(out, &pvol, sizeof pvol);

    /* A dam-break scene, written whole. */
    const float tank[3] = {0.4f, 0.3f, 0.3f};
    const uint32_t block[3] = {9, 8, 7};
    Scene s;
    OK(dam_scene(&c, tank, block, 0.02f, 11, &s));
    uint32_t sh[3] = {s.n, s.nf, s.ns};
    put(out, sh, sizeof sh);
    put(out, s.pos, 3 * (size_t)s.n * sizeof(float));
    put(out, s.vel, 3 * (size_t)s.n * sizeof(float));
    put(out, s.pmass, (size_t)s.n * sizeof(float));
    put(out, s.spans, (size_t)s.ns * sizeof(VxlSphSpan));
    scene_free(&s);

    /* A served stream, each request's configuration and solo footprint. */
    VxlSphRequest req[6];
    OK(vxl_sph_sample_requests(3, 6, 20.0, 4096, 65536, 10, 40, 4, req));
    for (int r = 0; r < 6; ++r) {
        VxlSphFluidConfig rc;
        uint64_t bytes = 0;
        OK(vxl_sph_request_config(&req[r], 5e-4f, 1ull << 21, &rc));
        OK(vxl_sph_request_bytes(&req[r], 5e-4f, 1ull << 21, &bytes));
        uint32_t nb = 0;
        OK(vxl_sph_request_bodies(&req[r], 0, NULL, &nb, NULL));
        uint64_t rec[3] = {req[r].steps, bytes, nb};
        put(out, rec, sizeof rec);
    }
    uint64_t slots = 0, cells = 0;
    uint32_t rb = 0;
    OK(vxl_sph_pool_limits(&slots, &cells, &rb));
    EXPECT(vxl_sph_pool_limits(NULL, &cells, &rb), VXL_SPH_NULL_POINTER);
    put(out, &slots, sizeof slots);
    put(out, &cells, sizeof cells);
    put(out, &rb, sizeof rb);
    return 0;
}

/* ---- stepper: the FluidStepper drop-in over a few ticks ---- */
static int run_stepper(FILE *out) {
    int32_t devices = 0;
    OK(vxl_sph_device_count(&devices));
    if (devices < 1 || !vxl_sph_backend_available()) {
        fprintf(stderr, "abi_harness: no device (%d), backend %s\n", devices, vxl_sph_backend_name());
        return 3;
    }
    VxlSphFluidConfig c;
    OK(dam_config(0.02f, 5e-4f, 2, &c));
    const float tank[3] = {0.4f, 0.3f, 0.3f};
    const uint32_t block[3] = {10, 9, 8};
    Scene s;
    OK(dam_scene(&c, tank, block, 0.02f, 7, &s));

    VxlSphStepper *st = NULL;
    EXPECT(vxl_sph_stepper_create(VXL_SPH_ABI_VERSION + 1, &c, 0, &st), VXL_SPH_VERSION_MISMATCH);
    OK(vxl_sph_stepper_create(VXL_SPH_ABI_VERSION, &c, 0, &st));
    float lo[3], hi[3];
    EXPECT(vxl_sph_stepper_bounds(st, lo, hi), VXL_SPH_NOT_READY); /* no state before the first step */

    const uint32_t ticks = 8;
    uint32_t hdr[4] = {0x50455453u /* "STEP" */, s.n, s.nf, ticks};
    put(out, hdr, sizeof hdr);
    VxlSphReaction *rx = calloc(s.ns ? s.ns : 1, sizeof *rx);
    if (!rx) return 2;
    for (uint32_t t = 0; t < ticks; ++t) {
        OK(vxl_sph_stepper_step(st, c.dt_tick, s.nf, s.n, s.pos, s.vel, s.pmass, s.spans, s.ns));
        uint32_t cnt = 0;
        OK(vxl_sph_stepper_reactions(st, 0, NULL, &cnt));
        if (cnt != s.ns) {
            fprintf(stderr, "abi_harness: %u reactions for %u spans\n", cnt, s.ns);
            return 1;
        }
        OK(vxl_sph_stepper_reactions(st, cnt, rx, &cnt));
        OK(vxl_sph_stepper_bounds(st, lo, hi));
        put(out, &cnt, sizeof cnt);
        put(out, rx, (size_t)cnt * sizeof *rx);
        put(out, lo, sizeof lo);
        put(out, hi, sizeof hi);
    }
    /* GpuFluidStepper asserts the particle set stays the same (pipeline/stepper.rs:84-96). */
    EXPECT(vxl_sph_stepper_step(st, c.dt_tick, s.nf, s.n - 1, s.pos, s.vel, s.pmass, s.spans, s.ns),
           VXL_SPH_SIZE_MISMATCH);
    EXPECT(vxl_sph_stepper_step(st, c.dt_tick, s.nf, s.n, NULL, s.vel, s.pmass, s.spans, s.ns),
           VXL_SPH_NULL_POINTER);
    float *pos = calloc(3 * (size_t)s.n, sizeof(float)), *vel = calloc(3 * (size_t)s.n, sizeof(float));
    if (!pos || !vel) return 2;
    OK(vxl_sph_stepper_read_state(st, s.n, pos, vel));
    put(out, pos, 3 * (size_t)s.n * sizeof(float));
    put(out, vel, 3 * (size_t)s.n * sizeof(float));
    OK(vxl_sph_stepper_destroy(st));
    free(pos);
    free(vel);
    free(rx);
    scene_free(&s);
    return 0;
}

/* ---- serve: continuous batching over a small stream ---- */
static int run_serve(FILE *out) {
    const uint32_t nreq = 10;
    VxlSphRequest req[10];
    OK(vxl_sph_sample_requests(21, nreq, 400.0, 2048, 12000, 3, 12, 2, req));
    VxlSphServeConfig sc;
    memset(&sc, 0, sizeof sc);
    sc.hdr.size = sizeof sc;
    sc.hdr.version = VXL_SPH_ABI_VERSION;
    sc.budget_bytes = 4ull << 30;
    sc.time_scale = 1.0;
    sc.max_wall_s = 120.0;
    sc.max_resident = 4;
    sc.builders = 3;
    sc.max_bins = 1ull << 21;
    sc.dt = 5e-4f;
    sc.device = 0;
    sc.admit_every = 2;
    VxlSphServeStats stats;
    VxlSphServeResult res[10];
    OK(vxl_sph_serve(&sc, nreq, req, &stats, res));
    if (stats.completed != nreq) {
        fprintf(stderr, "abi_harness: %u of %u requests completed\n", stats.completed, nreq);
        return 1;
    }
    uint32_t hdr[2] = {0x56524553u /* "SERV" */, nreq};
    put(out, hdr, sizeof hdr);
    for (uint32_t r = 0; r < nreq; ++r) {
        uint64_t rec[4] = {res[r].index, res[r].n, res[r].steps, res[r].state_hash};
        put(out, rec, sizeof rec);
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s cpu|stepper|serve OUT\n", argv[0]);
        return 2;
    }
    FILE *out = fopen(argv[2], "wb");
    if (!out) {
        perror(argv[2]);
        return 2;
    }
    int rc;
    if (!strcmp(argv[1], "cpu"))
        rc = run_cpu(out);
    else if (!strcmp(argv[1], "stepper"))
        rc = run_stepper(out);
    else if (!strcmp(argv[1], "serve"))
        rc = run_serve(out);
    else
        rc = 2;
    if (fclose(out) != 0) rc = rc ? rc : 2;
    if (!rc) printf("abi_harness %s ok\n", argv[1]);
    fflush(stdout);
    fflush(stderr);
    /* After a device was used, end without running static destructors: under
     * AddressSanitizer the HIP runtime's teardown at exit frees device memory
     * through the sanitizer's device allocator after the HSA runtime has
     * gone, which the sanitizer reports as a failed internal check. The
     * CPU-only mode exits normally (leak checking runs at exit). */
    if (strcmp(argv[1], "cpu") != 0) _exit(rc);
    return rc;
}
