// extern "C" boundary of the WCSPH library (include/vxl_sph.h). Every entry
// catches C++ exceptions and turns them into status codes; nothing throws
// across the boundary (ffi/src/lib.rs:1-8).
#include "../../include/vxl_sph.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "sph_device_step.h"
#include "sph_engine.h"
#include "sph_serve.h"
#include "sph_vxld.h"

using namespace vxl;

struct VxlSphEngine {
    Engine eng;
    explicit VxlSphEngine(int dev) : eng(dev) {}
};

struct VxlSphStepper {
    Engine eng;
    SceneConfig cfg;
    uint32_t scene = 0;
    bool started = false;
    uint32_t n = 0, nf = 0;
    std::vector<Span> spans;
    std::vector<VxlSphReaction> reactions;
    // Walls handed over before the first step (the scene is created by it).
    bool pend_halfspace = false;
    std::vector<WallProvider> pend_prov;
    struct PendTable {
        bool set = false;
        std::vector<uint32_t> ids, start;
        std::vector<float> planes;
    } pend_tab[2];
    VxlSphStepper(int dev, const SceneConfig &c) : eng(dev), cfg(c) {}
    void apply_pending_walls() {
        if (pend_halfspace) eng.set_halfspace_walls(scene, pend_prov);
        for (uint32_t side = 0; side < 2; ++side) {
            const PendTable &t = pend_tab[side];
            if (t.set)
                eng.set_wall_tables(scene, side, t.ids.data(), (uint32_t)t.ids.size(), t.start.data(), t.planes.data(),
                                    (uint32_t)(t.planes.size() / 6));
        }
        pend_halfspace = false;
        pend_prov.clear();
        pend_tab[0] = pend_tab[1] = PendTable{};
    }
};

namespace {

thread_local std::string g_last_error;

// Provider array + faces (listed provider by provider) -> engine providers.
std::vector<WallProvider> to_wall_providers(uint32_t nprov, const VxlSphWallProvider *provs, uint32_t nface,
                                            const VxlSphWallFace *faces) {
    std::vector<WallProvider> out(nprov);
    uint32_t last = 0;
    for (uint32_t k = 0; k < nprov; ++k) {
        out[k].id = provs[k].id;
        out[k].has_bounds = provs[k].has_bounds != 0;
        for (int a = 0; a < 3; ++a) {
            out[k].lo[a] = provs[k].lo[a];
            out[k].hi[a] = provs[k].hi[a];
        }
    }
    for (uint32_t f = 0; f < nface; ++f) {
        const uint32_t k = faces[f].provider;
        if (k >= nprov) throw std::invalid_argument("wall face names no provider");
        if (k < last) throw std::invalid_argument("wall faces must be listed provider by provider");
        last = k;
        out[k].faces.push_back(WallProvider::Face{faces[f].axis, faces[f].plane, faces[f].sign});
    }
    return out;
}

VxlSphStatus fail(VxlSphStatus s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

template <class F>
VxlSphStatus guard(F &&f) {
    try {
        f();
        return VXL_SPH_OK;
    } catch (const CapacityError &e) {
        return fail(VXL_SPH_CAPACITY, e.what());
    } catch (const DeviceError &e) {
        if (e.code == hipErrorOutOfMemory) return fail(VXL_SPH_OUT_OF_MEMORY, e.what());
        return fail(VXL_SPH_DEVICE_ERROR, e.what());
    } catch (const std::length_error &e) {
        return fail(VXL_SPH_SIZE_MISMATCH, e.what());
    } catch (const std::out_of_range &e) {
        return fail(VXL_SPH_NOT_FOUND, e.what());
    } catch (const std::domain_error &e) {
        return fail(VXL_SPH_UNSUPPORTED, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(VXL_SPH_INVALID_ARGUMENT, e.what());
    } catch (const std::logic_error &e) {
        return fail(VXL_SPH_NOT_READY, e.what());
    } catch (const std::bad_alloc &e) {
        return fail(VXL_SPH_OUT_OF_MEMORY, e.what());
    } catch (const std::exception &e) {
        return fail(VXL_SPH_PANIC_CAUGHT, e.what());
    } catch (...) {
        return fail(VXL_SPH_PANIC_CAUGHT, "unknown exception");
    }
}

VxlSphStatus check_header(const VxlSphHeader &h, uint32_t size) {
    if (h.version != VXL_SPH_ABI_VERSION) return fail(VXL_SPH_VERSION_MISMATCH, "block version mismatch");
    if (h.size != size) return fail(VXL_SPH_SIZE_MISMATCH, "block size mismatch");
    return VXL_SPH_OK;
}

BoxBody to_box_body(const VxlSphBoxBody &b) {
    BoxBody o;
    o.id = b.id;
    for (int a = 0; a < 3; ++a) {
        o.half[a] = b.half[a];
        o.pos[a] = b.pos[a];
        o.linvel[a] = b.linvel[a];
        o.angvel[a] = b.angvel[a];
    }
    for (int a = 0; a < 4; ++a) o.rot[a] = b.rot[a];
    return o;
}

SceneConfig to_scene_config(const VxlSphFluidConfig &c) {
    SceneConfig s;
    s.rest_density = c.rest_density;
    s.smoothing_radius = c.smoothing_radius;
    s.xsph_viscosity = c.xsph_viscosity;
    s.sound_speed = c.sound_speed;
    s.gamma_tait = c.gamma_tait;
    s.substeps = c.substeps;
    s.tensile_clamp = c.tensile_clamp;
    s.boundary_layers = c.boundary_layers;
    for (int a = 0; a < 3; ++a) s.gravity[a] = c.gravity[a];
    s.max_speed_frac = c.max_speed_frac;
    s.artificial_viscosity = c.artificial_viscosity;
    s.particle_mass = c.particle_mass;
    s.spacing = c.spacing;
    s.max_bins = c.max_bins ? c.max_bins : (1ull << 21);
    s.dt_tick = c.dt_tick;
    return s;
}

std::vector<Span> to_spans(const VxlSphSpan *spans, uint32_t nspans) {
    std::vector<Span> v(nspans);
    for (uint32_t k = 0; k < nspans; ++k) {
        v[k].body = spans[k].body;
        for (int a = 0; a < 3; ++a) v[k].origin[a] = spans[k].origin[a];
        v[k].start = spans[k].start;
        v[k].end = spans[k].end;
    }
    return v;
}

uint64_t fnv_state(const float *pos, const float *vel, uint32_t nf) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        h ^= (uint64_t)u;
        h *= 0x00000100000001b3ull;
    };
    for (uint32_t i = 0; i < nf; ++i) {
        mix(pos[3 * i]); mix(pos[3 * i + 1]); mix(pos[3 * i + 2]);
        mix(vel[3 * i]); mix(vel[3 * i + 1]); mix(vel[3 * i + 2]);
    }
    return h;
}

}  // namespace

namespace {

// Arguments of the device-array step: a valid config, a supported scene, and
// arrays that are device memory long enough for n particles.
SceneConfig device_step_args(const VxlSphFluidConfig &cfg, uint32_t n, uint32_t n_fluid, const float *pos,
                             const float *vel, const float *pmass, uint32_t max_n) {
    const SceneConfig sc = to_scene_config(cfg);
    if (n == 0) throw std::invalid_argument("empty scene");
    if (n_fluid > n) throw std::invalid_argument("n_fluid exceeds particle count");
    if (n > max_n) throw CapacityError("more particles than the step holds");
    if (fabsf(sc.gamma_tait - 7.0f) >= 1e-6f)
        throw std::domain_error("only the gamma = 7 Tait branch (fluid_force.rs:17-21) is supported");
    check_device_range(pos, 12ull * n, "pos");
    check_device_range(vel, 12ull * n, "vel");
    check_device_range(pmass, 4ull * n, "pmass");
    return sc;
}

}  // namespace

extern "C" {

uint32_t vxl_sph_abi_version(void) { return VXL_SPH_ABI_VERSION; }

const char *vxl_sph_backend_name(void) { return "hip-gfx950"; }

const char *vxl_sph_last_error(void) { return g_last_error.c_str(); }

VxlSphStatus vxl_sph_device_count(int32_t *out) {
    if (!out) return fail(VXL_SPH_NULL_POINTER, "out is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *out = n;
    return VXL_SPH_OK;
}

int32_t vxl_sph_backend_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int32_t vxl_sph_backend_submit(const char *tag) {
    if (!tag || std::strcmp(tag, "fluid") != 0) return 0;
    return vxl_sph_backend_available();
}

VxlSphStatus vxl_sph_default_config(VxlSphFluidConfig *out) {
    if (!out) return fail(VXL_SPH_NULL_POINTER, "out is null");
    SceneConfig d;
    std::memset(out, 0, sizeof(*out));
    out->hdr.size = sizeof(VxlSphFluidConfig);
    out->hdr.version = VXL_SPH_ABI_VERSION;
    out->rest_density = d.rest_density;
    out->smoothing_radius = d.smoothing_radius;
    out->xsph_viscosity = d.xsph_viscosity;
    out->sound_speed = d.sound_speed;
    out->gamma_tait = d.gamma_tait;
    out->substeps = d.substeps;
    out->tensile_clamp = d.tensile_clamp;
    out->boundary_layers = d.boundary_layers;
    for (int a = 0; a < 3; ++a) out->gravity[a] = d.gravity[a];
    out->max_speed_frac = d.max_speed_frac;
    out->artificial_viscosity = d.artificial_viscosity;
    out->particle_mass = d.particle_mass;
    out->spacing = d.spacing;
    out->max_bins = d.max_bins;
    out->dt_tick = d.dt_tick;
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_engine_create(uint32_t abi_version, int32_t device, VxlSphEngine **out) {
    if (!out) return fail(VXL_SPH_NULL_POINTER, "out is null");
    *out = nullptr;
    if (abi_version != VXL_SPH_ABI_VERSION) return fail(VXL_SPH_VERSION_MISMATCH, "ABI version mismatch");
    return guard([&] { *out = new VxlSphEngine(device); });
}

VxlSphStatus vxl_sph_engine_destroy(VxlSphEngine *e) {
    delete e;
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_engine_set_diagnostics(VxlSphEngine *e, int32_t on) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] { e->eng.set_diagnostics(on != 0); });
}

VxlSphStatus vxl_sph_engine_set_graphs(VxlSphEngine *e, int32_t on) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] { e->eng.set_graphs(on != 0); });
}

VxlSphStatus vxl_sph_engine_set_tuning(VxlSphEngine *e, const VxlSphTuning *t) {
    if (!e || !t) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(t->hdr, sizeof(VxlSphTuning));
    if (st != VXL_SPH_OK) return st;
    if (t->force_kernel > VXL_SPH_FORCE_STAGED) return fail(VXL_SPH_INVALID_ARGUMENT, "unknown kernel variant");
    return guard([&] {
        Engine::Tuning tu;
        tu.persist_blocks = t->persist_blocks;
        tu.force_kernel = t->force_kernel;
        tu.block_threads = t->block_threads;
        tu.stage_records = t->stage_records;
        tu.scan_tile_cells = t->scan_tile_cells;
        e->eng.set_tuning(tu);
    });
}

VxlSphStatus vxl_sph_engine_get_tuning(VxlSphEngine *e, VxlSphTuning *out) {
    if (!e || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    const Engine::Tuning &tu = e->eng.tuning();
    std::memset(out, 0, sizeof(*out));
    out->hdr.size = sizeof(VxlSphTuning);
    out->hdr.version = VXL_SPH_ABI_VERSION;
    out->persist_blocks = tu.persist_blocks;
    out->force_kernel = tu.force_kernel;
    out->block_threads = tu.block_threads;
    out->stage_records = tu.stage_records;
    out->scan_tile_cells = tu.scan_tile_cells;
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_engine_autotune(VxlSphEngine *e, uint32_t substeps, VxlSphTuning *out) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    VxlSphStatus st = guard([&] { e->eng.autotune(substeps); });
    if (st == VXL_SPH_OK && out) return vxl_sph_engine_get_tuning(e, out);
    return st;
}

VxlSphStatus vxl_sph_engine_step(VxlSphEngine *e, uint32_t ticks) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] { e->eng.step(ticks); });
}

VxlSphStatus vxl_sph_engine_sync(VxlSphEngine *e) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] { e->eng.sync(); });
}

VxlSphStatus vxl_sph_engine_device_bytes(VxlSphEngine *e, uint64_t *out) {
    if (!e || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    *out = e->eng.device_bytes();
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_engine_graph_stats(VxlSphEngine *e, uint64_t *built, uint64_t *updated) {
    if (!e || !built || !updated) return fail(VXL_SPH_NULL_POINTER, "null argument");
    *built = e->eng.graph_builds();
    *updated = e->eng.graph_updates();
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_engine_profile(VxlSphEngine *e, uint32_t substeps, uint32_t cap, VxlSphKernelTime *out,
                                    uint32_t *count) {
    if (!e || !count || (cap && !out)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const auto t = e->eng.profile(substeps);
        *count = (uint32_t)t.size();
        for (uint32_t k = 0; k < cap && k < t.size(); ++k) {
            std::memset(out[k].name, 0, sizeof(out[k].name));
            std::strncpy(out[k].name, t[k].first.c_str(), sizeof(out[k].name) - 1);
            out[k].ms_total = t[k].second;
        }
    });
}

VxlSphStatus vxl_sph_engine_step_timed(VxlSphEngine *e, uint32_t ticks, float *ms_out) {
    if (!e || (ticks && !ms_out)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const auto t = e->eng.step_timed(ticks);
        std::memcpy(ms_out, t.data(), 4ull * t.size());
    });
}

VxlSphStatus vxl_sph_scene_add(VxlSphEngine *e, const VxlSphFluidConfig *cfg, uint32_t n, uint32_t n_fluid,
                               const float *pos, const float *vel, const float *pmass, const VxlSphSpan *spans,
                               uint32_t nspans, uint32_t *out_scene) {
    if (!e || !cfg || !pos || !out_scene || (nspans && !spans)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        const SceneConfig c = to_scene_config(*cfg);
        const std::vector<Span> sp = to_spans(spans, nspans);
        *out_scene = e->eng.add_scene(c, n, n_fluid, pos, vel, pmass, sp.data(), nspans);
    });
}

VxlSphStatus vxl_sph_scene_remove(VxlSphEngine *e, uint32_t scene) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] { e->eng.remove_scene(scene); });
}

VxlSphStatus vxl_sph_scene_upload_boundary(VxlSphEngine *e, uint32_t scene, uint32_t nb, const float *pos,
                                           const float *vel) {
    if (!e || (nb && (!pos || !vel))) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] { e->eng.upload_boundary(scene, nb, pos, vel); });
}

VxlSphStatus vxl_sph_scene_set_halfspace_walls(VxlSphEngine *e, uint32_t scene, uint32_t nprov,
                                               const VxlSphWallProvider *provs, uint32_t nface,
                                               const VxlSphWallFace *faces) {
    if (!e || (nprov && !provs) || (nface && !faces)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] { e->eng.set_halfspace_walls(scene, to_wall_providers(nprov, provs, nface, faces)); });
}

VxlSphStatus vxl_sph_scene_set_wall_tables(VxlSphEngine *e, uint32_t scene, uint32_t side, uint32_t ne,
                                           const uint32_t *ids, const uint32_t *start, uint32_t np,
                                           const float *planes) {
    if (!e || !start || (ne && !ids) || (np && !planes)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] { e->eng.set_wall_tables(scene, side, ids, ne, start, planes, np); });
}

VxlSphStatus vxl_sph_scene_clear_walls(VxlSphEngine *e, uint32_t scene) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] { e->eng.clear_walls(scene); });
}

VxlSphStatus vxl_sph_scene_read_state(VxlSphEngine *e, uint32_t scene, uint32_t n, float *pos, float *vel) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] {
        if (n != e->eng.scene(scene).params.n) throw std::length_error("n differs from the scene's particle count");
        e->eng.read_state(scene, pos, vel);
    });
}

VxlSphStatus vxl_sph_scene_read_fields(VxlSphEngine *e, uint32_t scene, uint32_t n, float *dens, float *press,
                                       uint32_t *ncount, float *acc, float *xsph) {
    if (!e) return fail(VXL_SPH_NULL_POINTER, "engine is null");
    return guard([&] {
        if (n != e->eng.scene(scene).params.n) throw std::length_error("n differs from the scene's particle count");
        e->eng.read_fields(scene, dens, press, ncount, acc, xsph);
    });
}

VxlSphStatus vxl_sph_scene_read_grid(VxlSphEngine *e, uint32_t scene, VxlSphGridInfo *info, uint32_t start_cap,
                                     uint32_t *start, uint32_t n, uint32_t *items, uint32_t *keys) {
    if (!e || !info) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const SceneSlot &sc = e->eng.scene(scene);
        std::vector<uint32_t> st, it;
        const SceneGrid g = e->eng.read_grid(scene, &st, &it);
        for (int a = 0; a < 3; ++a) {
            info->min[a] = g.gmin[a];
            info->dims[a] = g.dims[a];
        }
        info->inv = g.inv;
        info->bin = g.bin;
        info->total = g.total;
        if (start) {
            if (start_cap < g.total + 1) throw std::length_error("start buffer smaller than total + 1");
            std::memcpy(start, st.data(), 4ull * (g.total + 1));
        }
        if (items || keys) {
            if (n != sc.params.n) throw std::length_error("n differs from the scene's particle count");
            if (items) std::memcpy(items, it.data(), 4ull * n);
            if (keys)
                for (uint32_t c = 0; c < g.total; ++c)
                    for (uint32_t k = st[c]; k < st[c + 1]; ++k) keys[it[k]] = c;
        }
    });
}

VxlSphStatus vxl_sph_scene_read_bounds(VxlSphEngine *e, uint32_t scene, float lo[3], float hi[3]) {
    if (!e || !lo || !hi) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] { e->eng.read_bounds(scene, lo, hi); });
}

VxlSphStatus vxl_sph_scene_state_hash(VxlSphEngine *e, uint32_t scene, uint64_t *out) {
    if (!e || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const SceneSlot &sc = e->eng.scene(scene);
        std::vector<float> p(3ull * sc.params.n), v(3ull * sc.params.n);
        e->eng.read_state(scene, p.data(), v.data());
        *out = fnv_state(p.data(), v.data(), sc.params.nf);
    });
}

VxlSphStatus vxl_sph_scene_state_hash_all(VxlSphEngine *e, uint32_t scene, uint64_t *out) {
    if (!e || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const SceneSlot &sc = e->eng.scene(scene);
        std::vector<float> p(3ull * sc.params.n), v(3ull * sc.params.n);
        e->eng.read_state(scene, p.data(), v.data());
        uint64_t h = 0xcbf29ce484222325ull;
        auto bytes = [&](const std::vector<float> &a) {
            const unsigned char *b = reinterpret_cast<const unsigned char *>(a.data());
            for (size_t k = 0; k < 4 * a.size(); ++k) {
                h ^= (uint64_t)b[k];
                h *= 0x00000100000001b3ull;
            }
        };
        bytes(p);
        bytes(v);
        *out = h;
    });
}

// ---- one scene in caller-owned device arrays ----


VxlSphStatus vxl_sph_step_device(const VxlSphFluidConfig *cfg, uint32_t n, uint32_t n_fluid, float *pos, float *vel,
                                 const float *pmass, uint32_t ticks, void *stream) {
    if (!cfg || !pos || !vel || !pmass) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        const SceneConfig sc = device_step_args(*cfg, n, n_fluid, pos, vel, pmass, kMaxPoolSlots - kBlock);
        if (ticks == 0) return;
        // The grid rule keeps a scene within max_bins cells: no capacity retry.
        const uint32_t ccap = (uint32_t)std::min<uint64_t>(sc.max_bins, (uint64_t)kMaxCells - 1u) + 1u;
        const size_t bytes = device_step_workspace(n, ccap, nullptr, nullptr, nullptr, nullptr, nullptr);
        char *ws = nullptr;
        const hipError_t e = hipMalloc((void **)&ws, bytes);
        if (e != hipSuccess) throw DeviceError(e, std::string("hipMalloc: ") + hipGetErrorString(e));
        try {
            device_step(sc, n, n_fluid, pos, vel, pmass, ticks, (hipStream_t)stream, ws, bytes, ccap);
        } catch (...) {
            (void)hipFree(ws);
            throw;
        }
        (void)hipFree(ws);
    });
}

VxlSphStatus vxl_sph_step_device_noalloc(const VxlSphFluidConfig *cfg, uint32_t n, uint32_t n_fluid, float *pos,
                                         float *vel, const float *pmass, uint32_t ticks, void *stream) {
    if (!cfg || !pos || !vel || !pmass) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        const SceneConfig sc = device_step_args(*cfg, n, n_fluid, pos, vel, pmass, ~0u);
        device_step_noalloc(sc, n, n_fluid, pos, vel, pmass, ticks, (hipStream_t)stream);
    });
}

// ---- FluidStepper drop-in ----

VxlSphStatus vxl_sph_stepper_create(uint32_t abi_version, const VxlSphFluidConfig *cfg, int32_t device,
                                    VxlSphStepper **out) {
    if (!out || !cfg) return fail(VXL_SPH_NULL_POINTER, "null argument");
    *out = nullptr;
    if (abi_version != VXL_SPH_ABI_VERSION) return fail(VXL_SPH_VERSION_MISMATCH, "ABI version mismatch");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] { *out = new VxlSphStepper(device, to_scene_config(*cfg)); });
}

VxlSphStatus vxl_sph_stepper_destroy(VxlSphStepper *s) {
    delete s;
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_stepper_step(VxlSphStepper *s, float dt_tick, uint32_t n_fluid, uint32_t n, const float *pos,
                                  const float *vel, const float *pmass, const VxlSphSpan *spans, uint32_t nspans) {
    if (!s || !pos || !vel || !pmass || (nspans && !spans)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        if (!s->started) {
            s->cfg.dt_tick = dt_tick;
            s->spans = to_spans(spans, nspans);
            s->scene = s->eng.add_scene(s->cfg, n, n_fluid, pos, vel, pmass, s->spans.data(), nspans);
            try {
                s->apply_pending_walls();
            } catch (...) {
                s->eng.remove_scene(s->scene);
                throw;
            }
            s->n = n;
            s->nf = n_fluid;
            s->started = true;
        } else {
            // pipeline/stepper.rs:80-96: fixed particle counts; pmass length-checked only.
            if (n_fluid != s->nf) throw std::length_error("fluid prefix length differs from the stepper's");
            if (n != s->n) throw std::length_error("particle count differs from the stepper's");
            if (dt_tick != s->cfg.dt_tick) throw std::invalid_argument("tick length differs from the first step's");
            // Spans come every tick (tick_average takes them, stepper.rs:97-109):
            // same ranges, torque origins that follow the bodies.
            std::vector<Span> sp = to_spans(spans, nspans);
            s->eng.set_span_origins(s->scene, sp.data(), nspans);
            s->spans = std::move(sp);
            s->eng.upload_boundary(s->scene, n - n_fluid, pos + 3ull * n_fluid, vel + 3ull * n_fluid);
        }
        s->eng.step(1);
        const std::vector<float> r = s->eng.read_reactions(s->scene);
        s->reactions.resize(s->spans.size());
        for (size_t k = 0; k < s->spans.size(); ++k) {
            s->reactions[k].body = s->spans[k].body;
            for (int a = 0; a < 3; ++a) {
                s->reactions[k].force[a] = r[6 * k + a];
                s->reactions[k].torque[a] = r[6 * k + 3 + a];
            }
        }
    });
}

VxlSphStatus vxl_sph_stepper_reactions(VxlSphStepper *s, uint32_t cap, VxlSphReaction *out, uint32_t *count) {
    if (!s || !count || (cap && !out)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    *count = (uint32_t)s->reactions.size();
    for (uint32_t k = 0; k < cap && k < *count; ++k) out[k] = s->reactions[k];
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_stepper_bounds(VxlSphStepper *s, float lo[3], float hi[3]) {
    if (!s || !lo || !hi) return fail(VXL_SPH_NULL_POINTER, "null argument");
    if (!s->started) return fail(VXL_SPH_NOT_READY, "no step taken yet");
    return guard([&] { s->eng.read_bounds(s->scene, lo, hi); });
}

VxlSphStatus vxl_sph_stepper_read_state(VxlSphStepper *s, uint32_t n, float *pos, float *vel) {
    if (!s) return fail(VXL_SPH_NULL_POINTER, "null argument");
    if (!s->started) return fail(VXL_SPH_NOT_READY, "no step taken yet");
    return guard([&] {
        if (n != s->n) throw std::length_error("n differs from the stepper's particle count");
        s->eng.read_state(s->scene, pos, vel);
    });
}

VxlSphStatus vxl_sph_stepper_set_wall_tables(VxlSphStepper *s, uint32_t side, uint32_t ne, const uint32_t *ids,
                                             const uint32_t *start, uint32_t np, const float *planes) {
    if (!s || !start || (ne && !ids) || (np && !planes)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    if (side > 1) return fail(VXL_SPH_INVALID_ARGUMENT, "wall table side must be 0 (mirror) or 1 (project)");
    return guard([&] {
        if (s->started) {
            s->eng.set_wall_tables(s->scene, side, ids, ne, start, planes, np);
            return;
        }
        auto &t = s->pend_tab[side];
        t.set = true;
        t.ids.assign(ids, ids + ne);
        t.start.assign(start, start + ne + 1);
        t.planes.assign(planes, planes + 6ull * np);
        s->pend_halfspace = false;
        s->pend_prov.clear();
    });
}

VxlSphStatus vxl_sph_stepper_set_halfspace_walls(VxlSphStepper *s, uint32_t nprov, const VxlSphWallProvider *provs,
                                                 uint32_t nface, const VxlSphWallFace *faces) {
    if (!s || (nprov && !provs) || (nface && !faces)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        std::vector<WallProvider> pv = to_wall_providers(nprov, provs, nface, faces);
        if (s->started) {
            s->eng.set_halfspace_walls(s->scene, pv);
            return;
        }
        s->pend_halfspace = true;
        s->pend_prov = std::move(pv);
        s->pend_tab[0] = s->pend_tab[1] = VxlSphStepper::PendTable{};
    });
}

// ---- scene construction ----

VxlSphStatus vxl_sph_box_lattice(const float half[3], float sp, float h, uint32_t layers, uint32_t cap,
                                 float *out_pts, uint32_t *out_n, uint32_t *out_n0, float *out_volume) {
    if (!half || !out_n) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        std::vector<float> pts;
        uint32_t n0 = 0;
        float vol = 0.0f;
        const uint32_t n = box_lattice(half, sp, h, layers, pts, n0, vol);
        *out_n = n;
        if (out_n0) *out_n0 = n0;
        if (out_volume) *out_volume = vol;
        if (out_pts) std::memcpy(out_pts, pts.data(), 12ull * std::min(n, cap));
    });
}

VxlSphStatus vxl_sph_default_rigid_config(VxlSphRigidConfig *out) {
    if (!out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    const RigidConfig d;
    std::memset(out, 0, sizeof(*out));
    out->hdr.size = sizeof(VxlSphRigidConfig);
    out->hdr.version = VXL_SPH_ABI_VERSION;
    out->substeps = d.substeps;
    for (int a = 0; a < 3; ++a) {
        out->gravity[a] = d.gravity[a];
        out->tank[a] = d.tank[a];
    }
    out->max_linear = d.max_linear;
    out->max_angular = d.max_angular;
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_scene_add_coupled(VxlSphEngine *e, const VxlSphFluidConfig *cfg, const VxlSphRigidConfig *rc,
                                       uint32_t n_fluid, const float *fluid_pos, const float *fluid_vel,
                                       uint32_t nbodies, const VxlSphRigidBox *bodies, uint32_t *out_scene) {
    if (!e || !cfg || !rc || !out_scene || (n_fluid && !fluid_pos) || (nbodies && !bodies))
        return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    st = check_header(rc->hdr, sizeof(VxlSphRigidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        RigidConfig r;
        r.substeps = rc->substeps;
        for (int a = 0; a < 3; ++a) {
            r.gravity[a] = rc->gravity[a];
            r.tank[a] = rc->tank[a];
        }
        r.max_linear = rc->max_linear;
        r.max_angular = rc->max_angular;
        std::vector<BoxBody> bs(nbodies);
        std::vector<float> dens(nbodies);
        for (uint32_t k = 0; k < nbodies; ++k) {
            bs[k] = to_box_body(bodies[k].box);
            dens[k] = bodies[k].density;
        }
        *out_scene = e->eng.add_coupled_scene(to_scene_config(*cfg), n_fluid, fluid_pos, fluid_vel, bs, dens, r);
    });
}

VxlSphStatus vxl_sph_scene_read_bodies(VxlSphEngine *e, uint32_t scene, uint32_t cap, VxlSphRigidBox *out,
                                       uint32_t *count) {
    if (!e || !count) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const SceneSlot &sc = e->eng.scene(scene);
        *count = (uint32_t)sc.rigid.size();
        if (!out) return;
        if (cap < *count) throw std::length_error("output buffer too small");
        const std::vector<RigidDev> cur = e->eng.read_bodies(scene);
        for (size_t k = 0; k < cur.size(); ++k) {
            VxlSphRigidBox &o = out[k];
            std::memset(&o, 0, sizeof(o));
            o.box.id = sc.rigid[k].body_id;
            for (int a = 0; a < 3; ++a) {
                o.box.half[a] = cur[k].half[a];
                o.box.pos[a] = cur[k].pos[a];
                o.box.linvel[a] = cur[k].linvel[a];
                o.box.angvel[a] = cur[k].angvel[a];
            }
            for (int a = 0; a < 4; ++a) o.box.rot[a] = cur[k].rot[a];
            o.density = sc.rigid[k].density;
        }
    });
}

VxlSphStatus vxl_sph_box_particles(const VxlSphFluidConfig *cfg, uint32_t nbodies, const VxlSphBoxBody *bodies,
                                   uint32_t first, uint32_t *n, uint32_t *nspans, float *pos, float *vel,
                                   float *pmass, VxlSphSpan *spans) {
    if (!cfg || !n || !nspans || (nbodies && !bodies)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        std::vector<BoxBody> bs(nbodies);
        for (uint32_t k = 0; k < nbodies; ++k) bs[k] = to_box_body(bodies[k]);
        std::vector<float> p, v, m;
        std::vector<Span> sp;
        p.resize(3ull * first);  // spans index the whole particle array
        v.resize(3ull * first);
        m.resize(first);
        append_boundary_boxes(to_scene_config(*cfg), bs, p, v, m, sp);
        const uint32_t count = (uint32_t)m.size() - first;
        const bool fill = pos != nullptr;
        if (fill && (*n < count || *nspans < sp.size())) throw std::length_error("output buffers too small");
        *n = count;
        *nspans = (uint32_t)sp.size();
        if (!fill) return;
        std::memcpy(pos, p.data() + 3ull * first, 12ull * count);
        if (vel) std::memcpy(vel, v.data() + 3ull * first, 12ull * count);
        if (pmass) std::memcpy(pmass, m.data() + first, 4ull * count);
        if (spans)
            for (size_t k = 0; k < sp.size(); ++k) {
                spans[k].body = sp[k].body;
                for (int a = 0; a < 3; ++a) spans[k].origin[a] = sp[k].origin[a];
                spans[k].start = sp[k].start;
                spans[k].end = sp[k].end;
            }
    });
}

VxlSphStatus vxl_sph_fluid_constants(const VxlSphFluidConfig *cfg, float out[8]) {
    if (!cfg || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        const FluidConstants c = fluid_constants(to_scene_config(*cfg));
        out[0] = c.h; out[1] = c.h2; out[2] = c.k6; out[3] = c.ks;
        out[4] = c.w0; out[5] = c.b_tait; out[6] = c.mass; out[7] = c.alpha_c;
    });
}

VxlSphStatus vxl_sph_dam_break(const VxlSphDamBreak *spec, const VxlSphFluidConfig *cfg, uint32_t *n,
                               uint32_t *n_fluid, uint32_t *nspans, float *pos, float *vel, float *pmass,
                               VxlSphSpan *spans) {
    if (!spec || !cfg || !n || !n_fluid || !nspans) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(spec->hdr, sizeof(VxlSphDamBreak));
    if (st != VXL_SPH_OK) return st;
    st = check_header(cfg->hdr, sizeof(VxlSphFluidConfig));
    if (st != VXL_SPH_OK) return st;
    return guard([&] {
        DamBreakSpec ds;
        for (int a = 0; a < 3; ++a) {
            ds.tank[a] = spec->tank[a];
            ds.block[a] = spec->block[a];
        }
        ds.d = spec->d;
        ds.jitter = spec->jitter;
        ds.seed = spec->seed;
        ds.walls = spec->walls != 0;
        std::vector<float> p, v, m;
        std::vector<Span> sp;
        uint32_t nf = 0;
        dam_break(ds, to_scene_config(*cfg), p, v, m, sp, nf);
        const uint32_t count = (uint32_t)(m.size());
        const bool fill = pos != nullptr;
        if (fill && (*n < count || *nspans < sp.size())) throw std::length_error("output buffers too small");
        *n = count;
        *n_fluid = nf;
        *nspans = (uint32_t)sp.size();
        if (!fill) return;
        std::memcpy(pos, p.data(), 4 * p.size());
        if (vel) std::memcpy(vel, v.data(), 4 * v.size());
        if (pmass) std::memcpy(pmass, m.data(), 4 * m.size());
        if (spans)
            for (size_t k = 0; k < sp.size(); ++k) {
                spans[k].body = sp[k].body;
                for (int a = 0; a < 3; ++a) spans[k].origin[a] = sp[k].origin[a];
                spans[k].start = sp[k].start;
                spans[k].end = sp[k].end;
            }
    });
}

}  // extern "C"

namespace {

void to_request(const SceneSample &s, VxlSphRequest &r) {
    std::memset(&r, 0, sizeof(r));
    r.dam.hdr.size = sizeof(VxlSphDamBreak);
    r.dam.hdr.version = VXL_SPH_ABI_VERSION;
    for (int a = 0; a < 3; ++a) {
        r.dam.tank[a] = s.dam.tank[a];
        r.dam.block[a] = s.dam.block[a];
    }
    r.dam.d = s.dam.d;
    r.dam.jitter = s.dam.jitter;
    r.dam.seed = s.dam.seed;
    r.dam.walls = s.dam.walls ? 1u : 0u;
    r.sound_speed = s.sound_speed;
    r.artificial_viscosity = s.alpha;
    r.steps = s.steps;
    r.nboxes = s.nboxes;
    r.arrival_s = s.arrival_s;
    r.fill = s.fill;
    for (uint32_t b = 0; b < 4; ++b) {
        r.boxes[b].id = s.boxes[b].id;
        for (int a = 0; a < 3; ++a) {
            r.boxes[b].half[a] = s.boxes[b].half[a];
            r.boxes[b].pos[a] = s.boxes[b].pos[a];
            r.boxes[b].linvel[a] = s.boxes[b].linvel[a];
            r.boxes[b].angvel[a] = s.boxes[b].angvel[a];
        }
        for (int a = 0; a < 4; ++a) r.boxes[b].rot[a] = s.boxes[b].rot[a];
        r.box_density[b] = s.box_density[b];
    }
}

SceneSample from_request(const VxlSphRequest &r) {
    SceneSample s;
    for (int a = 0; a < 3; ++a) {
        s.dam.tank[a] = r.dam.tank[a];
        s.dam.block[a] = r.dam.block[a];
    }
    s.dam.d = r.dam.d;
    s.dam.jitter = r.dam.jitter;
    s.dam.seed = r.dam.seed;
    s.dam.walls = r.dam.walls != 0;
    s.sound_speed = r.sound_speed;
    s.alpha = r.artificial_viscosity;
    s.steps = r.steps;
    s.nboxes = r.nboxes;
    s.arrival_s = r.arrival_s;
    s.fill = r.fill;
    for (uint32_t b = 0; b < 4; ++b) {
        s.boxes[b].id = r.boxes[b].id;
        for (int a = 0; a < 3; ++a) {
            s.boxes[b].half[a] = r.boxes[b].half[a];
            s.boxes[b].pos[a] = r.boxes[b].pos[a];
            s.boxes[b].linvel[a] = r.boxes[b].linvel[a];
            s.boxes[b].angvel[a] = r.boxes[b].angvel[a];
        }
        for (int a = 0; a < 4; ++a) s.boxes[b].rot[a] = r.boxes[b].rot[a];
        s.box_density[b] = r.box_density[b];
    }
    return s;
}

}  // namespace

extern "C" {

VxlSphStatus vxl_sph_sample_batch(uint64_t seed, uint32_t count, uint32_t nmin, uint32_t nmax, VxlSphRequest *out) {
    if (!out && count) return fail(VXL_SPH_NULL_POINTER, "out is null");
    if (nmin == 0 || nmax < nmin) return fail(VXL_SPH_INVALID_ARGUMENT, "bad particle-count range");
    return guard([&] {
        for (uint32_t k = 0; k < count; ++k) to_request(sample_batch_scene(seed, k, nmin, nmax), out[k]);
    });
}

VxlSphStatus vxl_sph_sample_requests(uint64_t seed, uint32_t count, double rate, uint32_t nmin, uint32_t nmax,
                                     uint32_t smin, uint32_t smax, uint32_t max_boxes, VxlSphRequest *out) {
    if (!out && count) return fail(VXL_SPH_NULL_POINTER, "out is null");
    if (!(rate > 0.0) || nmin == 0 || nmax < nmin || smin == 0 || smax < smin || max_boxes > 4)
        return fail(VXL_SPH_INVALID_ARGUMENT, "bad request stream parameters");
    return guard([&] {
        const auto v = sample_requests(seed, count, rate, nmin, nmax, smin, smax, max_boxes);
        for (uint32_t k = 0; k < count; ++k) to_request(v[k], out[k]);
    });
}

// ---- frame dumps ----

struct VxlSphVxld {
    vxl::VxldWriter w;
    VxlSphVxld(const char *path, uint32_t ticks, uint32_t tpf, float dt, uint32_t nfluid)
        : w(path, ticks, tpf, dt, nfluid) {}
};

namespace {
VxldBody to_vxld(const BoxBody &b, uint8_t awake) {
    VxldBody o;
    o.awake = awake;
    for (int a = 0; a < 3; ++a) {
        o.pos[a] = b.pos[a];
        o.half[a] = b.half[a];
    }
    for (int a = 0; a < 4; ++a) o.rot[a] = b.rot[a];
    return o;
}
}  // namespace

VxlSphStatus vxl_sph_vxld_open(const char *path, uint32_t ticks, uint32_t ticks_per_frame, float dt,
                               uint32_t nfluid, VxlSphVxld **out) {
    if (!path || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    *out = nullptr;
    return guard([&] { *out = new VxlSphVxld(path, ticks, ticks_per_frame, dt, nfluid); });
}

VxlSphStatus vxl_sph_vxld_frame(VxlSphVxld *w, VxlSphEngine *e, uint32_t tick, float ms, uint32_t nscenes,
                                const uint32_t *scenes) {
    if (!w || !e || (nscenes && !scenes)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        std::vector<VxldBody> bodies;
        std::vector<std::vector<float>> pos(nscenes);
        std::vector<std::pair<const float *, uint32_t>> fl;
        for (uint32_t k = 0; k < nscenes; ++k) {
            const SceneSlot &sc = e->eng.scene(scenes[k]);
            for (const BoxBody &b : sc.statics) bodies.push_back(to_vxld(b, 0u));
            const std::vector<RigidDev> dyn = e->eng.read_bodies(scenes[k]);
            for (const RigidDev &d : dyn) {
                VxldBody o;
                o.awake = 1u;
                for (int a = 0; a < 3; ++a) {
                    o.pos[a] = d.pos[a];
                    o.half[a] = d.half[a];
                }
                for (int a = 0; a < 4; ++a) o.rot[a] = d.rot[a];
                bodies.push_back(o);
            }
            pos[k].resize(3ull * sc.params.n);
            e->eng.read_state(scenes[k], pos[k].data(), nullptr);
            fl.push_back({pos[k].data(), sc.params.nf});
        }
        w->w.frame(tick, ms, bodies, fl);
    });
}

VxlSphStatus vxl_sph_vxld_frame_raw(VxlSphVxld *w, uint32_t tick, float ms, uint32_t nbodies,
                                    const VxlSphRigidBox *bodies, uint32_t nfluid, const uint32_t *counts,
                                    const float *pos) {
    if (!w || (nbodies && !bodies) || (nfluid && !counts)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        std::vector<VxldBody> bs;
        for (uint32_t k = 0; k < nbodies; ++k)
            bs.push_back(to_vxld(to_box_body(bodies[k].box), bodies[k].density > 0.0f ? 1u : 0u));
        std::vector<std::pair<const float *, uint32_t>> fl;
        size_t at = 0;
        for (uint32_t k = 0; k < nfluid; ++k) {
            if (counts[k] && !pos) throw std::invalid_argument("positions are null");
            fl.push_back({pos ? pos + 3 * at : nullptr, counts[k]});
            at += counts[k];
        }
        w->w.frame(tick, ms, bs, fl);
    });
}

VxlSphStatus vxl_sph_vxld_close(VxlSphVxld *w) {
    if (!w) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = guard([&] { w->w.close(); });
    delete w;
    return st;
}

VxlSphStatus vxl_sph_request_bodies(const VxlSphRequest *req, uint32_t cap, VxlSphRigidBox *out, uint32_t *count,
                                    VxlSphRigidConfig *rc) {
    if (!req || !count) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        const SceneSample s = from_request(*req);
        std::vector<BoxBody> bodies;
        std::vector<float> dens;
        RigidConfig r;
        if (s.nboxes) request_bodies(s, bodies, dens, r);
        *count = (uint32_t)bodies.size();
        if (rc) {
            vxl_sph_default_rigid_config(rc);
            for (int a = 0; a < 3; ++a) rc->tank[a] = r.tank[a];
        }
        if (!out) return;
        if (cap < *count) throw std::length_error("output buffer too small");
        for (size_t k = 0; k < bodies.size(); ++k) {
            const BoxBody &b = bodies[k];
            VxlSphRigidBox &o = out[k];
            std::memset(&o, 0, sizeof(o));
            o.box.id = b.id;
            for (int a = 0; a < 3; ++a) {
                o.box.half[a] = b.half[a];
                o.box.pos[a] = b.pos[a];
                o.box.linvel[a] = b.linvel[a];
                o.box.angvel[a] = b.angvel[a];
            }
            for (int a = 0; a < 4; ++a) o.box.rot[a] = b.rot[a];
            o.density = dens[k];
        }
    });
}

VxlSphStatus vxl_sph_request_bytes(const VxlSphRequest *req, float dt, uint64_t max_bins, uint64_t *out) {
    if (!req || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        ServeConfig sc;
        sc.dt = dt > 0 ? dt : 5e-4f;
        sc.max_bins = max_bins ? max_bins : (1ull << 21);
        *out = request_solo_bytes(from_request(*req), sc);
    });
}

VxlSphStatus vxl_sph_pool_limits(uint64_t *max_slots, uint64_t *max_cells, uint32_t *record_bytes) {
    if (!max_slots || !max_cells || !record_bytes) return fail(VXL_SPH_NULL_POINTER, "null argument");
    *max_slots = kMaxPoolSlots;
    *max_cells = kMaxCells;
    *record_bytes = kAtRecordBytes;
    return VXL_SPH_OK;
}

VxlSphStatus vxl_sph_request_config(const VxlSphRequest *req, float dt, uint64_t max_bins, VxlSphFluidConfig *out) {
    if (!req || !out) return fail(VXL_SPH_NULL_POINTER, "null argument");
    return guard([&] {
        ServeConfig sc;
        sc.dt = dt;
        sc.max_bins = max_bins ? max_bins : (1ull << 21);
        SceneSample s = from_request(*req);
        s.nboxes = 0;
        SceneConfig c;
        std::vector<float> p, v, m;
        std::vector<Span> sp;
        uint32_t nf = 0;
        // Only the configuration is needed: build a one-particle block.
        s.dam.block[0] = s.dam.block[1] = s.dam.block[2] = 1;
        s.dam.walls = false;
        build_request(s, sc, c, p, v, m, sp, nf);
        vxl_sph_default_config(out);
        out->rest_density = c.rest_density;
        out->smoothing_radius = c.smoothing_radius;
        out->xsph_viscosity = c.xsph_viscosity;
        out->sound_speed = c.sound_speed;
        out->gamma_tait = c.gamma_tait;
        out->substeps = c.substeps;
        out->tensile_clamp = c.tensile_clamp;
        out->boundary_layers = c.boundary_layers;
        for (int a = 0; a < 3; ++a) out->gravity[a] = c.gravity[a];
        out->max_speed_frac = c.max_speed_frac;
        out->artificial_viscosity = c.artificial_viscosity;
        out->particle_mass = c.particle_mass;
        out->spacing = c.spacing;
        out->max_bins = c.max_bins;
        out->dt_tick = c.dt_tick;
    });
}

VxlSphStatus vxl_sph_serve(const VxlSphServeConfig *cfg, uint32_t n, const VxlSphRequest *reqs,
                           VxlSphServeStats *stats, VxlSphServeResult *results) {
    if (!cfg || !stats || (n && !reqs)) return fail(VXL_SPH_NULL_POINTER, "null argument");
    VxlSphStatus st = check_header(cfg->hdr, sizeof(VxlSphServeConfig));
    if (st != VXL_SPH_OK) return st;
    uint32_t rejected = 0;
    const VxlSphStatus rc = guard([&] {
        ServeConfig sc;
        sc.budget_bytes = cfg->budget_bytes;
        sc.time_scale = cfg->time_scale > 0 ? cfg->time_scale : 1.0;
        sc.max_wall_s = cfg->max_wall_s > 0 ? cfg->max_wall_s : 3600.0;
        sc.max_resident = cfg->max_resident;
        sc.builders = cfg->builders ? cfg->builders : 4;
        sc.max_bins = cfg->max_bins ? cfg->max_bins : (1ull << 21);
        sc.dt = cfg->dt > 0 ? cfg->dt : 5e-4f;
        sc.max_particles = cfg->max_particles;
        sc.admit_every = cfg->admit_every ? cfg->admit_every : 128u;
        std::vector<SceneSample> v(n);
        for (uint32_t k = 0; k < n; ++k) {
            v[k] = from_request(reqs[k]);
        }
        Engine eng(cfg->device);
        std::vector<ServeResult> res;
        const ServeStats s = serve(eng, v, sc, res);
        std::memset(stats, 0, sizeof(*stats));
        stats->requests = s.requests;
        stats->completed = s.completed;
        stats->pool_steps = s.pool_steps;
        stats->fluid_updates = s.fluid_updates;
        stats->wall_s = s.wall_s;
        stats->updates_per_s = s.updates_per_s;
        stats->step_ms_p50 = s.step_ms_p50;
        stats->step_ms_p99 = s.step_ms_p99;
        stats->step_ms_mean = s.step_ms_mean;
        stats->loop_ms_p50 = s.loop_ms_p50;
        stats->loop_ms_p99 = s.loop_ms_p99;
        stats->latency_s_p50 = s.latency_s_p50;
        stats->latency_s_p99 = s.latency_s_p99;
        stats->wait_s_p50 = s.wait_s_p50;
        stats->wait_s_p99 = s.wait_s_p99;
        stats->peak_resident = s.peak_resident;
        stats->rejected = s.rejected;
        rejected = s.rejected;
        stats->peak_bytes = s.peak_bytes;
        stats->layouts = s.layouts;
        stats->layout_ms_p50 = s.layout_ms_p50;
        stats->layout_ms_p99 = s.layout_ms_p99;
        for (int k = 0; k < 5; ++k) stats->layout_phase_ms[k] = s.layout_phase_ms[k];
        if (results)
            for (uint32_t k = 0; k < n; ++k) {
                results[k].index = res[k].index;
                results[k].n = res[k].n;
                results[k].n_fluid = res[k].nf;
                results[k].steps = res[k].steps;
                results[k].arrival_s = res[k].arrival_s;
                results[k].admit_s = res[k].admit_s;
                results[k].finish_s = res[k].finish_s;
                results[k].state_hash = res[k].state_hash;
            }
    });
    if (rc == VXL_SPH_OK && rejected)
        return fail(VXL_SPH_CAPACITY, std::to_string(rejected) + " request(s) exceed the HBM budget on their own");
    return rc;
}

}  // extern "C"
