// Host runtime of the WCSPH pool (see sph_engine.h).
#include "sph_engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace vxl {

namespace {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(e, std::string(what) + ": " + hipGetErrorString(e));
}

uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// Cell-table entries to reserve for a scene: the grid of its initial box
// with room for the box to grow (splashes raise the free surface), capped by
// the scene's own budget.
uint32_t cell_reserve(uint32_t cells_hint, uint64_t max_bins, bool full = false) {
    if (full) return (uint32_t)std::min<uint64_t>(max_bins, (uint64_t)kMaxCells - 1u) + 1u;
    uint64_t want = std::max<uint64_t>((uint64_t)cells_hint * 4ull, 4096ull);
    want = std::min<uint64_t>(want, max_bins);
    return (uint32_t)want + 1u;  // + pad cell
}

}  // namespace

uint64_t Engine::pool_bytes(uint64_t slots, uint64_t ccap, bool diagnostics, bool reactions) {
    // Per slot: cur + spare state 36 + 36, grid-ordered A + gid 20, density
    // record 32, sort 16, neighbour record 40 (alloc_pool); diagnostics 36,
    // reaction buffers and slot list 36. Per cell: counts, start, fluid count (alloc_cells).
    const uint64_t per = 180ull + (diagnostics ? 36ull : 0ull) + (reactions ? 36ull : 0ull);
    return per * slots + 16ull * kReadPad + sizeof(SceneBox) * (slots / kPartSlots + 1) + 12ull * ccap + 4ull +
           4ull * ((ccap + 2047) / 2048 + 1);
}

uint64_t Engine::footprint(uint32_t n, uint32_t cells_hint, uint64_t max_bins) {
    // cur + spare (36 B each), grid-ordered A + gid 20, density record 32,
    // sort 16, neighbour record 40 B per slot; counts, start and fluid count
    // 12 B per reserved cell.
    const uint64_t slot_bytes = 36 + 36 + 20 + 32 + 16 + 40;
    return slot_bytes * round_up(n, kBlock) + 12ull * cell_reserve(cells_hint, max_bins);
}

Engine::Engine(int device) : device_(device) {
    hip_check(hipSetDevice(device_), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipHostMalloc((void **)&h_ctl_, sizeof(Control), hipHostMallocDefault), "hipHostMalloc");
    std::memset(h_ctl_, 0, sizeof(Control));
    d_ctl_ = dalloc<Control>(1);
    hip_check(hipMemcpy(d_ctl_, h_ctl_, sizeof(Control), hipMemcpyHostToDevice), "ctl init");
}

Engine::~Engine() {
    try {
        if (stream_) (void)hipStreamSynchronize(stream_);
        invalidate_graph();
        free_all();
        dfree(d_ctl_);
        if (h_ctl_) (void)hipHostFree(h_ctl_);
        if (stream_) (void)hipStreamDestroy(stream_);
    } catch (...) {
    }
}

template <class T>
T *Engine::dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw DeviceError(e, std::string("hipMalloc: ") + hipGetErrorString(e));
    bytes_ += n * sizeof(T);
    peak_bytes_ = std::max(peak_bytes_, bytes_);
    alloc_size_[p] = n * sizeof(T);
    return (T *)p;
}

void Engine::dfree(void *p) {
    if (!p) return;
    auto it = alloc_size_.find(p);
    if (it != alloc_size_.end()) {
        bytes_ -= it->second;
        alloc_size_.erase(it);
    }
    (void)hipFree(p);
}

void Engine::free_all() {
    free_walls();
    for (auto &sc : scenes_) free_wall_tables(sc);
    for (void *p : pool_allocs_) dfree(p);
    pool_allocs_.clear();
    dfree(d_params_);
    dfree(d_grid_);
    dfree(d_box_);
    dfree(d_block_scene_);
    dfree(d_tile_sums_);
    dfree(d_spans_);
    dfree(d_react_);
    dfree(d_rigid_);
    dfree(d_rlat_);
    d_spans_ = nullptr;
    d_react_ = nullptr;
    d_rigid_ = nullptr;
    d_rlat_ = nullptr;
    nspans_ = 0;
    d_params_ = nullptr;
    d_grid_ = nullptr;
    d_box_ = nullptr;
    d_block_scene_ = nullptr;
    d_tile_sums_ = nullptr;
    dfree(staging_);
    staging_ = nullptr;
    staging_bytes_ = 0;
    spare_cA_ = spare_cV_ = nullptr;
    spare_cg_ = nullptr;
    std::memset(&P_, 0, sizeof(P_));
    slots_ = 0;
    ccap_ = 0;
}

void Engine::invalidate_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    gexec_ = nullptr;
    graph_ = nullptr;
    graph_stale_ = false;
}

SceneSlot Engine::prepare_scene(const SceneConfig &cfg, uint32_t n, uint32_t nf, const float *pos, const float *vel,
                                const float *pmass, const Span *spans, uint32_t nspans, bool reactions) {
    if (nf > n) throw std::invalid_argument("n_fluid exceeds particle count");
    if (n == 0) throw std::invalid_argument("empty scene");
    if (n > kMaxPoolSlots - kBlock) throw CapacityError("a pool holds at most 2^27 particle slots");
    if (fabsf(cfg.gamma_tait - 7.0f) >= 1e-6f)
        throw std::domain_error("only the gamma = 7 Tait branch (fluid_force.rs:17-21) is supported");
    SceneSlot sc;
    sc.params = scene_params(cfg, n, nf);
    sc.substeps = std::max<uint32_t>(cfg.substeps, 1);
    // Pack the records now (the layout only copies them): A = (x, y, z, m)
    // with m = +pmass for fluid and -pmass for boundary particles, V = (v, 0).
    sc.stage_a.resize(n);
    sc.stage_v.resize(n);
    sc.params.fmass_uniform = 1u;
    for (uint32_t i = 0; i < n; ++i) {
        const float m = pmass ? pmass[i] : sc.params.mass;
        if (!(m > 0.0f) || !std::isfinite(m)) throw std::invalid_argument("particle masses must be positive and finite");
        if (i < nf && m != sc.params.mass) sc.params.fmass_uniform = 0u;
        sc.stage_a[i] = make_float4(pos[3ull * i], pos[3ull * i + 1], pos[3ull * i + 2], i < nf ? m : -m);
        sc.stage_v[i] = vel ? make_float4(vel[3ull * i], vel[3ull * i + 1], vel[3ull * i + 2], 0.0f)
                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    for (uint32_t k = 0; k < nspans; ++k) {
        if (spans[k].start < nf || spans[k].end > n || spans[k].start > spans[k].end)
            throw std::invalid_argument("span outside the boundary segment");
        sc.spans.push_back(spans[k]);
    }
    sc.want_reactions = reactions && nspans > 0;
    sc.params.react = sc.want_reactions ? kReactAll : 0u;
    // Initial grid size (host grid_box over the initial AABB) sizes the tables.
    float lo[3] = {pos[0], pos[1], pos[2]}, hi[3] = {pos[0], pos[1], pos[2]};
    for (uint32_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], pos[3 * i + a]);
            hi[a] = fmaxf(hi[a], pos[3 * i + a]);
        }
    float gmin[3], bin;
    uint32_t dims[3];
    grid_box(lo, hi, sc.params.h, sc.params.max_bins, gmin, bin, dims);
    sc.cells_hint = dims[0] * dims[1] * dims[2];
    return sc;
}

uint32_t Engine::add_prepared(SceneSlot &&sc) {
    uint64_t slots = round_up(sc.params.n, kBlock);
    for (const auto &s : scenes_)
        if (!s.frozen) slots += round_up(s.params.n, kBlock);  // retired scenes leave at the next layout
    if (slots > kMaxPoolSlots - kBlock) throw CapacityError("a pool holds at most 2^27 particle slots");
    sc.id = next_id_++;
    sc.resident = false;
    sc.frozen = false;
    scenes_.push_back(std::move(sc));
    dirty_ = true;
    return scenes_.back().id;
}

uint32_t Engine::add_scene(const SceneConfig &cfg, uint32_t n, uint32_t nf, const float *pos, const float *vel,
                           const float *pmass, const Span *spans, uint32_t nspans, bool reactions) {
    return add_prepared(prepare_scene(cfg, n, nf, pos, vel, pmass, spans, nspans, reactions));
}

SceneSlot Engine::prepare_coupled(const SceneConfig &cfg, uint32_t nf, const float *fpos, const float *fvel,
                                  const std::vector<BoxBody> &bodies, const std::vector<float> &density,
                                  const RigidConfig &rc) {
    if (density.size() != bodies.size()) throw std::invalid_argument("one density per body");
    std::vector<float> pos(fpos, fpos + 3ull * nf), vel, pmass(nf, 0.0f);
    if (fvel)
        vel.assign(fvel, fvel + 3ull * nf);
    else
        vel.assign(3ull * nf, 0.0f);
    const SceneParams probe = scene_params(cfg, nf ? nf : 1, nf);
    for (uint32_t i = 0; i < nf; ++i) pmass[i] = probe.mass;  // FluidSystem::mass (explicit or calibrated)
    std::vector<Span> spans;
    append_boundary_boxes(cfg, bodies, pos, vel, pmass, spans);
    std::vector<SceneSlot::Rigid> rigid;
    const float h = fmaxf(cfg.smoothing_radius, 1e-5f);
    for (size_t b = 0; b < bodies.size(); ++b) {
        if (!(density[b] > 0.0f)) continue;
        const BoxBody &bd = bodies[b];
        // append_boundary_boxes emits one span per body with particles, in order.
        uint32_t span_local = 0;
        bool found = false;
        for (uint32_t k = 0; k < spans.size(); ++k)
            if (spans[k].body == bd.id) {
                if (found) throw std::invalid_argument("body ids must be unique");
                span_local = k;
                found = true;
            }
        if (!found) continue;  // no particles: nothing couples
        SceneSlot::Rigid r;
        r.span_local = span_local;
        r.body_id = bd.id;
        r.density = density[b];
        uint32_t n0 = 0;
        float vol = 0.0f;
        box_lattice(bd.half, cfg.spacing, h, cfg.boundary_layers, r.lat, n0, vol);
        RigidDev &d = r.dev;
        for (int a = 0; a < 3; ++a) {
            d.pos[a] = bd.pos[a];
            d.linvel[a] = bd.linvel[a];
            d.angvel[a] = bd.angvel[a];
            d.half[a] = bd.half[a];
            d.gravity[a] = rc.gravity[a];
            d.tank[a] = rc.tank[a];
        }
        for (int a = 0; a < 4; ++a) d.rot[a] = bd.rot[a];
        box_mass_props(bd.half, density[b], d.inv_mass, d.inv_inertia);
        d.dt_tick = cfg.dt_tick;
        d.max_linear = rc.max_linear;
        d.max_angular = rc.max_angular;
        d.substeps = std::max<uint32_t>(rc.substeps, 1u);
        d.gstart = spans[span_local].start;
        d.gend = spans[span_local].end;
        if ((size_t)(d.gend - d.gstart) * 3 != r.lat.size()) throw std::logic_error("body lattice mismatch");
        rigid.push_back(std::move(r));
    }
    const uint32_t n = (uint32_t)pmass.size();
    SceneSlot sc = prepare_scene(cfg, n, nf, pos.data(), vel.data(), pmass.data(), spans.data(),
                                 (uint32_t)spans.size(), !rigid.empty());
    // Only the dynamic bodies consume a reaction; static bodies' spans are skipped.
    if (!rigid.empty()) sc.params.react = kReactDynamic;
    sc.rigid = std::move(rigid);
    for (size_t b = 0; b < bodies.size(); ++b)
        if (!(density[b] > 0.0f)) sc.statics.push_back(bodies[b]);
    return sc;
}

uint32_t Engine::add_coupled_scene(const SceneConfig &cfg, uint32_t nf, const float *fpos, const float *fvel,
                                   const std::vector<BoxBody> &bodies, const std::vector<float> &density,
                                   const RigidConfig &rc) {
    return add_prepared(prepare_coupled(cfg, nf, fpos, fvel, bodies, density, rc));
}

std::vector<RigidDev> Engine::read_bodies(uint32_t id) {
    if (dirty_) layout();
    sync();
    const SceneSlot &sc = scene(id);
    std::vector<RigidDev> out(sc.rigid.size());
    if (!out.empty())
        hip_check(hipMemcpy(out.data(), d_rigid_ + sc.params.rbody0, sizeof(RigidDev) * out.size(),
                            hipMemcpyDeviceToHost),
                  "read bodies");
    return out;
}

void Engine::remove_scene(uint32_t id) {
    auto it = std::find_if(scenes_.begin(), scenes_.end(), [&](const SceneSlot &s) { return s.id == id; });
    if (it == scenes_.end()) throw std::out_of_range("unknown scene");
    // Only the slot map changes; device state of the remaining scenes is moved
    // by the next layout. Kernels may still read the scene's wall tables.
    hip_check(hipStreamSynchronize(stream_), "sync before remove");
    free_wall_tables(*it);
    scenes_.erase(it);
    dirty_ = true;
}

bool Engine::has_scene(uint32_t id) const {
    return std::any_of(scenes_.begin(), scenes_.end(), [&](const SceneSlot &s) { return s.id == id; });
}

const SceneSlot &Engine::scene(uint32_t id) const {
    for (const auto &s : scenes_)
        if (s.id == id) return s;
    throw std::out_of_range("unknown scene");
}

void Engine::alloc_pool(uint32_t slots) {
    slots_ = slots;
    auto U = [&](uint32_t *&p) {
        p = dalloc<uint32_t>(slots);
        pool_allocs_.push_back(p);
    };
    auto F4 = [&](float4 *&p) {
        p = dalloc<float4>(slots);
        pool_allocs_.push_back(p);
    };
    F4(P_.cA); F4(P_.cV); U(P_.cg);
    F4(spare_cA_); F4(spare_cV_); U(spare_cg_);
    P_.sA = dalloc<float4>((size_t)slots + kReadPad);
    pool_allocs_.push_back(P_.sA);
    U(P_.sg);
    P_.part = dalloc<SceneBox>(slots / kPartSlots + 1);
    pool_allocs_.push_back(P_.part);
    P_.fr = dalloc<float4>(2ull * slots);
    pool_allocs_.push_back(P_.fr);
    U(P_.key); U(P_.slot);
    P_.items = dalloc<uint2>(slots);
    pool_allocs_.push_back(P_.items);
    P_.nbm = dalloc<uint32_t>(9ull * slots);
    pool_allocs_.push_back(P_.nbm);
    U(P_.nbs);
    if (diagnostics_) {
        U(P_.ncount);
        F4(P_.acc); F4(P_.xsph);
    } else {
        P_.ncount = nullptr;
        P_.acc = P_.xsph = nullptr;
    }
    if (want_reaction_buffers_) {
        F4(P_.bf); F4(P_.bp);
        U(P_.rlist);
    } else {
        P_.bf = P_.bp = nullptr;
        P_.rlist = nullptr;
    }
}

void Engine::alloc_cells(uint32_t ccap) {
    ccap_ = ccap;
    P_.counts = dalloc<uint32_t>(ccap);
    pool_allocs_.push_back(P_.counts);
    P_.start = dalloc<uint32_t>((size_t)ccap + 1);
    pool_allocs_.push_back(P_.start);
    P_.fcount = dalloc<uint32_t>(ccap);
    pool_allocs_.push_back(P_.fcount);
    hip_check(hipMemset(P_.counts, 0, sizeof(uint32_t) * ccap), "memset counts");
    tiles_cap_ = (ccap + kScanMinTile - 1) / kScanMinTile + 1;
    d_tile_sums_ = dalloc<uint32_t>(tiles_cap_);
    P_.ccap = ccap;
}

void Engine::free_cells() {
    void *cell_arrays[3] = {P_.counts, P_.start, P_.fcount};
    pool_allocs_.erase(std::remove_if(pool_allocs_.begin(), pool_allocs_.end(),
                                      [&](void *p) { return p == cell_arrays[0] || p == cell_arrays[1] ||
                                                            p == cell_arrays[2]; }),
                       pool_allocs_.end());
    for (void *p : cell_arrays) dfree(p);
    dfree(d_tile_sums_);
    P_.counts = P_.start = P_.fcount = nullptr;
    d_tile_sums_ = nullptr;
}

void *Engine::staging(size_t bytes) {
    if (bytes > staging_bytes_) {
        dfree(staging_);
        staging_ = nullptr;
        staging_bytes_ = 0;
        staging_ = dalloc<char>(bytes);
        staging_bytes_ = bytes;
    }
    return staging_;
}

// Assign slot ranges; grow the pool geometrically when it is too small; move
// resident scenes device-to-device (into the spare state arrays, which then
// become current); upload new scenes; rebuild the small scene tables.
void Engine::layout() {
    using LClock = std::chrono::steady_clock;
    const auto t_begin = LClock::now();
    auto lap = [&](double &acc, LClock::time_point &t) {
        const auto now = LClock::now();
        acc += std::chrono::duration<double, std::milli>(now - t).count();
        t = now;
    };
    auto t_phase = t_begin;
    hip_check(hipSetDevice(device_), "hipSetDevice");
    hip_check(hipStreamSynchronize(stream_), "sync before layout");
    lap(layout_phase_ms_[0], t_phase);
    // Finish (replay) any abandoned substep of the old layout first: the reset
    // of the control block below would otherwise drop it.
    if (P_.nblocks) check_error_state();
    // The tick graph's launches take the pool by value: it is captured again
    // for the new layout and, when the launch sequence is unchanged, updated
    // in place instead of instantiated again (build_graph).
    graph_stale_ = gexec_ != nullptr;
    // Retired scenes leave the pool (their readbacks completed with the sync above).
    for (size_t k = 0; k < scenes_.size();) {
        if (scenes_[k].frozen) {
            free_wall_tables(scenes_[k]);
            scenes_.erase(scenes_.begin() + (long)k);
        } else {
            ++k;
        }
    }
    // Dynamic bodies of resident scenes continue from their device state.
    if (P_.nrigid) {
        std::vector<RigidDev> cur(P_.nrigid);
        hip_check(hipMemcpy(cur.data(), d_rigid_, sizeof(RigidDev) * P_.nrigid, hipMemcpyDeviceToHost),
                  "rigid readback");
        for (auto &sc : scenes_) {
            if (!sc.resident) continue;
            for (size_t k = 0; k < sc.rigid.size(); ++k) sc.rigid[k].dev = cur[sc.params.rbody0 + k];
        }
    }
    const std::vector<uint32_t> old_base = [&] {
        std::vector<uint32_t> v;
        for (auto &sc : scenes_) v.push_back(sc.dev_base);
        return v;
    }();
    uint64_t total = 0;
    for (const auto &sc : scenes_) total += round_up(sc.params.n, kBlock);
    if (total > kMaxPoolSlots - kBlock) throw CapacityError("a pool holds at most 2^27 particle slots");
    uint32_t used = 0;
    uint64_t cells = 0;
    std::vector<uint32_t> block_scene;
    bool reactions = false;
    for (uint32_t k = 0; k < scenes_.size(); ++k) {
        SceneParams &p = scenes_[k].params;
        p.pbase = used;
        p.alloc = round_up(p.n, kBlock);
        used += p.alloc;
        for (uint32_t b = 0; b < p.alloc / kBlock; ++b) block_scene.push_back(k);
        cells += cell_reserve(scenes_[k].cells_hint, p.max_bins, full_cells_);
        reactions = reactions || scenes_[k].want_reactions;
    }
    used_slots_ = used;
    const uint32_t need = std::max<uint32_t>(used, kBlock);
    const bool grow = need > slots_ || (reactions && !want_reaction_buffers_) || diag_alloc_ != diagnostics_;
    float4 *srcA = P_.cA, *srcV = P_.cV;
    uint32_t *srcG = P_.cg;
    std::vector<void *> old_allocs;
    if (grow) {
        // New pool; the old one stays alive until the resident state is copied.
        old_allocs.swap(pool_allocs_);
        const uint32_t ccap_keep = ccap_;
        const uint64_t new_cells = std::min<uint64_t>(std::max<uint64_t>({cells, (uint64_t)ccap_keep, 4096ull}),
                                                      (uint64_t)kMaxCells);
        const bool react = want_reaction_buffers_ || reactions;
        const uint32_t cap = grown_capacity(need, new_cells, react);
        want_reaction_buffers_ = react;
        diag_alloc_ = diagnostics_;
        alloc_pool(cap);
        ccap_ = 0;
        dfree(d_tile_sums_);
        d_tile_sums_ = nullptr;
        alloc_cells((uint32_t)new_cells);
    } else if (cells > ccap_) {
        const uint32_t c = grown_cells(cells);
        free_cells();
        alloc_cells(c);
    }
    // Move resident scenes: old position (src arrays) -> new position (dst),
    // all of them in one launch.
    float4 *dstA = grow ? P_.cA : spare_cA_, *dstV = grow ? P_.cV : spare_cV_;
    uint32_t *dstG = grow ? P_.cg : spare_cg_;
    {
        std::vector<uint32_t> moves;  // (from, to, first index of this move) per resident scene
        uint32_t total_moved = 0;
        for (uint32_t k = 0; k < scenes_.size(); ++k) {
            const SceneSlot &sc = scenes_[k];
            if (!sc.resident) continue;
            moves.push_back(old_base[k]);
            moves.push_back(sc.params.pbase);
            moves.push_back(total_moved);
            total_moved += sc.params.n;
        }
        if (total_moved) {
            uint32_t *d_moves = dalloc<uint32_t>(moves.size() + 3);
            moves.push_back(0);
            moves.push_back(0);
            moves.push_back(total_moved);  // sentinel: the end of the last move
            hip_check(hipMemcpy(d_moves, moves.data(), 4 * moves.size(), hipMemcpyHostToDevice), "moves");
            const uint32_t nmoves = (uint32_t)(moves.size() / 3 - 1);
            const uint32_t grid = std::min<uint32_t>((total_moved + kBlock - 1) / kBlock, 16384u);
            hipLaunchKernelGGL(k_move_records, dim3(grid), dim3(kBlock), 0, stream_, (const float4 *)srcA,
                               (const float4 *)srcV, (const uint32_t *)srcG, dstA, dstV, dstG,
                               (const uint32_t *)d_moves, nmoves, total_moved);
            hip_check(hipGetLastError(), "k_move_records");
            hip_check(hipStreamSynchronize(stream_), "move sync");
            dfree(d_moves);
        }
    }
    lap(layout_phase_ms_[1], t_phase);
    if (!grow) {
        std::swap(P_.cA, spare_cA_);
        std::swap(P_.cV, spare_cV_);
        std::swap(P_.cg, spare_cg_);
    }
    for (void *p : old_allocs) dfree(p);
    // New scenes: index order (records packed when the scene was added); the
    // first grid build sorts them.
    for (auto &sc : scenes_) {
        if (sc.resident) continue;
        const uint32_t n = sc.params.n;
        hip_check(hipMemcpyAsync(P_.cA + sc.params.pbase, sc.stage_a.data(), 16ull * n, hipMemcpyHostToDevice, stream_),
                  "A");
        hip_check(hipMemcpyAsync(P_.cV + sc.params.pbase, sc.stage_v.data(), 16ull * n, hipMemcpyHostToDevice, stream_),
                  "V");
        hipLaunchKernelGGL(k_iota, dim3(std::min<uint32_t>((n + kBlock - 1) / kBlock, 4096u)), dim3(kBlock), 0, stream_,
                           P_.cg + sc.params.pbase, n);
        hip_check(hipGetLastError(), "k_iota");
    }
    hip_check(hipStreamSynchronize(stream_), "upload sync");
    for (auto &sc : scenes_) {
        if (sc.resident) continue;
        sc.resident = true;
        std::vector<float4>().swap(sc.stage_a);
        std::vector<float4>().swap(sc.stage_v);
    }
    for (auto &sc : scenes_) sc.dev_base = sc.params.pbase;
    lap(layout_phase_ms_[2], t_phase);
    // Scene tables.
    dfree(d_params_);
    dfree(d_grid_);
    dfree(d_box_);
    dfree(d_block_scene_);
    dfree(d_spans_);
    dfree(d_react_);
    d_spans_ = nullptr;
    d_react_ = nullptr;
    d_params_ = dalloc<SceneParams>(std::max<size_t>(scenes_.size(), 1));
    d_grid_ = dalloc<SceneGrid>(std::max<size_t>(scenes_.size(), 1));
    // No grid until the first substep of this layout (bin = 0).
    hip_check(hipMemset(d_grid_, 0, sizeof(SceneGrid) * std::max<size_t>(scenes_.size(), 1)), "memset grid");
    d_box_ = dalloc<SceneBox>(2 * std::max<size_t>(scenes_.size(), 1));
    d_block_scene_ = dalloc<uint32_t>(std::max<size_t>(block_scene.size(), 1));
    std::vector<SpanDev> spans;
    span_first_.assign(scenes_.size(), 0);
    for (uint32_t k = 0; k < scenes_.size(); ++k) {
        span_first_[k] = (uint32_t)spans.size();
        if (!scenes_[k].want_reactions) continue;
        for (const Span &s : scenes_[k].spans) {
            SpanDev d{};
            d.scene = k;
            d.start = scenes_[k].params.pbase + s.start;
            d.end = scenes_[k].params.pbase + s.end;
            d.substeps = scenes_[k].substeps;
            for (int a = 0; a < 3; ++a) d.origin[a] = s.origin[a];
            d.active = scenes_[k].params.react == kReactAll ? 1u : 0u;
            spans.push_back(d);
        }
        // A dynamic body's torque origin is its current position.
        for (const auto &r : scenes_[k].rigid) {
            SpanDev &d = spans[span_first_[k] + r.span_local];
            for (int a = 0; a < 3; ++a) d.origin[a] = r.dev.pos[a];
            d.active = 1u;
        }
    }
    // Rigid table (scene order) and the bodies' body-frame lattices.
    std::vector<RigidDev> rig;
    std::vector<float4> lat;
    for (uint32_t k = 0; k < scenes_.size(); ++k) {
        SceneSlot &sc = scenes_[k];
        sc.params.rbody0 = (uint32_t)rig.size();
        sc.params.nrbody = (uint32_t)sc.rigid.size();
        for (auto &r : sc.rigid) {
            r.dev.scene = k;
            r.dev.span = span_first_[k] + r.span_local;
            r.dev.lat = (uint32_t)lat.size();
            rig.push_back(r.dev);
            for (size_t q = 0; q + 2 < r.lat.size(); q += 3)
                lat.push_back(make_float4(r.lat[q], r.lat[q + 1], r.lat[q + 2], 0.0f));
        }
    }
    dfree(d_rigid_);
    dfree(d_rlat_);
    d_rigid_ = nullptr;
    d_rlat_ = nullptr;
    if (!rig.empty()) {
        d_rigid_ = dalloc<RigidDev>(rig.size());
        hip_check(hipMemcpy(d_rigid_, rig.data(), sizeof(RigidDev) * rig.size(), hipMemcpyHostToDevice), "rigid");
        d_rlat_ = dalloc<float4>(lat.size());
        hip_check(hipMemcpy(d_rlat_, lat.data(), sizeof(float4) * lat.size(), hipMemcpyHostToDevice), "lattice");
    }
    P_.rigid = d_rigid_;
    P_.nrigid = (uint32_t)rig.size();
    P_.rlat = d_rlat_;
    nspans_ = (uint32_t)spans.size();
    if (nspans_) {
        d_spans_ = dalloc<SpanDev>(nspans_);
        hip_check(hipMemcpy(d_spans_, spans.data(), sizeof(SpanDev) * nspans_, hipMemcpyHostToDevice), "spans");
        d_react_ = dalloc<float>(6ull * nspans_);
        hip_check(hipMemset(d_react_, 0, 24ull * nspans_), "react");
    }
    P_.nspans = nspans_;
    P_.spans = d_spans_;
    P_.react = d_react_;
    // Provider walls: scene table, per-slot arrays, device tables.
    free_walls();
    nwall_ = 0;
    wall_halfspace_ = false;
    wall_table_ = false;
    std::vector<WallSceneDev> walls;
    for (auto &sc : scenes_) {
        for (uint32_t side = 0; side < 2 && sc.wall_mode == kWallsTable; ++side) upload_wall_table(sc, side);
        walls.push_back(wall_scene(sc));
        nwall_ += sc.wall_mode != kWallsOff ? 1u : 0u;
        wall_halfspace_ = wall_halfspace_ || sc.wall_mode == kWallsHalfspace;
        wall_table_ = wall_table_ || sc.wall_mode == kWallsTable;
    }
    if (nwall_) {
        d_walls_ = dalloc<WallSceneDev>(walls.size());
        hip_check(hipMemcpy(d_walls_, walls.data(), sizeof(WallSceneDev) * walls.size(), hipMemcpyHostToDevice),
                  "walls");
        P_.went = dalloc<uint32_t>(2ull * slots_);
        P_.wsep = dalloc<uint64_t>(2ull * slots_);
        P_.wn = dalloc<float4>(slots_);
        P_.wband = dalloc<uint32_t>(slots_);
        P_.wband_n = dalloc<uint32_t>(1);
        hip_check(hipMemset(P_.wband_n, 0, 4), "band counter");
        P_.sep = dalloc<SepState>(walls.size() + 1);
        hip_check(hipMemset(P_.sep, 0, sizeof(SepState) * (walls.size() + 1)), "separation state");
        P_.wslots = slots_;
        if (wall_halfspace_) {
            P_.fpos = dalloc<float4>(2ull * slots_);
            P_.hhead = dalloc<uint32_t>(6ull * slots_);
            P_.hnext = dalloc<uint32_t>(3ull * slots_);
            hip_check(hipMemset(P_.hhead, 0xff, 24ull * slots_), "separation buckets");
        }
    }
    P_.walls = d_walls_;
    std::vector<SceneParams> hp;
    for (auto &sc : scenes_) hp.push_back(sc.params);
    if (!hp.empty())
        hip_check(hipMemcpy(d_params_, hp.data(), sizeof(SceneParams) * hp.size(), hipMemcpyHostToDevice), "params");
    if (!block_scene.empty())
        hip_check(hipMemcpy(d_block_scene_, block_scene.data(), 4 * block_scene.size(), hipMemcpyHostToDevice),
                  "block_scene");
    P_.params = d_params_;
    P_.grid = d_grid_;
    P_.box = d_box_;
    P_.ctl = d_ctl_;
    P_.block_scene = d_block_scene_;
    P_.nblocks = used / kBlock;
    P_.nscenes = (uint32_t)scenes_.size();
    // Control block + AABB of the current state into box[step & 1].
    std::memset(h_ctl_, 0, sizeof(Control));
    hip_check(hipMemcpy(d_ctl_, h_ctl_, sizeof(Control), hipMemcpyHostToDevice), "ctl");
    ctl_step_base_ = 0;
    std::vector<SceneBox> init(2 * std::max<size_t>(scenes_.size(), 1));
    for (auto &b : init) {
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = 0xffffffffu;
            b.hi[a] = 0u;
        }
    }
    hip_check(hipMemcpy(d_box_, init.data(), sizeof(SceneBox) * init.size(), hipMemcpyHostToDevice), "box");
    if (P_.nblocks) {
        hipLaunchKernelGGL(k_box, dim3(P_.nblocks), dim3(kBlock), 0, stream_, P_);
        hipLaunchKernelGGL(k_box_reduce, dim3(box_reduce_grid()), dim3(kBlock), 0, stream_, P_);
        hip_check(hipGetLastError(), "k_box");
    }
    for (uint32_t k = 0; k < scenes_.size(); ++k)
        if (scenes_[k].wall_mode == kWallsTable) {
            index_wall_table(k, 0);
            index_wall_table(k, 1);
        }
    hip_check(hipStreamSynchronize(stream_), "layout sync");
    lap(layout_phase_ms_[3], t_phase);
    layout_ms_.push_back((float)std::chrono::duration<double, std::milli>(LClock::now() - t_begin).count());
    substeps_done_ = 0;  // substep index of the new layout (tick phase, replay)
    ++layouts_;
    dirty_ = false;
}

uint32_t Engine::grown_capacity(uint32_t need, uint64_t cells, bool reactions) const {
    // Grow geometrically (x1.5) so that a stream of admissions relayouts
    // rarely; under a memory limit, only as far as the old pool (alive while
    // the resident scenes move) and the new one fit together.
    const uint32_t lim = kMaxPoolSlots - kBlock;
    uint32_t cap = round_up(std::min<uint32_t>(std::max<uint32_t>(need, slots_ + slots_ / 2), lim), kBlock);
    const uint32_t exact = round_up(std::min<uint32_t>(need, lim), kBlock);
    if (mem_limit_ && cap > exact && bytes_ + pool_bytes(cap, cells, diagnostics_, reactions) > mem_limit_)
        cap = exact;
    return cap;
}

uint32_t Engine::grown_cells(uint64_t cells) const {
    // x1.5 headroom (fewer reallocations as scenes arrive) unless it would not
    // fit under the memory limit.
    const uint64_t exact = std::min<uint64_t>(cells, (uint64_t)kMaxCells);
    const uint64_t grown = std::min<uint64_t>(std::max<uint64_t>(cells, (uint64_t)ccap_ + ccap_ / 2), (uint64_t)kMaxCells);
    if (mem_limit_ && grown > exact) {
        const uint64_t cells_now = ccap_ ? 12ull * ccap_ : 0ull;
        if (bytes_ - cells_now + 12ull * grown > mem_limit_) return (uint32_t)exact;
    }
    return (uint32_t)grown;
}

uint64_t Engine::projected_peak(uint32_t n, uint32_t cells_hint, uint64_t max_bins, bool reactions) const {
    uint64_t need = round_up(n, kBlock), cells = cell_reserve(cells_hint, max_bins, full_cells_);
    bool react = reactions || want_reaction_buffers_;
    uint64_t maxn = n, nscenes = 1;
    for (const auto &sc : scenes_) {
        if (sc.frozen) continue;  // dropped by the next layout
        need += round_up(sc.params.n, kBlock);
        cells += cell_reserve(sc.cells_hint, sc.params.max_bins, full_cells_);
        react = react || sc.want_reactions;
        maxn = std::max<uint64_t>(maxn, sc.params.n);
        ++nscenes;
    }
    if (need > kMaxPoolSlots - kBlock) return ~0ull;
    // Everything allocated now except the readback staging, which grows to the
    // largest scene read back (24 bytes per particle); the scene tables the
    // layout rebuilds (params, grid, boxes, block map, spans, rigid bodies and
    // their lattices; freed before they are reallocated, bounded generously).
    const uint64_t base = bytes_ - staging_bytes_;
    const uint64_t stage = std::max<uint64_t>(staging_bytes_, 24ull * maxn);
    const uint64_t tables = table_allowance(nscenes, need, react);
    const uint64_t cells_now = ccap_ ? 12ull * ccap_ + 4ull * tiles_cap_ + 4ull : 0ull;
    const bool grow = need > slots_ || (react && !want_reaction_buffers_) || diag_alloc_ != diagnostics_;
    if (grow) {
        const uint64_t new_cells = std::min<uint64_t>(std::max<uint64_t>({cells, (uint64_t)ccap_, 4096ull}),
                                                      (uint64_t)kMaxCells);
        const uint32_t cap = grown_capacity((uint32_t)need, new_cells, react);
        // old pool and cells alive until the resident scenes have moved
        return base + pool_bytes(cap, new_cells, diagnostics_, react) + stage + tables;
    }
    if (cells > ccap_) {
        const uint64_t c = grown_cells(cells);
        return base - cells_now + 12ull * c + 4ull * ((c + kScanMinTile - 1) / kScanMinTile + 1) + 4ull + stage + tables;
    }
    return base + stage + tables;
}

uint64_t Engine::table_allowance(uint64_t nscenes, uint64_t slots, bool reactions) {
    return 256ull * nscenes + slots / 32ull + 4096ull + (reactions ? 16ull * slots + 4096ull : 0ull);
}

uint64_t Engine::solo_peak(uint32_t n, uint32_t cells_hint, uint64_t max_bins, bool reactions, bool full_cells) {
    // projected_peak of an engine holding nothing but its control block
    const uint64_t need = round_up(n, kBlock);
    const uint64_t cells = std::min<uint64_t>(std::max<uint64_t>(cell_reserve(cells_hint, max_bins, full_cells), 4096ull),
                                              (uint64_t)kMaxCells);
    return sizeof(Control) + pool_bytes(need, cells, false, reactions) + 24ull * n +
           table_allowance(1, need, reactions);
}

void Engine::release_if_empty() {
    for (const auto &sc : scenes_)
        if (!sc.frozen) return;
    hip_check(hipStreamSynchronize(stream_), "sync before release");
    for (auto &sc : scenes_) free_wall_tables(sc);
    scenes_.clear();
    hip_check(hipStreamSynchronize(stream_), "sync before release");
    invalidate_graph();
    free_all();
    want_reaction_buffers_ = false;
    diag_alloc_ = false;
    used_slots_ = 0;
    P_.nblocks = 0;
    dirty_ = true;
}

SceneSlot &Engine::slot_of(uint32_t id) {
    for (auto &s : scenes_)
        if (s.id == id) return s;
    throw std::out_of_range("unknown scene");
}

void Engine::free_walls() {
    dfree(d_walls_);
    dfree(P_.went);
    dfree(P_.wsep);
    dfree(P_.wn);
    dfree(P_.sep);
    dfree(P_.wband);
    dfree(P_.wband_n);
    P_.wband = P_.wband_n = nullptr;
    dfree(P_.fpos);
    dfree(P_.hhead);
    dfree(P_.hnext);
    d_walls_ = nullptr;
    P_.walls = nullptr;
    P_.went = nullptr;
    P_.wsep = nullptr;
    P_.wn = nullptr;
    P_.sep = nullptr;
    P_.fpos = nullptr;
    P_.hhead = nullptr;
    P_.hnext = nullptr;
    P_.wslots = 0;
}

void Engine::free_wall_tables(SceneSlot &sc) {
    for (auto &t : sc.wall_tab) {
        dfree(t.d_ids);
        dfree(t.d_start);
        dfree(t.d_planes);
        t.d_ids = t.d_start = nullptr;
        t.d_planes = nullptr;
        t.cap_e = t.cap_p = 0;
    }
}

// Device copy of one table (grown geometrically; the caller has synchronised
// the stream, so no kernel reads the old buffers).
void Engine::upload_wall_table(SceneSlot &sc, uint32_t side) {
    auto &t = sc.wall_tab[side];
    if (t.start.empty()) t.start.assign(1, 0u);
    const size_t ne = t.ids.size(), np = t.planes.size() / 2;
    if (ne > t.cap_e || !t.d_start) {
        dfree(t.d_ids);
        dfree(t.d_start);
        t.cap_e = std::max<size_t>(ne, t.cap_e + t.cap_e / 2);
        t.d_ids = dalloc<uint32_t>(t.cap_e);
        t.d_start = dalloc<uint32_t>(t.cap_e + 1);
    }
    if (np > t.cap_p || !t.d_planes) {
        dfree(t.d_planes);
        t.cap_p = std::max<size_t>(np, t.cap_p + t.cap_p / 2);
        t.d_planes = dalloc<float4>(2 * std::max<size_t>(t.cap_p, 1));
    }
    if (ne) hip_check(hipMemcpy(t.d_ids, t.ids.data(), 4 * ne, hipMemcpyHostToDevice), "wall ids");
    hip_check(hipMemcpy(t.d_start, t.start.data(), 4 * (ne + 1), hipMemcpyHostToDevice), "wall start");
    if (np) hip_check(hipMemcpy(t.d_planes, t.planes.data(), 32 * np, hipMemcpyHostToDevice), "wall planes");
}

WallSceneDev Engine::wall_scene(const SceneSlot &sc) const {
    WallSceneDev w;
    std::memset(&w, 0, sizeof(w));
    w.mode = sc.wall_mode;
    if (sc.wall_mode == kWallsHalfspace) {
        w.nprov = (uint32_t)sc.wall_prov.size();
        std::copy(sc.wall_prov.begin(), sc.wall_prov.end(), w.prov);
        std::copy(sc.wall_face.begin(), sc.wall_face.end(), w.face);
    } else if (sc.wall_mode == kWallsTable) {
        for (uint32_t side = 0; side < 2; ++side) {
            w.ne[side] = (uint32_t)sc.wall_tab[side].ids.size();
            w.tstart[side] = sc.wall_tab[side].d_start;
            w.tplane[side] = (const float *)sc.wall_tab[side].d_planes;
        }
    }
    return w;
}

void Engine::index_wall_table(uint32_t k, uint32_t side) {
    const SceneSlot &sc = scenes_[k];
    const uint32_t nf = sc.params.nf, ne = (uint32_t)sc.wall_tab[side].ids.size();
    const uint32_t gc = std::max<uint32_t>(1, std::min<uint32_t>((nf + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_wall_index_clear, dim3(gc), dim3(kBlock), 0, stream_, P_, k, side);
    if (ne) {
        const uint32_t gs = std::min<uint32_t>((ne + kBlock - 1) / kBlock, 4096);
        hipLaunchKernelGGL(k_wall_index_set, dim3(gs), dim3(kBlock), 0, stream_, P_, k, side,
                           (const uint32_t *)sc.wall_tab[side].d_ids, ne);
    }
    hip_check(hipGetLastError(), "wall index");
}

void Engine::set_halfspace_walls(uint32_t id, const std::vector<WallProvider> &providers) {
    SceneSlot &sc = slot_of(id);
    if (providers.size() > kWallMaxProviders) throw std::invalid_argument("at most 8 wall providers per scene");
    std::vector<WallProvDev> prov;
    std::vector<WallFaceDev> face;
    for (const auto &p : providers) {
        WallProvDev d{};
        d.has_bounds = p.has_bounds ? 1u : 0u;
        for (int a = 0; a < 3; ++a) {
            d.lo[a] = p.lo[a];
            d.hi[a] = p.hi[a];
        }
        d.face0 = (uint32_t)face.size();
        d.nface = (uint32_t)p.faces.size();
        for (const auto &f : p.faces) {
            if (f.axis > 2) throw std::invalid_argument("wall face axis must be 0, 1 or 2");
            if (f.sign != 1.0f && f.sign != -1.0f) throw std::invalid_argument("wall face sign must be +1 or -1");
            if (!std::isfinite(f.plane)) throw std::invalid_argument("wall face plane must be finite");
            face.push_back(WallFaceDev{f.axis, f.plane, f.sign, 0u});
        }
        prov.push_back(d);
    }
    if (face.size() > kWallMaxFaces) throw std::invalid_argument("at most 16 wall faces per scene");
    hip_check(hipStreamSynchronize(stream_), "sync before walls");
    free_wall_tables(sc);
    for (auto &t : sc.wall_tab) t = SceneSlot::WallTable{};
    sc.wall_prov = std::move(prov);
    sc.wall_face = std::move(face);
    sc.wall_mode = sc.wall_prov.empty() ? kWallsOff : kWallsHalfspace;
    dirty_ = true;
}

void Engine::set_wall_tables(uint32_t id, uint32_t side, const uint32_t *ids, uint32_t ne, const uint32_t *start,
                             const float *planes, uint32_t np) {
    SceneSlot &sc = slot_of(id);
    if (side > 1) throw std::invalid_argument("wall table side must be 0 (mirror) or 1 (project)");
    if (ne && (!ids || !start)) throw std::invalid_argument("null wall table");
    if (np && !planes) throw std::invalid_argument("null wall planes");
    if ((start ? start[0] : 0u) != 0u || (start ? start[ne] : 0u) != np)
        throw std::invalid_argument("wall table start must run from 0 to the plane count");
    for (uint32_t e = 0; e < ne; ++e) {
        if (ids[e] >= sc.params.nf) throw std::invalid_argument("wall table id is not a fluid index");
        if (e && ids[e] <= ids[e - 1]) throw std::invalid_argument("wall table ids must be strictly ascending");
        if (start[e + 1] < start[e]) throw std::invalid_argument("wall table start must not decrease");
        if (start[e + 1] - start[e] > kWallMirrorPlanes)
            throw std::invalid_argument("at most 8 planes per wall table entry (gather_in keeps 8)");
    }
    for (size_t q = 0; q < 6ull * np; ++q)
        if (!std::isfinite(planes[q])) throw std::invalid_argument("wall planes must be finite");
    SceneSlot::WallTable t;
    t.ids.assign(ids, ids + ne);
    t.start.assign(1, 0u);
    if (ne) t.start.assign(start, start + ne + 1);
    t.planes.resize(2ull * np);
    for (uint32_t q = 0; q < np; ++q) {
        const float *o = planes + 6ull * q;
        t.planes[2 * q] = make_float4(o[0], o[1], o[2], 0.0f);
        t.planes[2 * q + 1] = make_float4(o[3], o[4], o[5], 0.0f);
    }
    hip_check(hipStreamSynchronize(stream_), "sync before wall table");
    auto &dst = sc.wall_tab[side];
    std::swap(dst.ids, t.ids);
    std::swap(dst.start, t.start);
    std::swap(dst.planes, t.planes);
    if (sc.wall_mode != kWallsTable) {
        sc.wall_prov.clear();
        sc.wall_face.clear();
        sc.wall_mode = kWallsTable;
        dirty_ = true;  // the layout uploads and indexes both tables
        return;
    }
    if (dirty_ || !sc.resident || !d_walls_) return;
    upload_wall_table(sc, side);
    const uint32_t k = (uint32_t)(&sc - scenes_.data());
    const WallSceneDev w = wall_scene(sc);
    hip_check(hipMemcpy(d_walls_ + k, &w, sizeof(w), hipMemcpyHostToDevice), "wall scene");
    index_wall_table(k, side);
}

void Engine::clear_walls(uint32_t id) {
    SceneSlot &sc = slot_of(id);
    hip_check(hipStreamSynchronize(stream_), "sync before walls");
    free_wall_tables(sc);
    for (auto &t : sc.wall_tab) t = SceneSlot::WallTable{};
    sc.wall_prov.clear();
    sc.wall_face.clear();
    if (sc.wall_mode != kWallsOff) dirty_ = true;
    sc.wall_mode = kWallsOff;
}

namespace {

// Template instance of a tuned kernel for its runtime knobs.
using KernelFn = void (*)(Pool);
KernelFn density_kernel(uint32_t bs, uint32_t stage) {
    switch (bs * 1000u + stage) {
        case 64128: return k_density<64, 128>;
        case 64256: return k_density<64, 256>;
        case 64512: return k_density<64, 512>;
        case 128128: return k_density<128, 128>;
        case 128256: return k_density<128, 256>;
        case 128512: return k_density<128, 512>;
        case 256128: return k_density<256, 128>;
        case 256512: return k_density<256, 512>;
        default: return k_density<256, 256>;
    }
}
KernelFn density_walls_kernel(uint32_t bs) {
    return bs == 64 ? k_density_walls<64> : bs == 128 ? k_density_walls<128> : k_density_walls<256>;
}
KernelFn force_kernel(uint32_t variant, uint32_t bs) {
    const uint32_t w = variant == Engine::kForceScan        ? 0u
                       : variant == Engine::kForceWalkAhead ? 2u
                       : variant == Engine::kForcePairs     ? 3u
                       : variant == Engine::kForceQueue     ? 4u
                       : variant == Engine::kForcePaired    ? 5u
                       : variant == Engine::kForceStaged    ? 6u
                                                            : 1u;
    switch (w * 1000u + bs) {
        case 6064: return k_force_t<6, 64>;
        case 6128: return k_force_t<6, 128>;
        case 6256: return k_force_t<6, 256>;
        case 5064: return k_force_t<5, 64>;
        case 5128: return k_force_t<5, 128>;
        case 5256: return k_force_t<5, 256>;
        case 4064: return k_force_t<4, 64>;
        case 4128: return k_force_t<4, 128>;
        case 4256: return k_force_t<4, 256>;
        case 3064: return k_force_t<3, 64>;
        case 3128: return k_force_t<3, 128>;
        case 3256: return k_force_t<3, 256>;
        case 64: return k_force_t<0, 64>;
        case 128: return k_force_t<0, 128>;
        case 256: return k_force_t<0, 256>;
        case 1064: return k_force_t<1, 64>;
        case 1128: return k_force_t<1, 128>;
        case 2064: return k_force_t<2, 64>;
        case 2128: return k_force_t<2, 128>;
        case 2256: return k_force_t<2, 256>;
        default: return k_force_t<1, 256>;
    }
}

}  // namespace

void Engine::enqueue_substep(uint32_t tick_sub, uint32_t tick_len, const std::function<void(const char *)> *mark) {
    const uint32_t nb = P_.nblocks;
    const uint32_t grid_slots = tuning_.persist_blocks ? std::min(nb, tuning_.persist_blocks) : nb;
    // density and force: workgroups of block_threads slots
    const uint32_t nsub = nb * (kBlock / tuning_.block_threads);
    const uint32_t grid_nb = tuning_.persist_blocks ? std::min(nsub, tuning_.persist_blocks) : nsub;
    const uint32_t bs = tuning_.block_threads;
    const uint32_t grid_scan = std::min<uint32_t>(tiles_cap_, 4096);
    auto M = [&](const char *name) {
        if (mark) (*mark)(name);
    };
    hipLaunchKernelGGL(k_grid_box, dim3(1), dim3(1024), 0, stream_, P_, tick_sub, tick_len);
    M("grid_box");
    hipLaunchKernelGGL(k_keys, dim3((grid_slots + kSortBlocks - 1) / kSortBlocks), dim3(kBlock), 0, stream_, P_);
    M("keys");
    switch (tuning_.scan_tile_cells) {
        case 1024:
            hipLaunchKernelGGL(k_scan_reduce<4>, dim3(grid_scan), dim3(kBlock), 0, stream_, P_, d_tile_sums_);
            M("scan_reduce");
            hipLaunchKernelGGL(k_scan_tiles<4>, dim3(1), dim3(1024), 0, stream_, P_, d_tile_sums_);
            M("scan_tiles");
            hipLaunchKernelGGL(k_scan_apply<4>, dim3(grid_scan), dim3(kBlock), 0, stream_, P_,
                               (const uint32_t *)d_tile_sums_);
            break;
        case 4096:
            hipLaunchKernelGGL(k_scan_reduce<16>, dim3(grid_scan), dim3(kBlock), 0, stream_, P_, d_tile_sums_);
            M("scan_reduce");
            hipLaunchKernelGGL(k_scan_tiles<16>, dim3(1), dim3(1024), 0, stream_, P_, d_tile_sums_);
            M("scan_tiles");
            hipLaunchKernelGGL(k_scan_apply<16>, dim3(grid_scan), dim3(kBlock), 0, stream_, P_,
                               (const uint32_t *)d_tile_sums_);
            break;
        default:
            hipLaunchKernelGGL(k_scan_reduce<8>, dim3(grid_scan), dim3(kBlock), 0, stream_, P_, d_tile_sums_);
            M("scan_reduce");
            hipLaunchKernelGGL(k_scan_tiles<8>, dim3(1), dim3(1024), 0, stream_, P_, d_tile_sums_);
            M("scan_tiles");
            hipLaunchKernelGGL(k_scan_apply<8>, dim3(grid_scan), dim3(kBlock), 0, stream_, P_,
                               (const uint32_t *)d_tile_sums_);
            break;
    }
    M("scan_apply");
    hipLaunchKernelGGL(k_place, dim3((grid_slots + kSortBlocks - 1) / kSortBlocks), dim3(kBlock), 0, stream_, P_);
    M("place");
    hipLaunchKernelGGL(k_canon_gather, dim3(grid_slots), dim3(kBlock), 0, stream_, P_);
    M("canon_gather");
    if (wall_halfspace_)
        hipLaunchKernelGGL(density_walls_kernel(bs), dim3(grid_nb), dim3(bs), 0, stream_, P_);
    else
        hipLaunchKernelGGL(density_kernel(bs, tuning_.stage_records), dim3(grid_nb), dim3(bs), 0, stream_, P_);
    M("density");
    if (wall_table_) {
        hipLaunchKernelGGL(k_wall_band, dim3(grid_slots), dim3(kBlock), 0, stream_, P_);
        hipLaunchKernelGGL(k_wall_density, dim3(std::min<uint32_t>(nb, 2048)), dim3(kBlock), 0, stream_, P_);
        M("wall_density");
    }
    hipLaunchKernelGGL(force_kernel(tuning_.force_kernel, bs), dim3(grid_nb), dim3(bs), 0, stream_, P_);
    M("force");
    if (nwall_) {
        hipLaunchKernelGGL(k_wall_project, dim3(grid_slots), dim3(kBlock), 0, stream_, P_);
        M("wall_project");
        if (wall_halfspace_) {
            const uint32_t gs = std::min<uint32_t>(nb, 8192);
            hipLaunchKernelGGL(k_sep_begin, dim3(gs), dim3(kBlock), 0, stream_, P_);
            M("sep_begin");
            for (uint32_t m = 0; m < kSepIters; ++m) {
                hipLaunchKernelGGL(k_sep_round, dim3(gs), dim3(kBlock), 0, stream_, P_, m);
                M(m == 0 ? "sep_round0" : "sep_rounds");
            }
            hipLaunchKernelGGL(k_sep_finish, dim3(gs), dim3(kBlock), 0, stream_, P_);
            M("sep_finish");
            hipLaunchKernelGGL(k_wall_separate, dim3(P_.nscenes), dim3(kBlock), 0, stream_, P_);
            M("wall_separate");
        }
        hipLaunchKernelGGL(k_wall_box, dim3(grid_slots), dim3(kBlock), 0, stream_, P_);
        M("wall_box");
    }
    hipLaunchKernelGGL(k_box_reduce, dim3(box_reduce_grid()), dim3(kBlock), 0, stream_, P_);
    M("box_reduce");
    if (nspans_) {
        hipLaunchKernelGGL(k_bforce_list, dim3((nb + kListBlocks - 1) / kListBlocks), dim3(kBlock), 0, stream_, P_);
        hipLaunchKernelGGL(k_bforce, dim3(std::min<uint32_t>(grid_slots * (kBlock / kBforceThreads), 4096u)),
                           dim3(kBforceThreads), 0, stream_, P_);
        M("bforce");
        hipLaunchKernelGGL(k_reactions, dim3(nspans_), dim3(kReactThreads), 0, stream_, P_);
        M("reactions");
    }
    hip_check(hipGetLastError(), "substep launch");
}

void Engine::enqueue_at(uint64_t i, const std::function<void(const char *)> *mark, bool head) {
    const uint32_t sub = pool_substeps();
    auto M = [&](const char *name) {
        if (mark) (*mark)(name);
    };
    if (P_.nrigid && head && i % sub == 0) {
        const uint32_t gb = (P_.nrigid + 63) / 64;
        hipLaunchKernelGGL(k_rigid, dim3(gb), dim3(64), 0, stream_, P_);
        M("rigid");
        hipLaunchKernelGGL(k_body_refresh, dim3(P_.nblocks), dim3(kBlock), 0, stream_, P_);
        hipLaunchKernelGGL(k_box_reduce, dim3(box_reduce_grid()), dim3(kBlock), 0, stream_, P_);
        M("body_refresh");
    }
    enqueue_substep((uint32_t)(i % sub), sub, mark);
    if (P_.nrigid && i % sub == sub - 1) {
        hipLaunchKernelGGL(k_rigid_collect, dim3((P_.nrigid + 63) / 64), dim3(64), 0, stream_, P_);
        M("rigid");
    }
    hip_check(hipGetLastError(), "tick launch");
}

uint32_t Engine::pool_substeps() const {
    uint32_t sub = 1;
    for (auto &s : scenes_) sub = std::max(sub, s.substeps);
    return sub;
}

std::vector<std::pair<std::string, float>> Engine::profile(uint32_t substeps) {
    sync();
    if (scenes_.empty()) return {};
    std::vector<hipEvent_t> evs;
    std::vector<std::string> names;
    auto ev = [&]() {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventRecord(e, stream_), "hipEventRecord");
        evs.push_back(e);
    };
    std::function<void(const char *)> mark = [&](const char *name) {
        names.push_back(name);
        ev();
    };
    for (uint32_t k = 0; k < substeps; ++k) {
        ev();
        names.push_back("");
        enqueue_at(substeps_done_ + k, &mark);
    }
    substeps_done_ += substeps;
    hip_check(hipStreamSynchronize(stream_), "profile sync");
    std::vector<std::pair<std::string, float>> out;
    for (size_t k = 1; k < evs.size(); ++k) {
        if (names[k].empty()) continue;
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, evs[k - 1], evs[k]), "hipEventElapsedTime");
        auto it = std::find_if(out.begin(), out.end(), [&](const auto &p) { return p.first == names[k]; });
        if (it == out.end())
            out.push_back({names[k], ms});
        else
            it->second += ms;
    }
    for (auto e : evs) (void)hipEventDestroy(e);
    check_error_state();
    return out;
}

Engine::Tuning Engine::autotune(uint32_t substeps) {
    if (dirty_) layout();
    if (scenes_.empty() || P_.nblocks == 0) return tuning_;
    static std::mutex mu;
    static std::map<std::tuple<std::string, uint32_t, uint32_t>, Tuning> cache;
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
    uint32_t bucket = 0;
    for (uint32_t v = used_slots_; v > 1; v >>= 1) ++bucket;
    // The grid's coarseness is part of the key: cells of many particles (a bin
    // of several h under a small grid budget) favour other walks. It is read
    // from the largest scene's last grid (one substep first if none ran yet).
    uint32_t big = 0;
    for (uint32_t k = 1; k < scenes_.size(); ++k)
        if (scenes_[k].params.n > scenes_[big].params.n) big = k;
    SceneGrid gb{};
    hip_check(hipMemcpy(&gb, d_grid_ + big, sizeof(SceneGrid), hipMemcpyDeviceToHost), "grid");
    if (!(gb.bin > 0.0f)) {
        step(1);
        sync();
        hip_check(hipMemcpy(&gb, d_grid_ + big, sizeof(SceneGrid), hipMemcpyDeviceToHost), "grid");
    }
    const float hb = scenes_[big].params.h;
    const float ratio = hb > 0.0f ? gb.bin / hb : 1.0f;  // (h = 0: the grid rule's 1e-6 bin)
    const uint32_t coarse =
        ratio > 1.0f && ratio < 1e30f ? std::min<uint32_t>(7u, (uint32_t)std::lround(std::log2((double)ratio))) : 0u;
    const auto key = std::make_tuple(std::string(prop.gcnArchName), bucket, coarse);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            set_tuning(it->second);
            return tuning_;
        }
    }
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    const uint32_t n = std::max<uint32_t>(substeps, 1);
    // Device time of n real substeps with tuning t (every variant gives the
    // same results, so tuning never perturbs the trajectory).
    auto cost = [&](const Tuning &t) {
        set_tuning(t);
        (void)profile(1);  // first launch of a variant loads its code object: not timed
        float ms = 0.0f;
        for (const auto &kv : profile(n)) ms += kv.second;
        return ms;
    };
    // Coordinate descent: each knob in turn over its values with the others
    // held, until a pass changes nothing (each knob acts mostly on one
    // kernel, so the knobs hardly interact). A change must win by 1 % to
    // stand over run-to-run noise, against the current choice timed right
    // after it: the scene moves on while it is tuned, and early in a run its
    // cost per substep drifts by more than that.
    struct Axis {
        uint32_t Tuning::*field;
        std::vector<uint32_t> values;
    };
    const std::vector<Axis> axes = {
        {&Tuning::force_kernel,
         {(uint32_t)kForceScan, (uint32_t)kForceWalk, (uint32_t)kForceWalkAhead, (uint32_t)kForcePairs,
          (uint32_t)kForceQueue, (uint32_t)kForcePaired, (uint32_t)kForceStaged}},
        {&Tuning::block_threads, {64u, 128u, 256u}},
        {&Tuning::stage_records, {128u, 256u, 512u}},
        {&Tuning::scan_tile_cells, {1024u, 2048u, 4096u}},
        {&Tuning::persist_blocks, {0u, 8u * cus}},
    };
    Tuning best = tuning_;
    for (int pass = 0; pass < 3; ++pass) {
        bool changed = false;
        for (const Axis &ax : axes)
            for (uint32_t v : ax.values) {
                if (best.*(ax.field) == v) continue;
                Tuning t = best;
                t.*(ax.field) = v;
                const float ms = cost(t);
                const float ref = cost(best);
                if (ms < 0.99f * ref) {
                    best = t;
                    changed = true;
                }
            }
        if (!changed) break;
    }
    set_tuning(best);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = best;
    return best;
}

std::vector<float> Engine::step_timed(uint32_t ticks) {
    if (dirty_) layout();
    std::vector<hipEvent_t> evs(ticks + 1);
    for (auto &e : evs) hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipEventRecord(evs[0], stream_), "hipEventRecord");
    for (uint32_t k = 0; k < ticks; ++k) {
        step(1);
        hip_check(hipEventRecord(evs[k + 1], stream_), "hipEventRecord");
    }
    hip_check(hipStreamSynchronize(stream_), "step_timed sync");
    std::vector<float> ms(ticks);
    for (uint32_t k = 0; k < ticks; ++k) hip_check(hipEventElapsedTime(&ms[k], evs[k], evs[k + 1]), "elapsed");
    for (auto e : evs) (void)hipEventDestroy(e);
    check_error_state();
    return ms;
}

uint32_t Engine::box_reduce_grid() const {
    const uint32_t npart = P_.nblocks * (kBlock / kPartSlots);
    return std::max<uint32_t>(1u, std::min<uint32_t>((npart + kBoxReduceSpan - 1) / kBoxReduceSpan, 4096u));
}

void Engine::set_tuning(const Tuning &t) {
    Tuning u = t;
    const Tuning d{};
    if (u.force_kernel == 0) u.force_kernel = d.force_kernel;
    if (u.block_threads == 0) u.block_threads = d.block_threads;
    if (u.stage_records == 0) u.stage_records = d.stage_records;
    if (u.scan_tile_cells == 0) u.scan_tile_cells = d.scan_tile_cells;
    if (u.force_kernel > kForceStaged) throw std::invalid_argument("unknown kernel variant");
    if (u.block_threads != 64 && u.block_threads != 128 && u.block_threads != 256)
        throw std::invalid_argument("block_threads must be 64, 128 or 256");
    if (u.stage_records != 128 && u.stage_records != 256 && u.stage_records != 512)
        throw std::invalid_argument("stage_records must be 128, 256 or 512");
    if (u.scan_tile_cells != 1024 && u.scan_tile_cells != 2048 && u.scan_tile_cells != 4096)
        throw std::invalid_argument("scan_tile_cells must be 1024, 2048 or 4096");
    tuning_ = u;
    invalidate_graph();
}

// One whole tick (rigid leg, substeps) as a graph; launched at tick starts.
void Engine::build_graph() {
    const auto t0 = std::chrono::steady_clock::now();
    if (!graph_stale_) invalidate_graph();
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
    for (uint32_t i = 0; i < pool_substeps(); ++i) enqueue_at(i);
    hip_check(hipStreamEndCapture(stream_, &g), "end capture");
    bool updated = false;
    if (graph_stale_ && gexec_) {
        // Same launch sequence (kernels, order) with new arguments and grid
        // sizes: update the executable graph. Any other change (a scene with
        // walls, bodies or reactions joined or left) fails the update, and
        // the graph is instantiated afresh.
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(gexec_, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        (void)hipGetLastError();  // a refused update is not an error of the next call
    }
    if (updated) {
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_ = g;
        ++graph_updates_;
    } else {
        invalidate_graph();
        graph_ = g;
        hip_check(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0), "graph instantiate");
        ++graph_builds_;
    }
    graph_stale_ = false;
    layout_phase_ms_[4] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Engine::step(uint32_t ticks) {
    if (dirty_) layout();
    if (scenes_.empty() || P_.nblocks == 0) return;
    hip_check(hipSetDevice(device_), "hipSetDevice");
    // A pool tick is as many launches as the scene with the most substeps
    // takes; every scene steps in the last of them (substep_active).
    const uint32_t sub = pool_substeps();
    uint64_t n = (uint64_t)ticks * sub;
    // A tick left open by profile() is finished launch by launch.
    while (n && substeps_done_ % sub) {
        enqueue_at(substeps_done_++);
        --n;
    }
    if (use_graph_) {
        if (!gexec_ || graph_stale_) build_graph();
        for (uint64_t t = 0; t < n / sub; ++t) hip_check(hipGraphLaunch(gexec_, stream_), "graph launch");
        substeps_done_ += n;
    } else {
        for (uint64_t k = 0; k < n; ++k) enqueue_at(substeps_done_++);
    }
}

void Engine::check_error_state() {
    hip_check(hipMemcpyAsync(h_ctl_, d_ctl_, sizeof(Control), hipMemcpyDeviceToHost, stream_), "ctl read");
    hip_check(hipStreamSynchronize(stream_), "sync");
    if (h_ctl_->err & kErrGrid)
        throw std::domain_error("a scene's particle box is not finite (NaN or infinite positions)");
    if (h_ctl_->err & kErrCellCapacity) {
        // The failed substep and every later one left the state untouched;
        // grow the cell tables and replay the missing substeps.
        const uint64_t done = h_ctl_->step - ctl_step_base_;
        const uint64_t missing = substeps_done_ - done;
        const uint32_t new_cap = (uint32_t)std::min<uint64_t>((uint64_t)ccap_ * 2ull, (uint64_t)kMaxCells);
        if (new_cap <= ccap_) throw CapacityError("cell table capacity exhausted");
        invalidate_graph();
        free_cells();
        alloc_cells(new_cap);
        h_ctl_->err = 0;
        hip_check(hipMemcpy(&d_ctl_->err, &h_ctl_->err, sizeof(uint32_t), hipMemcpyHostToDevice), "ctl reset");
        // Replay launch by launch. The rigid leg of the tick holding the failed
        // substep already ran (it precedes the failure); later ticks' legs and
        // the failed tick's closing collect returned early and run again.
        substeps_done_ = done;
        for (uint64_t k = 0; k < missing; ++k) enqueue_at(done + k, nullptr, k != 0);
        substeps_done_ += missing;
        check_error_state();
    }
}

void Engine::sync() {
    if (dirty_) layout();
    check_error_state();
}

void Engine::set_diagnostics(bool on) {
    if (on == diagnostics_) return;
    // Diagnostics arrays are part of the pool: force a relayout.
    diagnostics_ = on;
    dirty_ = true;
}

void Engine::upload_boundary(uint32_t id, uint32_t nb, const float *pos, const float *vel) {
    if (dirty_) layout();
    sync();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    const SceneSlot &sc = scenes_[k];
    const uint32_t n = sc.params.n, nf = sc.params.nf;
    if (nb != n - nf) throw std::length_error("boundary segment length differs from the scene's");
    if (nb == 0) return;
    if (!pos || !vel) throw std::invalid_argument("null boundary segment");
    // Only the segment crosses PCIe (24 B per boundary particle); the device
    // scatters it into the grid-ordered state through the slot map and
    // rebuilds the scene's AABB from the new positions.
    float *d = (float *)staging(24ull * nb);
    hip_check(hipMemcpyAsync(d, pos, 12ull * nb, hipMemcpyHostToDevice, stream_), "boundary pos");
    hip_check(hipMemcpyAsync(d + 3ull * nb, vel, 12ull * nb, hipMemcpyHostToDevice, stream_), "boundary vel");
    const uint32_t nblk = sc.params.alloc / kBlock;
    hipLaunchKernelGGL(k_boundary_scatter, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(nblk, 8192u))),
                       dim3(kBlock), 0, stream_, P_, k, (const float *)d, (const float *)(d + 3ull * nb));
    hipLaunchKernelGGL(k_box_reduce, dim3(box_reduce_grid()), dim3(kBlock), 0, stream_, P_);
    hip_check(hipGetLastError(), "k_boundary_scatter");
    // The caller may reuse its buffers once this returns.
    hip_check(hipStreamSynchronize(stream_), "sync");
}

void Engine::set_span_origins(uint32_t id, const Span *spans, uint32_t nspans) {
    if (dirty_) layout();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    SceneSlot &sc = scenes_[k];
    if (nspans != sc.spans.size()) throw std::length_error("span count differs from the scene's");
    for (uint32_t q = 0; q < nspans; ++q)
        if (spans[q].start != sc.spans[q].start || spans[q].end != sc.spans[q].end)
            throw std::length_error("span ranges differ from the scene's");
    bool moved = false;
    for (uint32_t q = 0; q < nspans; ++q) {
        sc.spans[q].body = spans[q].body;
        for (int a = 0; a < 3; ++a) {
            moved = moved || std::memcmp(&sc.spans[q].origin[a], &spans[q].origin[a], sizeof(float)) != 0;
            sc.spans[q].origin[a] = spans[q].origin[a];
        }
    }
    if (!moved || !sc.want_reactions || !d_spans_) return;
    // The reaction kernel reads the torque origins from the span table at run
    // time, so the captured tick stays valid.
    std::vector<SpanDev> dev(nspans);
    hip_check(hipMemcpy(dev.data(), d_spans_ + span_first_[k], sizeof(SpanDev) * nspans, hipMemcpyDeviceToHost),
              "spans");
    for (uint32_t q = 0; q < nspans; ++q)
        for (int a = 0; a < 3; ++a) dev[q].origin[a] = spans[q].origin[a];
    hip_check(hipMemcpy(d_spans_ + span_first_[k], dev.data(), sizeof(SpanDev) * nspans, hipMemcpyHostToDevice),
              "spans");
}

void Engine::freeze_scene(uint32_t id) {
    if (dirty_) layout();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    if (scenes_[k].frozen) return;
    scenes_[k].frozen = true;
    // Stream-ordered after every step enqueued so far; the captured tick reads
    // the parameters at run time, so it stays valid.
    hipLaunchKernelGGL(k_freeze, dim3(1), dim3(64), 0, stream_, d_params_, k);
    hip_check(hipGetLastError(), "k_freeze");
}

void Engine::read_state_async(uint32_t id, float *pos, float *vel) {
    if (dirty_) layout();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    const SceneSlot &sc = scenes_[k];
    if (sc.frozen) throw std::logic_error("the scene is retired");
    const uint32_t n = sc.params.n;
    // Stream order keeps one staging buffer safe for consecutive readbacks;
    // growing it frees the old one, which waits for the device.
    float *op = (float *)staging(24ull * n), *ov = op + 3ull * n;
    const uint32_t nblk = sc.params.alloc / kBlock;
    hipLaunchKernelGGL(k_unsort, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(nblk, 8192u))), dim3(kBlock), 0,
                       stream_, P_, k, op, ov);
    hip_check(hipGetLastError(), "k_unsort");
    hip_check(hipMemcpyAsync(pos, op, 12ull * n, hipMemcpyDeviceToHost, stream_), "pos");
    hip_check(hipMemcpyAsync(vel, ov, 12ull * n, hipMemcpyDeviceToHost, stream_), "vel");
}

void Engine::read_state(uint32_t id, float *pos, float *vel) {
    sync();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    const SceneSlot &sc = scenes_[k];
    if (sc.frozen) throw std::logic_error("the scene is retired");
    const uint32_t n = sc.params.n;
    float *op = (float *)staging(24ull * n), *ov = op + 3ull * n;
    const uint32_t nblk = sc.params.alloc / kBlock;
    hipLaunchKernelGGL(k_unsort, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(nblk, 8192u))), dim3(kBlock), 0,
                       stream_, P_, k, op, ov);
    hip_check(hipGetLastError(), "k_unsort");
    if (pos) hip_check(hipMemcpyAsync(pos, op, 12ull * n, hipMemcpyDeviceToHost, stream_), "pos");
    if (vel) hip_check(hipMemcpyAsync(vel, ov, 12ull * n, hipMemcpyDeviceToHost, stream_), "vel");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

void Engine::read_fields(uint32_t id, float *dens, float *press, uint32_t *ncount, float *acc, float *xsph) {
    sync();
    if ((ncount || acc || xsph) && !diagnostics_)
        throw std::logic_error("per-particle diagnostics are off (enable before stepping)");
    const SceneSlot &sc = scene(id);
    const uint32_t n = sc.params.n, nf = sc.params.nf, pb = sc.params.pbase;
    float *f = (float *)staging(36ull * slots_);
    float *od = f, *op = f + slots_, *oa = f + 2ull * slots_, *ox = f + 5ull * slots_;  // acc/xsph AoS
    uint32_t *ou = (uint32_t *)(f + 8ull * slots_);
    hipLaunchKernelGGL(k_unsort_fields, dim3(P_.nblocks), dim3(kBlock), 0, stream_, P_, od, op,
                       diagnostics_ ? ou : nullptr, diagnostics_ ? oa : nullptr, diagnostics_ ? ox : nullptr);
    hip_check(hipGetLastError(), "k_unsort_fields");
    if (dens) hip_check(hipMemcpyAsync(dens, od + pb, 4ull * n, hipMemcpyDeviceToHost, stream_), "dens");
    if (press) hip_check(hipMemcpyAsync(press, op + pb, 4ull * n, hipMemcpyDeviceToHost, stream_), "press");
    if (ncount) hip_check(hipMemcpyAsync(ncount, ou + pb, 4ull * n, hipMemcpyDeviceToHost, stream_), "ncount");
    if (acc) hip_check(hipMemcpyAsync(acc, oa + 3ull * pb, 12ull * nf, hipMemcpyDeviceToHost, stream_), "acc");
    if (xsph) hip_check(hipMemcpyAsync(xsph, ox + 3ull * pb, 12ull * nf, hipMemcpyDeviceToHost, stream_), "xsph");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

SceneGrid Engine::read_grid(uint32_t id, std::vector<uint32_t> *start, std::vector<uint32_t> *items) {
    sync();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    const SceneSlot &sc = scenes_[k];
    SceneGrid g{};
    hip_check(hipMemcpy(&g, d_grid_ + k, sizeof(SceneGrid), hipMemcpyDeviceToHost), "grid");
    const uint32_t pb = sc.params.pbase, n = sc.params.n;
    if (start || items) {
        // Device tables are x fastest; rebuild the reference layout
        // (cell = (x*ny + y)*nz + z, fluid grid.rs:80-89).
        std::vector<uint32_t> st((size_t)g.total + 1), it(n);
        hip_check(hipMemcpy(st.data(), P_.start + g.cbase, 4ull * (g.total + 1), hipMemcpyDeviceToHost), "start");
        hip_check(hipMemcpy(it.data(), P_.sg + pb, 4ull * n, hipMemcpyDeviceToHost), "items");
        const uint32_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
        std::vector<uint32_t> rs((size_t)g.total + 1), ri(n);
        uint32_t w = 0;
        for (uint32_t x = 0; x < nx; ++x)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t z = 0; z < nz; ++z) {
                    const uint32_t c = (x * ny + y) * nz + z;
                    const uint32_t ci = (z * ny + y) * nx + x;
                    rs[c] = w;
                    for (uint32_t k = st[ci]; k < st[ci + 1]; ++k) ri[w++] = it[k - pb];
                }
        rs[g.total] = w;
        if (start) *start = std::move(rs);
        if (items) *items = std::move(ri);
    }
    return g;
}

void Engine::read_bounds(uint32_t id, float lo[3], float hi[3]) {
    sync();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    SceneBox b{};
    const uint32_t par = h_ctl_->step & 1u;
    hip_check(hipMemcpy(&b, d_box_ + par * scenes_.size() + k, sizeof(SceneBox), hipMemcpyDeviceToHost), "box");
    for (int a = 0; a < 3; ++a) {
        lo[a] = o2f(b.lo[a]);
        hi[a] = o2f(b.hi[a]);
    }
}

std::vector<float> Engine::read_reactions(uint32_t id) {
    sync();
    uint32_t k = 0;
    for (; k < scenes_.size(); ++k)
        if (scenes_[k].id == id) break;
    if (k == scenes_.size()) throw std::out_of_range("unknown scene");
    const SceneSlot &sc = scenes_[k];
    std::vector<float> out(6 * sc.spans.size(), 0.0f);
    if (!sc.want_reactions || sc.spans.empty() || substeps_done_ == 0) return out;
    hip_check(hipMemcpy(out.data(), d_react_ + 6ull * span_first_[k], 4 * out.size(), hipMemcpyDeviceToHost),
              "reactions");
    return out;
}

}  // namespace vxl
