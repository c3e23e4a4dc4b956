// Host runtime of the WCSPH pool: device buffers, scene table, the substep
// launch sequence (captured once as a hipGraph and replayed), readbacks.
#ifndef VXL_SPH_ENGINE_H
#define VXL_SPH_ENGINE_H
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "sph_kernels.h"
#include "sph_scene.h"

namespace vxl {

struct DeviceError : std::runtime_error {
    hipError_t code;
    DeviceError(hipError_t c, const std::string &w) : std::runtime_error(w), code(c) {}
};
struct CapacityError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// One scene as the host knows it.
struct SceneSlot {
    uint32_t id = 0;
    SceneParams params{};  // pbase/alloc assigned by the layout
    uint32_t substeps = 1;
    std::vector<Span> spans;
    // host staging for a scene that is not on the device yet
    bool resident = false;
    bool frozen = false;    // retired in place (no particles stepped); dropped by the next layout
    uint32_t dev_base = 0;  // slot the scene's state starts at on the device
    std::vector<float4> stage_a, stage_v;  // packed records (A, V) until the layout uploads them
    uint32_t cells_hint = 0;             // grid cells of the initial box
    // reactions of the last tick (tick average), per span: F[3], tau[3]
    std::vector<float> reactions;
    bool want_reactions = false;
    // dynamic rigid bodies: device record (state as of the last layout or
    // readback), body-frame lattice (3 floats per particle), local span index
    struct Rigid {
        RigidDev dev{};
        std::vector<float> lat;
        uint32_t span_local = 0;
        uint32_t body_id = 0;
        float density = 0.0f;
    };
    std::vector<Rigid> rigid;
    std::vector<BoxBody> statics;  // static Box bodies of a coupled scene (for frame dumps)
    // Provider walls: half-space providers (faces grouped by provider) or
    // plane tables (side 0 mirror, 1 project) with their device copies.
    struct WallTable {
        std::vector<uint32_t> ids, start;
        std::vector<float4> planes;  // (point, normal) per plane
        uint32_t *d_ids = nullptr, *d_start = nullptr;
        float4 *d_planes = nullptr;
        size_t cap_e = 0, cap_p = 0;  // device capacity (entries, planes)
    };
    uint32_t wall_mode = kWallsOff;
    std::vector<WallProvDev> wall_prov;
    std::vector<WallFaceDev> wall_face;
    WallTable wall_tab[2];
};

// A half-space wall provider as the host describes it: the AABB of bounds()
// (has_bounds = 0: none) and its axis-aligned faces.
struct WallProvider {
    uint32_t id = 0;
    bool has_bounds = false;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    struct Face {
        uint32_t axis;
        float plane, sign;
    };
    std::vector<Face> faces;
};

// Rigid configuration of a coupled scene (PhysConfig defaults,
// crates/vxl-phys-core/src/config.rs:206-233) plus the tank contact planes.
struct RigidConfig {
    uint32_t substeps = 2;
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    float max_linear = 100.0f;
    float max_angular = 50.0f;
    float tank[3] = {0.0f, 0.0f, 0.0f};  // open tank [0,L0] x [0,inf) x [0,L2]; L0 <= 0: no contacts
};

class Engine {
public:
    explicit Engine(int device);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // A scene built on the host (records packed, grid size measured, bodies
    // and their lattices ready), so that adding it to the pool costs no more
    // than moving it in; any thread may prepare scenes.
    static SceneSlot prepare_scene(const SceneConfig &cfg, uint32_t n, uint32_t nf, const float *pos,
                                   const float *vel, const float *pmass, const Span *spans, uint32_t nspans,
                                   bool reactions = true);
    static SceneSlot prepare_coupled(const SceneConfig &cfg, uint32_t nf, const float *fpos, const float *fvel,
                                     const std::vector<BoxBody> &bodies, const std::vector<float> &density,
                                     const RigidConfig &rc);
    uint32_t add_prepared(SceneSlot &&sc);

    // Scene management (takes effect at the next step or readback).
    uint32_t add_scene(const SceneConfig &cfg, uint32_t n, uint32_t nf, const float *pos,
                       const float *vel, const float *pmass, const Span *spans, uint32_t nspans,
                       bool reactions = true);
    // A scene whose boundary is sampled from Box bodies (set_boundary_particles,
    // fluid_step.rs:13-53): bodies with density > 0 are dynamic and coupled
    // both ways (advanced each tick by gravity and the fluid reaction, their
    // particles rebuilt from the pose); the others are static.
    uint32_t add_coupled_scene(const SceneConfig &cfg, uint32_t nf, const float *fpos, const float *fvel,
                               const std::vector<BoxBody> &bodies, const std::vector<float> &density,
                               const RigidConfig &rc);
    // Current state of a scene's dynamic bodies (in the order given at creation).
    std::vector<RigidDev> read_bodies(uint32_t id);
    void remove_scene(uint32_t id);
    bool has_scene(uint32_t id) const;
    const SceneSlot &scene(uint32_t id) const;
    size_t scene_count() const { return scenes_.size(); }

    // Provider walls of a scene (see WallSceneDev): half-space providers
    // queried on the device every substep (FluidSystem::set_boundaries with
    // providers, fluid_access.rs:184-186), or sparse plane tables of side 0
    // (mirror) / 1 (project) that the host gathers from any provider and
    // replaces once per tick (WallStage::upload, walls.rs:104-116): ids
    // strictly ascending fluid indices, start[ne + 1] plane ranges, planes as
    // point xyz + normal xyz. Switching a scene's wall mode takes effect at
    // the next step; replacing a table of a resident scene is immediate.
    void set_halfspace_walls(uint32_t id, const std::vector<WallProvider> &providers);
    void set_wall_tables(uint32_t id, uint32_t side, const uint32_t *ids, uint32_t ne, const uint32_t *start,
                         const float *planes, uint32_t np);
    void clear_walls(uint32_t id);

    // Overwrite the boundary segment pos/vel of a resident scene (index order).
    void upload_boundary(uint32_t id, uint32_t nb, const float *pos, const float *vel);
    // Replace the torque origins of a scene's reaction spans (same count and
    // ranges as at creation; the FluidStepper passes them every tick).
    void set_span_origins(uint32_t id, const Span *spans, uint32_t nspans);

    // Advance every scene by `ticks` ticks of its own substeps (enqueue only).
    void step(uint32_t ticks);
    // Wait for the device and surface deferred errors.
    void sync();

    // Readbacks (synchronous), index order.
    void read_state(uint32_t id, float *pos, float *vel);
    void read_fields(uint32_t id, float *dens, float *press, uint32_t *ncount, float *acc, float *xsph);
    SceneGrid read_grid(uint32_t id, std::vector<uint32_t> *start, std::vector<uint32_t> *items);
    void read_bounds(uint32_t id, float lo[3], float hi[3]);
    // Tick-averaged (F, tau) per span of the scene: 6 floats each.
    std::vector<float> read_reactions(uint32_t id);

    // Per-kernel device time: runs `substeps` eager substeps with a HIP event
    // between consecutive kernels on the engine stream. Returns (name, ms total).
    std::vector<std::pair<std::string, float>> profile(uint32_t substeps);
    // Runs `ticks` ticks, recording the device time of each tick (ms).
    std::vector<float> step_timed(uint32_t ticks);

    void set_diagnostics(bool on);
    void set_graphs(bool on) { use_graph_ = on; invalidate_graph(); }
    hipStream_t stream() const { return stream_; }
    uint64_t substeps_done() const { return substeps_done_; }

    // Launch variants (autotuned; every variant gives identical results).
    enum : uint32_t {
        kForceScan = 1,
        kForceWalk = 2,
        kForceWalkAhead = 3,
        kForcePairs = 4,
        kForceQueue = 5,
        kForcePaired = 6,
        kForceStaged = 7
    };
    struct Tuning {
        uint32_t persist_blocks = 0;         // grid-stride grid size cap for slot kernels (0 = none)
        uint32_t force_kernel = kForceStaged;  // kForceScan ... kForceStaged
        uint32_t block_threads = 256;        // workgroup size of the density and force kernels: 64 | 128 | 256
        uint32_t stage_records = 256;        // density LDS staging depth per wave and row: 128 | 256 | 512
        uint32_t scan_tile_cells = 2048;     // cells per scan tile: 1024 | 2048 | 4096
    };
    // 0 in a kernel field selects its default.
    void set_tuning(const Tuning &t);
    // Time every launch variant on the current pool for `substeps` substeps
    // each (they are real substeps: all variants give identical results) and
    // keep the fastest. Results are cached per (device arch, log2 particle
    // count) for the life of the process.
    Tuning autotune(uint32_t substeps);
    const Tuning &tuning() const { return tuning_; }
    uint64_t device_bytes() const { return bytes_; }
    uint64_t layouts() const { return layouts_; }
    // Host time of every layout so far (ms), and the summed time of its phases:
    // 0 waiting for the device, 1 moving resident scenes (and growing the
    // pool), 2 uploading new scenes, 3 scene tables and boxes; 4 building the
    // step's graph afterwards.
    const std::vector<float> &layout_ms() const { return layout_ms_; }
    const double *layout_phase_ms() const { return layout_phase_ms_; }
    uint64_t graph_builds() const { return graph_builds_; }    // tick graphs instantiated
    uint64_t graph_updates() const { return graph_updates_; }  // tick graphs updated in place
    uint32_t capacity_slots() const { return slots_; }
    uint32_t used_slots() const { return used_slots_; }
    // Device bytes a scene of n particles (cells_hint grid cells) adds to the
    // pool: per-slot state and sort arrays plus its share of the cell tables.
    static uint64_t footprint(uint32_t n, uint32_t cells_hint, uint64_t max_bins);
    // Device bytes of a pool of `slots` slots and `ccap` cells (state, sort,
    // neighbour and cell arrays as the layout allocates them).
    static uint64_t pool_bytes(uint64_t slots, uint64_t ccap, bool diagnostics, bool reactions);
    // Peak device bytes of this engine while the next layout adds a scene of
    // n particles (the old and the new pool are both alive while resident
    // scenes move), including the readback staging its read_state needs.
    uint64_t projected_peak(uint32_t n, uint32_t cells_hint, uint64_t max_bins, bool reactions) const;
    // projected_peak for an engine that holds no scene.
    static uint64_t solo_peak(uint32_t n, uint32_t cells_hint, uint64_t max_bins, bool reactions,
                              bool full_cells = false);
    // Reserve cell tables for each scene's whole grid budget (max_bins + 1
    // cells) instead of a multiple of its initial grid: no capacity replay can
    // ever happen, so readbacks may be enqueued without a check.
    void set_full_cell_reserve(bool on) { full_cells_ = on; }
    // Retire a scene in place: it is not stepped any more (stream-ordered after
    // the steps enqueued so far) and leaves the pool at the next layout, with
    // no relayout now. Its state can no longer be read.
    void freeze_scene(uint32_t id);
    // Enqueue the readback of a scene's state (index order) into host memory
    // that stays valid until the stream has passed this point (pinned memory
    // for an asynchronous copy); no synchronisation.
    void read_state_async(uint32_t id, float *pos, float *vel);
    // Apply pending scene changes (the layout) now.
    void prepare() {
        if (dirty_) layout();
        if (use_graph_ && (!gexec_ || graph_stale_) && P_.nblocks) build_graph();
    }
    // Scene-table bytes a layout allocates (bounded above).
    static uint64_t table_allowance(uint64_t nscenes, uint64_t slots, bool reactions);
    // Pool growth stays within `bytes` (0 = no limit): the x1.5 headroom is
    // dropped when it would not fit beside the old pool.
    void set_memory_limit(uint64_t bytes) { mem_limit_ = bytes; }
    uint64_t peak_bytes() const { return peak_bytes_; }
    // Free every device array when no scene is left but retired ones (the
    // pool never shrinks otherwise).
    void release_if_empty();

private:
    void layout();            // assign slot ranges, (re)allocate, move + upload
    uint32_t box_reduce_grid() const;
    uint32_t grown_capacity(uint32_t need, uint64_t cells, bool reactions) const;
    uint32_t grown_cells(uint64_t cells) const;
    uint32_t pool_substeps() const;  // launches per pool tick: the most substeps of any scene
    // The launch sequence of one substep, launch tick_sub of a pool tick of
    // tick_len launches; `mark` (if set) is called after each kernel launch
    // with the kernel's name.
    void enqueue_substep(uint32_t tick_sub, uint32_t tick_len, const std::function<void(const char *)> *mark = nullptr);
    // Substep i (counted from the last layout) with the rigid leg's launches at
    // the tick's start and end when the pool has dynamic bodies.
    // `head` = false skips the tick's rigid leg (a replay of a substep whose
    // tick already ran it).
    void enqueue_at(uint64_t i, const std::function<void(const char *)> *mark = nullptr, bool head = true);
    void build_graph();
    void invalidate_graph();
    void check_error_state();  // read the control block; grow tables on demand
    void alloc_pool(uint32_t slots);
    void alloc_cells(uint32_t ccap);
    void free_cells();  // counts, start, fcount and the scan's tile sums
    void free_walls();  // wall pool arrays and the wall scene table
    void free_wall_tables(SceneSlot &sc);
    void upload_wall_table(SceneSlot &sc, uint32_t side);
    WallSceneDev wall_scene(const SceneSlot &sc) const;
    void index_wall_table(uint32_t k, uint32_t side);  // went of scene position k (stream-ordered)
    SceneSlot &slot_of(uint32_t id);
    void free_all();
    template <class T>
    T *dalloc(size_t n);
    void dfree(void *p);

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    std::vector<SceneSlot> scenes_;
    uint32_t next_id_ = 1;
    bool dirty_ = true;        // layout must be rebuilt
    bool diagnostics_ = false;
    bool use_graph_ = true;
    Tuning tuning_;

    uint32_t slots_ = 0;       // pool slots allocated
    uint32_t used_slots_ = 0;  // slots assigned to scenes
    uint32_t ccap_ = 0;
    uint32_t tiles_cap_ = 0;
    Pool P_{};
    std::vector<void *> pool_allocs_;
    SceneParams *d_params_ = nullptr;
    SceneGrid *d_grid_ = nullptr;
    SceneBox *d_box_ = nullptr;
    Control *d_ctl_ = nullptr;
    Control *h_ctl_ = nullptr;  // pinned mirror
    uint32_t *d_block_scene_ = nullptr;
    uint32_t *d_tile_sums_ = nullptr;
    SpanDev *d_spans_ = nullptr;
    float *d_react_ = nullptr;
    RigidDev *d_rigid_ = nullptr;
    float4 *d_rlat_ = nullptr;
    uint32_t nspans_ = 0;
    WallSceneDev *d_walls_ = nullptr;
    uint32_t nwall_ = 0;           // scenes with walls (wall kernels launched when > 0)
    bool wall_halfspace_ = false;  // some scene has half-space walls (ghosts in the density pass, separation)
    bool wall_table_ = false;      // some scene has plane tables (ghost kernels after the density pass)
    std::vector<uint32_t> span_first_;  // first global span of each scene (by scene position)
    // spare state arrays: resident scenes move into them at a layout change
    float4 *spare_cA_ = nullptr, *spare_cV_ = nullptr;
    uint32_t *spare_cg_ = nullptr;
    bool want_reaction_buffers_ = false;
    bool diag_alloc_ = false;
    // index-order staging for readbacks (grown on demand)
    void *staging_ = nullptr;
    size_t staging_bytes_ = 0;
    void *staging(size_t bytes);
    uint64_t layouts_ = 0;
    std::unordered_map<void *, size_t> alloc_size_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    bool graph_stale_ = false;     // gexec_ holds the previous layout's launches (update it before use)
    uint64_t graph_builds_ = 0;    // tick graphs instantiated
    uint64_t graph_updates_ = 0;   // tick graphs updated in place after a layout
    uint64_t substeps_done_ = 0;
    uint32_t ctl_step_base_ = 0;
    uint64_t bytes_ = 0;
    uint64_t peak_bytes_ = 0;
    uint64_t mem_limit_ = 0;
    bool full_cells_ = false;
    std::vector<float> layout_ms_;
    double layout_phase_ms_[5] = {0, 0, 0, 0, 0};
};

}  // namespace vxl
#endif
