"""GPU parity: the gfx950 WCSPH step against the oracle on identical seeded
inputs. Integer outputs (cell keys, permutation, neighbour counts) must be
bit-exact; float state is compared bit for bit as well (strict f32, same
operation order), with the tolerance contract (max relative error <= 1e-5 after
one step) asserted beside it so a failure says which bar broke."""
import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import (GOLDEN_DIR, GOLDEN_SCENES, bits, golden_spans, grid_table_hash, load_golden, max_rel,
                    oracle_dam_break, tank_bodies)

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5  # per-quantity floors: rho0 for density, h for position, c0*1e-3 for velocity


def small_scene(seed=7, block=(12, 10, 8), tank=(0.6, 0.4, 0.4), d=0.02, max_bins=1 << 21, **kw):
    cfg = V.dam_break_config(d, max_bins=max_bins, **kw)
    data = V.dam_break(cfg, tank, block, d, seed=seed)
    return cfg, data, tank, d


def assert_grid_equal(eng, sid, orc):
    g = eng.read_grid(sid)
    omin, oinv, obin, odims = orc.grid_info()
    ostart, oitems, okeys = orc.grid_tables()
    assert g["dims"] == odims
    assert bits(g["min"]).tolist() == bits(omin).tolist()
    assert bits([g["inv"]])[0] == bits([oinv])[0]
    assert bits([g["bin"]])[0] == bits([obin])[0]
    np.testing.assert_array_equal(g["start"], ostart)
    np.testing.assert_array_equal(g["items"], oitems)
    np.testing.assert_array_equal(g["keys"], okeys)


def assert_state_equal(gpu, orc, what, floor):
    rel = max_rel(gpu, orc, floor)
    assert rel <= TOL_REL, f"{what}: max rel error {rel:.3e} > {TOL_REL}"
    diff = int(np.count_nonzero(bits(gpu) != bits(orc)))
    assert diff == 0, f"{what}: {diff} values differ in bits (max rel {rel:.3e})"


def test_backend_identity():
    b = V.Backend()
    assert b.name() == "hip-gfx950"
    assert b.available()
    assert b.submit("fluid")
    assert not b.submit("cloth")


def test_one_step_bit_exact():
    cfg, data, tank, d = small_scene()
    eng = V.Engine(diagnostics=True)
    sid = eng.add(cfg, data, reactions=False)
    orc = oracle_dam_break(cfg, data, tank, d)
    eng.step(1)
    orc.step(cfg.dt_tick)
    assert_grid_equal(eng, sid, orc)
    f = eng.read_fields(sid, diagnostics=True)
    np.testing.assert_array_equal(f["ncount"], orc.ncount())
    assert_state_equal(f["dens"], orc.dens(), "density", cfg.rest_density)
    assert_state_equal(f["press"], orc.press(), "pressure", 1.0)
    assert_state_equal(f["acc"], orc.acc(), "acceleration", 1e-4)
    assert_state_equal(f["xsph"], orc.xsph(), "xsph", 1e-9)
    pos, vel = eng.read_state(sid)
    assert_state_equal(pos, orc.pos(), "position", cfg.smoothing_radius)
    assert_state_equal(vel, orc.vel(), "velocity", cfg.sound_speed * 1e-3)


def test_boundary_density_from_fluid_heads_bit_exact():
    """Without neighbour-count diagnostics a boundary particle scans only the
    fluid items heading each cell (fluid indices precede boundary ones);
    density and pressure of every particle, boundary included, equal the
    oracle's bit for bit over several steps."""
    cfg, data, tank, d = small_scene(seed=21, block=(14, 12, 10), tank=(0.5, 0.4, 0.35))
    eng = V.Engine(diagnostics=False)
    sid = eng.add(cfg, data, reactions=False)
    orc = oracle_dam_break(cfg, data, tank, d)
    for _ in range(5):
        eng.step(1)
        orc.step(cfg.dt_tick)
        f = eng.read_fields(sid)
        np.testing.assert_array_equal(bits(f["dens"]), bits(orc.dens()))
        np.testing.assert_array_equal(bits(f["press"]), bits(orc.press()))
    assert eng.state_hash(sid) == orc.state_hash()


def test_many_steps_bit_exact_and_rerun_hash():
    cfg, data, tank, d = small_scene(seed=11)
    orc = oracle_dam_break(cfg, data, tank, d)
    for _ in range(25):
        orc.step(cfg.dt_tick)
    hashes = []
    for run in range(3):
        eng = V.Engine()
        sid = eng.add(cfg, data, reactions=False)
        eng.step(25)
        hashes.append(eng.state_hash(sid))
        if run == 0:
            pos, vel = eng.read_state(sid)
            assert_state_equal(pos, orc.pos(), "position after 25 steps", cfg.smoothing_radius)
            assert_state_equal(vel, orc.vel(), "velocity after 25 steps", cfg.sound_speed * 1e-3)
            assert_grid_equal(eng, sid, orc)
        eng.close()
    assert hashes[0] == hashes[1] == hashes[2]
    assert hashes[0] == orc.state_hash()


def test_graph_replay_equals_eager():
    cfg, data, tank, d = small_scene(seed=3)
    out = []
    for graphs in (True, False):
        eng = V.Engine(graphs=graphs)
        sid = eng.add(cfg, data, reactions=False)
        for _ in range(4):  # replays interleaved with host synchronisations
            eng.step(3)
            eng.sync()
        out.append(eng.state_hash(sid))
        eng.close()
    assert out[0] == out[1]


def test_pure_fluid_reference_lattice_scene():
    """FluidConfig::default lattice block (the reference's grid-table pin
    scene, n = 12): GPU tables after 2 ticks x 4 substeps equal the oracle's."""
    cfg = V.fluid_config()
    n = 12
    sp = 0.05
    ocfg = O.default_config()
    orc = O.System.lattice(ocfg, (-n * sp * 0.5, 0.5, -n * sp * 0.5), (n, n, n), sp)
    pos = orc.pos()
    data = V.SceneData(pos, np.zeros_like(pos), np.full(len(pos), orc.mass, np.float32), len(pos), [])
    eng = V.Engine(diagnostics=True)
    sid = eng.add(cfg, data)
    for _ in range(2):
        eng.step(1)
        orc.step(1.0 / 60.0)
    assert_grid_equal(eng, sid, orc)
    pos_g, vel_g = eng.read_state(sid)
    assert_state_equal(pos_g, orc.pos(), "position", 0.1)
    assert_state_equal(vel_g, orc.vel(), "velocity", 0.01)
    assert eng.state_hash(sid) == orc.state_hash()


def test_reactions_match_oracle():
    """Akinci reactions per body (tick average over 2 substeps)."""
    cfg, data, tank, d = small_scene(seed=5)
    cfg.substeps = 2
    cfg.dt_tick = 1e-3
    orc = oracle_dam_break(cfg, data, tank, d)
    st = V.FluidStepper(cfg)
    for _ in range(3):
        st.step(cfg.dt_tick, data.n_fluid, data.pos, data.vel, data.pmass, data.spans)
        orc.step(cfg.dt_tick)
    body, ft = orc.reactions()
    got = st.reactions()
    assert [g[0] for g in got] == body.tolist()
    gft = np.array([np.concatenate([g[1], g[2]]) for g in got], np.float32)
    np.testing.assert_array_equal(bits(gft), bits(ft))
    pos, vel = st.read_state()
    assert_state_equal(pos, orc.pos(), "position", cfg.smoothing_radius)


def test_reactions_of_a_larger_tank_match_oracle():
    """Reactions on every span of a tank large enough that the reacting-slot
    list is filled by many workgroups (each reserving its range of the list
    with one atomic), and that the span fold runs over several chunks of its
    producer/consumer pipeline: 48 x 40 x 36 fluid in a 1.2 x 0.9 x 0.9 m
    tank, ~98k particles, 2 ticks of 2 substeps, bit-exact against the
    oracle."""
    cfg, data, tank, d = small_scene(seed=11, block=(48, 40, 36), tank=(1.2, 0.9, 0.9))
    cfg.substeps = 2
    cfg.dt_tick = 1e-3
    assert data.n > 16 * 256 * 4  # more than four workgroups of the list
    assert max(s[3] - s[2] for s in data.spans) > 2 * 1536  # several fold chunks
    orc = oracle_dam_break(cfg, data, tank, d, threads=8)
    st = V.FluidStepper(cfg)
    for t in range(2):
        st.step(cfg.dt_tick, data.n_fluid, data.pos, data.vel, data.pmass, data.spans)
        orc.step(cfg.dt_tick)
        body, ft = orc.reactions()
        got = st.reactions()
        assert [g[0] for g in got] == body.tolist()
        gft = np.array([np.concatenate([g[1], g[2]]) for g in got], np.float32)
        np.testing.assert_array_equal(bits(gft), bits(ft), err_msg=f"tick {t} reactions")
    pos, vel = st.read_state()
    assert_state_equal(pos, orc.pos(), "position", cfg.smoothing_radius)
    assert_state_equal(vel, orc.vel(), "velocity", cfg.sound_speed * 1e-3)


def test_batched_scenes_equal_alone():
    """Several scenes of uneven size and physics in one pool: each result is
    bit-identical to the scene run alone and to the oracle."""
    specs = [
        dict(seed=1, block=(6, 5, 4), tank=(0.3, 0.3, 0.3), c0=15.0, alpha=0.01),
        dict(seed=2, block=(14, 9, 7), tank=(0.5, 0.4, 0.3), c0=25.0, alpha=0.05),
        dict(seed=3, block=(9, 12, 10), tank=(0.4, 0.5, 0.4), c0=30.0, alpha=0.03),
    ]
    scenes = []
    for sp in specs:
        cfg = V.dam_break_config(0.02, c0=sp["c0"], alpha=sp["alpha"])
        data = V.dam_break(cfg, sp["tank"], sp["block"], 0.02, seed=sp["seed"])
        scenes.append((cfg, data, sp["tank"]))
    eng = V.Engine()
    ids = [eng.add(c, dt, reactions=False) for c, dt, _ in scenes]
    eng.step(8)
    mixed = [eng.state_hash(i) for i in ids]
    for k, (cfg, data, tank) in enumerate(scenes):
        alone = V.Engine()
        sid = alone.add(cfg, data, reactions=False)
        alone.step(8)
        assert alone.state_hash(sid) == mixed[k]
        orc = oracle_dam_break(cfg, data, tank, 0.02)
        for _ in range(8):
            orc.step(cfg.dt_tick)
        assert orc.state_hash() == mixed[k]


def test_scene_removal_keeps_others():
    cfg = V.dam_break_config(0.02)
    a = V.dam_break(cfg, (0.3, 0.3, 0.3), (6, 6, 6), 0.02, seed=21)
    b = V.dam_break(cfg, (0.4, 0.3, 0.3), (8, 6, 6), 0.02, seed=22)
    eng = V.Engine()
    ia = eng.add(cfg, a, reactions=False)
    ib = eng.add(cfg, b, reactions=False)
    eng.step(3)
    eng.remove(ia)
    eng.step(3)
    ref = V.Engine()
    rb = ref.add(cfg, b, reactions=False)
    ref.step(6)
    assert eng.state_hash(ib) == ref.state_hash(rb)


def test_raised_grid_budget_and_coarse_rule():
    """Coarse rule (tiny budget forces bin = 2h, 4h) stays bit-exact."""
    for budget in (512, 4096):
        cfg, data, tank, d = small_scene(seed=9, max_bins=budget)
        orc = oracle_dam_break(cfg, data, tank, d)
        eng = V.Engine()
        sid = eng.add(cfg, data, reactions=False)
        for _ in range(3):
            eng.step(1)
            orc.step(cfg.dt_tick)
        assert_grid_equal(eng, sid, orc)
        assert eng.state_hash(sid) == orc.state_hash()


def test_config1_dam_break_200_steps():
    """32^3 fluid + one-layer Akinci tank (2 x 1 x 1 m), 200 steps of 5e-4 s."""
    d = 0.02
    cfg = V.dam_break_config(d)
    tank = (2.0, 1.0, 1.0)
    data = V.dam_break(cfg, tank, (32, 32, 32), d, seed=1)
    orc = oracle_dam_break(cfg, data, tank, d, threads=8)
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    orc.step(cfg.dt_tick)
    eng.step(1)
    pos, vel = eng.read_state(sid)
    assert_state_equal(pos, orc.pos(), "position after 1 step", cfg.smoothing_radius)
    assert_state_equal(vel, orc.vel(), "velocity after 1 step", cfg.sound_speed * 1e-3)
    for _ in range(199):
        orc.step(cfg.dt_tick)
    eng.step(199)
    pos, vel = eng.read_state(sid)
    print("drift after 200 steps: max|dpos| %.3e m, max|dvel| %.3e m/s" % (
        np.abs(pos - orc.pos()).max(), np.abs(vel - orc.vel()).max()))
    assert eng.state_hash(sid) == orc.state_hash()


def test_reference_akinci_canary_on_gpu():
    """The reference's Akinci canary scene (packet_grid_canary.rs:27-57:
    8^3 lattice + floor Box, FluidConfig::default, 1 + 60 ticks x 4 substeps)
    through the FluidStepper drop-in: the fluid's lowest y must round to the
    reference's recorded CPU readout -1.81726 (docs/PLAN-gpu.md:2704), and the
    final state must equal the oracle's bit for bit."""
    N, SP = 8, np.float32(0.05)
    cfg = V.fluid_config()
    h = np.float32(cfg.smoothing_radius)
    orc = O.System.lattice(O.default_config(), (-N * SP * 0.5, 0.5, -N * SP * 0.5), (N, N, N), SP)
    fpos = orc.pos()
    nf = len(fpos)
    half = np.float32(N) * SP * np.float32(0.5) + np.float32(4.0) * h
    body = dict(id=0, half=(half, np.float32(2.0) * SP, half), pos=(0.0, np.float32(-2.0) * h, 0.0))
    bpos, bvel, bpm, spans = V.box_particles(cfg, [body], nf)
    orc.set_boundary_boxes([body])
    pos = np.concatenate([fpos, bpos])
    vel = np.concatenate([np.zeros_like(fpos), bvel])
    pm = np.concatenate([np.full(nf, orc.mass, np.float32), bpm])
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))
    np.testing.assert_array_equal(bits(pm), bits(orc.pmass()))
    st = V.FluidStepper(cfg)
    for _ in range(61):
        st.step(1.0 / 60.0, nf, pos, vel, pm, spans)
        orc.step(1.0 / 60.0)
    gpos, gvel = st.read_state()
    assert round(float(gpos[:nf, 1].min()), 5) == -1.81726
    np.testing.assert_array_equal(bits(gpos), bits(orc.pos()))
    np.testing.assert_array_equal(bits(gvel), bits(orc.vel()))
    body_ids, ft = orc.reactions()
    got = st.reactions()
    gft = np.array([np.concatenate([g[1], g[2]]) for g in got], np.float32)
    np.testing.assert_array_equal(bits(gft), bits(ft))


@pytest.mark.parametrize("force_kernel", [V.FORCE_SCAN, V.FORCE_WALK, V.FORCE_WALK_AHEAD, V.FORCE_PAIRS, V.FORCE_QUEUE,
                                          V.FORCE_PAIRED, V.FORCE_STAGED])
@pytest.mark.parametrize("persist", [0, 7])
@pytest.mark.parametrize("budget", [1 << 21, 1 << 9])
def test_kernel_variants_bit_exact(force_kernel, persist, budget):
    """Every autotuner choice (force: candidate scan, walk of the recorded
    neighbours directly or with loads one pair ahead, the pairs of each row
    spread over the wave's lanes, the walk from a queue of row chunks, that
    walk with records loaded by lane pairs or with each wave's own row staged
    in LDS; one block per
    256 slots or a grid-stride loop over a few blocks) gives the oracle's state
    bit for bit. The 2^9 budget doubles the bin: cells of ~64 particles, rows
    longer than 32 candidates (tail tests in the walk) and row unions longer than
    the density pass's LDS stage."""
    cfg, data, tank, d = small_scene(seed=13, block=(16, 12, 10), tank=(0.6, 0.5, 0.4), max_bins=budget)
    orc = oracle_dam_break(cfg, data, tank, d)
    eng = V.Engine(diagnostics=True)
    sid = eng.add(cfg, data, reactions=False)
    eng.set_tuning(persist_blocks=persist, force_kernel=force_kernel)
    for _ in range(6):
        eng.step(1)
        orc.step(cfg.dt_tick)
    f = eng.read_fields(sid, diagnostics=True)
    np.testing.assert_array_equal(bits(f["acc"]), bits(orc.acc()))
    np.testing.assert_array_equal(bits(f["xsph"]), bits(orc.xsph()))
    assert eng.state_hash(sid) == orc.state_hash()


@pytest.mark.parametrize("block_threads", [64, 128, 256])
@pytest.mark.parametrize("budget", [1 << 21, 1 << 9])
def test_tuning_knobs_bit_exact(block_threads, budget):
    """Every value of every autotuner knob — workgroup size of the density and
    force kernels, LDS staging depth of the density pass, cells per scan tile —
    with every force kernel gives the oracle's acceleration, XSPH and state
    bit for bit (at the coarse 2^9 budget unions outgrow even the 512-record
    stage and rows run past 32 candidates)."""
    cfg, data, tank, d = small_scene(seed=19, block=(16, 12, 10), tank=(0.6, 0.5, 0.4), max_bins=budget)
    orc = oracle_dam_break(cfg, data, tank, d)
    for _ in range(4):
        orc.step(cfg.dt_tick)
    want_acc, want_xs, want_hash = bits(orc.acc()), bits(orc.xsph()), orc.state_hash()
    for stage in (128, 256, 512):
        for scan in (1024, 2048, 4096):
            for fk in (V.FORCE_SCAN, V.FORCE_WALK, V.FORCE_WALK_AHEAD, V.FORCE_PAIRS, V.FORCE_QUEUE, V.FORCE_PAIRED,
                   V.FORCE_STAGED):
                eng = V.Engine(diagnostics=True)
                sid = eng.add(cfg, data, reactions=False)
                eng.set_tuning(force_kernel=fk, block_threads=block_threads, stage_records=stage,
                               scan_tile_cells=scan)
                assert eng.get_tuning()["block_threads"] == block_threads
                eng.step(4)
                f = eng.read_fields(sid, diagnostics=True)
                tag = f"stage {stage} scan {scan} force {fk}"
                np.testing.assert_array_equal(bits(f["acc"]), want_acc, err_msg=tag)
                np.testing.assert_array_equal(bits(f["xsph"]), want_xs, err_msg=tag)
                assert eng.state_hash(sid) == want_hash, tag
                eng.close()


def test_autotuner_picks_allowed_values_and_keeps_results():
    """The autotuner's choice lies in each knob's value set, and the state it
    leaves (it tunes on real substeps) equals the oracle after as many steps."""
    cfg, data, tank, d = small_scene(seed=21, block=(16, 12, 10), tank=(0.6, 0.5, 0.4))
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    t = eng.autotune(1)
    assert t["block_threads"] in (64, 128, 256) and t["stage_records"] in (128, 256, 512)
    assert t["scan_tile_cells"] in (1024, 2048, 4096) and t["force_kernel"] in (1, 2, 3, 4, 5, 6, 7)
    assert eng.get_tuning() == t
    # the tuning substeps advanced the scene; count them by replaying on the oracle
    orc = oracle_dam_break(cfg, data, tank, d)
    steps = 0
    while orc.state_hash() != eng.state_hash(sid) and steps < 200:
        orc.step(cfg.dt_tick)
        steps += 1
    assert orc.state_hash() == eng.state_hash(sid)  # (no steps when the choice came from the cache)
    eng.step(3)
    for _ in range(3):
        orc.step(cfg.dt_tick)
    assert eng.state_hash(sid) == orc.state_hash()


def test_autotune_cache_keys_the_grid_coarseness():
    """The autotuner's cache is keyed by the grid's coarseness as well as the
    pool size: a scene of the same size under a small grid budget (bin 2h,
    cells of ~64 particles, where other walks win) is tuned afresh rather than
    handed the bin = h choice, and a second such scene takes the cached
    choice. Tuning advances a scene by real substeps (one substep when the
    choice comes from the cache, to read the grid); the substeps are counted
    by stepping an untuned copy of the scene until its state hash matches."""
    def steps_taken(eng, sid, cfg, data):
        ref = V.Engine()
        sr = ref.add(cfg, data, reactions=False)
        k = 0
        while ref.state_hash(sr) != eng.state_hash(sid) and k < 400:
            ref.step(1)
            k += 1
        ref.close()
        return k

    cfg, data, _, _ = small_scene(seed=23, block=(16, 12, 10), tank=(0.6, 0.5, 0.4))
    fine = V.Engine()
    fine.add(cfg, data, reactions=False)
    fine.autotune(1)
    cfg2, data2, _, _ = small_scene(seed=23, block=(16, 12, 10), tank=(0.6, 0.5, 0.4), max_bins=1 << 9)
    coarse = V.Engine()
    sc = coarse.add(cfg2, data2, reactions=False)
    coarse.autotune(1)
    assert steps_taken(coarse, sc, cfg2, data2) > 2
    again = V.Engine()
    sa = again.add(cfg2, data2, reactions=False)
    again.autotune(1)
    assert steps_taken(again, sa, cfg2, data2) == 1
    for e in (fine, coarse, again):
        e.close()


def test_unclamped_pressure_bit_exact():
    """tensile_instability_suppression off (config.rs:16): negative pressures
    and negative c2 in the density record, every force kernel, 6 steps against
    the oracle bit for bit."""
    cfg, data, tank, d = small_scene(seed=17, block=(14, 10, 9), tank=(0.5, 0.4, 0.4))
    cfg.tensile_clamp = 0
    for fk in (V.FORCE_SCAN, V.FORCE_WALK, V.FORCE_WALK_AHEAD, V.FORCE_PAIRS, V.FORCE_QUEUE, V.FORCE_PAIRED,
                   V.FORCE_STAGED):
        orc = oracle_dam_break(cfg, data, tank, d)
        eng = V.Engine(diagnostics=True)
        sid = eng.add(cfg, data, reactions=False)
        eng.set_tuning(force_kernel=fk)
        for _ in range(6):
            eng.step(1)
            orc.step(cfg.dt_tick)
        f = eng.read_fields(sid, diagnostics=True)
        assert (orc.press() < 0).any()  # the branch the clamp would cut
        np.testing.assert_array_equal(bits(f["press"]), bits(orc.press()))
        np.testing.assert_array_equal(bits(f["acc"]), bits(orc.acc()))
        assert eng.state_hash(sid) == orc.state_hash()
        eng.close()


def test_coincident_particles_bit_exact():
    """Particles at one point (r = 0: the 1e-9 floor of the force denominator,
    fluid_force.rs:165-217; ties in the cell sort broken by index): a fluid
    pair, a cluster of four, and a fluid particle on a tank particle, through
    every force kernel, 3 steps against the oracle bit for bit, neighbour
    counts included."""
    cfg, data, tank, d = small_scene(seed=23, block=(10, 9, 8), tank=(0.5, 0.4, 0.4))
    pos = data.pos.copy()
    nf = data.n_fluid
    pos[5] = pos[6]
    pos[20:23] = pos[23]
    near_b = nf + int(np.argmin(np.linalg.norm(data.pos[nf:] - pos[40], axis=1)))
    pos[40] = pos[near_b]
    data = V.SceneData(pos, data.vel, data.pmass, nf, data.spans)
    for fk in (V.FORCE_SCAN, V.FORCE_WALK, V.FORCE_WALK_AHEAD, V.FORCE_PAIRS, V.FORCE_QUEUE, V.FORCE_PAIRED,
                   V.FORCE_STAGED):
        orc = oracle_dam_break(cfg, data, tank, d)
        eng = V.Engine(diagnostics=True)
        sid = eng.add(cfg, data, reactions=False)
        eng.set_tuning(force_kernel=fk)
        eng.step(1)
        orc.step(cfg.dt_tick)
        f = eng.read_fields(sid, diagnostics=True)
        np.testing.assert_array_equal(f["ncount"], orc.ncount())
        np.testing.assert_array_equal(bits(f["dens"]), bits(orc.dens()))
        np.testing.assert_array_equal(bits(f["acc"]), bits(orc.acc()))
        np.testing.assert_array_equal(bits(f["xsph"]), bits(orc.xsph()))
        for _ in range(2):
            eng.step(1)
            orc.step(cfg.dt_tick)
        assert eng.state_hash(sid) == orc.state_hash()
        gpos, _ = eng.read_state(sid)
        assert np.all(np.isfinite(gpos))
        eng.close()


def test_speed_clamp_bit_exact():
    """The CFL clamp of fluid_step.rs:110-123 (|v| > vmax = max_speed_frac * h
    / dt -> rescaled to vmax): a block thrown at ~41 m/s against a limit of
    32 m/s (h = 0.04 m, dt = 5e-4 s), 4 steps against the oracle bit for bit."""
    cfg, data, tank, d = small_scene(seed=19, block=(10, 8, 8), tank=(0.6, 0.5, 0.5))
    vel = data.vel.copy()
    vel[:data.n_fluid] = np.array([36.0, -9.0, 18.0], np.float32)
    data = V.SceneData(data.pos, vel, data.pmass, data.n_fluid, data.spans)
    vmax = np.float32(np.float32(cfg.max_speed_frac) * np.float32(cfg.smoothing_radius)) / np.float32(cfg.dt_tick)
    assert np.float32(np.linalg.norm(vel[0])) > vmax
    orc = oracle_dam_break(cfg, data, tank, d)
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    for _ in range(4):
        eng.step(1)
        orc.step(cfg.dt_tick)
    pos, gvel = eng.read_state(sid)
    speed = np.linalg.norm(gvel[:data.n_fluid].astype(np.float64), axis=1)
    assert speed.max() <= vmax * (1 + 1e-5)  # clamped
    np.testing.assert_array_equal(bits(gvel), bits(orc.vel()))
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))


def test_ragged_scene_sizes_in_one_pool():
    """Scenes of 1, 2, 63, 64, 65, 255, 256, 257 and 300 fluid particles (the
    first n points of a jittered lattice), a scene of boundary particles only
    and one fluid particle on a floor, stepped together: each equals its own
    oracle bit for bit (positions, velocities, densities)."""
    rng = np.random.default_rng(23)
    S = np.float32(0.05)
    grid = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)
    lattice = ((grid + 0.5) * S + rng.uniform(-0.004, 0.004, grid.shape)).astype(np.float32)
    cfg = V.fluid_config(substeps=1, particle_mass=0.125, spacing=float(S), dt_tick=1.0 / 120.0)
    mass = np.float32(cfg.particle_mass)
    ocfg = O.default_config(substeps=1)
    floor = dict(id=3, half=(0.3, 0.05, 0.3), pos=(0.2, -0.05, 0.2))
    scenes = []
    for nf in (1, 2, 63, 64, 65, 255, 256, 257, 300):
        pos = lattice[:nf]
        vel = rng.uniform(-0.2, 0.2, (nf, 3)).astype(np.float32)
        scenes.append((V.SceneData(pos, vel, np.full(nf, mass, np.float32), nf, []), None))
    for nf in (0, 1):
        fpos, fvel = lattice[:nf] + np.float32(0.1), np.zeros((nf, 3), np.float32)
        bpos, bvel, bpm, spans = V.box_particles(cfg, [floor], nf)
        data = V.SceneData(np.concatenate([fpos, bpos]), np.concatenate([fvel, bvel]),
                           np.concatenate([np.full(nf, mass, np.float32), bpm]), nf, spans)
        scenes.append((data, floor))
    eng = V.Engine()
    ids, orcs = [], []
    for data, body in scenes:
        ids.append(eng.add(cfg, data, reactions=False))
        nf = data.n_fluid
        o = O.System.particles(ocfg, data.pos[:nf], data.vel[:nf], mass, S)
        if body:
            o.set_boundary_boxes([body])
        np.testing.assert_array_equal(bits(o.pos()), bits(data.pos))
        orcs.append(o)
    eng.step(5)
    for o in orcs:
        for _ in range(5):
            o.step(float(cfg.dt_tick))
    for sid, o, (data, _) in zip(ids, orcs, scenes):
        pos, vel = eng.read_state(sid)
        dens = eng.read_fields(sid)["dens"]
        np.testing.assert_array_equal(bits(pos), bits(o.pos()), err_msg=f"n_fluid {data.n_fluid}")
        np.testing.assert_array_equal(bits(vel), bits(o.vel()), err_msg=f"n_fluid {data.n_fluid}")
        np.testing.assert_array_equal(bits(dens), bits(o.dens()), err_msg=f"n_fluid {data.n_fluid}")


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_positions_abandon_the_substep(bad):
    """A NaN or infinite position makes the scene's particle box non-finite;
    the substep is abandoned on the device and the host reports it (rather
    than binning every particle into one cell)."""
    cfg, data, tank, d = small_scene(seed=3)
    pos = data.pos.copy()
    pos[5, 1] = bad
    eng = V.Engine()
    eng.add(cfg, V.SceneData(pos, data.vel, data.pmass, data.n_fluid, data.spans), reactions=False)
    eng.step(1)
    with pytest.raises(V.VxlSphError, match="not finite"):
        eng.sync()
    eng.close()


def test_batch_equal_alone_and_oracle():
    reqs = V.sample_batch(11, 5, nmin=2000, nmax=30000)
    built = [V.request_scene(r) + (tuple(r.dam.tank), r.dam.d) for r in reqs]
    eng = V.Engine()
    ids = [eng.add(cfg, data, reactions=False) for cfg, data, _, _ in built]
    eng.step(4)
    for sid, (cfg, data, tank, d) in zip(ids, built):
        orc = oracle_dam_break(cfg, data, tank, d, threads=8)
        for _ in range(4):
            orc.step(cfg.dt_tick)
        assert eng.state_hash(sid) == orc.state_hash()


def test_serve_results_equal_alone():
    """Served form: requests arriving over time, admitted as they arrive and
    co-stepped with whatever else is resident; every result is bit-identical
    to the same request stepped alone and to the oracle (requests with boxes
    through the reference's coupled tick order: rigid substeps, Akinci
    particles from the new poses, fluid tick)."""
    from test_gpu_coupled import OracleCoupled

    reqs = V.sample_requests(5, 8, rate=200.0, nmin=1500, nmax=12000, smin=3, smax=12, max_boxes=3)
    assert any(r.nboxes for r in reqs) and any(not r.nboxes for r in reqs)
    stats, results = V.serve(reqs, budget_bytes=1 << 30, time_scale=1.0, builders=2, admit_every=2)
    assert stats["completed"] == len(reqs) and stats["status"] == "OK"
    assert stats["peak_resident"] >= 2  # requests overlapped on the device
    for r, res in zip(reqs, results):
        cfg, data = V.request_scene(r)
        rc, bodies = V.request_bodies(r)
        tank, d = tuple(r.dam.tank), r.dam.d
        eng = V.Engine()
        if bodies:
            sid = eng.add_coupled(cfg, rc, data.pos[:data.n_fluid], bodies)
            statics = [b for b in bodies if not b["density"] > 0]
            boxes = [b for b in bodies if b["density"] > 0]
            orc = OracleCoupled(cfg, data, statics, boxes, tank)
            for _ in range(r.steps):
                orc.tick()
            ohash = orc.sys.state_hash()
        else:
            sid = eng.add(cfg, data, reactions=False)
            orc = oracle_dam_break(cfg, data, tank, d, threads=8)
            for _ in range(r.steps):
                orc.step(cfg.dt_tick)
            ohash = orc.state_hash()
        eng.step(r.steps)
        assert eng.state_hash(sid) == res["state_hash"]
        assert res["state_hash"] == ohash
        eng.close()


def test_serve_residency_limit_and_fifo_admission():
    # Requests long enough (150-200 steps) that a second one is built and
    # waiting while the first still runs, whatever the host's speed.
    reqs = V.sample_requests(6, 6, rate=1000.0, nmin=1500, nmax=4000, smin=150, smax=200)
    stats, results = V.serve(reqs, max_resident=2, builders=2, admit_every=4)
    assert stats["completed"] == len(reqs)
    assert stats["peak_resident"] == 2
    admits = [r["admit_s"] for r in sorted(results, key=lambda r: r["arrival_s"])]
    assert admits == sorted(admits)  # admitted in arrival order


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_golden_fixture_bit_exact(name):
    """The committed oracle fixtures (tests/golden) replayed through the C ABI
    from their stored inputs: grid tables, neighbour counts, density,
    pressure, acceleration, XSPH, positions, velocities and both hashes equal
    bit for bit after every recorded step."""
    meta, arrs = load_golden(name)
    cfg = V.dam_break_config(meta["d"], c0=meta["c0"], alpha=meta["alpha"], eps=meta["eps"], dt=meta["dt"],
                             layers=meta["layers"], max_bins=meta["max_bins"])
    assert bits([cfg.particle_mass])[0] == meta["mass_bits"]
    data = V.SceneData(arrs["in_pos"], arrs["in_vel"], arrs["in_pmass"], meta["n_fluid"], golden_spans(arrs))
    eng = V.Engine(diagnostics=True)
    sid = eng.add(cfg, data, reactions=False)
    last = max(meta["full_steps"] + meta["hash_steps"])
    for step in range(1, last + 1):
        eng.step(1)
        p = f"s{step}_"
        if step not in meta["full_steps"] and step not in meta["hash_steps"]:
            continue
        g = eng.read_grid(sid)
        assert grid_table_hash(g["start"], g["items"]) == int(arrs[p + "grid_hash"]), f"{name} step {step} grid"
        assert eng.state_hash(sid) == int(arrs[p + "state_hash"]), f"{name} step {step} state"
        if step not in meta["full_steps"]:
            continue
        assert bits(g["min"]).tolist() == arrs[p + "grid_min"].tolist()
        assert [int(bits([g["inv"]])[0]), int(bits([g["bin"]])[0])] == arrs[p + "grid_inv_bin"].tolist()
        assert g["dims"] == arrs[p + "grid_dims"].tolist()
        np.testing.assert_array_equal(g["start"], arrs[p + "start"])
        np.testing.assert_array_equal(g["items"], arrs[p + "items"])
        np.testing.assert_array_equal(g["keys"], arrs[p + "keys"])
        f = eng.read_fields(sid, diagnostics=True)
        np.testing.assert_array_equal(f["ncount"], arrs[p + "ncount"])
        pos, vel = eng.read_state(sid)
        got = dict(dens=f["dens"], press=f["press"], acc=f["acc"], xsph=f["xsph"], pos=pos, vel=vel)
        for k, v in got.items():
            np.testing.assert_array_equal(bits(v), arrs[p + k], err_msg=f"{name} step {step} {k}")


def test_grid_box_known_answers_on_device():
    """The device grid box (k_grid_box) on scenes whose particle box is each
    known-answer case of crates/vxl-phys-core/src/grid.rs:59-111 with h > 0
    (h = 0 has no finite kernel constants, so it is covered by the oracle
    test only)."""
    import json
    import os

    with open(os.path.join(GOLDEN_DIR, "grid_box_kat.json")) as fh:
        cases = [c for c in json.load(fh) if c["h"] > 0]
    assert len(cases) == 4
    for c in cases:
        pts = [c["lo"]] if c["lo"] == c["hi"] else [c["lo"], c["hi"]]
        pos = np.array(pts, np.float32)
        # One substep: the grid read back is the one built from these positions.
        cfg = V.fluid_config(smoothing_radius=c["h"], max_bins=c["max_bins"], particle_mass=1.0, spacing=c["h"],
                             substeps=1)
        data = V.SceneData(pos, np.zeros_like(pos), np.ones(len(pos), np.float32), len(pos), [])
        eng = V.Engine()
        sid = eng.add(cfg, data, reactions=False)
        eng.step(1)
        g = eng.read_grid(sid)
        assert g["dims"] == c["dims"], c["what"]
        assert int(bits([g["bin"]])[0]) == c["bin_bits"], c["what"]
        assert bits(g["min"]).tolist() == bits(np.array(c["lo"], np.float32)).tolist(), c["what"]


def test_relayout_updates_the_tick_graph_in_place():
    """Scenes joining and leaving a pool change the tick graph's arguments
    and grid sizes but not its launch sequence: the executable graph is
    updated in place (hipGraphExecUpdate) rather than instantiated again,
    and every scene still equals the same scene stepped alone."""
    specs = [dict(seed=51, block=(8, 6, 6)), dict(seed=52, block=(12, 9, 7)), dict(seed=53, block=(10, 10, 8))]
    scenes = []
    for sp in specs:
        cfg = V.dam_break_config(0.02)
        scenes.append((cfg, V.dam_break(cfg, (0.4, 0.3, 0.3), sp["block"], 0.02, seed=sp["seed"])))
    eng = V.Engine()
    a = eng.add(*scenes[0], reactions=False)
    eng.step(2)
    b = eng.add(*scenes[1], reactions=False)  # relayout: grows the pool
    eng.step(2)
    eng.remove(a)                            # relayout: the pool shrinks to one scene
    eng.step(2)
    c = eng.add(*scenes[2], reactions=False)
    eng.step(2)
    built, updated = eng.graph_stats()
    assert built >= 1 and updated >= 2, (built, updated)
    for sid, (cfg, data), ticks in ((b, scenes[1], 6), (c, scenes[2], 2)):
        alone = V.Engine()
        s = alone.add(cfg, data, reactions=False)
        alone.step(ticks)
        assert eng.state_hash(sid) == alone.state_hash(s)
