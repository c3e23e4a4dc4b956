"""GPU parity of provider walls against the oracle, bit for bit.

Half-space providers run the CPU engine's semantics on the device every
substep (mirror ghosts summed with each neighbour, projection, coincident
separation: fluid_density.rs:17-51, :76-121, fluid_boundary.rs:45-131).
Plane tables run the GPU pipeline's semantics with tables gathered by the host
once per tick (wall_ghost.wgsl:61-178, pipeline/stepper.rs:118-137). The
providers are the reference's test doubles FloorY and Tank
(fluid/src/tests.rs:11-110)."""
import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import bits, oracle_config

pytestmark = pytest.mark.gpu

F = np.float32
S = 0.05
PROVIDERS = V.Providers({0: V.TANK})


def lattice(origin, dims, spacing=S):
    """FluidSystem::new's lattice positions (fluid_access.rs:14-26)."""
    o = np.asarray(origin, F)
    idx = np.array(list(np.ndindex(*dims)), F)
    return (o + (idx + F(0.5)) * F(spacing)).astype(F)


def configs(**kw):
    """Library and oracle configs with the lattice-calibrated particle mass of
    spacing S (FluidSystem::new, fluid_access.rs:27-41), given explicitly."""
    cfg = V.fluid_config(spacing=S, **kw)
    ocfg = oracle_config(cfg, threads=4)
    mass = F(O.System.lattice(ocfg, (0.0, 0.0, 0.0), (2, 2, 2), S).mass)
    cfg.particle_mass = float(mass)
    return cfg, ocfg, mass


def tank_drop(substeps=4):
    """An 8 x 8 x 8 block filling the Tank's footprint, 0.1 above the floor,
    and a coincident pair about to hit the floor in a corner (projected onto
    one point, then separated)."""
    cfg, ocfg, mass = configs(substeps=substeps, dt_tick=1.0 / 60.0)
    pos = np.concatenate([lattice((-0.2, 0.1, -0.2), (8, 8, 8)),
                          np.array([(0.17, 0.004, -0.17), (0.17, 0.004, -0.17)], F)])
    vel = np.zeros_like(pos)
    vel[-2:] = (0.0, -1.0, 0.0)
    return cfg, ocfg, pos, vel, F(mass)


def make_pair(cfg, ocfg, pos, vel, mass, persist=0):
    eng = V.Engine(diagnostics=True)
    if persist:
        eng.set_tuning(persist_blocks=persist)
    data = V.SceneData(pos, vel, np.full(len(pos), mass, F), len(pos), [])
    sid = eng.add(cfg, data)
    orc = O.System.particles(ocfg, pos, vel, mass, F(S))
    return eng, sid, orc


def assert_same(eng, sid, orc, what):
    pos, vel = eng.read_state(sid)
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()), err_msg=f"{what}: positions")
    np.testing.assert_array_equal(bits(vel), bits(orc.vel()), err_msg=f"{what}: velocities")
    f = eng.read_fields(sid)
    np.testing.assert_array_equal(bits(f["dens"]), bits(orc.dens()), err_msg=f"{what}: density")
    np.testing.assert_array_equal(bits(f["press"]), bits(orc.press()), err_msg=f"{what}: pressure")


@pytest.mark.parametrize("persist", [0, 7])
def test_halfspace_walls_match_oracle_bit_for_bit(persist):
    cfg, ocfg, pos, vel, mass = tank_drop()
    eng, sid, orc = make_pair(cfg, ocfg, pos, vel, mass, persist)
    eng.set_halfspace_walls(sid, [V.TANK])
    orc.set_halfspace_walls([O.TANK])
    pushed = 0
    for chunk in (1, 9, 10, 20):
        eng.step(chunk)
        for _ in range(chunk):
            orc.step(1.0 / 60.0)
            pushed += orc.last_pushed()
        assert_same(eng, sid, orc, f"after {chunk}")
        assert eng.state_hash(sid) == orc.state_hash()
    assert pushed > 0
    p, _ = eng.read_state(sid)
    assert p[:, 1].min() > -0.01 and np.abs(p[:, 0]).max() < 0.21


def test_halfspace_walls_substep_one_and_reference_floor():
    """FloorY under a falling 6^3 block (the reference's rest test scene,
    tests.rs:273-289, shortened), one substep per tick."""
    cfg, ocfg, mass = configs(substeps=1, dt_tick=1.0 / 240.0, xsph_viscosity=0.05)
    pos = lattice((-0.15, 0.06, -0.15), (6, 6, 6))
    eng, sid, orc = make_pair(cfg, ocfg, pos, np.zeros_like(pos), mass)
    floor = V.HalfspaceProvider([(1, 0.0, 1.0)], bounds=((-50.0, -1.0, -50.0), (50.0, 0.0, 50.0)))
    eng.set_halfspace_walls(sid, [floor])
    orc.set_halfspace_walls([O.FLOOR_Y])
    eng.step(120)
    for _ in range(120):
        orc.step(1.0 / 240.0)
    assert_same(eng, sid, orc, "floor")


def test_many_pushed_particles_take_the_global_sort():
    """3600 particles start just below the floor: all are pushed in the first
    substep (more than the LDS list holds), then separated in index order."""
    cfg, ocfg, mass = configs(substeps=1, dt_tick=1.0 / 240.0)
    pos = lattice((-1.5, -0.035, -1.5), (60, 1, 60))
    # coincident partners for some of them, so separations trigger
    pos = np.concatenate([pos, pos[::97]])
    eng, sid, orc = make_pair(cfg, ocfg, pos, np.zeros_like(pos), mass)
    floor = V.HalfspaceProvider([(1, 0.0, 1.0)], bounds=((-50.0, -1.0, -50.0), (50.0, 0.0, 50.0)))
    eng.set_halfspace_walls(sid, [floor])
    orc.set_halfspace_walls([O.FLOOR_Y])
    eng.step(1)
    orc.step(1.0 / 240.0)
    assert orc.last_pushed() == len(pos)
    assert_same(eng, sid, orc, "first tick")
    eng.step(3)
    for _ in range(3):
        orc.step(1.0 / 240.0)
    assert_same(eng, sid, orc, "after 4")


@pytest.mark.parametrize("cluster", [4, 9])
def test_coincident_clusters_separate_like_the_serial_pass(cluster):
    """Clusters of `cluster` coincident particles pushed onto one point: each
    separation depends on the earlier ones' final positions. Clusters of 4
    settle in the parallel rounds; a cluster of 9 outlasts them and the scene
    takes the serial pass. Both must equal the reference's serial order."""
    cfg, ocfg, mass = configs(substeps=1, dt_tick=1.0 / 240.0)
    base = lattice((-0.6, -0.035, -0.6), (24, 1, 24))
    extra = [base[j] for j in range(0, len(base), 53) for _ in range(cluster - 1)]
    pos = np.concatenate([base, np.array(extra, F)])
    eng, sid, orc = make_pair(cfg, ocfg, pos, np.zeros_like(pos), mass)
    floor = V.HalfspaceProvider([(1, 0.0, 1.0)], bounds=((-50.0, -1.0, -50.0), (50.0, 0.0, 50.0)))
    eng.set_halfspace_walls(sid, [floor])
    orc.set_halfspace_walls([O.FLOOR_Y])
    for t in range(3):
        eng.step(1)
        orc.step(1.0 / 240.0)
        assert_same(eng, sid, orc, f"tick {t}")


def test_wall_tables_match_oracle_bit_for_bit():
    """Tables gathered on the host each tick from the device positions (the
    product's gather) and from the oracle's positions (the oracle's gather)."""
    cfg, ocfg, pos, vel, mass = tank_drop(substeps=2)
    eng, sid, orc = make_pair(cfg, ocfg, pos, vel, mass)
    orc.set_halfspace_walls([O.TANK])  # providers for the oracle's gather
    h = F(cfg.smoothing_radius)
    nf = len(pos)
    for tick in range(40):
        dpos, _ = eng.read_state(sid)
        for side, fn in ((0, V.gather_wall_contacts), (1, V.gather_wall_project_contacts)):
            eng.set_wall_tables(sid, side, *fn([0], h, dpos[:nf], PROVIDERS))
            orc.set_wall_tables(side, *orc.gather_wall_tables(side))
        eng.step(1)
        orc.step(1.0 / 60.0)
        if tick % 10 == 9:
            assert_same(eng, sid, orc, f"tick {tick}")
    p, _ = eng.read_state(sid)
    assert p[:, 1].min() > -0.02


def test_stepper_gather_walls_drop_in():
    """FluidStepper::gather_walls + step (world_step/fluid_stepper.rs:66-79)
    through the stepper ABI against the oracle's table path."""
    cfg, ocfg, pos, vel, mass = tank_drop(substeps=2)
    st = V.FluidStepper(cfg)
    st.set_boundaries([0])
    orc = O.System.particles(ocfg, pos, vel, mass, F(S))
    orc.set_halfspace_walls([O.TANK])
    pm = np.full(len(pos), mass, F)
    for tick in range(30):
        st.gather_walls(cfg.smoothing_radius, PROVIDERS)
        for side in (0, 1):
            orc.set_wall_tables(side, *orc.gather_wall_tables(side))
        st.step(cfg.dt_tick, len(pos), pos, vel, pm, [])
        orc.step(1.0 / 60.0)
    gp, gv = st.read_state()
    np.testing.assert_array_equal(bits(gp), bits(orc.pos()))
    np.testing.assert_array_equal(bits(gv), bits(orc.vel()))
    st.close()


def test_walls_in_a_mixed_pool_equal_alone():
    """A half-space scene, a table scene and a scene without walls share one
    pool; a scene is removed and one added mid-run (relayouts rebuild the wall
    tables): every scene ends where it ends alone."""
    cfg, ocfg, pos, vel, mass = tank_drop(substeps=2)
    pm = np.full(len(pos), mass, F)
    data = V.SceneData(pos, vel, pm, len(pos), [])
    h = F(cfg.smoothing_radius)
    plain_cfg = V.dam_break_config(0.02)
    plain_cfg.substeps = cfg.substeps  # one pool shares the substep count
    plain = V.dam_break(plain_cfg, (0.5, 0.4, 0.4), (10, 8, 8), 0.02, seed=3)

    def tables(eng, sid):
        dpos, _ = eng.read_state(sid)
        for side, fn in ((0, V.gather_wall_contacts), (1, V.gather_wall_project_contacts)):
            eng.set_wall_tables(sid, side, *fn([0], h, dpos[:len(pos)], PROVIDERS))

    alone = {}
    e = V.Engine()
    a = e.add(cfg, data)
    e.set_halfspace_walls(a, [V.TANK])
    e.step(30)
    alone["half"] = e.state_hash(a)
    e.close()
    e = V.Engine()
    b = e.add(cfg, data)
    for _ in range(30):
        tables(e, b)
        e.step(1)
    alone["table"] = e.state_hash(b)
    e.close()

    m = V.Engine()
    o1 = m.add(plain_cfg, plain, reactions=False)
    a = m.add(cfg, data)
    m.set_halfspace_walls(a, [V.TANK])
    b = m.add(cfg, data)
    for t in range(30):
        if t == 10:
            m.remove(o1)
        if t == 20:
            m.add(plain_cfg, plain, reactions=True)
        tables(m, b)
        m.step(1)
    assert m.state_hash(a) == alone["half"]
    assert m.state_hash(b) == alone["table"]
    m.clear_walls(a)
    m.step(1)  # walls removed: the pool steps on without them


def test_coupled_boxes_in_a_provider_tank_match_oracle():
    """Dynamic boxes (rigid leg on the device, both-way coupling) in fluid held
    by half-space provider walls instead of static Akinci boxes: the boxes'
    particles take no ghosts and are not projected (fluid_density.rs:53-69,
    fluid_boundary.rs:37-39); fluid, bodies and hashes equal the oracle's."""
    from test_gpu_coupled import OracleCoupled, assert_bodies_equal, coupled_scene

    cfg, fluid, _statics, boxes, tank, d = coupled_scene(seed=21)
    nf = fluid.n_fluid
    walls = V.HalfspaceProvider([(1, 0.0, 1.0), (0, 0.0, 1.0), (0, tank[0], -1.0), (2, 0.0, 1.0),
                                 (2, tank[2], -1.0)])
    eng = V.Engine()
    sid = eng.add_coupled(cfg, V.rigid_config(tank=tank), fluid.pos[:nf], boxes)
    eng.set_halfspace_walls(sid, [walls])
    orc = OracleCoupled(cfg, fluid, [], boxes, tank)
    orc.sys.set_halfspace_walls([dict(id=0, bounds=None, faces=walls.faces)])
    for chunk in (1, 39, 60):
        eng.step(chunk)
        for _ in range(chunk):
            orc.tick()
        pos, vel = eng.read_state(sid)
        np.testing.assert_array_equal(bits(pos), bits(orc.sys.pos()), err_msg="positions")
        np.testing.assert_array_equal(bits(vel), bits(orc.sys.vel()), err_msg="velocities")
        assert eng.state_hash(sid) == orc.sys.state_hash()
        assert_bodies_equal(eng.read_bodies(sid), orc)
    p = pos[:nf]
    assert p[:, 1].min() > -0.01 and p[:, 0].min() > -0.01 and p[:, 0].max() < tank[0] + 0.01
