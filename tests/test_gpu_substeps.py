"""Scenes of one pool with different substeps per tick. A pool tick runs as
many launches as the scene with the most substeps takes; a scene with s
substeps steps in the last s of them and is carried over unchanged in the
others. Each scene must come out as FluidSystem::step gives it alone
(fluid_step.rs:75-127: `substeps` substeps of dt_tick / substeps, reactions
averaged over them): positions, velocities, densities and pressures equal the
oracle's bit for bit, and a coupled scene's bodies (driven by the tick-averaged
reaction) equal the reference tick order's."""
import numpy as np
import pytest

import vxl_sph as V
from scenes import bits, oracle_dam_break
from test_gpu_coupled import OracleCoupled, assert_bodies_equal, coupled_scene
from test_gpu_parity import small_scene

pytestmark = pytest.mark.gpu


def mixed_pool(graphs):
    cfg_a, data_a, tank_a, d = small_scene(seed=41)                    # 1 substep
    cfg_b, data_b, tank_b, _ = small_scene(seed=43, block=(10, 9, 8))  # 2 substeps, reactions on
    cfg_b.substeps = 2
    cfg_b.dt_tick = 1e-3
    cfg_c, fluid_c, statics, boxes, tank_c, _ = coupled_scene(seed=47)  # 4 substeps, dynamic boxes
    cfg_c.substeps = 4
    cfg_c.dt_tick = 1e-3
    eng = V.Engine(graphs=graphs)
    a = eng.add(cfg_a, data_a, reactions=False)
    b = eng.add(cfg_b, data_b, reactions=True)
    nf_c = fluid_c.n_fluid
    c = eng.add_coupled(cfg_c, V.rigid_config(tank=tank_c), fluid_c.pos[:nf_c], statics + boxes)
    plain = [(a, cfg_a, data_a, tank_a, d), (b, cfg_b, data_b, tank_b, d)]
    coupled = (c, cfg_c, fluid_c, statics, boxes, tank_c)
    return eng, plain, coupled


def test_mixed_substep_pool_matches_oracle():
    eng, plain, (c, cfg_c, fluid_c, statics, boxes, tank_c) = mixed_pool(graphs=True)
    ticks = 3
    eng.step(ticks)
    for sid, cfg, data, tank, d in plain:
        orc = oracle_dam_break(cfg, data, tank, d)
        for _ in range(ticks):
            orc.step(cfg.dt_tick)
        pos, vel = eng.read_state(sid)
        what = f"scene with {cfg.substeps} substep(s)"
        np.testing.assert_array_equal(bits(pos), bits(orc.pos()), err_msg=f"{what}: positions")
        np.testing.assert_array_equal(bits(vel), bits(orc.vel()), err_msg=f"{what}: velocities")
        f = eng.read_fields(sid)
        np.testing.assert_array_equal(bits(f["dens"]), bits(orc.dens()), err_msg=f"{what}: densities")
        np.testing.assert_array_equal(bits(f["press"]), bits(orc.press()), err_msg=f"{what}: pressures")
    orc = OracleCoupled(cfg_c, fluid_c, statics, boxes, tank_c)
    for _ in range(ticks):
        orc.tick()
    pos, vel = eng.read_state(c)
    np.testing.assert_array_equal(bits(pos), bits(orc.sys.pos()), err_msg="coupled scene (4 substeps): positions")
    np.testing.assert_array_equal(bits(vel), bits(orc.sys.vel()), err_msg="coupled scene (4 substeps): velocities")
    assert_bodies_equal(eng.read_bodies(c), orc)


def test_mixed_substep_pool_eager_equals_graph():
    hashes = []
    for graphs in (True, False):
        eng, plain, coupled = mixed_pool(graphs)
        eng.step(2)
        hashes.append([eng.state_hash(sid) for sid, *_ in plain] + [eng.state_hash(coupled[0])])
    assert hashes[0] == hashes[1]
