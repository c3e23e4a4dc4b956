"""GPU parity of the coupled step: a dam break with dynamic boxes dropped into
the tank. Each tick the oracle follows the reference's order
(crates/vxl-phys/src/world_step.rs:13-123): rigid substeps driven by gravity
and the previous fluid tick's reaction (oracle/rigid_oracle.cpp), the Akinci
particles rebuilt from the new poses (set_boundary_particles,
fluid_step.rs:13-53), then the fluid tick. The device runs the same loop
(k_rigid, k_body_refresh, substeps, k_rigid_collect); fluid state, reactions
and body states must agree bit for bit."""
import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import bits, oracle_config, tank_bodies

pytestmark = pytest.mark.gpu

F = np.float32


def coupled_scene(seed=5, block=(12, 10, 8), tank=(0.6, 0.5, 0.4), d=0.02):
    cfg = V.dam_break_config(d)
    fluid = V.dam_break(cfg, tank, block, d, seed=seed, walls=False)
    statics = [dict(b, density=0.0) for b in tank_bodies(tank, d)]
    c, s = np.cos(0.15), np.sin(0.15)
    boxes = [
        dict(id=100, half=(0.05, 0.04, 0.05), pos=(0.12, 0.27, 0.2), rot=(0.0, 0.0, float(s), float(c)),
             density=500.0),
        dict(id=101, half=(0.06, 0.05, 0.04), pos=(0.42, 0.12, 0.22), rot=(float(s), 0.0, 0.0, float(c)),
             angvel=(0.0, 2.0, 0.0), density=1800.0),
    ]
    return cfg, fluid, statics, boxes, tank, d


class OracleCoupled:
    """The reference tick order on the oracle."""

    def __init__(self, cfg, fluid, statics, boxes, tank):
        nf = fluid.n_fluid
        self.cfg = cfg
        self.sys = O.System.particles(oracle_config(cfg, threads=8), fluid.pos[:nf], fluid.vel[:nf],
                                      F(cfg.particle_mass), F(cfg.spacing))
        self.statics = statics
        self.ids = [b["id"] for b in boxes]
        self.rigid = O.RigidBoxes(boxes)
        self.rcfg = O.rigid_config(F(cfg.dt_tick), 2, (0.0, -9.81, 0.0), tank=tank)
        self.react = np.zeros((len(boxes), 6), F)
        self.halves = [b["half"] for b in boxes]

    def poses(self):
        out = []
        for k, bid in enumerate(self.ids):
            s = self.rigid.state(k)
            out.append(dict(id=bid, half=self.halves[k], pos=tuple(s["pos"]), rot=tuple(s["rot"]),
                            linvel=tuple(s["linvel"]), angvel=tuple(s["angvel"])))
        return out

    def tick(self):
        self.rigid.tick(self.rcfg, self.react)
        self.sys.set_boundary_boxes(self.statics + self.poses())
        self.sys.step(F(self.cfg.dt_tick))
        body, ft = self.sys.reactions()
        for k, bid in enumerate(self.ids):
            self.react[k] = ft[list(body).index(bid)]


def assert_bodies_equal(got, orc):
    for k, g in enumerate(got):
        s = orc.rigid.state(k)
        for f in ("pos", "rot", "linvel", "angvel"):
            np.testing.assert_array_equal(bits(g[f]), bits(s[f]), err_msg=f"body {g['id']} {f}")


def test_coupled_boxes_match_oracle_bit_for_bit():
    cfg, fluid, statics, boxes, tank, d = coupled_scene()
    nf = fluid.n_fluid
    rc = V.rigid_config(tank=tank)
    eng = V.Engine()
    sid = eng.add_coupled(cfg, rc, fluid.pos[:nf], statics + boxes)
    orc = OracleCoupled(cfg, fluid, statics, boxes, tank)
    start = {b["id"]: np.array(b["pos"], F) for b in boxes}
    for chunk in (1, 49, 100, 150):
        eng.step(chunk)
        for _ in range(chunk):
            orc.tick()
        pos, vel = eng.read_state(sid)
        np.testing.assert_array_equal(bits(pos), bits(orc.sys.pos()), err_msg="positions")
        np.testing.assert_array_equal(bits(vel), bits(orc.sys.vel()), err_msg="velocities")
        assert eng.state_hash(sid) == orc.sys.state_hash()
        assert_bodies_equal(eng.read_bodies(sid), orc)
    # The coupling did something: the light box is lifted/pushed by the fluid,
    # the heavy box has fallen onto the floor and rests on it.
    got = {b["id"]: b for b in eng.read_bodies(sid)}
    assert np.linalg.norm(got[100]["pos"] - start[100]) > 1e-3
    heavy = got[101]
    assert heavy["pos"][1] < 0.12 and heavy["pos"][1] > 0.0
    assert np.all(np.abs(orc.react[0]) >= 0) and np.any(orc.react[0] != 0)


def test_coupled_scene_is_unchanged_by_relayouts_and_neighbours():
    """A coupled scene stepped alone equals the same scene in a pool that gains
    and loses another scene mid-run (relayouts move the body state and keep
    the last reaction)."""
    cfg, fluid, statics, boxes, tank, d = coupled_scene(seed=9)
    nf = fluid.n_fluid
    rc = V.rigid_config(tank=tank)
    alone = V.Engine()
    a = alone.add_coupled(cfg, rc, fluid.pos[:nf], statics + boxes)
    alone.step(120)
    mixed = V.Engine()
    other_cfg = V.dam_break_config(0.02, c0=25.0)
    other = V.dam_break(other_cfg, (0.5, 0.4, 0.4), (10, 8, 8), 0.02, seed=3)
    o1 = mixed.add(other_cfg, other, reactions=False)
    m = mixed.add_coupled(cfg, rc, fluid.pos[:nf], statics + boxes)
    mixed.step(40)
    mixed.remove(o1)
    mixed.step(40)
    mixed.add(other_cfg, other, reactions=True)
    mixed.step(40)
    assert mixed.state_hash(m) == alone.state_hash(a)
    for ga, gm in zip(alone.read_bodies(a), mixed.read_bodies(m)):
        for f in ("pos", "rot", "linvel", "angvel"):
            np.testing.assert_array_equal(bits(ga[f]), bits(gm[f]))


def test_vxld_frames_of_a_coupled_scene_and_byte_hash(tmp_path):
    """VXLD v3 frames written from the engine hold the scene's bodies (static
    tank asleep, dynamic boxes awake at their current pose) and its fluid
    positions; the byte-wise state hash equals a Python restatement of
    packet_grid_canary.rs:102-111 over the read-back arrays."""
    from test_vxld import read_vxld

    cfg, fluid, statics, boxes, tank, d = coupled_scene(seed=13)
    nf = fluid.n_fluid
    eng = V.Engine()
    sid = eng.add_coupled(cfg, V.rigid_config(tank=tank), fluid.pos[:nf], statics + boxes)
    path = tmp_path / "coupled.bin"
    w = V.VxldWriter(path, ticks=20, ticks_per_frame=10, dt=cfg.dt_tick, nfluid=1)
    for t in (10, 20):
        eng.step(10)
        w.frame(eng, t, 0.0, [sid])
    w.close()
    dump = read_vxld(path)
    assert [f["tick"] for f in dump["frames"]] == [10, 20]
    last = dump["frames"][-1]
    pos, vel = eng.read_state(sid)
    np.testing.assert_array_equal(last["fluids"][0], pos[:nf])
    assert [b["awake"] for b in last["bodies"]] == [0] * len(statics) + [1] * len(boxes)
    for got, want in zip(last["bodies"][len(statics):], eng.read_bodies(sid)):
        np.testing.assert_array_equal(np.array(got["pos"], F), want["pos"])
        np.testing.assert_array_equal(np.array(got["rot"], F), want["rot"])
        np.testing.assert_array_equal(np.array(got["half"], F), want["half"])
    h = 0xCBF29CE484222325
    for b in np.concatenate([pos.ravel(), vel.ravel()]).astype("<f4").tobytes():
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    assert eng.state_hash_all(sid) == h
