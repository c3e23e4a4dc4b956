"""The reference's Akinci boundary-particle tests (crates/vxl-phys-fluid/src/
tests.rs:387-604, crates/vxl-phys-gpu/tests/frozen_boundary.rs) through the
library: floor density restored, fluid column held, buoyant reaction on a
submerged box in the Tank provider, coupling deterministic, boundary particles
frozen. Each scene also matches the oracle bit for bit."""
import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import bits
from test_reference_boundary_oracle import (BOX_HALF, S, assert_buoyant, buoyancy_scene, floor,
                                            window_mean_reaction)

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def lattice_scene(ocfg, origin, dims, bodies):
    """The oracle lattice (FluidSystem::new) plus Box boundary particles, and
    the same particles as a library scene (mass of the lattice)."""
    orc = O.System.lattice(ocfg, origin, dims, S)
    fpos = orc.pos()
    nf = len(fpos)
    cfg = V.fluid_config(particle_mass=float(orc.mass), spacing=S, dt_tick=float(DT))
    pos, vel, pm, spans = fpos, np.zeros_like(fpos), np.full(nf, orc.mass, np.float32), []
    if bodies:
        orc.set_boundary_boxes(bodies)
        bpos, bvel, bpm, spans = V.box_particles(cfg, bodies, nf)
        pos, vel, pm = np.concatenate([pos, bpos]), np.concatenate([vel, bvel]), np.concatenate([pm, bpm])
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))
    np.testing.assert_array_equal(bits(pm), bits(orc.pmass()))
    return orc, cfg, V.SceneData(pos, vel, pm, nf, spans)


def low_mean(pos, dens, nf, y_max=0.06):
    sel = pos[:nf, 1] < y_max
    assert sel.any()
    return float(dens[:nf][sel].astype(np.float64).mean())


def test_reference_boundary_particles_restore_near_floor_density():
    """tests.rs:387-437, the free and the floored block in one pool."""
    ocfg = O.default_config(threads=2)
    free_o, cfg, free_d = lattice_scene(ocfg, (-0.15, 0.0, -0.15), (7, 7, 7), [])
    on_o, _, on_d = lattice_scene(ocfg, (-0.15, 0.0, -0.15), (7, 7, 7), [floor(0.3)])
    eng = V.Engine()
    a = eng.add(cfg, free_d, reactions=False)
    b = eng.add(cfg, on_d, reactions=False)
    eng.step(1)
    free_o.step(float(DT))
    on_o.step(float(DT))
    nf = free_d.n_fluid
    rho = {}
    for sid, orc, name in ((a, free_o, "free"), (b, on_o, "on")):
        pos, _ = eng.read_state(sid)
        dens = eng.read_fields(sid)["dens"]
        np.testing.assert_array_equal(bits(pos), bits(orc.pos()))
        np.testing.assert_array_equal(bits(dens[:nf]), bits(orc.dens()[:nf]))
        rho[name] = low_mean(pos, dens, nf)
    rho0 = ocfg.rest_density
    assert rho["free"] < 0.85 * rho0
    assert rho["on"] > rho["free"] + 0.10 * rho0
    assert rho["on"] < 1.3 * rho0


def test_reference_boundary_particles_hold_fluid_column():
    """tests.rs:439-490: 60 ticks on a floor of boundary particles."""
    orc, cfg, data = lattice_scene(O.default_config(threads=4), (-0.3, 0.0, -0.3), (12, 12, 10), [floor(0.5)])
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    eng.step(60)
    for _ in range(60):
        orc.step(float(DT))
    pos, _ = eng.read_state(sid)
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))
    p = pos[:data.n_fluid]
    assert np.all(np.isfinite(p))
    core = (np.abs(p[:, 0]) <= 0.25) & (np.abs(p[:, 2]) <= 0.25)
    assert p[core, 1].min() > -0.06


def buoyancy_stepper():
    """The buoyancy scene through the FluidStepper drop-in: the Tank as
    half-space provider walls, the box's particles re-sent every tick as the
    reference re-sets them (tests.rs:511-514)."""
    ocfg = O.default_config(xsph_viscosity=0.05, threads=2)
    orc = buoyancy_scene(ocfg)
    nf = len(O.System.lattice(ocfg, (-0.15, 0.05, -0.15), (7, 7, 7), S).pos())
    cfg = V.fluid_config(xsph_viscosity=0.05)
    body = dict(id=7, half=(BOX_HALF,) * 3, pos=(0.0, 0.12, 0.0))
    bpos, bvel, bpm, spans = V.box_particles(cfg, [body], nf)
    o_pos = orc.pos()
    pos = np.concatenate([o_pos[:nf], bpos])
    vel = np.concatenate([np.zeros((nf, 3), np.float32), bvel])
    pm = np.concatenate([np.full(nf, orc.mass, np.float32), bpm])
    np.testing.assert_array_equal(bits(pos), bits(o_pos))
    st = V.FluidStepper(cfg)
    st.set_halfspace_walls([V.TANK])
    return orc, st, (nf, pos, vel, pm, spans)


def test_reference_submerged_box_gets_buoyant_reaction():
    orc, st, args = buoyancy_stepper()

    def step():
        st.step(float(DT), *args)
        orc.step(float(DT))

    def reaction():
        got = {g[0]: np.concatenate([g[1], g[2]]) for g in st.reactions()}
        body, ft = orc.reactions()
        np.testing.assert_array_equal(bits(got[7]), bits(ft[list(body).index(7)]))
        return got[7]

    ft = window_mean_reaction(step, reaction)
    assert_buoyant(ft, 1000.0)
    pos, _ = st.read_state()
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))


def test_reference_boundary_coupling_is_deterministic():
    """tests.rs:563-604 on the device: two runs of 60 ticks are bit-identical."""
    def digest():
        _, st, args = buoyancy_stepper()
        for _ in range(60):
            st.step(float(DT), *args)
        pos, vel = st.read_state()
        ft = np.array([np.concatenate([g[1], g[2]]) for g in st.reactions()], np.float32)
        st.close()
        return bits(pos), bits(vel), bits(ft)
    a, b = digest(), digest()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_reference_frozen_boundary():
    """crates/vxl-phys-gpu/tests/frozen_boundary.rs: a 13^3 lattice (2197 fluid
    particles, 21 past a multiple of 64) stepped 5 ticks, then a Box of half 2
    spacings submerged at the origin; 3 more ticks of 2 substeps leave every
    boundary particle's position and velocity bit for bit as set, and the
    fluid equal to the oracle."""
    n = 13
    ocfg = O.default_config(substeps=2, threads=2)
    orc = O.System.lattice(ocfg, (-n * S * 0.5, 0.0, -n * S * 0.5), (n, n, n), S)
    for _ in range(5):
        orc.step(float(DT))
    nf = len(orc.pos())
    assert nf % 64 != 0
    body = dict(id=0, half=(2 * S, 2 * S, 2 * S), pos=(0.0, 0.0, 0.0))
    assert orc.set_boundary_boxes([body]) > 0
    cfg = V.fluid_config(substeps=2, particle_mass=float(orc.mass), spacing=S, dt_tick=float(DT))
    pos, vel = orc.pos(), orc.vel()
    pm = orc.pmass()
    eng = V.Engine()
    sid = eng.add(cfg, V.SceneData(pos, vel, pm, nf, []), reactions=False)
    eng.step(3)
    for _ in range(3):
        orc.step(float(DT))
    gpos, gvel = eng.read_state(sid)
    np.testing.assert_array_equal(bits(gpos[nf:]), bits(pos[nf:]))
    np.testing.assert_array_equal(bits(gvel[nf:]), bits(vel[nf:]))
    assert not np.array_equal(bits(gpos[:nf]), bits(pos[:nf]))  # the fluid moved
    np.testing.assert_array_equal(bits(gpos), bits(orc.pos()))
    np.testing.assert_array_equal(bits(gvel), bits(orc.vel()))
