"""The reference's Akinci boundary-particle tests restated on the oracle
(crates/vxl-phys-fluid/src/tests.rs:387-604): a floor of boundary particles
restores the density of the fluid layer on it, holds a fluid column, and a
submerged box in the Tank provider receives a buoyant reaction. Same scenes,
same thresholds as the reference. No GPU."""
import numpy as np

import oracle as O
from scenes import bits

S = 0.05  # lattice spacing of FluidSystem::new in these tests; h = 2 S by default


def floor(half_xz):
    """The reference's static floor Box (top face at y = 0)."""
    return dict(id=0, half=(half_xz, 0.05, half_xz), pos=(0.0, -0.05, 0.0))


def mean_low_density(s, nf, y_max=0.06):
    p, d = s.pos()[:nf], s.dens()[:nf]
    sel = p[:, 1] < y_max
    assert sel.any()
    return float(d[sel].astype(np.float64).mean())


def restore_near_floor(cfg):
    """tests.rs:387-437: one tick without and with a floor of boundary particles."""
    free = O.System.lattice(cfg, (-0.15, 0.0, -0.15), (7, 7, 7), S)
    on = O.System.lattice(cfg, (-0.15, 0.0, -0.15), (7, 7, 7), S)
    nf = len(free.pos())
    assert on.set_boundary_boxes([floor(0.3)]) > 0
    free.step(1.0 / 60.0)
    on.step(1.0 / 60.0)
    return free, on, nf


def test_reference_boundary_particles_restore_near_floor_density():
    cfg = O.default_config(threads=2)
    free, on, nf = restore_near_floor(cfg)
    rho0 = cfg.rest_density
    rho_free, rho_on = mean_low_density(free, nf), mean_low_density(on, nf)
    assert rho_free < 0.85 * rho0
    assert rho_on > rho_free + 0.10 * rho0
    assert rho_on < 1.3 * rho0


def test_reference_boundary_particles_hold_fluid_column():
    """tests.rs:439-490: a 12 x 12 x 10 block on a floor of boundary particles,
    60 ticks; inside |x|, |z| <= 0.25 no particle sinks below -0.06."""
    cfg = O.default_config(threads=4)
    s = O.System.lattice(cfg, (-0.3, 0.0, -0.3), (12, 12, 10), S)
    nf = len(s.pos())
    assert s.set_boundary_boxes([floor(0.5)]) > 0
    for _ in range(60):
        s.step(1.0 / 60.0)
    p = s.pos()[:nf]
    assert np.all(np.isfinite(p))
    core = (np.abs(p[:, 0]) <= 0.25) & (np.abs(p[:, 2]) <= 0.25)
    assert p[core, 1].min() > -0.06


BOX_HALF = 0.06


def buoyancy_scene(cfg):
    """tests.rs:492-520: a 7^3 block in the Tank provider with a fully
    submerged static box (id 7, half 0.06) at (0, 0.12, 0)."""
    s = O.System.lattice(cfg, (-0.15, 0.05, -0.15), (7, 7, 7), S)
    s.set_halfspace_walls([O.TANK])
    assert s.set_boundary_boxes([dict(id=7, half=(BOX_HALF,) * 3, pos=(0.0, 0.12, 0.0))]) > 0
    return s


def window_mean_reaction(step, reaction, settle=180, win=60):
    """The reference's window mean: settle, then average the box's (F, tau)
    over win ticks."""
    for _ in range(settle):
        step()
    acc = np.zeros(6)
    for _ in range(win):
        step()
        acc += reaction().astype(np.float64)
    return acc / win


def assert_buoyant(ft, rho0):
    expect = rho0 * (2.0 * BOX_HALF) ** 3 * 9.81
    fx, fy, fz = ft[:3]
    assert 0.5 * expect < fy < 2.0 * expect, (fy, expect)
    assert abs(fx) < 0.5 * expect and abs(fz) < 0.5 * expect
    assert np.linalg.norm(ft[3:]) < 0.5 * expect * BOX_HALF


def test_reference_submerged_box_gets_buoyant_reaction():
    cfg = O.default_config(xsph_viscosity=0.05, threads=2)
    s = buoyancy_scene(cfg)

    def reaction():
        body, ft = s.reactions()
        return ft[list(body).index(7)]

    ft = window_mean_reaction(lambda: s.step(1.0 / 60.0), reaction)
    assert_buoyant(ft, cfg.rest_density)


def test_reference_boundary_coupling_is_deterministic():
    """tests.rs:563-604: the buoyancy scene twice, 60 ticks: positions,
    densities and reactions bit-identical."""
    def digest():
        s = buoyancy_scene(O.default_config(xsph_viscosity=0.05, threads=2))
        for _ in range(60):
            s.step(1.0 / 60.0)
        return bits(s.pos()), bits(s.dens()), bits(s.reactions()[1])
    a, b = digest(), digest()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
