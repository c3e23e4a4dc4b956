"""Pins of the oracle's provider walls (mirror ghosts, projection, coincident
separation, plane tables) against closed forms, the reference's own provider
tests and the two restatements against each other. No GPU.

Providers are the reference's test doubles FloorY and Tank
(crates/vxl-phys-fluid/src/tests.rs:11-110)."""
import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import bits

F = np.float32
S = 0.05  # lattice spacing; h = 2 S in FluidConfig::default


def lattice_in_tank(cfg, dims=(8, 7, 8)):
    """dims lattice filling the Tank's 0.4 x 0.4 footprint at spacing S, the
    bottom layer at S / 2 above the floor (every wall at half a spacing; the
    lattice origin is its corner, particles sit at (i + 1/2) S)."""
    return O.System.lattice(cfg, (-0.2, 0.0, -0.2), dims, S)


def closed_form_density(idx, dims, h, k6, mass, planes):
    """Density of lattice particle idx when every band plane sits half a
    spacing beyond the lattice: the real neighbours plus, for each plane within
    h of the particle, the mirror images of the neighbours, which are exactly
    the lattice points one step beyond that plane and no other plane, except
    the particle's own image. planes: per axis, (low wall?, high wall?)."""
    i = np.array(idx)
    n = np.array(dims)
    pos = (i + 0.5) * S  # distance of the particle to the low walls
    band = [(planes[a][0] and pos[a] < h, planes[a][1] and (n[a] * S - pos[a]) < h) for a in range(3)]
    tot = 0.0
    r = int(np.ceil(h / S)) + 1
    for o in np.ndindex(2 * r + 1, 2 * r + 1, 2 * r + 1):
        o = np.array(o) - r
        r2 = float(np.sum((o * S) ** 2))
        if r2 > h * h:
            continue
        t = i + o
        beyond = []
        ok = True
        for a in range(3):
            if t[a] < 0:
                if band[a][0]:
                    beyond.append((a, 0))
                else:
                    ok = False
            elif t[a] >= n[a]:
                if band[a][1]:
                    beyond.append((a, 1))
                else:
                    ok = False
        if not ok or len(beyond) > 1:
            continue
        if len(beyond) == 1:
            a, side = beyond[0]
            q = t.copy()
            q[a] = -1 - t[a] if side == 0 else 2 * n[a] - 1 - t[a]
            if np.array_equal(q, i):
                continue  # the particle's own image is not a neighbour's ghost
        tt = h * h - r2
        tot += k6 * tt ** 3
    return mass * tot


def test_mirror_ghost_density_closed_form_in_tank():
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), threads=2)
    dims = (8, 7, 8)
    s = lattice_in_tank(cfg, dims)
    s.set_halfspace_walls([O.TANK])
    s.step(0.0)  # zero step: positions stay, the density pass runs
    c = s.constants()
    h, k6, mass = float(c["h"]), float(c["k6"]), float(s.mass)
    dens = s.dens()
    # Lattice index order of FluidSystem::new (x outer, then y, then z).
    planes = [(True, True), (True, False), (True, True)]
    worst = 0.0
    for q, idx in enumerate(np.ndindex(*dims)):
        want = closed_form_density(idx, dims, h, k6, mass, planes)
        worst = max(worst, abs(float(dens[q]) - want) / want)
    assert worst < 1e-5, worst
    # The interior of the lattice reads rest density; the wall layers too
    # (up to the missing own image): the ghosts did fill the kernel support.
    assert abs(float(dens[dims[1] * dims[2] * 3 + 3 * dims[2] + 3]) - 1000.0) < 0.05


# Tank faces as (axis, plane, fluid-side sign).
TANK_FACES = [(1, 0.0, 1.0), (0, -0.2, 1.0), (0, 0.2, -1.0), (2, -0.2, 1.0), (2, 0.2, -1.0)]


def reflected_density(pos, i, h, k6, mass, faces=TANK_FACES):
    """Density of fluid particle i (all particles fluid) in f64 geometry: its
    neighbours within h, and for every wall within h of it (0 < sdf < h) the
    neighbours' mirror images x[axis] -> 2 plane - x[axis] that fall within h."""
    p = pos[i].astype(np.float64)
    d2 = np.sum((pos.astype(np.float64) - p) ** 2, axis=1)
    nb = np.nonzero((d2 <= h * h) & (np.arange(len(pos)) != i))[0]
    band = [(a, pl) for (a, pl, sg) in TANK_FACES if 0.0 < (p[a] - pl) * sg < h]
    W = lambda r2: k6 * (h * h - r2) ** 3
    tot = W(0.0) + sum(W(d2[j]) for j in nb)
    for (a, pl) in band:
        g = pos[nb].astype(np.float64)
        g[:, a] = 2.0 * pl - g[:, a]
        r2 = np.sum((g - p) ** 2, axis=1)
        tot += sum(W(x) for x in r2 if x <= h * h)
    return mass * tot


@pytest.mark.parametrize("fault", [0, 4])
def test_mirror_ghosts_are_reflections_in_the_wall(fault):
    """A lattice off the half-spacing grid (walls at 0.3 and 0.7 spacings), so
    a ghost reflected the wrong way lands elsewhere: the density of every
    particle against the geometric mirror images. The planted wrong reflection
    (fault 4) must fail."""
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), threads=2)
    s = O.System.lattice(cfg, (-0.2 - 0.2 * S, -0.2 * S, -0.2 - 0.2 * S), (8, 6, 8), S)
    s.set_halfspace_walls([O.TANK])
    s.set_fault(fault)
    s.step(0.0)
    c = s.constants()
    h, k6, mass = float(c["h"]), float(c["k6"]), float(s.mass)
    pos, dens = s.pos(), s.dens()
    worst = max(abs(float(dens[i]) - reflected_density(pos, i, h, k6, mass)) / 1000.0 for i in range(len(pos)))
    if fault == 0:
        assert worst < 1e-5, worst
    else:
        assert worst > 1e-3, worst


def one_particle(cfg, p, v):
    return O.System.particles(cfg, np.array([p], F), np.array([v], F), F(0.01), F(S))


def test_projection_pushes_to_the_rest_line_and_stops_inward_motion():
    """boundary_pass (fluid_boundary.rs:45-104): a particle that ends the
    integration below the floor is put on the rest line sdf = skin = 0.15 h
    with its inward normal velocity removed; tangential velocity is kept."""
    cfg = O.default_config(substeps=1)
    s = one_particle(cfg, (0.0, 0.01, 0.0), (0.3, -1.0, 0.2))
    s.set_halfspace_walls([O.FLOOR_Y])
    s.step(0.02)
    p, v = s.pos()[0], s.vel()[0]
    skin = F(0.15) * F(0.1)
    assert abs(p[1] - skin) < 2e-8
    assert v[1] == 0.0 and v[0] == F(0.3) and v[2] == F(0.2)
    assert s.last_pushed() == 1


def test_projection_of_a_deep_penetration_is_capped_at_h():
    cfg = O.default_config(substeps=1)
    s = one_particle(cfg, (0.0, -0.5, 0.0), (0.0, 0.0, 0.0))
    s.set_halfspace_walls([O.FLOOR_Y])
    s.step(0.02)
    g = F(-9.81) * F(0.02)
    y_int = F(F(-0.5) + F(g * F(0.02)))
    assert s.pos()[0][1] == F(y_int + F(0.1))
    assert s.vel()[0][1] == 0.0


def test_projection_keeps_outward_velocity_and_pushes_each_wall_of_a_corner():
    cfg = O.default_config(substeps=1, gravity=(0.0, 0.0, 0.0))
    # Below the floor, beyond the left wall, moving up and further left.
    s = one_particle(cfg, (-0.205, -0.004, 0.0), (-0.05, 0.5, 0.0))
    s.set_halfspace_walls([O.TANK])
    s.step(0.001)
    p, v = s.pos()[0], s.vel()[0]
    skin = F(0.015)
    assert abs(p[0] - F(-0.2 + skin)) < 2e-8 and abs(p[1] - skin) < 2e-8
    assert v[0] == 0.0 and v[1] == F(0.5)  # upward (outward) motion is kept


def test_coincident_separation_moves_the_later_particle_by_r_min():
    """separate_coincident (fluid_boundary.rs:110-131): two particles projected
    onto the same point; the higher index moves 0.2 h along its normal sum."""
    cfg = O.default_config(substeps=1)
    pos = np.array([(0.0, 0.01, 0.0), (0.0, 0.01, 0.0)], F)
    vel = np.array([(0.0, -1.0, 0.0), (0.0, -1.0, 0.0)], F)
    s = O.System.particles(cfg, pos, vel, F(0.01), F(S))
    s.set_halfspace_walls([O.FLOOR_Y])
    s.step(0.02)
    p = s.pos()
    skin = F(0.015)
    assert abs(p[0][1] - skin) < 2e-8
    assert abs((p[1][1] - p[0][1]) - F(0.2) * F(0.1)) < 2e-8
    assert p[1][0] == p[0][0] and p[1][2] == p[0][2]
    s2 = O.System.particles(cfg, pos, vel, F(0.01), F(S))
    s2.set_halfspace_walls([O.FLOOR_Y])
    s2.set_fault(5)
    s2.step(0.02)
    assert s2.pos()[0][1] > s2.pos()[1][1]  # the planted order moves the wrong one


def test_tables_gathered_from_the_providers_match_the_provider_path():
    """The GPU path's plane tables (gathered once per tick) against the CPU
    path's per-substep queries, one substep per tick: the ghost sums differ
    only in summation order (wall_ghost.wgsl:1-6), the projection only where a
    plane's push changes another plane's sdf (never for orthogonal faces)."""
    cfg = O.default_config(substeps=1, threads=2)
    a = O.System.lattice(cfg, (-0.2, 0.01, -0.2), (8, 6, 8), S)
    b = O.System.lattice(cfg, (-0.2, 0.01, -0.2), (8, 6, 8), S)
    a.set_halfspace_walls([O.TANK])
    b.set_halfspace_walls([O.TANK])
    for tick in range(30):
        for side in (0, 1):
            b.set_wall_tables(side, *b.gather_wall_tables(side))
        a.step(1.0 / 240.0)
        b.step(1.0 / 240.0)
        da, db = a.dens(), b.dens()
        assert np.max(np.abs(da - db) / 1000.0) < 1e-4
    assert np.max(np.abs(a.pos() - b.pos())) < 1e-4
    assert a.state_hash() != 0


def test_gather_is_the_same_in_the_product_and_the_oracle():
    """vxl_sph.gather_wall_contacts (host code of the product) against the
    oracle's gather_in restatement, on a lattice straddling the Tank walls."""
    cfg = O.default_config(threads=1)
    s = O.System.lattice(cfg, (-0.23, -0.02, -0.22), (10, 5, 9), S)
    s.set_halfspace_walls([O.TANK])
    ps = s.pos()[:s.n_fluid]
    h = s.constants()["h"]
    provs = V.Providers({0: V.TANK})
    for side, fn in ((0, V.gather_wall_contacts), (1, V.gather_wall_project_contacts)):
        oi, ost, opl = s.gather_wall_tables(side)
        pi, pst, ppl = fn([0], h, ps, provs)
        assert len(oi) > 0
        np.testing.assert_array_equal(oi, pi)
        np.testing.assert_array_equal(ost, pst)
        np.testing.assert_array_equal(bits(opl), bits(ppl))
    # the project table also holds penetrating particles (sdf <= 0)
    assert len(s.gather_wall_tables(1)[0]) > len(s.gather_wall_tables(0)[0])


def test_reference_particles_rest_on_boundary_without_penetration():
    """tests.rs:273-289: a 5^3 block dropped on FloorY, 300 ticks."""
    cfg = O.default_config(xsph_viscosity=0.05, threads=2)
    s = O.System.lattice(cfg, (-0.1, 1.2, -0.1), (5, 5, 5), S)
    s.set_halfspace_walls([O.FLOOR_Y])
    for _ in range(300):
        s.step(1.0 / 60.0)
    y = s.pos()[:, 1]
    assert np.all(y > -0.01) and np.all(y < 0.6)


def test_reference_deterministic_bitwise_on_floor():
    """tests.rs:252-270: two runs of 300 ticks (substeps 2) are bit-identical."""
    def run():
        cfg = O.default_config(substeps=2, threads=2)
        s = O.System.lattice(cfg, (-0.1, 1.0, -0.1), (5, 5, 5), S)
        s.set_halfspace_walls([O.FLOOR_Y])
        for _ in range(300):
            s.step(1.0 / 60.0)
        return bits(s.pos())
    np.testing.assert_array_equal(run(), run())


@pytest.mark.slow
def test_reference_hydrostatic_pressure_increases_with_depth():
    """tests.rs:151-221: a 7^3 column in the Tank settles; pressure rises with
    depth within the reference's bands."""
    cfg = O.default_config(xsph_viscosity=0.05, threads=4)
    s = O.System.lattice(cfg, (-0.15, 0.05, -0.15), (7, 7, 7), S)
    s.set_halfspace_walls([O.TANK])
    for _ in range(240):
        s.step(1.0 / 60.0)
    p = s.pos()
    pr = s.press()
    ymin, ymax = p[:, 1].min(), p[:, 1].max()
    height = ymax - ymin
    assert 0.20 < height < 0.32
    t = (p[:, 1] - ymin) / height
    lo, mid, hi = pr[t < 1 / 3].mean(), pr[(t >= 1 / 3) & (t < 2 / 3)].mean(), pr[t >= 2 / 3].mean()
    assert lo > mid > hi
    mid_ref = 1000.0 * 9.81 * 0.5 * height
    assert mid_ref * 0.4 < mid < mid_ref * 1.1
    assert lo < 2.0 * 1000.0 * 9.81 * height * 5.0 / 6.0
    assert hi * 3.0 < mid
    assert np.all(np.sum(s.vel() ** 2, axis=1) < 9.0)
