"""Pins of the oracle (the CPU restatement of the reference WCSPH step) against
the reference's own recorded values, known-answer tests and invariants
(crates/vxl-phys-fluid/src/tests.rs, boundary.rs tests, core/grid.rs tests,
docs/PLAN-gpu.md grid-table hashes). No GPU."""
import numpy as np
import pytest

import oracle as O
from scenes import bits

MB = 1 << 20


# --- core/grid.rs:59-111 known answers -------------------------------------

def test_grid_box_single_particle():
    p = (1.0, 2.0, 3.0)
    mn, b, d = O.grid_box(p, p, 0.05, MB)
    assert list(mn) == [1.0, 2.0, 3.0] and np.float32(b) == np.float32(0.05) and d == [1, 1, 1]


def test_grid_box_floor_plus_one():
    _, b, d = O.grid_box((0, 0, 0), (0.25, 0.25, 0.25), 0.05, MB)
    assert np.float32(b) == np.float32(0.05) and d == [6, 6, 6]


def test_grid_box_budget_doubles():
    _, b, d = O.grid_box((0, 0, 0), (5, 5, 5), 0.05, 10_000)
    assert d == [13, 13, 13] and np.float32(b) == np.float32(0.4)


def test_grid_box_dim_clamp():
    _, b, d = O.grid_box((0, 0, 0), (1e6, 1e6, 1e6), 0.05, MB)
    assert all(x <= 1 << 14 for x in d) and d[0] * d[1] * d[2] <= MB and b > 0.05


def test_grid_box_tiny_h():
    p = (1.0, 2.0, 3.0)
    _, b, d = O.grid_box(p, p, 0.0, MB)
    assert np.float32(b) == np.float32(1e-6) and d == [1, 1, 1]
    _, b2, _ = O.grid_box((0, 0, 0), (1, 1, 1), 0.0, MB)
    assert b2 > 1e-6 and np.isfinite(b2)


# --- reference grid-table hashes (docs/PLAN-gpu.md:2632, :260, :2824) -------
# FluidConfig::default, n^3 lattice at 0.05 from (-n*0.025, 0.5, -n*0.025),
# 6 ticks x 4 substeps; FNV-1a over start, items, per-particle bin
# (crates/vxl-phys-gpu/examples/gpu_grid_probe.rs:50-68, :163-214).

@pytest.mark.parametrize("n,expected", [(40, 0x69bd920ac578612d), (50, 0xb895c37c2fd663fb)])
def test_reference_grid_table_hash(n, expected):
    cfg = O.default_config(threads=4)
    sp = 0.05
    s = O.System.lattice(cfg, (-n * sp * 0.5, 0.5, -n * sp * 0.5), (n, n, n), sp)
    for _ in range(6):
        s.step(1.0 / 60.0)
    assert s.grid_hash() == expected


@pytest.mark.slow
def test_reference_grid_table_hash_n74():
    cfg = O.default_config(threads=8)
    n, sp = 74, 0.05
    s = O.System.lattice(cfg, (-n * sp * 0.5, 0.5, -n * sp * 0.5), (n, n, n), sp)
    for _ in range(6):
        s.step(1.0 / 60.0)
    assert s.grid_hash() == 0x6308f0e1e5bee72b


# --- reference Akinci canary (docs/PLAN-gpu.md:2704, :2819) -----------------
# crates/vxl-phys-gpu/tests/packet_grid_canary.rs:27-57, :136-145: 8^3 lattice
# + a floor Box of Akinci particles (2 layers), FluidConfig::default, 1 + 60
# ticks x 4 substeps; recorded CPU fluid y_min = -1.81726. Pressure, Monaghan
# viscosity, XSPH, the boundary density/force terms and the kernel constants
# all act in this splash, so it pins the float arithmetic the table hashes
# above cannot (those scenes stay a rest lattice in free fall).

def canary_ymin(fault=0):
    N, SP = 8, np.float32(0.05)
    cfg = O.default_config()
    h = np.float32(cfg.smoothing_radius)
    s = O.System.lattice(cfg, (-N * SP * 0.5, 0.5, -N * SP * 0.5), (N, N, N), SP)
    half = np.float32(N) * SP * np.float32(0.5) + np.float32(4.0) * h
    s.set_boundary_boxes([dict(id=0, half=(half, np.float32(2.0) * SP, half), pos=(0.0, np.float32(-2.0) * h, 0.0))])
    s.set_fault(fault)
    for _ in range(61):
        s.step(1.0 / 60.0)
    return float(s.pos()[:s.n_fluid, 1].min())


def test_reference_akinci_canary_ymin():
    assert round(canary_ymin(), 5) == -1.81726


@pytest.mark.parametrize("fault", [1, 2, 3])
def test_canary_detects_a_planted_mistake(fault):
    """Each planted mistake in the oracle (reassociated poly6 product, dropped
    XSPH term, transposed neighbour order) moves the canary readout."""
    assert round(canary_ymin(fault), 5) != -1.81726


# --- boundary.rs tests --------------------------------------------------------

def test_box_lattice_two_layers_inside():
    half, s = 0.06, 0.05
    pts, n0, vol = O.lattice_box((half,) * 3, s, 2 * s, 2)
    assert n0 == 24 and len(pts) == 48
    total = O.self_sum_bruteforce(pts, 2 * s)
    assert abs(float(vol) * float(total) - 1.0) < 1e-4
    a = np.abs(pts)
    assert np.all(a <= half)
    assert np.all(np.any(a < half, axis=1))


def test_lattice_deterministic():
    a = O.lattice_box((0.07,) * 3, 0.05, 0.1, 2)
    b = O.lattice_box((0.07,) * 3, 0.05, 0.1, 2)
    assert np.array_equal(bits(a[0]), bits(b[0])) and a[2] == b[2]


@pytest.mark.parametrize("half,s,h,layers", [((0.3, 0.05, 0.3), 0.05, 0.1, 2), ((0.5, 0.02, 0.27), 0.02, 0.04, 1),
                                             ((0.2, 0.2, 0.02), 0.02, 0.04, 1)])
def test_self_sum_grid_equals_bruteforce(half, s, h, layers):
    pts, _, _ = O.lattice_box(half, s, h, layers)
    assert bits([O.self_sum_grid(pts, h)])[0] == bits([O.self_sum_bruteforce(pts, h)])[0]


def test_self_sum_closed_form_isolated_points():
    """Points farther apart than h: every per-particle sum is W(0) exactly."""
    h = np.float32(0.1)
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0]], np.float32)
    k6 = np.float32(315.0) / (np.float32(64.0) * np.float32(np.pi) * h ** 9)
    w0 = k6 * (h * h) ** 3
    assert abs(float(O.self_sum_bruteforce(pts, h)) / float(w0) - 1.0) < 1e-6


# --- src/tests.rs invariants -------------------------------------------------

def still_block():
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), substeps=2)
    return O.System.lattice(cfg, (-0.1, 0.2, -0.1), (5, 5, 5), 0.05)


def test_rest_lattice_density_matches_rest_density():
    s = still_block()
    s.step(1.0 / 60.0)
    origin = np.array([-0.1, 0.2, -0.1], np.float32)
    q = (s.pos()[:s.n_fluid] - origin) * np.float32(20.0)
    inner = np.all((q >= 2.4) & (q <= 3.6), axis=1)
    assert inner.sum() == 8
    rel = np.abs(s.dens()[:s.n_fluid][inner] - 1000.0) / 1000.0
    assert np.all(rel < 0.02)


def test_grid_neighbors_equal_brute_force():
    s = still_block()
    s.step(1.0 / 60.0)
    p = s.pos()
    h2 = s.constants()["h2"]
    for i in range(s.n_total):
        grid = np.sort(s.neighbors(i))
        d = p - p[i]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        brute = np.nonzero(r2 <= h2)[0]
        brute = brute[brute != i]
        np.testing.assert_array_equal(grid, brute)


def test_momentum_conserved_without_gravity():
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), xsph_viscosity=0.0, substeps=2)
    s = O.System.lattice(cfg, (-0.125, 0.2, -0.125), (5, 5, 5), 0.05)
    a = np.arange(s.n_fluid, dtype=np.float32)
    v = np.stack([0.3 * np.sin(a), 0.3 * np.cos(a * np.float32(1.3)), 0.3 * np.sin(a * np.float32(0.7))], 1)
    s.set_velocities(v.astype(np.float32))
    p0 = s.vel()[:s.n_fluid].astype(np.float64).sum(0)
    for _ in range(120):
        s.step(1.0 / 60.0)
    p1 = s.vel()[:s.n_fluid].astype(np.float64).sum(0)
    assert np.linalg.norm(p1 - p0) < 1e-3


def test_deterministic_and_parallel_equals_serial():
    def run(threads, boundary):
        cfg = O.default_config(threads=threads)
        s = O.System.lattice(cfg, (-0.15, 0.0, -0.15), (7, 7, 7), 0.05)
        if boundary:
            s.set_boundary_boxes([dict(id=0, half=(0.3, 0.05, 0.3), pos=(0.0, -0.05, 0.0))])
        for _ in range(20):
            s.step(1.0 / 60.0)
        body, ft = s.reactions()
        return np.concatenate([bits(s.pos()).ravel(), bits(s.vel()).ravel(), bits(s.dens()), bits(ft).ravel()])
    for b in (False, True):
        a1 = run(1, b)
        assert np.array_equal(a1, run(1, b))
        assert np.array_equal(a1, run(4, b))


def test_boundary_reaction_balance():
    """Sum m*a over fluid + sum of boundary forces ~ 0 (pairwise
    antisymmetric pressure + viscosity; tests/boundary_reaction_balance.rs)."""
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), xsph_viscosity=0.0, substeps=1)
    s = O.System.lattice(cfg, (-0.15, 0.02, -0.15), (6, 4, 6), 0.05)
    s.set_boundary_boxes([dict(id=0, half=(0.3, 0.05, 0.3), pos=(0.0, -0.05, 0.0))])
    rng = np.random.default_rng(3)
    s.set_velocities(rng.uniform(-0.2, 0.2, (s.n_fluid, 3)).astype(np.float32))
    s.step(1.0 / 240.0)
    ma = (s.pmass()[:s.n_fluid, None].astype(np.float64) * s.acc().astype(np.float64)).sum(0)
    bf = s.bforce().astype(np.float64).sum(0)
    scale = np.abs(s.pmass()[:s.n_fluid, None] * s.acc()).sum()
    assert np.linalg.norm(ma + bf) / scale < 1e-3


def test_reaction_is_sum_of_boundary_forces():
    """Reaction per span = serial sum of that span's bforce (fluid_boundary.rs:10-31)."""
    cfg = O.default_config(substeps=1)
    s = O.System.lattice(cfg, (-0.1, 0.05, -0.1), (4, 4, 4), 0.05)
    s.set_boundary_boxes([dict(id=7, half=(0.25, 0.05, 0.25), pos=(0.0, -0.05, 0.0))])
    s.step(1.0 / 240.0)
    body, ft = s.reactions()
    assert body.tolist() == [7]
    bf = s.bforce()
    acc = np.zeros(3, np.float32)
    for f in bf:
        acc = (acc + f).astype(np.float32)
    assert np.array_equal(bits(acc), bits(ft[0, :3]))


def test_single_pair_closed_form():
    """Two fluid particles at distance r < h: density, pressure, pressure
    acceleration and XSPH from the closed-form kernels."""
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), substeps=1, artificial_viscosity=0.0)
    r = np.float32(0.06)
    pos = np.array([[0, 0.5, 0], [r, 0.5, 0]], np.float32)
    vel = np.array([[0.1, 0, 0], [-0.1, 0, 0]], np.float32)
    m = np.float32(1.0)  # rho > rho0, so the pressure is not clamped
    s = O.System.particles(cfg, pos, vel, m, 0.05)
    s.step(1e-4)
    c = s.constants()
    h = np.float64(c["h"])
    k6 = 315.0 / (64 * np.pi * h ** 9)
    ks = 45.0 / (np.pi * h ** 6)
    w = k6 * (h * h - r * r) ** 3
    rho = m * (k6 * h ** 6 + w)
    assert abs(s.dens()[0] / rho - 1) < 1e-5
    q = rho / 1000.0
    p = max(c["b_tait"] * (q ** 7 - 1), 0.0)
    assert abs(s.press()[0] - p) <= 1e-5 * max(p, 1.0)
    a_expect = m * ks * (h - r) ** 2 * (2 * p / rho ** 2)  # along -x for particle 0 (d = p0 - p1)
    assert abs(s.acc()[0, 0] + a_expect) <= 1e-4 * max(a_expect, 1.0)
    xs = (-0.2) * (2 * m / (2 * rho)) * w
    assert abs(s.xsph()[0, 0] - xs) <= 1e-5 * abs(xs)


def pair_viscosity_error(fault=0):
    """Relative error of the Monaghan viscosity term of an approaching pair
    against its closed form (fluid_force.rs:190-197): the acceleration with
    alpha > 0 minus the one with alpha = 0 is d * (cv / r) with
    cv = m (alpha c mu / rho_bar) ks (h - r)^2, mu = vdn h / (r^2 + 0.01 h^2),
    vdn = -(v_ij . d) > 0 and rho_bar the mean density of the pair."""
    r = np.float32(0.06)
    pos = np.array([[0, 0.5, 0], [r, 0.5, 0]], np.float32)
    vel = np.array([[0.1, 0, 0], [-0.1, 0, 0]], np.float32)  # approaching
    m = np.float32(1.0)
    acc = {}
    for alpha in (0.0, 0.5):
        cfg = O.default_config(gravity=(0.0, 0.0, 0.0), substeps=1, artificial_viscosity=alpha)
        s = O.System.particles(cfg, pos, vel, m, 0.05)
        s.set_fault(fault)
        s.step(1e-4)
        acc[alpha] = s.acc()[0].astype(np.float64)
        rho = np.float64(s.dens()[0])
        c = s.constants()
    h = np.float64(c["h"])
    ks = 45.0 / (np.pi * h ** 6)
    alpha_c = 0.5 * np.float64(cfg.sound_speed)
    rr = np.float64(r)
    vdn = 0.2 * rr  # -(v_ij . d) with v_ij = (0.2, 0, 0), d = (-r, 0, 0)
    mu = vdn * h / (rr * rr + 0.01 * h * h)
    cv = np.float64(m) * (alpha_c * mu / rho) * ks * (h - rr) ** 2
    diff = acc[0.5] - acc[0.0]
    assert abs(diff[1]) == 0.0 and abs(diff[2]) == 0.0
    return abs(diff[0] - (-cv)) / cv


def test_approaching_pair_viscosity_closed_form():
    """Monaghan viscosity (alpha > 0) of an approaching pair against the
    closed form; a planted mistake (the density sum in place of its mean)
    fails it."""
    assert pair_viscosity_error() < 1e-4
    assert pair_viscosity_error(fault=6) > 0.4


def clamp_case(v0, fault=0):
    """One isolated fluid particle (no neighbours, no gravity) moving at v0:
    fluid_step.rs:110-123 keeps its velocity below vmax = frac h / dt."""
    cfg = O.default_config(gravity=(0.0, 0.0, 0.0), substeps=1)
    s = O.System.particles(cfg, np.array([[0.0, 0.5, 0.0]], np.float32), np.array([v0], np.float32),
                           np.float32(1.0), 0.05)
    s.set_fault(fault)
    dt = 1e-3
    s.step(dt)
    c = s.constants()
    vmax = float(cfg.max_speed_frac) * float(c["h"]) / dt
    return s.vel()[0].astype(np.float64), vmax


def test_speed_clamp_closed_form():
    """Faster than vmax: scaled to |v| = vmax along the same direction;
    slower: untouched. A planted mistake (clamp skipped) fails it."""
    v0 = np.array([30.0, 40.0, 0.0])
    v, vmax = clamp_case(v0)
    assert vmax == pytest.approx(40.0, rel=1e-6)
    assert abs(np.linalg.norm(v) / vmax - 1.0) < 1e-6
    assert np.allclose(v / np.linalg.norm(v), v0 / np.linalg.norm(v0), atol=1e-6)
    slow = np.array([3.0, -4.0, 1.0])
    v_slow, _ = clamp_case(slow)
    np.testing.assert_array_equal(v_slow.astype(np.float32), slow.astype(np.float32))
    v_bad, _ = clamp_case(v0, fault=7)
    assert abs(np.linalg.norm(v_bad) / vmax - 1.0) > 0.2


def test_time_weighted_reactions_over_substeps():
    """With substeps > 1 the reaction is the mean over substeps
    (fluid_step.rs:75-89; src/tests.rs:668-712)."""
    base = dict(gravity=(0.0, -9.81, 0.0))
    cfg4 = O.default_config(substeps=4, **base)
    s4 = O.System.lattice(cfg4, (-0.1, 0.02, -0.1), (4, 4, 4), 0.05)
    s4.set_boundary_boxes([dict(id=0, half=(0.25, 0.05, 0.25), pos=(0.0, -0.05, 0.0))])
    cfg1 = O.default_config(substeps=1, **base)
    s1 = O.System.lattice(cfg1, (-0.1, 0.02, -0.1), (4, 4, 4), 0.05)
    s1.set_boundary_boxes([dict(id=0, half=(0.25, 0.05, 0.25), pos=(0.0, -0.05, 0.0))])
    acc = np.zeros(6, np.float32)
    for k in range(4):
        s1.step(1.0 / 240.0)
        ft = s1.reactions()[1][0]
        acc = ft.copy() if k == 0 else (acc + ft).astype(np.float32)
    mean = (acc * np.float32(0.25)).astype(np.float32)
    s4.step(1.0 / 60.0)
    np.testing.assert_array_equal(bits(s4.reactions()[1][0]), bits(mean))


def test_skipped_reactions_leave_the_fluid_bit_identical():
    """The boundary-force pass only feeds the bodies' reactions: skipping it
    (the benchmark's static-tank line) changes no fluid bit and zeroes the
    reaction, while the full pass gives a nonzero one."""
    def run(skip):
        cfg = O.default_config(substeps=2, skip_reactions=skip)
        s = O.System.lattice(cfg, (-0.1, 0.02, -0.1), (4, 4, 4), 0.05)
        s.set_boundary_boxes([dict(id=3, half=(0.25, 0.05, 0.25), pos=(0.0, -0.05, 0.0))])
        for _ in range(3):
            s.step(1.0 / 240.0)
        return s
    full, skipped = run(0), run(1)
    for k in ("pos", "vel", "dens", "press"):
        assert np.array_equal(bits(getattr(full, k)()), bits(getattr(skipped, k)())), k
    assert np.any(full.reactions()[1][0] != 0)
    assert not np.any(skipped.reactions()[1][0])
