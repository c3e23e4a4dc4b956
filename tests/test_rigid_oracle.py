"""The rigid-box leg of the oracle (oracle/rigid_oracle.cpp) against closed
forms and invariants: box mass properties (mass.rs:35-48), semi-implicit
Euler free fall (integrate/src/lib.rs:18-64), exact force balance, principal
axis rotation (math.rs:228-240), the speed clamps (config.rs:232-233), and
the tank contact rule (no penetration, boxes come to rest on the floor and
stay inside the walls)."""
import numpy as np
import pytest

import oracle as O

F = np.float32
G = (0.0, -9.81, 0.0)


def test_box_mass_props_closed_form():
    im, ii = O.box_mass_props((0.05, 0.1, 0.15), 1000.0)
    m = F(F(F(F(1000.0) * F(0.1)) * F(0.2)) * F(0.3))
    assert F(1.0) / im == pytest.approx(6.0, rel=1e-6) and im == F(1.0) / m
    ix = F(m / F(12.0)) * F(F(0.2) * F(0.2) + F(0.3) * F(0.3))
    assert ii[0] == F(1.0) / ix
    np.testing.assert_allclose(1.0 / ii, [6 / 12 * 0.13, 6 / 12 * 0.10, 6 / 12 * 0.05], rtol=1e-6)


def test_free_fall_is_semi_implicit_euler():
    dt_tick, sub, n = F(5e-4), 2, 400
    body = O.RigidBoxes([dict(half=(0.1, 0.1, 0.1), pos=(0.0, 10.0, 0.0), density=700.0)])
    cfg = O.rigid_config(dt_tick, sub, G)
    for _ in range(n):
        body.tick(cfg)
    s = body.state(0)
    k = n * sub
    dt = float(dt_tick) / sub
    # v_k = -g k dt, y_k = y0 - g dt^2 k (k + 1) / 2
    assert s["linvel"][1] == pytest.approx(-9.81 * k * dt, rel=1e-5)
    # f32 accumulation over 800 substeps: about one ulp of the position per substep at most
    assert s["pos"][1] == pytest.approx(10.0 - 9.81 * dt * dt * k * (k + 1) / 2, abs=k * 1e-6)
    assert s["pos"][0] == 0.0 and s["pos"][2] == 0.0 and np.all(s["angvel"] == 0)
    assert np.all(s["rot"] == np.array([0, 0, 0, 1], F))


def test_reaction_equal_to_weight_holds_the_box():
    half, dens = (0.1, 0.05, 0.08), 1500.0
    im, _ = O.box_mass_props(half, dens)
    weight = np.array(G, F) * (F(1.0) / im)
    body = O.RigidBoxes([dict(half=half, pos=(1.0, 0.5, 1.0), density=dens)])
    cfg = O.rigid_config(F(5e-4), 2, G)
    react = np.concatenate([-weight, np.zeros(3, F)])[None, :]
    for _ in range(100):
        body.tick(cfg, react)
    s = body.state(0)
    assert np.all(s["linvel"] == 0) and np.all(s["pos"] == np.array([1.0, 0.5, 1.0], F))


def test_principal_axis_rotation_angle():
    w = 3.0
    body = O.RigidBoxes([dict(half=(0.1, 0.2, 0.3), pos=(0, 0, 0), angvel=(0.0, w, 0.0), density=1000.0)])
    cfg = O.rigid_config(F(1e-3), 2, (0.0, 0.0, 0.0))
    n = 500
    for _ in range(n):
        body.tick(cfg)
    q = body.state(0)["rot"]
    angle = 2.0 * np.arctan2(np.linalg.norm(q[:3]), q[3])
    assert angle == pytest.approx(w * n * 1e-3, rel=2e-3)
    assert abs(q[0]) < 1e-6 and abs(q[2]) < 1e-6 and q[1] > 0
    assert np.allclose(body.state(0)["angvel"], [0, w, 0])  # torque-free, principal axis


def test_torque_gives_inverse_inertia_times_dt():
    half, dens = (0.1, 0.2, 0.3), 800.0
    im, ii = O.box_mass_props(half, dens)
    body = O.RigidBoxes([dict(half=half, pos=(0, 0, 0), density=dens)])
    cfg = O.rigid_config(F(1e-3), 1, (0.0, 0.0, 0.0))
    tau = np.array([0.0, 0.0, 2.0], F)
    body.tick(cfg, np.concatenate([np.zeros(3, F), tau])[None, :])
    assert body.state(0)["angvel"][2] == F(F(tau[2] * ii[2]) * F(1e-3))


def test_speed_clamps():
    body = O.RigidBoxes([dict(half=(0.1, 0.1, 0.1), pos=(0, 0, 0), density=1000.0)])
    cfg = O.rigid_config(F(1e-2), 1, (0.0, 0.0, 0.0))
    body.tick(cfg, np.array([[1e9, 0, 0, 0, 1e9, 0]], F))
    s = body.state(0)
    assert np.linalg.norm(s["linvel"]) == pytest.approx(100.0, rel=1e-6)
    assert np.linalg.norm(s["angvel"]) == pytest.approx(50.0, rel=1e-6)


@pytest.mark.parametrize("tilt", [0.0, 0.4])
def test_box_settles_on_the_floor_without_penetration(tilt):
    half = (0.1, 0.05, 0.08)
    c, s = np.cos(tilt / 2), np.sin(tilt / 2)
    rot = (float(s / np.sqrt(2)), 0.0, float(s / np.sqrt(2)), float(c))
    body = O.RigidBoxes([dict(half=half, pos=(1.0, 0.6, 0.5), rot=rot, density=900.0)])
    cfg = O.rigid_config(F(5e-4), 2, G, tank=(2.0, 1.0, 1.0))
    lowest = []
    for _ in range(6000):
        body.tick(cfg)
        st = body.state(0)
        q = st["rot"]
        corners = []
        for k in range(8):
            lp = np.array([half[0] if k & 1 else -half[0], half[1] if k & 2 else -half[1],
                           half[2] if k & 4 else -half[2]])
            qv = q[:3].astype(np.float64)
            t = 2 * np.cross(qv, lp)
            corners.append(st["pos"] + lp + np.cross(qv, t) + q[3] * t)
        lowest.append(min(cc[1] for cc in corners))
    assert min(lowest) > -2e-3  # at most one substep of fall past the floor before the push
    assert lowest[-1] == pytest.approx(0.0, abs=1e-4)
    assert np.linalg.norm(body.state(0)["linvel"]) < 0.05


def test_walls_keep_the_box_inside():
    half = (0.1, 0.1, 0.1)
    body = O.RigidBoxes([dict(half=half, pos=(1.5, 0.1, 0.5), linvel=(5.0, 0.0, -4.0), density=1000.0)])
    cfg = O.rigid_config(F(5e-4), 2, G, tank=(2.0, 1.0, 1.0))
    for _ in range(2000):
        body.tick(cfg)
        p = body.state(0)["pos"]
        assert 0.1 - 1e-2 <= p[0] <= 1.9 + 1e-2 and 0.1 - 1e-2 <= p[2] <= 0.9 + 1e-2
    # Inelastic, frictionless corner impulses: part of the momentum turns into spin, the rest is lost.
    v = body.state(0)["linvel"]
    assert np.linalg.norm(v) < 0.5 * np.linalg.norm([5.0, 4.0])
