"""GPU parity at the stated workloads and through the drop-in paths the other
files do not reach:

* the batched form's own scene set (64 dam breaks, fluid counts log-uniform in
  [8192, 524288], the set the benchmark steps), every scene against the
  oracle after two pooled steps (the reference's per-tick CPU-vs-GPU check,
  crates/vxl-phys-gpu/examples/gpu_tick_probe.rs:256-321);
* the served form under an HBM budget that binds, with a request that is too
  large on its own;
* the FluidStepper with a boundary body that moves every tick (positions,
  velocities and span origins uploaded per tick, pipeline/stepper.rs:70-110)
  and its bounds() (pipeline/readback.rs:110-154);
* boundary reactions of particles with more fluid neighbours than one batch
  of the reaction pass holds.
"""
import numpy as np
import pytest

import oracle as O
import vxl_sph as V
from scenes import bits, oracle_config, oracle_dam_break

pytestmark = pytest.mark.gpu

F = np.float32
ORACLE_THREADS = 16


def test_batch_64_scene_set_each_scene_against_oracle():
    """The benchmark's 64-scene set in one pool, two pooled steps: every
    scene's grid tables, densities and state equal the oracle's bit for bit."""
    reqs = V.sample_batch(4, 64)
    built = [V.request_scene(r) + (tuple(r.dam.tank), r.dam.d) for r in reqs]
    nf = [data.n_fluid for _, data, _, _ in built]
    assert len(built) == 64 and min(nf) >= 8192 and max(nf) <= 524288 * 1.05
    eng = V.Engine()
    ids = [eng.add(cfg, data, reactions=False) for cfg, data, _, _ in built]
    eng.step(2)
    for k, (sid, (cfg, data, tank, d)) in enumerate(zip(ids, built)):
        print(f"scene {k}: {data.n_fluid} fluid + {data.n - data.n_fluid} tank", flush=True)
        orc = oracle_dam_break(cfg, data, tank, d, threads=ORACLE_THREADS, own_lattice=False)
        orc.step(cfg.dt_tick)
        orc.step(cfg.dt_tick)
        g = eng.read_grid(sid)
        ostart, oitems, okeys = orc.grid_tables()
        np.testing.assert_array_equal(g["start"], ostart, err_msg=f"scene {k} start")
        np.testing.assert_array_equal(g["items"], oitems, err_msg=f"scene {k} items")
        np.testing.assert_array_equal(g["keys"], okeys, err_msg=f"scene {k} keys")
        f = eng.read_fields(sid)
        np.testing.assert_array_equal(bits(f["dens"]), bits(orc.dens()), err_msg=f"scene {k} density")
        pos, vel = eng.read_state(sid)
        np.testing.assert_array_equal(bits(pos), bits(orc.pos()), err_msg=f"scene {k} positions")
        np.testing.assert_array_equal(bits(vel), bits(orc.vel()), err_msg=f"scene {k} velocities")
        assert eng.state_hash(sid) == orc.state_hash()
        del orc
    eng.close()


def served_alone_hash(r):
    cfg, data = V.request_scene(r)
    rc, bodies = V.request_bodies(r)
    eng = V.Engine()
    if bodies:
        sid = eng.add_coupled(cfg, rc, data.pos[:data.n_fluid], bodies)
    else:
        sid = eng.add(cfg, data, reactions=False)
    eng.step(r.steps)
    h = eng.state_hash(sid)
    eng.close()
    return h


def test_serve_budget_binds_and_oversized_request_is_rejected():
    """An HBM budget below what the stream would use unconstrained: requests
    wait for residents to retire, the device never holds more than the
    budget, a request too large on its own is rejected (CAPACITY, steps = 0)
    while every other request is served, each equal to the request alone."""
    reqs = V.sample_requests(12, 7, rate=400.0, nmin=1500, nmax=20000, smin=20, smax=40, max_boxes=2)
    sizes = [V.request_bytes(r) for r in reqs]
    order = sorted(range(len(reqs)), key=lambda k: sizes[k])
    big, second = order[-1], order[-2]
    assert sizes[big] > sizes[second]
    budget = (sizes[big] + sizes[second]) // 2  # every request but the largest fits alone
    free, _ = V.serve([r for k, r in enumerate(reqs) if k != big], budget_bytes=1 << 40, builders=2, admit_every=2)
    assert free["peak_bytes"] > budget  # unconstrained, the stream would hold more
    stats, results = V.serve(reqs, budget_bytes=budget, builders=2, admit_every=2)
    assert stats["status"] == "CAPACITY"
    assert stats["rejected"] == 1 and results[big]["steps"] == 0
    assert stats["completed"] == len(reqs) - 1
    assert stats["peak_bytes"] <= budget
    for k, (r, res) in enumerate(zip(reqs, results)):
        if k == big:
            continue
        assert res["steps"] == r.steps
        assert res["state_hash"] == served_alone_hash(r), f"request {k}"


def moving_wall_scene(seed=17):
    cfg = V.dam_break_config(0.02)
    tank = (0.5, 0.4, 0.4)
    data = V.dam_break(cfg, tank, (10, 10, 8), 0.02, seed=seed)
    return cfg, data, tank


def test_stepper_moving_body_matches_oracle_with_bounds():
    """The tank floor (body 0, whose particles are the lowest of the scene)
    rises every tick: the stepper gets the moved boundary segment, its
    velocity and the moved span origin each tick, as World::fluid_stepper_pass
    hands them over. State, reactions and bounds() equal the oracle given the
    same boundary each tick, bit for bit; bounds() is the tight box of every
    particle, so its low y follows the floor up (a box kept from the previous
    tick would hold the old floor)."""
    cfg, data, tank = moving_wall_scene()
    nf = data.n_fluid
    spans = list(data.spans)
    floor = [k for k, sp in enumerate(spans) if sp[0] == 0][0]
    _, org, s0, s1 = spans[floor]
    st = V.FluidStepper(cfg)
    orc = O.System.particles(oracle_config(cfg), data.pos[:nf], data.vel[:nf], F(cfg.particle_mass),
                             F(cfg.spacing))
    pos = data.pos.copy()
    vel = data.vel.copy()
    assert pos[s0:s1, 1].min() == pos[:, 1].min()
    step = F(0.002)
    for t in range(6):
        if t:
            pos[s0:s1, 1] = (pos[s0:s1, 1] + step).astype(F)
            vel[s0:s1, 1] = F(step / F(cfg.dt_tick))
            org = (org[0], float(F(F(org[1]) + step)), org[2])
            spans[floor] = (0, org, s0, s1)
        orc.set_boundary_raw(pos[nf:], vel[nf:], data.pmass[nf:], spans)
        orc.step(cfg.dt_tick)
        st.step(cfg.dt_tick, nf, pos, vel, data.pmass, spans)
        gp, gv = st.read_state()
        np.testing.assert_array_equal(bits(gp), bits(orc.pos()), err_msg=f"tick {t} positions")
        np.testing.assert_array_equal(bits(gv), bits(orc.vel()), err_msg=f"tick {t} velocities")
        body, ft = orc.reactions()
        got = st.reactions()
        assert [g[0] for g in got] == body.tolist()
        gft = np.array([np.concatenate([g[1], g[2]]) for g in got], F)
        np.testing.assert_array_equal(bits(gft), bits(ft), err_msg=f"tick {t} reactions")
        lo, hi = st.bounds()
        op = orc.pos()
        np.testing.assert_array_equal(bits(lo), bits(op.min(axis=0)), err_msg=f"tick {t} bounds lo")
        np.testing.assert_array_equal(bits(hi), bits(op.max(axis=0)), err_msg=f"tick {t} bounds hi")
    assert lo[1] == pos[s0:s1, 1].min() > data.pos[s0:s1, 1].min()  # the risen floor is the low y


def test_boundary_reactions_past_one_batch():
    """A plate of boundary particles inside fluid sampled at a third of h:
    each plate particle has about 100 fluid neighbours, more than one batch of
    the reaction pass (16), so its sum takes several passes in index order.
    Reactions and state equal the oracle bit for bit."""
    s = F(0.02)
    cfg = V.fluid_config(smoothing_radius=0.06, particle_mass=float(F(1000.0) * s * s * s), spacing=float(s),
                         substeps=1, dt_tick=2.5e-4, sound_speed=20.0, artificial_viscosity=0.05,
                         xsph_viscosity=0.5, boundary_layers=1)
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij"), -1).reshape(-1, 3)
    fpos = ((g.astype(F) + F(0.5)) * s + rng.uniform(-0.001, 0.001, g.shape).astype(F)).astype(F)
    fvel = rng.uniform(-0.05, 0.05, g.shape).astype(F)
    q = np.stack(np.meshgrid(np.arange(3), np.arange(3), indexing="ij"), -1).reshape(-1, 2)
    bpos = np.zeros((len(q), 3), F)
    bpos[:, 0] = F(0.11) + q[:, 0].astype(F) * s
    bpos[:, 1] = F(0.121)
    bpos[:, 2] = F(0.11) + q[:, 1].astype(F) * s
    bvel = np.zeros_like(bpos)
    nf, nb = len(fpos), len(bpos)
    pos = np.concatenate([fpos, bpos]).astype(F)
    vel = np.concatenate([fvel, bvel]).astype(F)
    pmass = np.full(nf + nb, F(cfg.particle_mass), F)
    pmass[nf:] = F(0.004)
    spans = [(7, (0.13, 0.121, 0.13), nf, nf + nb)]
    orc = O.System.particles(oracle_config(cfg), fpos, fvel, F(cfg.particle_mass), s)
    orc.set_boundary_raw(bpos, bvel, pmass[nf:], spans)
    st = V.FluidStepper(cfg)
    for t in range(3):
        st.step(cfg.dt_tick, nf, pos, vel, pmass, spans)
        orc.step(cfg.dt_tick)
    nbr = [len(orc.neighbors(nf + k)) for k in range(nb)]
    assert min(nbr) > 64, nbr  # several batches per plate particle
    body, ft = orc.reactions()
    got = st.reactions()
    gft = np.array([np.concatenate([g[1], g[2]]) for g in got], F)
    np.testing.assert_array_equal(bits(gft), bits(ft))
    gp, gv = st.read_state()
    np.testing.assert_array_equal(bits(gp), bits(orc.pos()))
    np.testing.assert_array_equal(bits(gv), bits(orc.vel()))
