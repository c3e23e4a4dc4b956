"""The 1M and 10M dam breaks at their full size: the first step(s) bit-exact
against the oracle (grid tables, neighbour counts, every particle's density,
pressure, position and velocity), then the full run on the GPU three times
with identical state hashes, finite state and the fluid inside the tank.
(The oracle would need minutes per hundred steps at these sizes, so the long
runs are checked through these size-independent properties.)"""
import numpy as np
import pytest

import vxl_sph as V
from scenes import bits, oracle_dam_break

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def check_against_oracle(cfg, data, tank, d, steps, diagnostics):
    eng = V.Engine(diagnostics=diagnostics)
    sid = eng.add(cfg, data, reactions=False)
    orc = oracle_dam_break(cfg, data, tank, d, threads=ORACLE_THREADS, own_lattice=False)
    for _ in range(steps):
        eng.step(1)
        orc.step(cfg.dt_tick)
    g = eng.read_grid(sid)
    ostart, oitems, okeys = orc.grid_tables()
    np.testing.assert_array_equal(g["start"], ostart)
    np.testing.assert_array_equal(g["items"], oitems)
    np.testing.assert_array_equal(g["keys"], okeys)
    f = eng.read_fields(sid, diagnostics=diagnostics)
    if diagnostics:
        np.testing.assert_array_equal(f["ncount"], orc.ncount())
    np.testing.assert_array_equal(bits(f["dens"]), bits(orc.dens()))
    np.testing.assert_array_equal(bits(f["press"]), bits(orc.press()))
    pos, vel = eng.read_state(sid)
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))
    np.testing.assert_array_equal(bits(vel), bits(orc.vel()))
    eng.close()


def long_runs(cfg, data, tank, steps, runs=3):
    hashes = []
    for _ in range(runs):
        eng = V.Engine()
        sid = eng.add(cfg, data, reactions=False)
        eng.step(steps)
        hashes.append(eng.state_hash(sid))
        pos, vel = eng.read_state(sid)
        eng.close()
    assert len(set(hashes)) == 1, hashes
    nf = data.n_fluid
    fp, fv = pos[:nf], vel[:nf]
    assert np.all(np.isfinite(fp)) and np.all(np.isfinite(fv))
    h = cfg.smoothing_radius
    lo = np.array([-h, -h, -h], np.float32)
    hi = np.array([tank[0] + h, tank[1] + 10.0, tank[2] + h], np.float32)  # open top
    assert np.all(fp >= lo) and np.all(fp <= hi)
    return hashes[0]


def test_1m_dam_break_full_size():
    """100^3 fluid particles in a 4 x 2 x 2 m tank (~0.2M Akinci particles),
    d = 0.02 m, dt = 5e-4 s: 2 steps against the oracle, 1000 steps x 3."""
    d, tank = 0.02, (4.0, 2.0, 2.0)
    cfg = V.dam_break_config(d)
    data = V.dam_break(cfg, tank, (100, 100, 100), d, seed=2)
    assert data.n_fluid == 1_000_000 and 150_000 < data.n - data.n_fluid < 250_000
    check_against_oracle(cfg, data, tank, d, steps=2, diagnostics=True)
    long_runs(cfg, data, tank, steps=1000)


def test_10m_dam_break_full_size():
    """215^3 fluid particles in a 4 x 2.5 x 2.5 m tank, d = 0.01 m,
    dt = 2.5e-4 s, the reference grid budget 2^21 (bin 2h at this size): one
    step against the oracle, 500 steps x 3."""
    d, tank = 0.01, (4.0, 2.5, 2.5)
    cfg = V.dam_break_config(d, dt=2.5e-4, max_bins=1 << 21)
    data = V.dam_break(cfg, tank, (215, 215, 215), d, seed=1)
    assert data.n_fluid == 9_938_375
    check_against_oracle(cfg, data, tank, d, steps=1, diagnostics=False)
    long_runs(cfg, data, tank, steps=500)


def test_10m_dam_break_benched_budget():
    """The benchmark's configuration of the same scene: grid budget 2^22, which
    keeps bin = h (202 x 126 x 127 cells, 3.2M cells). One step against the
    oracle, bit for bit: cell keys, start, items, density, pressure,
    positions, velocities."""
    d, tank = 0.01, (4.0, 2.5, 2.5)
    cfg = V.dam_break_config(d, dt=2.5e-4, max_bins=1 << 22)
    data = V.dam_break(cfg, tank, (215, 215, 215), d, seed=1)
    check_against_oracle(cfg, data, tank, d, steps=1, diagnostics=False)
