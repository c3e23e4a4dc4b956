"""Golden fixtures (tests/golden, written by make_golden.py): the oracle
replays every committed fixture bit for bit from its stored inputs, the
Python table hash equals the oracle's, and the grid box rule meets the known
answers of crates/vxl-phys-core/src/grid.rs:59-111. The GPU side of the same
fixtures is in test_gpu_parity.py."""
import json
import os

import numpy as np
import pytest

import oracle as O
from scenes import GOLDEN_DIR, GOLDEN_SCENES, golden_spans, grid_table_hash, load_golden

F = np.float32


def oracle_from_golden(meta, arrs, threads=4):
    d = meta["d"]
    cfg = O.default_config(smoothing_radius=float(F(2.0) * F(d)), sound_speed=meta["c0"],
                           artificial_viscosity=meta["alpha"], xsph_viscosity=meta["eps"], substeps=1,
                           boundary_layers=meta["layers"], max_bins=meta["max_bins"], threads=threads)
    nf = meta["n_fluid"]
    mass = np.array(meta["mass_bits"], np.uint32).view(F)
    s = O.System.particles(cfg, arrs["in_pos"][:nf], arrs["in_vel"][:nf], mass, F(d))
    s.set_boundary_raw(arrs["in_pos"][nf:], arrs["in_vel"][nf:], arrs["in_pmass"][nf:], golden_spans(arrs))
    return s


def u(a):
    return np.ascontiguousarray(a, F).view(np.uint32)


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_oracle_replays_golden_fixture(name):
    meta, arrs = load_golden(name)
    s = oracle_from_golden(meta, arrs)
    assert s.n_total == meta["n"] and s.n_fluid == meta["n_fluid"]
    np.testing.assert_array_equal(u(s.pmass()), u(arrs["in_pmass"]))
    last = max(meta["full_steps"] + meta["hash_steps"])
    for step in range(1, last + 1):
        s.step(F(meta["dt"]))
        p = f"s{step}_"
        if step in meta["full_steps"] or step in meta["hash_steps"]:
            assert s.grid_hash() == int(arrs[p + "grid_hash"])
            assert s.state_hash() == int(arrs[p + "state_hash"])
        if step in meta["full_steps"]:
            mn, inv, b, dims = s.grid_info()
            start, items, keys = s.grid_tables()
            np.testing.assert_array_equal(u(mn), arrs[p + "grid_min"])
            np.testing.assert_array_equal(u(np.array([inv, b], F)), arrs[p + "grid_inv_bin"])
            assert dims == arrs[p + "grid_dims"].tolist()
            np.testing.assert_array_equal(start, arrs[p + "start"])
            np.testing.assert_array_equal(items, arrs[p + "items"])
            np.testing.assert_array_equal(keys, arrs[p + "keys"])
            np.testing.assert_array_equal(s.ncount(), arrs[p + "ncount"])
            for f in ("dens", "press", "acc", "xsph", "pos", "vel"):
                np.testing.assert_array_equal(u(getattr(s, f)()), arrs[p + f], err_msg=f"{name} step {step} {f}")


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_python_table_hash_equals_oracle(name):
    meta, arrs = load_golden(name)
    step = meta["full_steps"][0]
    p = f"s{step}_"
    assert grid_table_hash(arrs[p + "start"], arrs[p + "items"]) == int(arrs[p + "grid_hash"])


def test_golden_fixture_sanity():
    """Fixture invariants a plausible generator mistake would break: masses
    (fluid = 1000 d^3, boundary = rho0 V_b > 0), the fluid prefix inside the
    tank, one span per tank body covering the boundary segment, neighbour
    counts of an interior jittered lattice near 4/3 pi (h/d)^3 = 33."""
    for name in GOLDEN_SCENES:
        meta, arrs = load_golden(name)
        nf, n, d = meta["n_fluid"], meta["n"], meta["d"]
        mass = np.array(meta["mass_bits"], np.uint32).view(F)
        assert mass == F(F(F(F(1000.0) * F(d)) * F(d)) * F(d))
        assert np.all(arrs["in_pmass"][:nf] == mass) and np.all(arrs["in_pmass"][nf:] > 0)
        fp = arrs["in_pos"][:nf]
        assert fp.min() > 0 and np.all(fp.max(0) < np.array(meta["tank"], F))
        r = arrs["in_span_range"]
        assert r[0, 0] == nf and r[-1, 1] == n and np.all(r[1:, 0] == r[:-1, 1]) and len(r) == 5
        nc = arrs[f"s{meta['full_steps'][0]}_ncount"][:nf]
        assert 25 <= int(np.median(nc)) <= 40


def test_grid_box_known_answers():
    with open(os.path.join(GOLDEN_DIR, "grid_box_kat.json")) as f:
        cases = json.load(f)
    assert len(cases) == 6
    for c in cases:
        mn, b, dims = O.grid_box(c["lo"], c["hi"], c["h"], c["max_bins"])
        assert int(np.array(b, F).view(np.uint32)) == c["bin_bits"] and dims == c["dims"], c["what"]
        e = c["expect"]
        if "bin" in e:
            assert F(b) == F(e["bin"]), c["what"]
        if "dims" in e:
            assert dims == e["dims"], c["what"]
        if "dim_max" in e:
            assert max(dims) <= e["dim_max"], c["what"]
        if "cells_max" in e:
            assert dims[0] * dims[1] * dims[2] <= e["cells_max"], c["what"]
        if "bin_gt" in e:
            assert b > e["bin_gt"] and np.isfinite(b), c["what"]
        assert list(mn) == [F(x) for x in c["lo"]], c["what"]
