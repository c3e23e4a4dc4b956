"""Writes the golden fixtures of the WCSPH step with the oracle (test
infrastructure; the product never reads this script).

Each fixture is a small dam break in the style of the 32^3 tank benchmark
(h = 2d, mass = 1000 d^3, c0 = 20, alpha = 0.02, eps = 0.5, dt = 5e-4, one
substep, one Akinci layer, tank = floor + 4 walls of thickness 2d): a
jittered fluid block in the tank corner. The file holds the inputs (every
particle's position, velocity and mass, the boundary spans) and, after the
listed steps, the grid tables (min, inv, bin, dims, start, items, keys in the
reference layout), neighbour counts, density, pressure, acceleration, XSPH,
positions and velocities as raw f32 bits, and the grid and state hashes.

Usage: python tests/golden/make_golden.py   (rewrites tests/golden/*.npz)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))
import oracle as O  # noqa: E402
from scenes import tank_bodies  # noqa: E402

F = np.float32

# name: (fluid block, tank extent, d, seed, steps with a full dump, steps with hashes only)
SCENES = {
    "dam_512": ((8, 8, 8), (0.4, 0.3, 0.2), 0.02, 101, (1, 10), ()),
    "dam_4096": ((16, 16, 16), (0.8, 0.5, 0.4), 0.02, 202, (1,), (10,)),
}
PHYS = dict(c0=20.0, alpha=0.02, eps=0.5, dt=5e-4, layers=1, max_bins=1 << 21, jitter=0.001)


def oracle_config(d, threads=4):
    return O.default_config(smoothing_radius=float(F(2.0) * F(d)), sound_speed=PHYS["c0"],
                            artificial_viscosity=PHYS["alpha"], xsph_viscosity=PHYS["eps"], substeps=1,
                            boundary_layers=PHYS["layers"], max_bins=PHYS["max_bins"], threads=threads)


def fluid_block(block, d, seed):
    """Lattice (i + 1/2) d, index x-major (i ny nz + j nz + k), plus an
    independent U(-jitter, jitter) offset per coordinate (PCG64, `seed`)."""
    d = F(d)
    bx, by, bz = block
    i, j, k = np.meshgrid(np.arange(bx), np.arange(by), np.arange(bz), indexing="ij")
    lat = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(F)
    base = (lat + F(0.5)) * d
    rng = np.random.Generator(np.random.PCG64(seed))
    jit = rng.uniform(-PHYS["jitter"], PHYS["jitter"], size=base.shape).astype(F)
    return (base + jit).astype(F)


def build(name):
    block, tank, d, seed, full, hashed = SCENES[name]
    cfg = oracle_config(d)
    mass = F(F(F(F(1000.0) * F(d)) * F(d)) * F(d))
    fpos = fluid_block(block, d, seed)
    s = O.System.particles(cfg, fpos, np.zeros_like(fpos), mass, F(d))
    bodies = tank_bodies(tank, d)
    s.set_boundary_boxes(bodies)
    # Spans: body order, each body's lattice contiguous.
    spans, at = [], len(fpos)
    h = float(F(2.0) * F(d))
    for b in bodies:
        pts, _, _ = O.lattice_box(b["half"], F(d), h, PHYS["layers"])
        spans.append((b["id"], tuple(float(x) for x in b["pos"]), at, at + len(pts)))
        at += len(pts)
    assert at == s.n_total
    return s, spans, mass, (full, hashed)


def dump(s, prefix, out, full):
    out[prefix + "grid_hash"] = np.array(s.grid_hash(), np.uint64)
    out[prefix + "state_hash"] = np.array(s.state_hash(), np.uint64)
    if not full:
        return
    mn, inv, b, dims = s.grid_info()
    start, items, keys = s.grid_tables()
    u = lambda a: np.ascontiguousarray(a, F).view(np.uint32)  # noqa: E731
    out[prefix + "grid_min"] = u(mn)
    out[prefix + "grid_inv_bin"] = u(np.array([inv, b], F))
    out[prefix + "grid_dims"] = np.array(dims, np.uint32)
    out[prefix + "start"] = start
    out[prefix + "items"] = items
    out[prefix + "keys"] = keys
    out[prefix + "ncount"] = s.ncount()
    out[prefix + "dens"] = u(s.dens())
    out[prefix + "press"] = u(s.press())
    out[prefix + "acc"] = u(s.acc())
    out[prefix + "xsph"] = u(s.xsph())
    out[prefix + "pos"] = u(s.pos())
    out[prefix + "vel"] = u(s.vel())


def make(name):
    s, spans, mass, (full, hashed) = build(name)
    block, tank, d, seed, _, _ = SCENES[name]
    out = {}
    meta = dict(name=name, block=block, tank=tank, d=d, seed=seed, mass_bits=int(np.array(mass).view(np.uint32)),
                n=int(s.n_total), n_fluid=int(s.n_fluid), full_steps=list(full), hash_steps=list(hashed), **PHYS)
    out["meta"] = np.array(json.dumps(meta))
    out["in_pos"] = s.pos()
    out["in_vel"] = s.vel()
    out["in_pmass"] = s.pmass()
    out["in_span_body"] = np.array([sp[0] for sp in spans], np.uint32)
    out["in_span_origin"] = np.array([sp[1] for sp in spans], F)
    out["in_span_range"] = np.array([(sp[2], sp[3]) for sp in spans], np.uint32)
    last = max(list(full) + list(hashed))
    for step in range(1, last + 1):
        s.step(F(PHYS["dt"]))
        if step in full or step in hashed:
            dump(s, f"s{step}_", out, step in full)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{path}: n = {s.n_total} ({s.n_fluid} fluid), {os.path.getsize(path)} bytes")


def grid_box_kat():
    """Known answers of the grid box rule: the cases and assertions of
    crates/vxl-phys-core/src/grid.rs:59-111 (`expect`), with the oracle's
    result beside them (`bin_bits`, `dims`)."""
    mb = 1 << 20
    p = [1.0, 2.0, 3.0]
    cases = [
        dict(what="single particle gets one cell", lo=p, hi=p, h=0.05, max_bins=mb,
             expect=dict(bin=0.05, dims=[1, 1, 1])),
        dict(what="dims are floor plus one", lo=[0.0] * 3, hi=[0.25] * 3, h=0.05, max_bins=mb,
             expect=dict(bin=0.05, dims=[6, 6, 6])),
        dict(what="oversize budget doubles cell", lo=[0.0] * 3, hi=[5.0] * 3, h=0.05, max_bins=10000,
             expect=dict(bin=0.4, dims=[13, 13, 13])),
        dict(what="per dim clamp to 14 bits", lo=[0.0] * 3, hi=[1.0e6] * 3, h=0.05, max_bins=mb,
             expect=dict(dim_max=1 << 14, cells_max=mb, bin_gt=0.05)),
        dict(what="tiny h is floored", lo=p, hi=p, h=0.0, max_bins=mb, expect=dict(bin=1e-6, dims=[1, 1, 1])),
        dict(what="h = 0 doubles into the budget", lo=[0.0] * 3, hi=[1.0] * 3, h=0.0, max_bins=mb,
             expect=dict(bin_gt=1e-6, cells_max=mb)),
    ]
    for c in cases:
        _, b, dims = O.grid_box(c["lo"], c["hi"], c["h"], c["max_bins"])
        c["bin_bits"] = int(np.array(b, F).view(np.uint32))
        c["dims"] = dims
    path = os.path.join(HERE, "grid_box_kat.json")
    with open(path, "w") as f:
        json.dump(cases, f, indent=1)
    print(path)


if __name__ == "__main__":
    O.build()
    for name in SCENES:
        make(name)
    grid_box_kat()
