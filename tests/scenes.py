"""Scene helpers shared by the tests: the dam-break tank spec restated with
numpy f32 arithmetic (an independent copy of the scene definition) and
constructors of oracle systems for a library-built scene."""
import numpy as np

import oracle as O

F = np.float32


def tank_bodies(L, d):
    """Floor + 4 walls of thickness 2d around the open tank [0,L]^3."""
    L = [F(x) for x in L]
    d = F(d)
    t = F(2.0) * d
    hx, hy, hz = L[0] * F(0.5), L[1] * F(0.5), L[2] * F(0.5)
    th = t * F(0.5)
    return [
        dict(id=0, half=(hx + t, th, hz + t), pos=(hx, -th, hz)),
        dict(id=1, half=(th, hy, hz + t), pos=(-th, hy, hz)),
        dict(id=2, half=(th, hy, hz + t), pos=(L[0] + th, hy, hz)),
        dict(id=3, half=(hx + t, hy, th), pos=(hx, hy, -th)),
        dict(id=4, half=(hx + t, hy, th), pos=(hx, hy, L[2] + th)),
    ]


def oracle_config(cfg, threads=1, max_bins=None):
    """OrConfig equal to a library VxlSphFluidConfig."""
    return O.default_config(
        rest_density=cfg.rest_density, smoothing_radius=cfg.smoothing_radius, xsph_viscosity=cfg.xsph_viscosity,
        sound_speed=cfg.sound_speed, gamma_tait=cfg.gamma_tait, substeps=cfg.substeps, threads=threads,
        tensile_clamp=cfg.tensile_clamp, boundary_layers=cfg.boundary_layers, gravity=tuple(cfg.gravity),
        max_speed_frac=cfg.max_speed_frac, artificial_viscosity=cfg.artificial_viscosity,
        max_bins=max_bins if max_bins is not None else cfg.max_bins)


def oracle_dam_break(cfg, data, tank, d, threads=1, own_lattice=True):
    """Oracle system for a library dam-break scene: the same fluid particles;
    the tank either rebuilt by the oracle's own lattice restatement or copied."""
    oc = oracle_config(cfg, threads)
    nf = data.n_fluid
    s = O.System.particles(oc, data.pos[:nf], data.vel[:nf], F(cfg.particle_mass), F(cfg.spacing))
    if data.n > nf:
        if own_lattice:
            s.set_boundary_boxes(tank_bodies(tank, d))
        else:
            s.set_boundary_raw(data.pos[nf:], data.vel[nf:], data.pmass[nf:], data.spans)
    return s


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def max_rel(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


# --- golden fixtures (tests/golden/make_golden.py) ---------------------------

GOLDEN_DIR = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)),
                                        "golden")
GOLDEN_SCENES = ("dam_512", "dam_4096")


def load_golden(name):
    """(meta dict, arrays dict) of one fixture; no pickles."""
    import json
    import os

    with np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False) as z:
        arrs = {k: z[k] for k in z.files}
    return json.loads(str(arrs.pop("meta"))), arrs


def golden_spans(arrs):
    return [(int(b), tuple(float(x) for x in o), int(r[0]), int(r[1]))
            for b, o, r in zip(arrs["in_span_body"], arrs["in_span_origin"], arrs["in_span_range"])]


def fnv_words(words, h=0xcbf29ce484222325):
    for v in np.asarray(words, np.uint32).tolist():
        h = ((h ^ v) * 0x00000100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def grid_table_hash(start, items):
    """FNV-1a-64 over start, 0xffffffff, items, 0xffffffff, per-particle cell
    (the table hash of crates/vxl-phys-gpu/examples/gpu_grid_probe.rs:50-68)."""
    start = np.asarray(start, np.int64)
    items = np.asarray(items, np.uint32)
    bins = np.full(len(items), 0xFFFFFFFF, np.uint32)
    counts = np.diff(start)
    bins[items] = np.repeat(np.arange(len(counts), dtype=np.uint32), counts)
    h = fnv_words(start.astype(np.uint32))
    h = fnv_words([0xFFFFFFFF], h)
    h = fnv_words(items, h)
    h = fnv_words([0xFFFFFFFF], h)
    return fnv_words(bins, h)
