"""The C consumer of include/vxl_sph.h on the device, plain and under host
sanitizers.

* `abi_harness stepper` drives the FluidStepper drop-in the way
  World::fluid_stepper_pass does (crates/vxl-phys/src/world_step/
  fluid_stepper.rs:66-79): step, reactions and bounds every tick, two
  substeps per tick. Reactions, bounds and the final state equal the oracle
  bit for bit; bounds() is the tight box of every particle
  (FluidStepper::bounds, crates/vxl-phys-gpu/src/pipeline/stepper.rs:139-146).
* The AddressSanitizer + UndefinedBehaviorSanitizer build runs the stepper and
  the threaded serve loop (loop, builder and hasher threads); each writes what
  the plain build writes and reports nothing. (Under ThreadSanitizer the HIP
  runtime finds no device, so the ThreadSanitizer build runs only the CPU-side
  entry points, in the CPU tests.)
"""

import numpy as np
import pytest

import vxl_sph as V
from scenes import bits, oracle_config
from test_harness_cpu import Reader, harness

import oracle as O

pytestmark = pytest.mark.gpu

F = np.float32
REACTION = np.dtype([("body", "<u4"), ("force", "<f4", 3), ("torque", "<f4", 3)])
ASAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def parse_stepper(raw):
    r = Reader(raw)
    magic, n, nf, ticks = (int(x) for x in r.take("<u4", 4))
    assert magic == 0x50455453
    per_tick = []
    for _ in range(ticks):
        cnt = int(r.take("<u4", 1)[0])
        rx = r.take(REACTION, cnt)
        lo, hi = r.take("<f4", 3), r.take("<f4", 3)
        per_tick.append((rx, lo, hi))
    pos = r.take("<f4", 3 * n).reshape(n, 3)
    vel = r.take("<f4", 3 * n).reshape(n, 3)
    assert r.done()
    return n, nf, per_tick, pos, vel


def parse_serve(raw):
    r = Reader(raw)
    magic, n = (int(x) for x in r.take("<u4", 2))
    assert magic == 0x56524553
    rec = r.take("<u8", 4 * n).reshape(n, 4)
    assert r.done()
    return rec


def test_c_harness_stepper_matches_oracle(tmp_path):
    n, nf, per_tick, pos, vel = parse_stepper(harness("", "stepper", tmp_path / "stepper.bin"))
    cfg = V.dam_break_config(0.02)
    cfg.substeps = 2
    data = V.dam_break(cfg, (0.4, 0.3, 0.3), (10, 9, 8), 0.02, seed=7)
    assert (n, nf) == (data.n, data.n_fluid)
    orc = O.System.particles(oracle_config(cfg), data.pos[:nf], data.vel[:nf], F(cfg.particle_mass),
                             F(cfg.spacing))
    orc.set_boundary_raw(data.pos[nf:], data.vel[nf:], data.pmass[nf:], data.spans)
    for t, (rx, lo, hi) in enumerate(per_tick):
        orc.step(cfg.dt_tick)
        body, ft = orc.reactions()
        assert rx["body"].tolist() == body.tolist(), f"tick {t}"
        got = np.concatenate([rx["force"], rx["torque"]], axis=1).astype(F)
        np.testing.assert_array_equal(bits(got), bits(ft), err_msg=f"tick {t} reactions")
        op = orc.pos()
        np.testing.assert_array_equal(bits(lo), bits(op.min(axis=0)), err_msg=f"tick {t} bounds lo")
        np.testing.assert_array_equal(bits(hi), bits(op.max(axis=0)), err_msg=f"tick {t} bounds hi")
    np.testing.assert_array_equal(bits(pos), bits(orc.pos()))
    np.testing.assert_array_equal(bits(vel), bits(orc.vel()))


def test_c_harness_stepper_under_asan_ubsan(tmp_path):
    plain = harness("", "stepper", tmp_path / "plain.bin")
    assert harness("_asan", "stepper", tmp_path / "asan.bin", env=ASAN_ENV, timeout=300) == plain


def test_c_harness_serve_under_asan_ubsan(tmp_path):
    plain = harness("", "serve", tmp_path / "plain.bin")
    rec = parse_serve(plain)
    assert sorted(rec[:, 0].tolist()) == list(range(len(rec))) and np.all(rec[:, 2] > 0)
    assert harness("_asan", "serve", tmp_path / "asan.bin", env=ASAN_ENV, timeout=300) == plain

