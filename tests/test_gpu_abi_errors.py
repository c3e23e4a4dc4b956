"""Argument checks of the entry points that need a device: each misuse the
reference's GpuFluidStepper asserts on (pipeline/stepper.rs:70-97) or the
FluidSystem accessors reject comes back as a status code with a message, and
the stepper or engine stays usable afterwards (crates/vxl-phys-ffi/src/
lib.rs:13-41: panics become status codes)."""
import numpy as np
import pytest

import vxl_sph as V
from test_gpu_parity import small_scene

pytestmark = pytest.mark.gpu

SIZE_MISMATCH, INVALID_ARGUMENT = 3, 16


def status(fn):
    with pytest.raises(V.VxlSphError) as e:
        fn()
    return e.value.code


def test_stepper_rejects_a_changed_particle_set_and_tick():
    cfg, data, tank, d = small_scene(seed=29)
    st = V.FluidStepper(cfg)
    nf, pos, vel, pm, spans = data.n_fluid, data.pos, data.vel, data.pmass, data.spans
    st.step(cfg.dt_tick, nf, pos, vel, pm, spans)
    # a different fluid prefix, a different particle count, a different tick
    assert status(lambda: st.step(cfg.dt_tick, nf - 1, pos, vel, pm, spans)) == SIZE_MISMATCH
    assert status(lambda: st.step(cfg.dt_tick, nf, pos[:-1], vel[:-1], pm[:-1], spans)) == SIZE_MISMATCH
    assert status(lambda: st.step(cfg.dt_tick * 2.0, nf, pos, vel, pm, spans)) == INVALID_ARGUMENT
    # still usable: the state after one more valid step is that of two steps
    st.step(cfg.dt_tick, nf, pos, vel, pm, spans)
    ref = V.FluidStepper(cfg)
    for _ in range(2):
        ref.step(cfg.dt_tick, nf, pos, vel, pm, spans)
    a, b = st.read_state(), ref.read_state()
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_engine_rejects_malformed_scenes_and_reads():
    cfg, data, tank, d = small_scene(seed=31)
    eng = V.Engine()
    bad_prefix = V.SceneData(data.pos, data.vel, data.pmass, data.n + 1, [])
    assert status(lambda: eng.add(cfg, bad_prefix, reactions=False)) == INVALID_ARGUMENT
    pm = data.pmass.copy()
    pm[3] = -1.0
    assert status(lambda: eng.add(cfg, V.SceneData(data.pos, data.vel, pm, data.n_fluid, []),
                                  reactions=False)) == INVALID_ARGUMENT
    empty = V.SceneData(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros(0, np.float32), 0, [])
    assert status(lambda: eng.add(cfg, empty, reactions=False)) == INVALID_ARGUMENT
    sid = eng.add(cfg, data, reactions=False)
    eng.step(1)
    n = data.n
    eng.scenes[sid] = (n - 1,) + eng.scenes[sid][1:]  # ask for one particle fewer than the scene holds
    assert status(lambda: eng.read_state(sid)) == SIZE_MISMATCH
    eng.scenes[sid] = (n,) + eng.scenes[sid][1:]
    pos, vel = eng.read_state(sid)
    assert pos.shape == (n, 3) and np.all(np.isfinite(pos))
