"""The device-array step (vxl_sph_step_device): a scene held in device
memory of the caller, advanced in place, equals the engine's result for the
same scene bit for bit; misuse comes back as status codes before any kernel
runs."""
import numpy as np
import pytest

import vxl_sph as V
from devmem import DeviceArray
from scenes import bits

pytestmark = pytest.mark.gpu

INVALID_ARGUMENT, SIZE_MISMATCH, UNSUPPORTED = 16, 3, 17


def engine_state(cfg, data, ticks):
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    eng.step(ticks)
    pos, vel = eng.read_state(sid)
    eng.close()
    return pos, vel


def step_device(cfg, data, ticks):
    p, v, m = DeviceArray(data.pos.astype(np.float32)), DeviceArray(data.vel.astype(np.float32)), \
        DeviceArray(data.pmass.astype(np.float32))
    try:
        V.step_device(cfg, data.n, data.n_fluid, p.addr, v.addr, m.addr, ticks)
        return p.get(), v.get()
    finally:
        for a in (p, v, m):
            a.close()


@pytest.mark.parametrize("substeps", [1, 2])
def test_device_step_equals_engine(substeps):
    cfg = V.dam_break_config(0.02)
    cfg.substeps = substeps
    data = V.dam_break(cfg, (0.6, 0.4, 0.4), (12, 10, 8), 0.02, seed=23)
    want = engine_state(cfg, data, 5)
    got = step_device(cfg, data, 5)
    np.testing.assert_array_equal(bits(got[0]), bits(want[0]))
    np.testing.assert_array_equal(bits(got[1]), bits(want[1]))


def test_device_step_rejects_misuse():
    cfg = V.dam_break_config(0.02)
    data = V.dam_break(cfg, (0.4, 0.3, 0.3), (6, 6, 6), 0.02, seed=2)
    n, nf = data.n, data.n_fluid
    p, v, m = DeviceArray(data.pos), DeviceArray(data.vel), DeviceArray(data.pmass)
    short = DeviceArray(data.pos[:n // 2])
    try:
        def code(fn):
            with pytest.raises(V.VxlSphError) as e:
                fn()
            return e.value.code

        host = data.pos.ctypes.data  # host memory: rejected before any kernel runs
        assert code(lambda: V.step_device(cfg, n, nf, host, v.addr, m.addr, 1)) == INVALID_ARGUMENT
        assert code(lambda: V.step_device(cfg, n, nf, short.addr, v.addr, m.addr, 1)) == SIZE_MISMATCH
        assert code(lambda: V.step_device(cfg, n, n + 1, p.addr, v.addr, m.addr, 1)) == INVALID_ARGUMENT
        bad = data.pmass.copy()
        bad[5] = 0.0
        mb = DeviceArray(bad)
        try:
            assert code(lambda: V.step_device(cfg, n, nf, p.addr, v.addr, mb.addr, 1)) == INVALID_ARGUMENT
        finally:
            mb.close()
        np.testing.assert_array_equal(bits(p.get()), bits(data.pos))  # untouched by the failed calls
        nanpos = data.pos.copy()
        nanpos[3, 1] = np.nan
        pn = DeviceArray(nanpos)
        try:
            assert code(lambda: V.step_device(cfg, n, nf, pn.addr, v.addr, m.addr, 1)) == UNSUPPORTED
        finally:
            pn.close()
        V.step_device(cfg, n, nf, p.addr, v.addr, m.addr, 2)  # still usable
        want = engine_state(cfg, data, 2)
        np.testing.assert_array_equal(bits(p.get()), bits(want[0]))
    finally:
        for a in (p, v, m, short):
            a.close()
