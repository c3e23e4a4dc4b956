"""The allocation-free device step (vxl_sph_step_device_noalloc) against the
allocating one (vxl_sph_step_device): two calls in a row, the second with a
smaller scene (the shared scratch keeps the first call's bytes), bit for bit;
a scene above the stated maximum is refused before any launch and leaves the
arrays untouched."""
import numpy as np
import pytest

import vxl_sph as V
from devmem import DeviceArray
from scenes import bits

pytestmark = pytest.mark.gpu

CAPACITY = 20


def run(fn, cfg, data, ticks):
    p, v, m = DeviceArray(data.pos), DeviceArray(data.vel), DeviceArray(data.pmass)
    try:
        fn(cfg, data.n, data.n_fluid, p.addr, v.addr, m.addr, ticks)
        return p.get(), v.get()
    finally:
        for a in (p, v, m):
            a.close()


def test_noalloc_equals_allocating_step_on_two_calls():
    # the 32^3 dam break of the reference's CPU-sized case (~75k particles with its tank)
    big_cfg = V.dam_break_config(0.02)
    big = V.dam_break(big_cfg, (2.0, 1.0, 1.0), (32, 32, 32), 0.02, seed=1)
    small_cfg = V.dam_break_config(0.02, c0=25.0, alpha=0.05)
    small = V.dam_break(small_cfg, (0.5, 0.4, 0.4), (10, 8, 8), 0.02, seed=4)
    assert 65536 < big.n <= 131072 and small.n < big.n // 4
    for cfg, data, ticks in ((big_cfg, big, 4), (small_cfg, small, 6)):
        print(f"scene of {data.n} particles, {ticks} ticks", flush=True)
        want = run(V.step_device, cfg, data, ticks)
        got = run(V.step_device_noalloc, cfg, data, ticks)
        np.testing.assert_array_equal(bits(got[0]), bits(want[0]))
        np.testing.assert_array_equal(bits(got[1]), bits(want[1]))
        assert not np.array_equal(bits(got[0]), bits(data.pos))  # it stepped


def test_noalloc_refuses_scenes_above_its_maximum():
    cfg = V.dam_break_config(0.02)
    data = V.dam_break(cfg, (2.0, 1.0, 1.0), (50, 50, 50), 0.02, seed=2)
    assert data.n > 131072
    p, v, m = DeviceArray(data.pos), DeviceArray(data.vel), DeviceArray(data.pmass)
    try:
        with pytest.raises(V.VxlSphError) as e:
            V.step_device_noalloc(cfg, data.n, data.n_fluid, p.addr, v.addr, m.addr, 1)
        assert e.value.code == CAPACITY
        np.testing.assert_array_equal(bits(p.get()), bits(data.pos))
    finally:
        for a in (p, v, m):
            a.close()
