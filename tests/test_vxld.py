"""VXLD v3 frame dumps (the reference demos' format: dam_break.rs:14-44,
:84-128, body records as showcase.rs:311-360) read back with an independent
parser written from that layout, and the second state-hash convention
(byte-wise FNV-1a over all positions then all velocities,
packet_grid_canary.rs:102-111) against a Python restatement."""
import struct

import numpy as np

import vxl_sph as V


def read_vxld(path):
    with open(path, "rb") as f:
        data = f.read()
    at = 0

    def take(fmt):
        nonlocal at
        vals = struct.unpack_from("<" + fmt, data, at)
        at += struct.calcsize("<" + fmt)
        return vals

    assert data[:4] == b"VXLD"
    at = 4
    version, ticks, tpf, dt = take("IIIf")
    dims = take("III")
    origin_step = take("4f")
    (nsplat,) = take("I")
    (nmesh,) = take("I")
    (nfluid,) = take("I")
    frames = []
    nbits = (dims[0] * dims[1] * dims[2] + 7) // 8
    while at < len(data):
        tick, ms = take("If")
        (nb,) = take("I")
        bodies = []
        for _ in range(nb):
            kind, awake = take("BB")
            pos, rot, half = take("3f"), take("4f"), take("3f")
            bodies.append(dict(kind=kind, awake=awake, pos=pos, rot=rot, half=half))
        at += nbits
        fluids = []
        for _ in range(nfluid):
            (cnt,) = take("I")
            fluids.append(np.frombuffer(data, np.float32, 3 * cnt, at).reshape(cnt, 3))
            at += 12 * cnt
        frames.append(dict(tick=tick, ms=ms, bodies=bodies, fluids=fluids))
    return dict(version=version, ticks=ticks, tpf=tpf, dt=dt, dims=dims, origin_step=origin_step, nsplat=nsplat,
                nmesh=nmesh, nfluid=nfluid, frames=frames)


def test_vxld_raw_frames_round_trip(tmp_path):
    path = tmp_path / "dump.bin"
    w = V.VxldWriter(path, ticks=6, ticks_per_frame=2, dt=1 / 60, nfluid=2)
    rng = np.random.default_rng(3)
    bodies = [dict(id=0, half=(1.0, 0.05, 0.5), pos=(0.5, -0.05, 0.25), density=0.0),
              dict(id=7, half=(0.1, 0.1, 0.1), pos=(0.3, 0.6, 0.2), rot=(0.0, 0.6, 0.0, 0.8), density=800.0)]
    frames = []
    for t in (2, 4, 6):
        fl = [rng.random((5 + t, 3), dtype=np.float32), rng.random((0, 3), dtype=np.float32)]
        w.frame_raw(t, 0.25 * t, bodies, fl)
        frames.append(fl)
    w.close()
    d = read_vxld(path)
    assert (d["version"], d["ticks"], d["tpf"], d["nfluid"]) == (3, 6, 2, 2)
    assert d["dt"] == np.float32(1 / 60) and d["dims"] == (0, 0, 0) and d["nsplat"] == 0 and d["nmesh"] == 0
    assert [f["tick"] for f in d["frames"]] == [2, 4, 6]
    for fr, fl in zip(d["frames"], frames):
        assert [b["awake"] for b in fr["bodies"]] == [0, 1] and all(b["kind"] == 0 for b in fr["bodies"])
        assert fr["bodies"][1]["rot"] == (0.0, np.float32(0.6), 0.0, np.float32(0.8))
        assert fr["bodies"][0]["half"] == (1.0, np.float32(0.05), 0.5)
        for got, want in zip(fr["fluids"], fl):
            np.testing.assert_array_equal(got, want)
