#!/usr/bin/env python3
"""WCSPH step benchmark on one MI355X.

Workload: the large single dam-break scene — 215^3 = 9,938,375 fluid particles
at d = 0.01 m (h = 0.02 m) in a 4 x 2.5 x 2.5 m open tank sampled as one layer
of Akinci boundary particles, dt = 2.5e-4 s, strict f32. One "step" = one
substep of every stage (grid box, keys, scan, place, canonical gather,
density + Tait, force + AV + XSPH + integration) for the whole scene, inputs
resident in HBM. Prints ONE JSON line.

value = fluid particle-updates/s (fluid particles x steps / wall time of the
timed region, bracketed by device synchronisation). roofline = the dominant
kernel's algorithmic bytes per launch / its mean HIP-event duration against the
8 TB/s HBM3E peak; issue_roofline = the two neighbour passes' vector-ALU
instructions per launch (rocprofv3 counters) / their durations against the
FP32 issue rate measured on the same device. cpu_baseline = the oracle (a
strict-f32 C++ restatement of the reference CPU step) timed on this
machine's host cores on a bounded sample of the same scene.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rust-wl_amd"))

METRIC = "WCSPH particle-updates/s at 10M (HBM roofline %); p50/p99 step latency, 64 scenes"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of each stage (compulsory SoA traffic, neighbour reads
# served on chip): per particle N, per fluid particle NF, per grid cell C.
STAGE_BYTES = {
    "keys": dict(N=16),            # R pos 12, W key 4
    "scan": dict(C=8),             # R counts 4, W start 4
    "place": dict(N=8),            # R key 4, W item 4
    "canon_gather": dict(N=60),    # R item 4, R pos/vel/pmass 28, W copies 28
    "density": dict(N=20, C=4),    # R pos 12 + pmass 4, W dens 4; R start
    "force": dict(NF=60, C=4),     # R pos, vel, pmass, dens 32 + item 4, W pos, vel 24; R start
}


def issue_rate(kern, tuning, args):
    """The neighbour passes against the device's FP32 vector-ALU issue rate:
    their wave-wide VALU instructions per launch (SQ_INSTS_VALU from the
    committed rocprofv3 counter pass, default workload only) over their
    HIP-event times, per SIMD per second, against the rate that
    rust-wl_amd/valu_rate (independent FMA chains) measures on this device
    now. Both passes are bound by instruction issue, not by HBM."""
    path = os.path.join(ROOT, "profiles", "wcsph_10M_valu.json")
    exe = os.path.join(ROOT, "rust-wl_amd", "valu_rate")
    if args.block != 215 or args.budget_log2 != 22 or not os.path.exists(path) or not os.path.exists(exe):
        return None
    import subprocess

    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        raise RuntimeError(f"valu_rate failed ({r.returncode}): {r.stderr.strip()}")
    cal = json.loads(r.stdout.strip().splitlines()[-1])
    with open(path) as f:
        counts = json.load(f)
    entries, probe = counts["entries"], counts["probe"]
    # The launch-geometry fields each pass's instruction count depends on.
    keys = {"force": ("force_kernel", "block_threads"), "density": ("block_threads", "stage_records")}
    peak = cal["fma_wave_instr_per_simd_s"]
    out = {"bound": "valu", "unit": "wave-instr/SIMD/s", "peak": round(peak), "peak_kernel_ms": cal["kernel_ms"],
           "simds": cal["simds"], "peak_per_cycle": probe["valu_per_simd_per_cycle"],
           "tuning": {k: tuning[k] for k in ("force_kernel", "block_threads", "stage_records")}, "kernels": {}}
    for name, ks in keys.items():
        counted = next((e for e in entries if all(e["tuning"][k] == tuning[k] for k in ks)), None)
        if counted is None:
            continue  # no counter pass for this launch geometry
        c = counted["kernels"][name]
        if name in kern and kern[name] > 0:
            ach = c["valu_wave_instr"] / (kern[name] * 1e-3) / cal["simds"]
            # per clock cycle (the counter pass's own cycles, kernel and probe
            # alike): the issue rate without the clock's dependence on load
            pc = c["valu_per_simd_per_cycle"]
            out["kernels"][name] = {"kernel": c["kernel"], "valu_wave_instr": c["valu_wave_instr"],
                                    "achieved": round(ach), "frac": round(ach / peak, 4),
                                    "per_cycle": pc, "frac_per_cycle": round(pc / probe["valu_per_simd_per_cycle"], 4)}
    if not out["kernels"]:
        return None
    print("# vector-ALU issue: peak %.3e wave-instr/SIMD/s; " % peak
          + ", ".join(f"{k} {v['frac']:.0%} ({v['frac_per_cycle']:.0%} per cycle)" for k, v in out["kernels"].items()),
          file=sys.stderr, flush=True)
    return out


def stage_bytes(name, N, NF, C):
    m = STAGE_BYTES.get(name, {})
    return m.get("N", 0) * N + m.get("NF", 0) * NF + m.get("C", 0) * C


def measured_traffic(kernel, args):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy), for the default
    workload only; None otherwise."""
    path = os.path.join(ROOT, "profiles", "wcsph_10M_traffic.json")
    if kernel != "force" or args.block != 215 or args.budget_log2 != 22 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return round(t["traffic_bytes"])


def build_scene(V, block, d, tank, max_bins, seed=1):
    cfg = V.dam_break_config(d, dt=2.5e-4, max_bins=max_bins)
    data = V.dam_break(cfg, tank, block, d, seed=seed)
    return cfg, data


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The host cores this job may use: the machine's count, capped by the
    host-core share of one GPU (OMP_NUM_THREADS, 16 when unset)."""
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(os.cpu_count() or 1, share))


def cpu_baseline(cfg, data, steps, threads, reactions=False):
    """Oracle (test infrastructure) on the same scene, bounded sample. With
    reactions=False the boundary-force pass is skipped, as the GPU headline
    line (a static tank, nothing consumes its reaction) skips it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    oc = O.default_config(
        rest_density=cfg.rest_density, smoothing_radius=cfg.smoothing_radius, xsph_viscosity=cfg.xsph_viscosity,
        sound_speed=cfg.sound_speed, gamma_tait=cfg.gamma_tait, substeps=cfg.substeps, threads=threads,
        tensile_clamp=cfg.tensile_clamp, boundary_layers=cfg.boundary_layers, gravity=tuple(cfg.gravity),
        max_speed_frac=cfg.max_speed_frac, artificial_viscosity=cfg.artificial_viscosity, max_bins=cfg.max_bins,
        skip_reactions=0 if reactions else 1)
    nf = data.n_fluid
    s = O.System.particles(oc, data.pos[:nf], data.vel[:nf], np.float32(cfg.particle_mass), np.float32(cfg.spacing))
    s.set_boundary_raw(data.pos[nf:], data.vel[nf:], data.pmass[nf:], data.spans)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(cfg.dt_tick)
    dt = time.perf_counter() - t0
    return nf * steps / dt, dt


def bench_stepper(V, cfg, data, ticks):
    """The drop-in path: FluidStepper (GpuFluidStepper semantics) with
    reactions on every span and the boundary segment uploaded every tick, as
    World::fluid_stepper_pass drives it; host-to-host time per tick."""
    st = V.FluidStepper(cfg)
    nf = data.n_fluid
    pos, vel = data.pos.copy(), data.vel.copy()  # the stepper writes the state back in place
    st.step(cfg.dt_tick, nf, pos, vel, data.pmass, data.spans)  # uploads the scene
    st.step(cfg.dt_tick, nf, pos, vel, data.pmass, data.spans)
    t0 = time.perf_counter()
    for _ in range(ticks):
        st.step(cfg.dt_tick, nf, pos, vel, data.pmass, data.spans)
        st.reactions()
    dt = time.perf_counter() - t0
    st.close()
    return {"ms_per_tick": 1e3 * dt / ticks, "value": nf * ticks / dt, "unit": "particle-updates/s", "ticks": ticks,
            "fluid_particles": nf, "boundary_particles": data.n - nf, "spans": len(data.spans)}


def bench_batched(V, args, one_at_a_time=True):
    """64 uneven dam-break scenes in one pool (one launch per stage, one graph
    per step) against the same scenes stepped one at a time."""
    reqs = V.sample_batch(4, args.batch_scenes)
    built = [V.request_scene(r) for r in reqs]
    nf_total = sum(d.n_fluid for _, d in built)
    eng = V.Engine()
    ids = [eng.add(c, d, reactions=False) for c, d in built]
    if args.tuning:
        eng.set_tuning(**json.loads(args.tuning))
        tuned = eng.get_tuning()
    elif args.force_kernel is not None:
        eng.set_tuning(force_kernel=args.force_kernel)
        tuned = eng.get_tuning()
    else:
        tuned = eng.autotune(2)
    eng.step(3)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.batch_steps)
    eng.sync()
    wall = time.perf_counter() - t0
    lat = eng.step_timed(args.batch_steps)
    prof = {k: round(v / 6, 4) for k, v in eng.profile(6).items()}
    eng.close()
    batched_ups = nf_total * args.batch_steps / wall
    pooled = {"scenes": len(built), "fluid_particles": nf_total,
              "total_particles": sum(d.n for _, d in built),
              "value": batched_ups, "unit": "particle-updates/s",
              "ms_per_step": 1e3 * wall / args.batch_steps,
              "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
              "tuning": tuned,
              "kernel_ms_per_step": prof,
              "workload": "seeded set: fluid counts log-uniform [8192, 524288], tank x,z ~ U[1,4] m, y = 1 m, "
                          "fill ~ U[0.1,0.5], c0 ~ U[15,30], alpha ~ U[0.01,0.05], d = 0.02 m, dt = 5e-4 s"}
    if not one_at_a_time:
        print(f"# batched: {len(built)} scenes, {nf_total} fluid, {1e3 * wall / args.batch_steps:.3f} ms/step "
              f"(p50 {pooled['step_latency_ms']['p50']:.3f} / p99 {pooled['step_latency_ms']['p99']:.3f} ms); "
              "kernels " + ", ".join(f"{k} {v:.3f}" for k, v in sorted(prof.items(), key=lambda x: -x[1])),
              file=sys.stderr, flush=True)
        return pooled
    # One scene at a time through the same engine code (plain single-scene path).
    seq_ms = 0.0
    one = V.Engine()
    for c, d in built:
        sid = one.add(c, d, reactions=False)
        one.step(2)
        one.sync()
        t0 = time.perf_counter()
        one.step(args.batch_steps // 3)
        one.sync()
        seq_ms += 1e3 * (time.perf_counter() - t0) / (args.batch_steps // 3)
        one.remove(sid)
    one.close()
    seq_ups = nf_total / (seq_ms * 1e-3)
    print(f"# batched: {len(built)} scenes, {nf_total} fluid, {1e3 * wall / args.batch_steps:.3f} ms/step "
          f"vs one-at-a-time {seq_ms:.3f} ms -> {batched_ups / seq_ups:.2f}x", file=sys.stderr, flush=True)
    pooled.update({"one_at_a_time": {"value": seq_ups, "ms_per_step_all_scenes": seq_ms},
                   "speedup_vs_one_at_a_time": batched_ups / seq_ups})
    return pooled


def bench_served(V, args):
    """Continuous batching over the seeded Poisson request stream (2000
    requests at 20 requests/s in real time by default, about 100 s)."""
    reqs = V.sample_requests(9, args.serve_requests, rate=20.0, nmin=4096, nmax=2097152, smin=100, smax=2000,
                             max_boxes=4)
    stats, results = V.serve(reqs, budget_bytes=200 << 30, time_scale=1.0, builders=8,
                             max_particles=args.serve_max_particles)
    print(f"# served: {stats['completed']}/{stats['requests']} requests, {stats['updates_per_s']:.3e} upd/s, "
          f"step p50 {stats['step_ms_p50']:.3f} ms p99 {stats['step_ms_p99']:.3f} ms, completion interval "
          f"p50 {stats['loop_ms_p50']:.3f} ms p99 {stats['loop_ms_p99']:.3f} ms, "
          f"peak resident {stats['peak_resident']}", file=sys.stderr, flush=True)
    keep = ("requests", "completed", "rejected", "pool_steps", "fluid_updates", "wall_s", "updates_per_s",
            "step_ms_p50", "step_ms_p99", "loop_ms_p50", "loop_ms_p99", "latency_s_p50", "latency_s_p99",
            "wait_s_p50", "wait_s_p99", "peak_resident", "peak_bytes", "layouts", "layout_ms_p50", "layout_ms_p99",
            "layout_phase_ms", "status")
    out = {k: stats[k] for k in keep}
    out["interval_ms_p50"] = out.pop("loop_ms_p50")
    out["interval_ms_p99"] = out.pop("loop_ms_p99")
    out["requests_with_boxes"] = sum(1 for r in reqs if r.nboxes)
    out["dynamic_boxes"] = sum(r.nboxes for r in reqs)
    out["workload"] = (f"{args.serve_requests} requests of the seeded Poisson stream (20 requests/s, fluid "
                       "log-uniform [4096, 2097152], steps U[100, 2000], 0-4 dynamic boxes of side 0.1-0.3 m and "
                       "density 500-2000 kg/m^3 coupled both ways), HBM budget 200 GB, real-time arrivals, "
                       f"at most {args.serve_max_particles} resident particles per pool step")
    return out


def tank_walls(V, tank):
    """The dam-break tank [0, L0] x [0, inf) x [0, L2] as one half-space
    provider: floor, then the x and z walls."""
    return V.HalfspaceProvider([(1, 0.0, 1.0), (0, 0.0, 1.0), (0, tank[0], -1.0), (2, 0.0, 1.0), (2, tank[2], -1.0)])


def bench_walls(V, args):
    """The 1M dam break (100^3 fluid, d = 0.02 m, tank 4 x 2 x 2 m) held by
    provider walls instead of Akinci particles. Half-space provider: the
    device queries the walls every substep (mirror ghosts, projection,
    coincident separation). Plane tables: the host reads the fluid back,
    gathers both tables and uploads them once per tick (the GpuFluidStepper
    contract); its device time and its host share are reported apart."""
    d, tank = 0.02, (4.0, 2.0, 2.0)
    cfg = V.dam_break_config(d)
    data = V.dam_break(cfg, tank, (100, 100, 100), d, seed=2, walls=False)
    nf = data.n_fluid
    walls = tank_walls(V, tank)
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    eng.set_halfspace_walls(sid, [walls])
    if args.force_kernel is not None:
        eng.set_tuning(force_kernel=args.force_kernel)
        tuned = eng.get_tuning()
    else:
        tuned = eng.autotune(2)
    eng.step(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.walls_steps)
    eng.sync()
    wall = time.perf_counter() - t0
    prof = eng.profile(args.profile_steps)
    kern = {k: round(v / args.profile_steps, 4) for k, v in prof.items()}
    wall_ms = sum(v for k, v in kern.items() if k.startswith(("wall_", "sep_")))
    eng.close()
    half = {"value": nf * args.walls_steps / wall, "unit": "particle-updates/s",
            "ms_per_step": 1e3 * wall / args.walls_steps, "wall_kernels_ms_per_step": round(wall_ms, 4),
            "kernel_ms_per_step": kern, "tuning": tuned}
    # Plane tables, one gather per tick.
    provs = V.Providers({0: walls})
    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    h = np.float32(cfg.smoothing_radius)
    host_s = dev_s = 0.0
    ticks = args.walls_table_ticks
    for t in range(ticks + 2):
        t0 = time.perf_counter()
        pos, _ = eng.read_state(sid)
        for side, fn in ((0, V.gather_wall_contacts), (1, V.gather_wall_project_contacts)):
            eng.set_wall_tables(sid, side, *fn([0], h, pos[:nf], provs))
        t1 = time.perf_counter()
        eng.step(1)
        eng.sync()
        t2 = time.perf_counter()
        if t >= 2:
            host_s += t1 - t0
            dev_s += t2 - t1
    eng.close()
    table = {"device_ms_per_step": 1e3 * dev_s / ticks, "host_gather_upload_ms_per_tick": 1e3 * host_s / ticks,
             "value_device_only": nf * ticks / dev_s, "value_with_host_gather": nf * ticks / (dev_s + host_s),
             "ticks": ticks}
    print(f"# walls: half-space {half['ms_per_step']:.3f} ms/step ({wall_ms:.3f} ms in wall kernels), "
          f"tables {table['device_ms_per_step']:.3f} ms/step + host {table['host_gather_upload_ms_per_tick']:.1f} "
          f"ms/tick", file=sys.stderr, flush=True)
    return {"workload": f"dam break 100^3 = {nf} fluid particles, d = {d} m, tank {tank[0]}x{tank[1]}x{tank[2]} m "
                        "as half-space provider walls (floor + 4 walls), dt = 5e-4 s, one substep per tick",
            "halfspace": half, "tables": table}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--block", type=int, default=215, help="fluid lattice edge (215 -> 9.94M particles)")
    ap.add_argument("--budget-log2", type=int, default=22,
                    help="grid budget GRID_MAX_BINS = 2^k (22 keeps bin = h at 10M; the reference rule is 21)")
    ap.add_argument("--cpu-steps", type=int, default=12,
                    help="oracle steps timed at the host-core share (about 10 s on 16 cores)")
    ap.add_argument("--cpu-single-steps", type=int, default=1,
                    help="steps of the single-thread CPU baseline (the reference's default threads = 1); 0 skips it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--force-kernel", type=int, default=None, choices=(1, 2, 3, 4, 5, 6, 7),
                    help="force kernel instead of the autotuner's choice: 1 candidate scan, 2 neighbour walk, "
                         "3 walk with loads one pair ahead, 4 pair-parallel rows, 5 queue walk, "
                         "6 queue walk with lane-pair record loads, 7 queue walk with each wave's own row staged in LDS")
    ap.add_argument("--sweep", action="store_true", help="print per-kernel times for every tuning variant")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched measurement")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the pooled 64-scene step (no 10M scene, no one-at-a-time comparison) and print "
                         "its figures as one JSON line; for profiling the batched kernels")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--only-serve", action="store_true",
                    help="run only the served stream and print its figures as one JSON line; for profiling the "
                         "served pool's kernels")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-max-particles", type=int, default=24_000_000,
                    help="resident particles per pool step at most (the step's size; the stream overloads one GPU, "
                         "so the excess waits in the queue)")
    ap.add_argument("--no-walls", action="store_true", help="skip the provider-wall measurement")
    ap.add_argument("--tuning", default=None,
                    help='launch geometry as JSON instead of the autotuner (e.g. \'{"force_kernel": 3, '
                         '"block_threads": 128}\'), for profiles of a fixed kernel set')
    ap.add_argument("--late-steps", type=int, default=480,
                    help="step the headline scene (afresh, with the headline's tuning) to this step and time 20 "
                         "more (0 skips)")
    ap.add_argument("--no-ref-rule", action="store_true", help="skip the 2^21 grid-budget figure")
    ap.add_argument("--no-stepper", action="store_true", help="skip the FluidStepper drop-in timings")
    ap.add_argument("--stepper-ticks", type=int, default=20)
    ap.add_argument("--walls-steps", type=int, default=200)
    ap.add_argument("--walls-table-ticks", type=int, default=5)
    ap.add_argument("--no-issue", action="store_true",
                    help="skip the vector-ALU issue-rate measurement (rust-wl_amd/valu_rate) of the neighbour passes")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run one plain device copy of BYTES (scale for the HBM byte counters) and exit")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n}))
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU (one MI355X); --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        sys.exit(2)

    import vxl_sph as V

    if args.only_batch:
        print(json.dumps({"batched": bench_batched(V, args, one_at_a_time=False)}), flush=True)
        return
    if args.only_serve:
        print(json.dumps({"served": bench_served(V, args)}), flush=True)
        return

    d = 0.01
    tank = (4.0, 2.5, 2.5)
    block = (args.block,) * 3
    t0 = time.perf_counter()
    cfg, data = build_scene(V, block, d, tank, 1 << args.budget_log2)
    t_build = time.perf_counter() - t0
    N, NF = data.n, data.n_fluid
    print(f"# scene: {NF} fluid + {N - NF} Akinci particles, built in {t_build:.1f}s", file=sys.stderr, flush=True)

    eng = V.Engine()
    sid = eng.add(cfg, data, reactions=False)
    if args.tuning:
        eng.set_tuning(**json.loads(args.tuning))
    elif args.force_kernel is not None:
        eng.set_tuning(force_kernel=args.force_kernel)
    else:
        tuned = eng.autotune(2)
        print(f"# autotuned: {tuned}", file=sys.stderr, flush=True)
    sweep = None
    if args.sweep:
        # Each knob over its values around the tuned point (the others held):
        # per-kernel device ms per substep, mean of 3 eager substeps.
        eng.step(2)
        saved = eng.get_tuning()
        sweep = []
        for field, values in (("force_kernel", (1, 2, 3, 4, 5, 6, 7)), ("block_threads", (64, 128, 256)),
                              ("stage_records", (128, 256, 512)), ("scan_tile_cells", (1024, 2048, 4096)),
                              ("persist_blocks", (0, 2048))):
            for v in values:
                t = dict(saved)
                t[field] = v
                eng.set_tuning(**t)
                eng.profile(1)
                pr = {k: round(x / 3, 4) for k, x in eng.profile(3).items()}
                sweep.append({"knob": field, "value": v, "kernel_ms": pr, "total_ms": round(sum(pr.values()), 4)})
                print(f"# sweep {field}={v}: total {sum(pr.values()):.3f} ms (force {pr.get('force', 0):.3f}, "
                      f"density {pr.get('density', 0):.3f})", file=sys.stderr, flush=True)
        eng.set_tuning(**saved)
    eng.step(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = NF * args.steps / wall
    print(f"# timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    lat = eng.step_timed(args.steps)
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))

    prof = eng.profile(args.profile_steps)
    tuning = eng.get_tuning()
    g = eng.read_grid(sid)
    # Late in the run: the same scene from its start, with the tuning above,
    # stepped to step late_steps, then 20 steps timed (the flow is far from the
    # initial lattice by then). A fresh engine, so the window is exact (the
    # autotuner's own substeps advanced the first one).
    late = None
    if args.late_steps:
        el = V.Engine()
        el.add(cfg, data, reactions=False)
        el.set_tuning(**tuning)
        el.step(args.late_steps)
        el.sync()
        t0 = time.perf_counter()
        el.step(20)
        el.sync()
        lw = time.perf_counter() - t0
        el.close()
        late = {"steps": f"{args.late_steps}-{args.late_steps + 20}",
                "ms_per_step": 1e3 * lw / 20, "value": NF * 20 / lw, "unit": "particle-updates/s"}
        print(f"# late run (steps {late['steps']}): {late['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    C = g["total"] + 1
    per_kernel = {k: v / args.profile_steps for k, v in prof.items()}
    scan_ms = sum(per_kernel.get(k, 0.0) for k in ("scan_reduce", "scan_tiles", "scan_apply"))
    kern = {k: v for k, v in per_kernel.items() if not k.startswith("scan_")}
    kern["scan"] = scan_ms
    dom = max(kern, key=kern.get)
    dom_bytes = stage_bytes(dom, N, NF, C)
    dom_gbs = dom_bytes / (kern[dom] * 1e-3) / 1e9
    stage_gbs = {k: round(stage_bytes(k, N, NF, C) / (v * 1e-3) / 1e9, 1) for k, v in kern.items() if v > 0}
    b_step = 104 * N + 60 * NF + 16 * C
    step_gbs = b_step / (ms_per_step * 1e-3) / 1e9
    print(f"# kernels (ms/step): " + ", ".join(f"{k} {v:.3f}" for k, v in sorted(kern.items(), key=lambda x: -x[1])),
          file=sys.stderr, flush=True)

    eng.close()
    # The reference's own grid rule at this size (GRID_MAX_BINS = 2^21, core/grid.rs:20): the bin doubles to 2h.
    ref_rule = None
    if args.budget_log2 != 21 and not args.no_ref_rule:
        cfg21, data21 = build_scene(V, block, d, tank, 1 << 21)
        e21 = V.Engine()
        s21 = e21.add(cfg21, data21, reactions=False)
        t21 = e21.autotune(2)
        e21.step(args.warmup)
        e21.sync()
        t0 = time.perf_counter()
        e21.step(args.steps)
        e21.sync()
        w21 = time.perf_counter() - t0
        g21 = e21.read_grid(s21)
        e21.close()
        ref_rule = {"grid_budget": "2^21", "bin_m": float(g21["bin"]), "dims": g21["dims"],
                    "ms_per_step": 1e3 * w21 / args.steps, "value": NF * args.steps / w21,
                    "unit": "particle-updates/s", "tuning": t21}
        print(f"# reference grid rule 2^21 (bin {float(g21['bin']):.3g} m): {ref_rule['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    stepper = None
    if not args.no_stepper:
        stepper = {}
        cfg2 = V.dam_break_config(0.02)
        data2 = V.dam_break(cfg2, (4.0, 2.0, 2.0), (100, 100, 100), 0.02, seed=2)
        stepper["dam_break_1M"] = bench_stepper(V, cfg2, data2, args.stepper_ticks)
        stepper["dam_break_10M"] = bench_stepper(V, cfg, data, max(args.stepper_ticks // 2, 1))
        print("# FluidStepper drop-in: " + ", ".join(f"{k} {v['ms_per_tick']:.3f} ms/tick"
                                                    for k, v in stepper.items()), file=sys.stderr, flush=True)
    batched = None if args.no_batch else bench_batched(V, args)
    served = None if args.no_serve else bench_served(V, args)
    walls = None if args.no_walls else bench_walls(V, args)

    issue = None if args.no_issue else issue_rate(kern, tuning, args)
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        print(f"# cpu baseline: oracle, {threads} threads, {args.cpu_steps} steps", file=sys.stderr, flush=True)
        cv, cdt = cpu_baseline(cfg, data, args.cpu_steps, threads)
        cpu = {"value": cv, "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "sample": f"{args.cpu_steps} steps of the same {NF}-fluid scene ({cdt:.1f} s), strict-f32 C++ oracle, "
                         f"phase-parallel over {threads} threads (the host share of one GPU; {os.cpu_count()} "
                         "cores visible), no boundary-force pass (as the GPU line)"}
        if args.cpu_single_steps > 0:
            print(f"# cpu baseline: oracle, 1 thread, {args.cpu_single_steps} step(s)", file=sys.stderr, flush=True)
            c1, c1dt = cpu_baseline(cfg, data, args.cpu_single_steps, 1)
            cpu["single_thread"] = {"value": c1, "unit": "particle-updates/s", "cores": 1,
                                    "sample": f"{args.cpu_single_steps} step(s) of the same scene ({c1dt:.1f} s), "
                                              "threads = 1 (the reference's default)"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered dam-break lattice + one-layer Akinci tank)",
        "config": {"workload": f"dam break {block[0]}^3 = {NF} fluid + {N - NF} Akinci tank particles, "
                               f"d = {d} m, h = {2 * d} m, tank {tank[0]}x{tank[1]}x{tank[2]} m, dt = 2.5e-4 s, "
                               f"grid budget 2^{args.budget_log2} (bin {float(g['bin']):.3g} m, dims {g['dims']})",
                   "fluid_particles": NF, "total_particles": N, "grid_cells": C - 1},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom, args),
                     "algorithmic_bytes": dom_bytes, "kernel_ms": round(kern[dom], 4)},
        "cpu_baseline": cpu,
        "issue_roofline": issue,
        "step_latency_ms": {"p50": p50, "p99": p99},
        "step_roofline": {"algorithmic_bytes_per_step": b_step, "achieved_GBs": round(step_gbs, 1),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "kernel_ms_per_step": {k: round(v, 4) for k, v in kern.items()},
        "kernel_algorithmic_GBs": stage_gbs,
        "tuning": tuning,
        "tuning_sweep": sweep,
        "late_run": late,
        "reference_grid_rule": ref_rule,
        "stepper": stepper,
        "batched": batched,
        "served": served,
        "walls": walls,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
