"""Python host mirror of the MI355X WCSPH C ABI (include/vxl_sph.h).

Mirrors the reference engine's interfaces for the fluid path:
  * ``Backend``      — PhysGpuBackend (crates/vxl-phys-gpu/src/lib.rs:20-29)
  * ``FluidStepper`` — FluidStepper / GpuFluidStepper
                       (crates/vxl-phys-core/src/interop.rs:293-316,
                        crates/vxl-phys-gpu/src/pipeline/stepper.rs:70-147)
  * ``Engine`` / ``Scene`` — FluidSystem step + accessors over a pool of
                       independent scenes (crates/vxl-phys-fluid/src/fluid_step.rs,
                       fluid_access.rs)

Everything calls libvxl_sph.so (hand-written gfx950 kernels). There is no CPU
fallback: if the library is missing or no device is visible the calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VXL_SPH_LIB names another build of the same library (the host-sanitizer
# builds the tests load); by default the one beside this file.
LIB_PATH = os.environ.get("VXL_SPH_LIB") or os.path.join(_HERE, "libvxl_sph.so")
ABI_VERSION = 1
# Force kernel variants of VxlSphTuning (vxl_sph.h VXL_SPH_FORCE_*).
FORCE_DEFAULT, FORCE_SCAN, FORCE_WALK, FORCE_WALK_AHEAD, FORCE_PAIRS, FORCE_QUEUE, FORCE_PAIRED, FORCE_STAGED = \
    0, 1, 2, 3, 4, 5, 6, 7

STATUS = {
    0: "OK", 1: "NULL_POINTER", 2: "VERSION_MISMATCH", 3: "SIZE_MISMATCH", 4: "NOT_READY",
    5: "PANIC_CAUGHT", 16: "INVALID_ARGUMENT", 17: "UNSUPPORTED", 18: "DEVICE_ERROR",
    19: "OUT_OF_MEMORY", 20: "CAPACITY", 21: "NOT_FOUND",
}


class VxlSphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code
        self.status = STATUS.get(code, str(code))


class Header(C.Structure):
    _fields_ = [("size", C.c_uint32), ("version", C.c_uint32)]


class FluidConfig(C.Structure):
    """VxlSphFluidConfig = FluidConfig (fluid/src/config.rs:6-69) + scene values."""
    _fields_ = [
        ("hdr", Header),
        ("rest_density", C.c_float),
        ("smoothing_radius", C.c_float),
        ("xsph_viscosity", C.c_float),
        ("sound_speed", C.c_float),
        ("gamma_tait", C.c_float),
        ("substeps", C.c_uint32),
        ("tensile_clamp", C.c_uint32),
        ("boundary_layers", C.c_uint32),
        ("gravity", C.c_float * 3),
        ("max_speed_frac", C.c_float),
        ("artificial_viscosity", C.c_float),
        ("particle_mass", C.c_float),
        ("spacing", C.c_float),
        ("max_bins", C.c_uint64),
        ("dt_tick", C.c_float),
        ("reserved", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [("hdr", Header), ("persist_blocks", C.c_uint32), ("force_kernel", C.c_uint32),
                ("block_threads", C.c_uint32), ("stage_records", C.c_uint32), ("scan_tile_cells", C.c_uint32),
                ("reserved", C.c_uint32)]


TUNING_FIELDS = ("persist_blocks", "force_kernel", "block_threads", "stage_records", "scan_tile_cells")


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("ms_total", C.c_float)]


class Span(C.Structure):
    _fields_ = [("body", C.c_uint32), ("origin", C.c_float * 3), ("start", C.c_uint32), ("end", C.c_uint32)]


class Reaction(C.Structure):
    _fields_ = [("body", C.c_uint32), ("force", C.c_float * 3), ("torque", C.c_float * 3)]


class GridInfo(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("inv", C.c_float), ("bin", C.c_float), ("dims", C.c_uint32 * 3),
                ("total", C.c_uint32)]


class BoxBody(C.Structure):
    _fields_ = [("id", C.c_uint32), ("half", C.c_float * 3), ("pos", C.c_float * 3), ("rot", C.c_float * 4),
                ("linvel", C.c_float * 3), ("angvel", C.c_float * 3)]


class RigidConfig(C.Structure):
    _fields_ = [("hdr", Header), ("substeps", C.c_uint32), ("gravity", C.c_float * 3), ("max_linear", C.c_float),
                ("max_angular", C.c_float), ("tank", C.c_float * 3), ("reserved", C.c_uint32)]


class RigidBox(C.Structure):
    _fields_ = [("box", BoxBody), ("density", C.c_float)]


class WallProvider(C.Structure):
    _fields_ = [("id", C.c_uint32), ("has_bounds", C.c_uint32), ("lo", C.c_float * 3), ("hi", C.c_float * 3)]


class WallFace(C.Structure):
    _fields_ = [("axis", C.c_uint32), ("plane", C.c_float), ("sign", C.c_float), ("provider", C.c_uint32)]


class DamBreakSpec(C.Structure):
    _fields_ = [("hdr", Header), ("tank", C.c_float * 3), ("block", C.c_uint32 * 3), ("d", C.c_float),
                ("jitter", C.c_float), ("seed", C.c_uint64), ("walls", C.c_uint32), ("reserved", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("dam", DamBreakSpec), ("sound_speed", C.c_float), ("artificial_viscosity", C.c_float),
                ("steps", C.c_uint32), ("nboxes", C.c_uint32), ("arrival_s", C.c_double), ("fill", C.c_double),
                ("boxes", BoxBody * 4), ("box_density", C.c_float * 4)]


class ServeConfig(C.Structure):
    _fields_ = [("hdr", Header), ("budget_bytes", C.c_uint64), ("time_scale", C.c_double),
                ("max_wall_s", C.c_double), ("max_resident", C.c_uint32), ("builders", C.c_uint32),
                ("max_bins", C.c_uint64), ("dt", C.c_float), ("device", C.c_int32), ("max_particles", C.c_uint64),
                ("admit_every", C.c_uint32), ("reserved", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("requests", C.c_uint32), ("completed", C.c_uint32), ("pool_steps", C.c_uint64),
                ("fluid_updates", C.c_uint64), ("wall_s", C.c_double), ("updates_per_s", C.c_double),
                ("step_ms_p50", C.c_double), ("step_ms_p99", C.c_double), ("step_ms_mean", C.c_double),
                ("loop_ms_p50", C.c_double), ("loop_ms_p99", C.c_double), ("latency_s_p50", C.c_double),
                ("latency_s_p99", C.c_double), ("wait_s_p50", C.c_double), ("wait_s_p99", C.c_double),
                ("peak_resident", C.c_uint32), ("rejected", C.c_uint32), ("peak_bytes", C.c_uint64),
                ("layouts", C.c_uint64), ("layout_ms_p50", C.c_double), ("layout_ms_p99", C.c_double),
                ("layout_phase_ms", C.c_double * 5)]


class ServeResult(C.Structure):
    _fields_ = [("index", C.c_uint32), ("n", C.c_uint32), ("n_fluid", C.c_uint32), ("steps", C.c_uint32),
                ("arrival_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("state_hash", C.c_uint64)]


_lib = None

# Every symbol declared in include/vxl_sph.h.
EXPORTS = (
    "vxl_sph_abi_version", "vxl_sph_backend_name", "vxl_sph_backend_available", "vxl_sph_backend_submit",
    "vxl_sph_device_count", "vxl_sph_last_error", "vxl_sph_default_config",
    "vxl_sph_engine_create", "vxl_sph_engine_destroy", "vxl_sph_engine_set_diagnostics",
    "vxl_sph_engine_set_graphs", "vxl_sph_engine_step", "vxl_sph_engine_sync", "vxl_sph_engine_device_bytes",
    "vxl_sph_engine_graph_stats",
    "vxl_sph_engine_profile", "vxl_sph_engine_step_timed", "vxl_sph_engine_set_tuning",
    "vxl_sph_engine_get_tuning", "vxl_sph_engine_autotune",
    "vxl_sph_scene_add", "vxl_sph_scene_remove", "vxl_sph_scene_upload_boundary", "vxl_sph_scene_read_state",
    "vxl_sph_scene_read_fields", "vxl_sph_scene_read_grid", "vxl_sph_scene_read_bounds",
    "vxl_sph_scene_state_hash",
    "vxl_sph_stepper_create", "vxl_sph_stepper_destroy", "vxl_sph_stepper_step", "vxl_sph_stepper_reactions",
    "vxl_sph_stepper_bounds", "vxl_sph_stepper_read_state",
    "vxl_sph_box_lattice", "vxl_sph_box_particles", "vxl_sph_fluid_constants", "vxl_sph_dam_break",
    "vxl_sph_default_rigid_config", "vxl_sph_scene_add_coupled", "vxl_sph_scene_read_bodies",
    "vxl_sph_request_bodies", "vxl_sph_scene_state_hash_all", "vxl_sph_vxld_open", "vxl_sph_vxld_frame",
    "vxl_sph_vxld_frame_raw", "vxl_sph_vxld_close",
    "vxl_sph_sample_batch", "vxl_sph_sample_requests", "vxl_sph_request_config", "vxl_sph_serve",
    "vxl_sph_scene_set_halfspace_walls", "vxl_sph_scene_set_wall_tables", "vxl_sph_scene_clear_walls",
    "vxl_sph_stepper_set_wall_tables", "vxl_sph_stepper_set_halfspace_walls",
    "vxl_sph_request_bytes", "vxl_sph_pool_limits", "vxl_sph_step_device", "vxl_sph_step_device_noalloc",
)


def lib():
    """Load libvxl_sph.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    up = C.POINTER(C.c_uint32)
    st = C.c_int
    L.vxl_sph_abi_version.restype = C.c_uint32
    L.vxl_sph_backend_name.restype = C.c_char_p
    L.vxl_sph_backend_available.restype = C.c_int32
    L.vxl_sph_backend_submit.restype = C.c_int32
    L.vxl_sph_backend_submit.argtypes = [C.c_char_p]
    L.vxl_sph_last_error.restype = C.c_char_p
    L.vxl_sph_device_count.restype = st
    L.vxl_sph_device_count.argtypes = [C.POINTER(C.c_int32)]
    L.vxl_sph_default_config.restype = st
    L.vxl_sph_default_config.argtypes = [C.POINTER(FluidConfig)]
    L.vxl_sph_engine_create.restype = st
    L.vxl_sph_engine_create.argtypes = [C.c_uint32, C.c_int32, C.POINTER(P)]
    for name in ("vxl_sph_engine_destroy", "vxl_sph_engine_sync"):
        getattr(L, name).restype = st
        getattr(L, name).argtypes = [P]
    for name in ("vxl_sph_engine_set_diagnostics", "vxl_sph_engine_set_graphs"):
        getattr(L, name).restype = st
        getattr(L, name).argtypes = [P, C.c_int32]
    L.vxl_sph_engine_step.restype = st
    L.vxl_sph_engine_step.argtypes = [P, C.c_uint32]
    L.vxl_sph_engine_device_bytes.restype = st
    L.vxl_sph_engine_device_bytes.argtypes = [P, C.POINTER(C.c_uint64)]
    L.vxl_sph_engine_graph_stats.restype = st
    L.vxl_sph_engine_graph_stats.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.vxl_sph_engine_set_tuning.restype = st
    L.vxl_sph_engine_set_tuning.argtypes = [P, C.POINTER(Tuning)]
    L.vxl_sph_engine_get_tuning.restype = st
    L.vxl_sph_engine_get_tuning.argtypes = [P, C.POINTER(Tuning)]
    L.vxl_sph_engine_autotune.restype = st
    L.vxl_sph_engine_autotune.argtypes = [P, C.c_uint32, C.POINTER(Tuning)]
    L.vxl_sph_engine_profile.restype = st
    L.vxl_sph_engine_profile.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(KernelTime), up]
    L.vxl_sph_engine_step_timed.restype = st
    L.vxl_sph_engine_step_timed.argtypes = [P, C.c_uint32, fp]
    L.vxl_sph_scene_add.restype = st
    L.vxl_sph_scene_add.argtypes = [P, C.POINTER(FluidConfig), C.c_uint32, C.c_uint32, fp, fp, fp,
                                    C.POINTER(Span), C.c_uint32, up]
    L.vxl_sph_scene_remove.restype = st
    L.vxl_sph_scene_remove.argtypes = [P, C.c_uint32]
    L.vxl_sph_scene_upload_boundary.restype = st
    L.vxl_sph_scene_upload_boundary.argtypes = [P, C.c_uint32, C.c_uint32, fp, fp]
    L.vxl_sph_scene_read_state.restype = st
    L.vxl_sph_scene_read_state.argtypes = [P, C.c_uint32, C.c_uint32, fp, fp]
    L.vxl_sph_scene_read_fields.restype = st
    L.vxl_sph_scene_read_fields.argtypes = [P, C.c_uint32, C.c_uint32, fp, fp, up, fp, fp]
    L.vxl_sph_scene_read_grid.restype = st
    L.vxl_sph_scene_read_grid.argtypes = [P, C.c_uint32, C.POINTER(GridInfo), C.c_uint32, up, C.c_uint32, up, up]
    L.vxl_sph_scene_read_bounds.restype = st
    L.vxl_sph_scene_read_bounds.argtypes = [P, C.c_uint32, fp, fp]
    L.vxl_sph_scene_state_hash.restype = st
    L.vxl_sph_scene_state_hash.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint64)]
    L.vxl_sph_stepper_create.restype = st
    L.vxl_sph_stepper_create.argtypes = [C.c_uint32, C.POINTER(FluidConfig), C.c_int32, C.POINTER(P)]
    L.vxl_sph_stepper_destroy.restype = st
    L.vxl_sph_stepper_destroy.argtypes = [P]
    L.vxl_sph_stepper_step.restype = st
    L.vxl_sph_stepper_step.argtypes = [P, C.c_float, C.c_uint32, C.c_uint32, fp, fp, fp, C.POINTER(Span),
                                       C.c_uint32]
    L.vxl_sph_stepper_reactions.restype = st
    L.vxl_sph_stepper_reactions.argtypes = [P, C.c_uint32, C.POINTER(Reaction), up]
    L.vxl_sph_stepper_bounds.restype = st
    L.vxl_sph_stepper_bounds.argtypes = [P, fp, fp]
    L.vxl_sph_stepper_read_state.restype = st
    L.vxl_sph_stepper_read_state.argtypes = [P, C.c_uint32, fp, fp]
    L.vxl_sph_default_rigid_config.restype = st
    L.vxl_sph_default_rigid_config.argtypes = [C.POINTER(RigidConfig)]
    L.vxl_sph_scene_add_coupled.restype = st
    L.vxl_sph_scene_add_coupled.argtypes = [P, C.POINTER(FluidConfig), C.POINTER(RigidConfig), C.c_uint32, fp, fp,
                                            C.c_uint32, C.POINTER(RigidBox), up]
    L.vxl_sph_scene_read_bodies.restype = st
    L.vxl_sph_scene_read_bodies.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(RigidBox), up]
    L.vxl_sph_request_bodies.restype = st
    L.vxl_sph_request_bodies.argtypes = [C.POINTER(Request), C.c_uint32, C.POINTER(RigidBox), up,
                                         C.POINTER(RigidConfig)]
    L.vxl_sph_scene_state_hash_all.restype = st
    L.vxl_sph_scene_state_hash_all.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint64)]
    L.vxl_sph_vxld_open.restype = st
    L.vxl_sph_vxld_open.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.POINTER(P)]
    L.vxl_sph_vxld_frame.restype = st
    L.vxl_sph_vxld_frame.argtypes = [P, P, C.c_uint32, C.c_float, C.c_uint32, up]
    L.vxl_sph_vxld_frame_raw.restype = st
    L.vxl_sph_vxld_frame_raw.argtypes = [P, C.c_uint32, C.c_float, C.c_uint32, C.POINTER(RigidBox), C.c_uint32, up,
                                         fp]
    L.vxl_sph_vxld_close.restype = st
    L.vxl_sph_vxld_close.argtypes = [P]
    L.vxl_sph_box_lattice.restype = st
    L.vxl_sph_box_lattice.argtypes = [fp, C.c_float, C.c_float, C.c_uint32, C.c_uint32, fp, up, up, fp]
    L.vxl_sph_box_particles.restype = st
    L.vxl_sph_box_particles.argtypes = [C.POINTER(FluidConfig), C.c_uint32, C.POINTER(BoxBody), C.c_uint32, up, up,
                                        fp, fp, fp, C.POINTER(Span)]
    L.vxl_sph_sample_batch.restype = st
    L.vxl_sph_sample_batch.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Request)]
    L.vxl_sph_sample_requests.restype = st
    L.vxl_sph_sample_requests.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.POINTER(Request)]
    L.vxl_sph_request_config.restype = st
    L.vxl_sph_request_config.argtypes = [C.POINTER(Request), C.c_float, C.c_uint64, C.POINTER(FluidConfig)]
    L.vxl_sph_step_device.restype = st
    L.vxl_sph_step_device.argtypes = [C.POINTER(FluidConfig), C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_uint32, C.c_void_p]
    L.vxl_sph_step_device_noalloc.restype = st
    L.vxl_sph_step_device_noalloc.argtypes = L.vxl_sph_step_device.argtypes
    L.vxl_sph_request_bytes.restype = st
    L.vxl_sph_request_bytes.argtypes = [C.POINTER(Request), C.c_float, C.c_uint64, C.POINTER(C.c_uint64)]
    L.vxl_sph_pool_limits.restype = st
    L.vxl_sph_pool_limits.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), up]
    L.vxl_sph_serve.restype = st
    L.vxl_sph_serve.argtypes = [C.POINTER(ServeConfig), C.c_uint32, C.POINTER(Request), C.POINTER(ServeStats),
                                C.POINTER(ServeResult)]
    L.vxl_sph_scene_set_halfspace_walls.restype = st
    L.vxl_sph_scene_set_halfspace_walls.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(WallProvider), C.c_uint32,
                                                    C.POINTER(WallFace)]
    L.vxl_sph_scene_set_wall_tables.restype = st
    L.vxl_sph_scene_set_wall_tables.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, up, up, C.c_uint32, fp]
    L.vxl_sph_scene_clear_walls.restype = st
    L.vxl_sph_scene_clear_walls.argtypes = [P, C.c_uint32]
    L.vxl_sph_stepper_set_wall_tables.restype = st
    L.vxl_sph_stepper_set_wall_tables.argtypes = [P, C.c_uint32, C.c_uint32, up, up, C.c_uint32, fp]
    L.vxl_sph_stepper_set_halfspace_walls.restype = st
    L.vxl_sph_stepper_set_halfspace_walls.argtypes = [P, C.c_uint32, C.POINTER(WallProvider), C.c_uint32,
                                                      C.POINTER(WallFace)]
    L.vxl_sph_fluid_constants.restype = st
    L.vxl_sph_fluid_constants.argtypes = [C.POINTER(FluidConfig), fp]
    L.vxl_sph_dam_break.restype = st
    L.vxl_sph_dam_break.argtypes = [C.POINTER(DamBreakSpec), C.POINTER(FluidConfig), up, up, up, fp, fp, fp,
                                    C.POINTER(Span)]
    _lib = L
    return L


def _check(code):
    if code != 0:
        raise VxlSphError(code, (lib().vxl_sph_last_error() or b"").decode(errors="replace"))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


def fluid_config(**kw):
    """FluidConfig::default() (config.rs:51-69) with overrides."""
    c = FluidConfig()
    _check(lib().vxl_sph_default_config(C.byref(c)))
    for k, v in kw.items():
        if k == "gravity":
            c.gravity = (C.c_float * 3)(*v)
        else:
            setattr(c, k, v)
    return c


def dam_break_config(d, c0=20.0, alpha=0.02, eps=0.5, dt=5e-4, layers=1, max_bins=1 << 21):
    """Physics of the dam-break scenes: h = 2d, mass = 1000 d^3, Tait gamma 7,
    one substep per tick of length dt."""
    d = np.float32(d)
    rho0 = np.float32(1000.0)
    mass = np.float32(np.float32(np.float32(rho0 * d) * d) * d)
    return fluid_config(smoothing_radius=float(np.float32(2.0) * d), sound_speed=c0, artificial_viscosity=alpha,
                        xsph_viscosity=eps, substeps=1, boundary_layers=layers, particle_mass=float(mass),
                        spacing=float(d), max_bins=max_bins, dt_tick=dt)


def fluid_constants(cfg):
    o = np.zeros(8, np.float32)
    _check(lib().vxl_sph_fluid_constants(C.byref(cfg), _f(o)))
    return dict(zip(["h", "h2", "k6", "ks", "w0", "b_tait", "mass", "alpha_c"], o))


def box_lattice(half, s, h, layers):
    """boundary::lattice for Shape::Box (boundary.rs:67-88)."""
    half = np.asarray(half, np.float32)
    n = C.c_uint32()
    _check(lib().vxl_sph_box_lattice(_f(half), C.c_float(s), C.c_float(h), layers, 0, None, C.byref(n), None, None))
    pts = np.zeros((max(n.value, 1), 3), np.float32)
    n0 = C.c_uint32()
    vol = C.c_float()
    _check(lib().vxl_sph_box_lattice(_f(half), C.c_float(s), C.c_float(h), layers, n.value, _f(pts), C.byref(n),
                                     C.byref(n0), C.byref(vol)))
    return pts[:n.value], n0.value, np.float32(vol.value)


def _fill_box(o, b):
    o.id = b["id"]
    o.half = (C.c_float * 3)(*b["half"])
    o.pos = (C.c_float * 3)(*b["pos"])
    o.rot = (C.c_float * 4)(*b.get("rot", (0.0, 0.0, 0.0, 1.0)))
    o.linvel = (C.c_float * 3)(*b.get("linvel", (0.0, 0.0, 0.0)))
    o.angvel = (C.c_float * 3)(*b.get("angvel", (0.0, 0.0, 0.0)))


def rigid_config(substeps=None, gravity=None, max_linear=None, max_angular=None, tank=None):
    """Rigid configuration of a coupled scene (PhysConfig defaults where None)."""
    c = RigidConfig()
    _check(lib().vxl_sph_default_rigid_config(C.byref(c)))
    if substeps is not None:
        c.substeps = substeps
    if gravity is not None:
        c.gravity = (C.c_float * 3)(*gravity)
    if max_linear is not None:
        c.max_linear = max_linear
    if max_angular is not None:
        c.max_angular = max_angular
    if tank is not None:
        c.tank = (C.c_float * 3)(*tank)
    return c


def box_particles(cfg, bodies, first):
    """set_boundary_particles for Box bodies (fluid_step.rs:13-53): returns
    (pos, vel, pmass, spans) with span indices starting at `first`."""
    arr = (BoxBody * max(len(bodies), 1))()
    for k, b in enumerate(bodies):
        _fill_box(arr[k], b)
    n, ns = C.c_uint32(), C.c_uint32()
    _check(lib().vxl_sph_box_particles(C.byref(cfg), len(bodies), arr, first, C.byref(n), C.byref(ns),
                                       None, None, None, None))
    pos = np.zeros((n.value, 3), np.float32)
    vel = np.zeros((n.value, 3), np.float32)
    pm = np.zeros(n.value, np.float32)
    sp = (Span * max(ns.value, 1))()
    _check(lib().vxl_sph_box_particles(C.byref(cfg), len(bodies), arr, first, C.byref(n), C.byref(ns),
                                       _f(pos), _f(vel), _f(pm), sp))
    spans = [(sp[k].body, tuple(sp[k].origin), sp[k].start, sp[k].end) for k in range(ns.value)]
    return pos, vel, pm, spans


class SceneData:
    """All particles of a scene in index order (fluid prefix first)."""

    def __init__(self, pos, vel, pmass, n_fluid, spans):
        self.pos = pos
        self.vel = vel
        self.pmass = pmass
        self.n_fluid = n_fluid
        self.spans = spans  # list of (body, origin(3), start, end)

    @property
    def n(self):
        return len(self.pmass)


def dam_break(cfg, tank, block, d, jitter=0.05, seed=1, walls=True):
    """Seeded dam-break scene built by the library (jittered lattice + Akinci tank)."""
    spec = DamBreakSpec()
    spec.hdr.size = C.sizeof(DamBreakSpec)
    spec.hdr.version = ABI_VERSION
    spec.tank = (C.c_float * 3)(*tank)
    spec.block = (C.c_uint32 * 3)(*block)
    spec.d = d
    spec.jitter = jitter
    spec.seed = seed
    spec.walls = 1 if walls else 0
    n, nf, ns = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().vxl_sph_dam_break(C.byref(spec), C.byref(cfg), C.byref(n), C.byref(nf), C.byref(ns),
                                   None, None, None, None))
    pos = np.zeros((n.value, 3), np.float32)
    vel = np.zeros((n.value, 3), np.float32)
    pm = np.zeros(n.value, np.float32)
    sp = (Span * max(ns.value, 1))()
    _check(lib().vxl_sph_dam_break(C.byref(spec), C.byref(cfg), C.byref(n), C.byref(nf), C.byref(ns),
                                   _f(pos), _f(vel), _f(pm), sp))
    spans = [(sp[k].body, tuple(sp[k].origin), sp[k].start, sp[k].end) for k in range(ns.value)]
    return SceneData(pos, vel, pm, nf.value, spans)


def _vec3(a, n, what):
    """(n, 3) contiguous f32 copy of a per-particle vector array; anything of
    another length is a size mismatch (the C side reads exactly 3 n floats)."""
    a = np.ascontiguousarray(a, np.float32)
    if a.size != 3 * n or (a.ndim == 2 and a.shape[1] != 3) or a.ndim > 2:
        raise VxlSphError(3, f"{what} holds {a.size} floats, expected {n} x 3")
    return a.reshape(n, 3)


def _scalars(a, n, what):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    if a.size != n:
        raise VxlSphError(3, f"{what} holds {a.size} values, expected {n}")
    return a


def _spans_arr(spans):
    arr = (Span * max(len(spans), 1))()
    for k, (body, origin, s, e) in enumerate(spans):
        arr[k].body = body
        arr[k].origin = (C.c_float * 3)(*origin)
        arr[k].start = s
        arr[k].end = e
    return arr


class Backend:
    """PhysGpuBackend (gpu/src/lib.rs:20-29)."""

    def name(self):
        return lib().vxl_sph_backend_name().decode()

    def available(self):
        return bool(lib().vxl_sph_backend_available())

    def submit(self, tag):
        return bool(lib().vxl_sph_backend_submit(tag.encode()))


class Engine:
    """A pool of independent scenes on one device, stepped together."""

    def __init__(self, device=0, diagnostics=False, graphs=True):
        self.h = C.c_void_p()
        _check(lib().vxl_sph_engine_create(ABI_VERSION, device, C.byref(self.h)))
        if diagnostics:
            _check(lib().vxl_sph_engine_set_diagnostics(self.h, 1))
        if not graphs:
            _check(lib().vxl_sph_engine_set_graphs(self.h, 0))
        self.scenes = {}

    def close(self):
        if self.h:
            lib().vxl_sph_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, cfg, data, reactions=True):
        pm = np.ascontiguousarray(data.pmass, np.float32).reshape(-1)
        pos = _vec3(data.pos, len(pm), "pos")
        vel = _vec3(data.vel, len(pm), "vel")
        spans = data.spans if reactions else []
        sid = C.c_uint32()
        _check(lib().vxl_sph_scene_add(self.h, C.byref(cfg), data.n, data.n_fluid, _f(pos), _f(vel), _f(pm),
                                       _spans_arr(spans), len(spans), C.byref(sid)))
        self.scenes[sid.value] = (data.n, data.n_fluid, len(spans))
        return sid.value

    def remove(self, sid):
        _check(lib().vxl_sph_scene_remove(self.h, sid))
        self.scenes.pop(sid, None)

    def step(self, ticks=1):
        _check(lib().vxl_sph_engine_step(self.h, ticks))

    def sync(self):
        _check(lib().vxl_sph_engine_sync(self.h))

    def device_bytes(self):
        v = C.c_uint64()
        _check(lib().vxl_sph_engine_device_bytes(self.h, C.byref(v)))
        return v.value

    def graph_stats(self):
        """(tick graphs instantiated, tick graphs updated in place after a layout)."""
        b, u = C.c_uint64(), C.c_uint64()
        _check(lib().vxl_sph_engine_graph_stats(self.h, C.byref(b), C.byref(u)))
        return b.value, u.value

    def set_tuning(self, persist_blocks=0, force_kernel=FORCE_DEFAULT, block_threads=0, stage_records=0,
                   scan_tile_cells=0):
        """force_kernel: FORCE_SCAN | FORCE_WALK | FORCE_WALK_AHEAD | FORCE_PAIRS | FORCE_QUEUE |
        FORCE_PAIRED | FORCE_STAGED; persist_blocks: grid size cap of the
        per-slot kernels (0 = one block per 256 slots); block_threads: workgroup size of the density and
        force kernels (64, 128, 256); stage_records: density LDS staging depth (128, 256, 512);
        scan_tile_cells: cells per scan tile (1024, 2048, 4096). 0 keeps a field's default. Every
        choice gives identical results."""
        t = Tuning()
        t.hdr.size = C.sizeof(Tuning)
        t.hdr.version = ABI_VERSION
        t.persist_blocks = persist_blocks
        t.force_kernel = force_kernel
        t.block_threads = block_threads
        t.stage_records = stage_records
        t.scan_tile_cells = scan_tile_cells
        _check(lib().vxl_sph_engine_set_tuning(self.h, C.byref(t)))

    def autotune(self, substeps=2):
        """Pick the fastest launch variant on the current pool (advances it by
        real substeps; results are identical for every variant)."""
        t = Tuning()
        _check(lib().vxl_sph_engine_autotune(self.h, substeps, C.byref(t)))
        return {f: getattr(t, f) for f in TUNING_FIELDS}

    def get_tuning(self):
        t = Tuning()
        _check(lib().vxl_sph_engine_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f in TUNING_FIELDS}

    def profile(self, substeps):
        """{kernel name: total ms} over `substeps` eager substeps (HIP events)."""
        arr = (KernelTime * 32)()
        cnt = C.c_uint32()
        _check(lib().vxl_sph_engine_profile(self.h, substeps, 32, arr, C.byref(cnt)))
        return {arr[k].name.decode(): float(arr[k].ms_total) for k in range(min(cnt.value, 32))}

    def step_timed(self, ticks):
        """Device ms of each of `ticks` ticks."""
        ms = np.zeros(max(ticks, 1), np.float32)
        _check(lib().vxl_sph_engine_step_timed(self.h, ticks, _f(ms)))
        return ms[:ticks]

    def upload_boundary(self, sid, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = _vec3(vel, len(pos), "vel")
        _check(lib().vxl_sph_scene_upload_boundary(self.h, sid, len(pos), _f(pos), _f(vel)))

    def read_state(self, sid):
        n = self.scenes[sid][0]
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        _check(lib().vxl_sph_scene_read_state(self.h, sid, n, _f(pos), _f(vel)))
        return pos, vel

    def read_fields(self, sid, diagnostics=False):
        n, nf, _ = self.scenes[sid]
        dens = np.zeros(n, np.float32)
        press = np.zeros(n, np.float32)
        if not diagnostics:
            _check(lib().vxl_sph_scene_read_fields(self.h, sid, n, _f(dens), _f(press), None, None, None))
            return dict(dens=dens, press=press)
        ncount = np.zeros(n, np.uint32)
        acc = np.zeros((nf, 3), np.float32)
        xs = np.zeros((nf, 3), np.float32)
        _check(lib().vxl_sph_scene_read_fields(self.h, sid, n, _f(dens), _f(press), _u(ncount), _f(acc), _f(xs)))
        return dict(dens=dens, press=press, ncount=ncount, acc=acc, xsph=xs)

    def read_grid(self, sid):
        n = self.scenes[sid][0]
        info = GridInfo()
        _check(lib().vxl_sph_scene_read_grid(self.h, sid, C.byref(info), 0, None, 0, None, None))
        start = np.zeros(info.total + 1, np.uint32)
        items = np.zeros(n, np.uint32)
        keys = np.zeros(n, np.uint32)
        _check(lib().vxl_sph_scene_read_grid(self.h, sid, C.byref(info), len(start), _u(start), n, _u(items),
                                             _u(keys)))
        return dict(min=np.array(info.min, np.float32), inv=np.float32(info.inv), bin=np.float32(info.bin),
                    dims=[int(x) for x in info.dims], total=int(info.total), start=start, items=items, keys=keys)

    def add_coupled(self, cfg, rigid_cfg, fluid_pos, bodies, fluid_vel=None):
        """Scene of fluid particles and Box bodies (dicts with id, half, pos,
        rot, linvel, angvel, density; density > 0 = dynamic, coupled both ways)."""
        fp_ = np.ascontiguousarray(fluid_pos, np.float32).reshape(-1, 3)
        fv = None if fluid_vel is None else np.ascontiguousarray(fluid_vel, np.float32).reshape(-1, 3)
        arr = (RigidBox * max(len(bodies), 1))()
        for k, b in enumerate(bodies):
            _fill_box(arr[k].box, b)
            arr[k].density = b.get("density", 0.0)
        sid = C.c_uint32()
        _check(lib().vxl_sph_scene_add_coupled(self.h, C.byref(cfg), C.byref(rigid_cfg), len(fp_), _f(fp_),
                                               _f(fv) if fv is not None else None, len(bodies), arr,
                                               C.byref(sid)))
        # particle count: fluid + every body's lattice (one span per body with particles)
        counts = [len(box_lattice(b["half"], cfg.spacing, max(cfg.smoothing_radius, 1e-5), cfg.boundary_layers)[0])
                  for b in bodies]
        self.scenes[sid.value] = (len(fp_) + sum(counts), len(fp_), sum(1 for c in counts if c))
        return sid.value

    def read_bodies(self, sid):
        """Current state of the scene's dynamic bodies: list of dicts."""
        cnt = C.c_uint32()
        _check(lib().vxl_sph_scene_read_bodies(self.h, sid, 0, None, C.byref(cnt)))
        arr = (RigidBox * max(cnt.value, 1))()
        _check(lib().vxl_sph_scene_read_bodies(self.h, sid, cnt.value, arr, C.byref(cnt)))
        out = []
        for k in range(cnt.value):
            b = arr[k].box
            out.append(dict(id=b.id, half=np.array(b.half[:], np.float32), pos=np.array(b.pos[:], np.float32),
                            rot=np.array(b.rot[:], np.float32), linvel=np.array(b.linvel[:], np.float32),
                            angvel=np.array(b.angvel[:], np.float32), density=arr[k].density))
        return out

    def read_bounds(self, sid):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        _check(lib().vxl_sph_scene_read_bounds(self.h, sid, _f(lo), _f(hi)))
        return lo, hi

    def state_hash_all(self, sid):
        """Byte-wise FNV-1a over all positions then all velocities (packet_grid_canary.rs:102-111)."""
        v = C.c_uint64()
        _check(lib().vxl_sph_scene_state_hash_all(self.h, sid, C.byref(v)))
        return v.value

    def state_hash(self, sid):
        v = C.c_uint64()
        _check(lib().vxl_sph_scene_state_hash(self.h, sid, C.byref(v)))
        return v.value

    def set_halfspace_walls(self, sid, providers):
        """Half-space providers (HalfspaceProvider) queried on the device every
        substep with the CPU engine's semantics; [] removes them."""
        provs, np_, faces, nf = _wall_arrays(providers)
        _check(lib().vxl_sph_scene_set_halfspace_walls(self.h, sid, np_, provs, nf, faces))

    def set_wall_tables(self, sid, side, ids, start, planes):
        """Plane table of side 0 (mirror) / 1 (project), as gather_wall_contacts
        returns it; held until replaced."""
        t = _Table(ids, start, planes)
        _check(lib().vxl_sph_scene_set_wall_tables(self.h, sid, side, *t.args()))

    def clear_walls(self, sid):
        _check(lib().vxl_sph_scene_clear_walls(self.h, sid))


class FluidStepper:
    """FluidStepper drop-in with GpuFluidStepper semantics (pipeline/stepper.rs:70-147)."""

    def __init__(self, cfg, device=0):
        self.h = C.c_void_p()
        _check(lib().vxl_sph_stepper_create(ABI_VERSION, C.byref(cfg), device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().vxl_sph_stepper_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, dt_tick, n_fluid, pos, vel, pmass, spans):
        # GpuFluidStepper::step asserts these lengths (pipeline/stepper.rs:84-96).
        pmass = np.ascontiguousarray(pmass, np.float32).reshape(-1)
        pos = _vec3(pos, len(pmass), "pos")
        vel = _vec3(vel, len(pmass), "vel")
        if getattr(self, "_gather", None) is not None:
            # gather_walls before the first step: the positions this step starts from.
            h, providers = self._gather
            self._gather = None
            self._upload_walls(h, providers, pos.reshape(-1, 3)[:n_fluid])
        self._n = len(pmass)
        self._nf = n_fluid
        _check(lib().vxl_sph_stepper_step(self.h, C.c_float(dt_tick), n_fluid, len(pmass), _f(pos), _f(vel),
                                          _f(pmass), _spans_arr(spans), len(spans)))

    def set_boundaries(self, ids):
        """Provider ids whose walls gather_walls collects (GpuFluidStepper::new's
        `boundaries`, pipeline/stepper.rs:34-56)."""
        self._boundaries = [int(b) for b in ids]

    def gather_walls(self, h, providers):
        """FluidStepper::gather_walls (interop.rs:313-315; pipeline/stepper.rs:118-137):
        both plane tables from the stepper's current fluid positions; before the
        first step, from the positions that step starts from."""
        if not getattr(self, "_boundaries", None):
            return
        if not hasattr(self, "_n"):
            self._gather = (h, providers)
            return
        pos, _ = self.read_state()
        self._upload_walls(h, providers, pos[:self._nf])

    def _upload_walls(self, h, providers, ps):
        for side, fn in ((0, gather_wall_contacts), (1, gather_wall_project_contacts)):
            t = _Table(*fn(self._boundaries, h, ps, providers))
            _check(lib().vxl_sph_stepper_set_wall_tables(self.h, side, *t.args()))

    def set_halfspace_walls(self, providers):
        """Half-space providers queried on the device every substep (CPU semantics)."""
        provs, np_, faces, nf = _wall_arrays(providers)
        _check(lib().vxl_sph_stepper_set_halfspace_walls(self.h, np_, provs, nf, faces))

    def reactions(self):
        cnt = C.c_uint32()
        _check(lib().vxl_sph_stepper_reactions(self.h, 0, None, C.byref(cnt)))
        arr = (Reaction * max(cnt.value, 1))()
        _check(lib().vxl_sph_stepper_reactions(self.h, cnt.value, arr, C.byref(cnt)))
        return [(arr[k].body, np.array(arr[k].force, np.float32), np.array(arr[k].torque, np.float32))
                for k in range(cnt.value)]

    def bounds(self):
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        _check(lib().vxl_sph_stepper_bounds(self.h, _f(lo), _f(hi)))
        return lo, hi

    def read_state(self):
        pos = np.zeros((self._n, 3), np.float32)
        vel = np.zeros((self._n, 3), np.float32)
        _check(lib().vxl_sph_stepper_read_state(self.h, self._n, _f(pos), _f(vel)))
        return pos, vel


# --- provider walls (host side) ----------------------------------------------

class HalfspaceProvider:
    """A wall provider made of axis-aligned half-space faces, with the
    ProviderColliders point queries (interop.rs:166-224): face (axis, plane,
    sign) has its fluid side at sign * (x[axis] - plane) > 0; a query point p
    within 2 * probe of it gets the contact (p with p[axis] = plane, normal
    sign * e_axis). The reference's test doubles FloorY and Tank
    (fluid/src/tests.rs:11-110) are such providers; the device evaluates the
    same queries itself when the provider is handed to set_halfspace_walls."""

    def __init__(self, faces, bounds=None, id=0):
        self.id = int(id)
        self.faces = [(int(a), np.float32(p), np.float32(sg)) for (a, p, sg) in faces]
        self._bounds = None if bounds is None else (np.asarray(bounds[0], np.float32),
                                                    np.asarray(bounds[1], np.float32))

    def bounds(self, id=None):
        return self._bounds

    def contacts_point(self, id, p, skin):
        """[(point, normal)] for one query point (the reference's per-point query)."""
        p = np.asarray(p, np.float32)
        out = []
        for (axis, plane, sign) in self.faces:
            if (p[axis] - plane) * sign > np.float32(2.0) * np.float32(skin):
                continue
            pt = p.copy()
            pt[axis] = plane
            n = np.zeros(3, np.float32)
            n[axis] = sign
            out.append((pt, n))
        return out

    def contacts_point_boundary(self, id, p, skin):
        return self.contacts_point(id, p, skin)

    def contacts_points(self, id, ps, probe):
        """The same query for many points at once: per face in order,
        (mask, contact points[n, 3], normal)."""
        for (axis, plane, sign) in self.faces:
            mask = ~((ps[:, axis] - plane) * sign > np.float32(2.0) * np.float32(probe))
            pt = ps.copy()
            pt[:, axis] = plane
            n = np.zeros(3, np.float32)
            n[axis] = sign
            yield mask, pt, n


FLOOR_Y = HalfspaceProvider([(1, 0.0, 1.0)], bounds=((-50.0, -1.0, -50.0), (50.0, 0.0, 50.0)))
TANK = HalfspaceProvider([(1, 0.0, 1.0), (0, -0.2, 1.0), (0, 0.2, -1.0), (2, -0.2, 1.0), (2, 0.2, -1.0)],
                         bounds=((-50.0, -50.0, -50.0), (50.0, 50.0, 50.0)))


class Providers:
    """A ProviderColliders set: provider objects by id."""

    def __init__(self, by_id):
        self.by_id = dict(by_id)

    def bounds(self, id):
        pv = self.by_id.get(id)
        return None if pv is None else pv.bounds(id)

    def contacts_point(self, id, p, skin):
        pv = self.by_id.get(id)
        return None if pv is None else pv.contacts_point(id, p, skin)

    def contacts_point_boundary(self, id, p, skin):
        pv = self.by_id.get(id)
        return None if pv is None else pv.contacts_point_boundary(id, p, skin)


def _gather(boundaries, h, ps, providers, project):
    # gather_in (pipeline/walls.rs:168-216): per particle, providers in order,
    # contacts in order; keep 0 < sdf < h (mirror) or sdf < h (project), at
    # most 8 per particle. Providers with a batched query (contacts_points)
    # are evaluated for all particles at once in the same f32 operations.
    ps = np.ascontiguousarray(ps, np.float32).reshape(-1, 3)
    n = len(ps)
    h = np.float32(h)
    count = np.zeros(n, np.int64)
    rows = []  # (particle, order, point, normal)
    order = 0
    for bid in boundaries:
        bb = providers.bounds(bid)
        live = np.ones(n, bool)
        if bb is not None:
            lo, hi = (np.asarray(b, np.float32) for b in bb)
            for a in range(3):
                live &= ~((ps[:, a] < lo[a] - h) | (ps[:, a] > hi[a] + h))
        pv = providers.by_id.get(bid) if isinstance(providers, Providers) else None
        batched = pv is not None and hasattr(pv, "contacts_points")
        if batched:
            contacts = pv.contacts_points(bid, ps, h)
        else:
            contacts = []
            per = [None] * n
            for i in np.nonzero(live)[0]:
                q = (providers.contacts_point_boundary if project else providers.contacts_point)(bid, ps[i], h)
                per[i] = q or []
            width = max((len(q) for q in per if q), default=0)
            for k in range(width):
                mask = np.array([bool(q) and len(q) > k for q in per], bool)
                pt = np.zeros((n, 3), np.float32)
                nr = np.zeros((n, 3), np.float32)
                for i in np.nonzero(mask)[0]:
                    pt[i], nr[i] = per[i][k]
                contacts.append((mask, pt, nr))
        for mask, pt, nr in contacts:
            nr2 = np.broadcast_to(np.asarray(nr, np.float32), (n, 3))
            d = ps - pt
            sdf = d[:, 0] * nr2[:, 0] + d[:, 1] * nr2[:, 1] + d[:, 2] * nr2[:, 2]
            keep = live & mask & (sdf < h)
            if not project:
                keep &= sdf > np.float32(0.0)
            keep &= count < 8
            idx = np.nonzero(keep)[0]
            count[idx] += 1
            if len(idx):
                rows.append((idx, np.full(len(idx), order), pt[idx], nr2[idx]))
            order += 1
    if not rows:
        return np.zeros(0, np.uint32), np.zeros(1, np.uint32), np.zeros((0, 6), np.float32)
    part = np.concatenate([r[0] for r in rows])
    ordk = np.concatenate([r[1] for r in rows])
    pts = np.concatenate([r[2] for r in rows])
    nrs = np.concatenate([r[3] for r in rows])
    srt = np.lexsort((ordk, part))
    part, pts, nrs = part[srt], pts[srt], nrs[srt]
    ids, counts = np.unique(part, return_counts=True)
    start = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    planes = np.concatenate([pts, nrs], axis=1).astype(np.float32)
    return ids.astype(np.uint32), start, planes


def gather_wall_contacts(boundaries, h, ps, providers):
    """Mirror table (ids, start, planes[np, 6]) of the fluid positions ps
    (pipeline/walls.rs:145-153; contacts_point, 0 < sdf < h)."""
    return _gather(boundaries, h, ps, providers, project=False)


def gather_wall_project_contacts(boundaries, h, ps, providers):
    """Project table (pipeline/walls.rs:155-166; contacts_point_boundary, sdf < h)."""
    return _gather(boundaries, h, ps, providers, project=True)


def _wall_arrays(providers):
    provs = (WallProvider * max(len(providers), 1))()
    flist = []
    for k, pv in enumerate(providers):
        provs[k].id = pv.id
        b = pv.bounds(pv.id)
        provs[k].has_bounds = 0 if b is None else 1
        if b is not None:
            provs[k].lo = (C.c_float * 3)(*[float(x) for x in b[0]])
            provs[k].hi = (C.c_float * 3)(*[float(x) for x in b[1]])
        for (axis, plane, sign) in pv.faces:
            flist.append((axis, float(plane), float(sign), k))
    faces = (WallFace * max(len(flist), 1))()
    for q, (axis, plane, sign, k) in enumerate(flist):
        faces[q].axis = axis
        faces[q].plane = plane
        faces[q].sign = sign
        faces[q].provider = k
    return provs, len(providers), faces, len(flist)


class _Table:
    """A plane table as C arguments (ne, ids, start, np, planes); holds the arrays."""

    def __init__(self, ids, start, planes):
        ids = np.ascontiguousarray(ids, np.uint32)
        self.start = np.ascontiguousarray(start, np.uint32)
        planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 6)
        self.ne, self.np = len(ids), len(planes)
        self.ids = ids if len(ids) else np.zeros(1, np.uint32)
        self.planes = planes if len(planes) else np.zeros((1, 6), np.float32)

    def args(self):
        return self.ne, _u(self.ids), _u(self.start), self.np, _f(self.planes)


def sample_batch(seed, count, nmin=8192, nmax=524288):
    """Batched-form scene set (count seeded dam breaks of uneven size)."""
    arr = (Request * max(count, 1))()
    _check(lib().vxl_sph_sample_batch(seed, count, nmin, nmax, arr))
    return [arr[k] for k in range(count)]


def sample_requests(seed, count, rate=20.0, nmin=4096, nmax=2097152, smin=100, smax=2000, max_boxes=0):
    """Served-form Poisson request stream."""
    arr = (Request * max(count, 1))()
    _check(lib().vxl_sph_sample_requests(seed, count, rate, nmin, nmax, smin, smax, max_boxes, arr))
    return [arr[k] for k in range(count)]


def request_config(req, dt=5e-4, max_bins=1 << 21):
    cfg = FluidConfig()
    _check(lib().vxl_sph_request_config(C.byref(req), C.c_float(dt), max_bins, C.byref(cfg)))
    return cfg


def request_scene(req, dt=5e-4, max_bins=1 << 21):
    """(config, SceneData) of one request, built exactly as the server builds it
    (for a request with boxes: the fluid only; see request_bodies)."""
    cfg = request_config(req, dt, max_bins)
    dam = req.dam
    walls = bool(dam.walls) and req.nboxes == 0
    data = dam_break(cfg, tuple(dam.tank), tuple(dam.block), dam.d, jitter=dam.jitter, seed=dam.seed, walls=walls)
    return cfg, data


def request_bodies(req):
    """(rigid config, body dicts) a request with boxes is served with; ([], None) without boxes."""
    cnt = C.c_uint32()
    rc = RigidConfig()
    _check(lib().vxl_sph_request_bodies(C.byref(req), 0, None, C.byref(cnt), C.byref(rc)))
    arr = (RigidBox * max(cnt.value, 1))()
    _check(lib().vxl_sph_request_bodies(C.byref(req), cnt.value, arr, C.byref(cnt), C.byref(rc)))
    bodies = []
    for k in range(cnt.value):
        b = arr[k].box
        bodies.append(dict(id=b.id, half=tuple(b.half), pos=tuple(b.pos), rot=tuple(b.rot), linvel=tuple(b.linvel),
                           angvel=tuple(b.angvel), density=arr[k].density))
    return (rc if cnt.value else None), bodies


def step_device(cfg, n, n_fluid, pos_ptr, vel_ptr, pmass_ptr, ticks, stream=None):
    """Advance a scene held in device arrays (integer device addresses: pos and
    vel n x 3 f32, pmass n f32, index order) by `ticks` ticks in place."""
    _check(lib().vxl_sph_step_device(C.byref(cfg), n, n_fluid, pos_ptr, vel_ptr, pmass_ptr, ticks, stream))


def step_device_noalloc(cfg, n, n_fluid, pos_ptr, vel_ptr, pmass_ptr, ticks, stream=None):
    """step_device without any allocation (the library's per-device scratch;
    at most 131072 particles)."""
    _check(lib().vxl_sph_step_device_noalloc(C.byref(cfg), n, n_fluid, pos_ptr, vel_ptr, pmass_ptr, ticks, stream))


def request_bytes(req, dt=5e-4, max_bins=1 << 21):
    """Peak device bytes of serving `req` alone (pool, staging, tables)."""
    v = C.c_uint64()
    _check(lib().vxl_sph_request_bytes(C.byref(req), C.c_float(dt), max_bins, C.byref(v)))
    return v.value


def pool_limits():
    """(max slots, max cells, bytes per slot of the widest 32-bit-addressed record)."""
    a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint32()
    _check(lib().vxl_sph_pool_limits(C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def serve(reqs, budget_bytes=200 << 30, time_scale=1.0, max_wall_s=3600.0, max_resident=0, builders=4,
          max_bins=1 << 21, dt=5e-4, device=0, max_particles=0, admit_every=0):
    """Continuous-batching run over a request list; returns (stats dict, results list)."""
    cfg = ServeConfig()
    cfg.hdr.size = C.sizeof(ServeConfig)
    cfg.hdr.version = ABI_VERSION
    cfg.budget_bytes = budget_bytes
    cfg.time_scale = time_scale
    cfg.max_wall_s = max_wall_s
    cfg.max_resident = max_resident
    cfg.builders = builders
    cfg.max_bins = max_bins
    cfg.dt = dt
    cfg.device = device
    cfg.max_particles = max_particles
    cfg.admit_every = admit_every
    n = len(reqs)
    arr = (Request * max(n, 1))(*reqs)
    st = ServeStats()
    res = (ServeResult * max(n, 1))()
    code = lib().vxl_sph_serve(C.byref(cfg), n, arr, C.byref(st), res)
    if code not in (0, 20):
        _check(code)
    # CAPACITY: some request alone exceeds the budget; it is not served (steps
    # = 0 in its result) and the other requests' results are complete.
    stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
    stats["layout_phase_ms"] = dict(zip(("wait", "move", "upload", "tables", "graph"), st.layout_phase_ms[:]))
    stats["status"] = STATUS[code]
    results = [dict(index=r.index, n=r.n, n_fluid=r.n_fluid, steps=r.steps, arrival_s=r.arrival_s,
                    admit_s=r.admit_s, finish_s=r.finish_s, state_hash=r.state_hash) for r in res[:n]]
    return stats, results


class VxldWriter:
    """VXLD v3 frame dump (the reference demos' format, read by its renderer)."""

    def __init__(self, path, ticks, ticks_per_frame, dt, nfluid):
        self.h = C.c_void_p()
        _check(lib().vxl_sph_vxld_open(str(path).encode(), ticks, ticks_per_frame, C.c_float(dt), nfluid,
                                       C.byref(self.h)))

    def frame(self, engine, tick, ms, scenes):
        arr = (C.c_uint32 * max(len(scenes), 1))(*scenes)
        _check(lib().vxl_sph_vxld_frame(self.h, engine.h, tick, C.c_float(ms), len(scenes), arr))

    def frame_raw(self, tick, ms, bodies, fluids):
        """bodies: dicts as Engine.add_coupled takes; fluids: list of (n, 3) arrays."""
        arr = (RigidBox * max(len(bodies), 1))()
        for k, b in enumerate(bodies):
            _fill_box(arr[k].box, b)
            arr[k].density = b.get("density", 0.0)
        counts = np.array([len(f) for f in fluids], np.uint32)
        pos = (np.concatenate([np.asarray(f, np.float32).reshape(-1, 3) for f in fluids])
               if fluids else np.zeros((0, 3), np.float32))
        pos = np.ascontiguousarray(pos, np.float32)
        _check(lib().vxl_sph_vxld_frame_raw(self.h, tick, C.c_float(ms), len(bodies), arr, len(fluids), _u(counts),
                                            _f(pos) if len(pos) else None))

    def close(self):
        if self.h:
            _check(lib().vxl_sph_vxld_close(self.h))
            self.h = C.c_void_p()
