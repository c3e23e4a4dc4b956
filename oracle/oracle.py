"""ctypes binding of the WCSPH oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or as the timed CPU baseline. It restates
the reference CPU path (crates/vxl-phys-fluid); see wcsph_oracle.h.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")  # ORACLE_LIB: a sanitizer build


class OrConfig(C.Structure):
    _fields_ = [
        ("rest_density", C.c_float),
        ("smoothing_radius", C.c_float),
        ("xsph_viscosity", C.c_float),
        ("sound_speed", C.c_float),
        ("gamma_tait", C.c_float),
        ("substeps", C.c_uint32),
        ("threads", C.c_uint32),
        ("tensile_clamp", C.c_uint32),
        ("boundary_layers", C.c_uint32),
        ("gravity", C.c_float * 3),
        ("max_speed_frac", C.c_float),
        ("artificial_viscosity", C.c_float),
        ("max_bins", C.c_uint64),
        ("skip_reactions", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class OrBoxBody(C.Structure):
    _fields_ = [
        ("id", C.c_uint32),
        ("half", C.c_float * 3),
        ("pos", C.c_float * 3),
        ("rot", C.c_float * 4),
        ("linvel", C.c_float * 3),
        ("angvel", C.c_float * 3),
    ]


class OrRigidBody(C.Structure):
    _fields_ = [
        ("half", C.c_float * 3),
        ("pos", C.c_float * 3),
        ("rot", C.c_float * 4),
        ("linvel", C.c_float * 3),
        ("angvel", C.c_float * 3),
        ("inv_mass", C.c_float),
        ("inv_inertia", C.c_float * 3),
    ]


class OrRigidConfig(C.Structure):
    _fields_ = [
        ("dt_tick", C.c_float),
        ("substeps", C.c_uint32),
        ("gravity", C.c_float * 3),
        ("max_linear", C.c_float),
        ("max_angular", C.c_float),
        ("tank", C.c_float * 3),
    ]


class OrWallProvider(C.Structure):
    _fields_ = [
        ("id", C.c_uint32),
        ("has_bounds", C.c_uint32),
        ("lo", C.c_float * 3),
        ("hi", C.c_float * 3),
    ]


class OrWallFace(C.Structure):
    _fields_ = [
        ("axis", C.c_uint32),
        ("plane", C.c_float),
        ("sign", C.c_float),
        ("provider", C.c_uint32),
    ]


_lib = None


def build():
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        fp = C.POINTER(C.c_float)
        up = C.POINTER(C.c_uint32)
        L.or_default_config.argtypes = [C.POINTER(OrConfig)]
        L.or_new_lattice.restype = P
        L.or_new_lattice.argtypes = [C.POINTER(OrConfig), fp, up, C.c_float]
        L.or_new_particles.restype = P
        L.or_new_particles.argtypes = [C.POINTER(OrConfig), C.c_uint32, fp, fp, C.c_float, C.c_float]
        L.or_free.argtypes = [P]
        L.or_set_boundary_boxes.restype = C.c_uint32
        L.or_set_boundary_boxes.argtypes = [P, C.c_uint32, C.POINTER(OrBoxBody)]
        L.or_set_boundary_raw.argtypes = [P, C.c_uint32, fp, fp, fp, C.c_uint32, up, fp, up, up]
        L.or_set_velocities.argtypes = [P, C.c_uint32, fp]
        L.or_set_threads.argtypes = [P, C.c_uint32]
        L.or_set_fault.argtypes = [P, C.c_int]
        L.or_step.argtypes = [P, C.c_float]
        for name in ("or_n_fluid", "or_n_total", "or_n_spans"):
            getattr(L, name).restype = C.c_uint32
            getattr(L, name).argtypes = [P]
        L.or_mass.restype = C.c_float
        L.or_mass.argtypes = [P]
        L.or_constants.argtypes = [P, fp]
        for name in ("or_get_pos", "or_get_vel", "or_get_pmass", "or_get_dens", "or_get_press",
                     "or_get_acc", "or_get_xsph", "or_get_bforce"):
            getattr(L, name).argtypes = [P, fp]
        L.or_get_ncount.argtypes = [P, up]
        L.or_get_reactions.argtypes = [P, up, fp]
        L.or_grid_info.argtypes = [P, fp, fp, fp, up]
        L.or_grid_total.restype = C.c_uint64
        L.or_grid_total.argtypes = [P]
        L.or_grid_tables.argtypes = [P, up, up, up]
        L.or_grid_hash.restype = C.c_uint64
        L.or_grid_hash.argtypes = [P]
        L.or_state_hash.restype = C.c_uint64
        L.or_state_hash.argtypes = [P]
        L.or_grid_box.argtypes = [fp, fp, C.c_float, C.c_uint64, fp, fp, up]
        L.or_lattice_box.restype = C.c_uint32
        L.or_lattice_box.argtypes = [fp, C.c_float, C.c_float, C.c_uint32, fp, C.c_uint32, up, fp]
        L.or_boundary_self_sum_bruteforce.restype = C.c_float
        L.or_boundary_self_sum_bruteforce.argtypes = [fp, C.c_uint32, C.c_float]
        L.or_boundary_self_sum_grid.restype = C.c_float
        L.or_boundary_self_sum_grid.argtypes = [fp, C.c_uint32, C.c_float]
        L.or_neighbors.restype = C.c_uint32
        L.or_neighbors.argtypes = [P, C.c_uint32, up, C.c_uint32]
        L.or_box_mass_props.restype = None
        L.or_box_mass_props.argtypes = [fp, C.c_float, fp, fp]
        L.or_rigid_tick.restype = None
        L.or_rigid_tick.argtypes = [C.POINTER(OrRigidConfig), C.c_uint32, C.POINTER(OrRigidBody), fp]
        L.or_set_halfspace_walls.restype = C.c_int
        L.or_set_halfspace_walls.argtypes = [P, C.c_uint32, C.POINTER(OrWallProvider), C.c_uint32,
                                             C.POINTER(OrWallFace)]
        L.or_gather_wall_tables.restype = C.c_uint32
        L.or_gather_wall_tables.argtypes = [P, C.c_uint32, C.c_uint32, up, up, C.c_uint32, fp, up]
        L.or_set_wall_tables.restype = C.c_int
        L.or_set_wall_tables.argtypes = [P, C.c_uint32, C.c_uint32, up, up, C.c_uint32, fp]
        L.or_clear_walls.argtypes = [P]
        L.or_last_pushed.restype = C.c_uint32
        L.or_last_pushed.argtypes = [P]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def default_config(**kw):
    c = OrConfig()
    lib().or_default_config(C.byref(c))
    for k, v in kw.items():
        if k == "gravity":
            c.gravity = (C.c_float * 3)(*v)
        else:
            setattr(c, k, v)
    return c


def grid_box(lo, hi, h, max_bins):
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    mn = np.zeros(3, np.float32)
    b = np.zeros(1, np.float32)
    d = np.zeros(3, np.uint32)
    lib().or_grid_box(_f(lo), _f(hi), C.c_float(h), max_bins, _f(mn), _f(b), _u(d))
    return mn, float(b[0]), [int(x) for x in d]


def lattice_box(half, s, h, layers):
    half = np.asarray(half, np.float32)
    n0 = np.zeros(1, np.uint32)
    vol = np.zeros(1, np.float32)
    n = lib().or_lattice_box(_f(half), C.c_float(s), C.c_float(h), layers, None, 0, None, None)
    pts = np.zeros((max(n, 1), 3), np.float32)
    lib().or_lattice_box(_f(half), C.c_float(s), C.c_float(h), layers, _f(pts), n, _u(n0), _f(vol))
    return pts[:n], int(n0[0]), np.float32(vol[0])


def self_sum_bruteforce(pts, h):
    pts = np.ascontiguousarray(pts, np.float32)
    return np.float32(lib().or_boundary_self_sum_bruteforce(_f(pts), len(pts), C.c_float(h)))


def self_sum_grid(pts, h):
    pts = np.ascontiguousarray(pts, np.float32)
    return np.float32(lib().or_boundary_self_sum_grid(_f(pts), len(pts), C.c_float(h)))


class System:
    """FluidSystem restated (fluid_access.rs, fluid_step.rs)."""

    def __init__(self, ptr):
        self.p = C.c_void_p(ptr)

    @classmethod
    def lattice(cls, cfg, origin, dims, spacing):
        o = np.asarray(origin, np.float32)
        d = np.asarray(dims, np.uint32)
        return cls(lib().or_new_lattice(C.byref(cfg), _f(o), _u(d), C.c_float(spacing)))

    @classmethod
    def particles(cls, cfg, pos, vel, mass, spacing):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        velp = None
        if vel is not None:
            vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
            velp = _f(vel)
        return cls(lib().or_new_particles(C.byref(cfg), len(pos), _f(pos), velp,
                                          C.c_float(mass), C.c_float(spacing)))

    def __del__(self):
        if getattr(self, "p", None) is not None and _lib is not None:
            _lib.or_free(self.p)
            self.p = None

    def set_boundary_boxes(self, bodies):
        arr = (OrBoxBody * len(bodies))()
        for k, b in enumerate(bodies):
            arr[k].id = b["id"]
            arr[k].half = (C.c_float * 3)(*b["half"])
            arr[k].pos = (C.c_float * 3)(*b["pos"])
            arr[k].rot = (C.c_float * 4)(*b.get("rot", (0.0, 0.0, 0.0, 1.0)))
            arr[k].linvel = (C.c_float * 3)(*b.get("linvel", (0.0, 0.0, 0.0)))
            arr[k].angvel = (C.c_float * 3)(*b.get("angvel", (0.0, 0.0, 0.0)))
        return lib().or_set_boundary_boxes(self.p, len(bodies), arr)

    def set_boundary_raw(self, pos, vel, pmass, spans):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        pmass = np.ascontiguousarray(pmass, np.float32)
        body = np.array([s[0] for s in spans], np.uint32)
        org = np.array([s[1] for s in spans], np.float32).reshape(-1, 3)
        st = np.array([s[2] for s in spans], np.uint32)
        en = np.array([s[3] for s in spans], np.uint32)
        lib().or_set_boundary_raw(self.p, len(pos), _f(pos), _f(vel), _f(pmass), len(spans),
                                  _u(body), _f(org), _u(st), _u(en))

    def set_velocities(self, vel):
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        lib().or_set_velocities(self.p, len(vel), _f(vel))

    def set_threads(self, t):
        lib().or_set_threads(self.p, t)

    def set_fault(self, code):
        lib().or_set_fault(self.p, code)

    def step(self, dt_tick):
        lib().or_step(self.p, C.c_float(dt_tick))

    @property
    def n_fluid(self):
        return lib().or_n_fluid(self.p)

    @property
    def n_total(self):
        return lib().or_n_total(self.p)

    @property
    def mass(self):
        return np.float32(lib().or_mass(self.p))

    def constants(self):
        o = np.zeros(8, np.float32)
        lib().or_constants(self.p, _f(o))
        return dict(zip(["h", "h2", "k6", "ks", "w0", "b_tait", "mass", "alpha_c"], o))

    def _vec(self, fn, n):
        o = np.zeros((n, 3), np.float32)
        fn(self.p, _f(o))
        return o

    def _sc(self, fn, n, dt=np.float32):
        o = np.zeros(n, dt)
        fn(self.p, _f(o) if dt == np.float32 else _u(o))
        return o

    def pos(self):
        return self._vec(lib().or_get_pos, self.n_total)

    def vel(self):
        return self._vec(lib().or_get_vel, self.n_total)

    def acc(self):
        return self._vec(lib().or_get_acc, self.n_fluid)

    def xsph(self):
        return self._vec(lib().or_get_xsph, self.n_fluid)

    def bforce(self):
        return self._vec(lib().or_get_bforce, self.n_total - self.n_fluid)

    def pmass(self):
        return self._sc(lib().or_get_pmass, self.n_total)

    def dens(self):
        return self._sc(lib().or_get_dens, self.n_total)

    def press(self):
        return self._sc(lib().or_get_press, self.n_total)

    def ncount(self):
        return self._sc(lib().or_get_ncount, self.n_total, np.uint32)

    def reactions(self):
        n = lib().or_n_spans(self.p)
        body = np.zeros(max(n, 1), np.uint32)
        ft = np.zeros((max(n, 1), 6), np.float32)
        lib().or_get_reactions(self.p, _u(body), _f(ft))
        return body[:n], ft[:n]

    def grid_info(self):
        mn = np.zeros(3, np.float32)
        inv = np.zeros(1, np.float32)
        b = np.zeros(1, np.float32)
        d = np.zeros(3, np.uint32)
        lib().or_grid_info(self.p, _f(mn), _f(inv), _f(b), _u(d))
        return mn, inv[0], b[0], [int(x) for x in d]

    def grid_tables(self):
        tot = lib().or_grid_total(self.p)
        start = np.zeros(tot + 1, np.uint32)
        items = np.zeros(self.n_total, np.uint32)
        keys = np.zeros(self.n_total, np.uint32)
        lib().or_grid_tables(self.p, _u(start), _u(items), _u(keys))
        return start, items, keys

    def grid_hash(self):
        return int(lib().or_grid_hash(self.p))

    def state_hash(self):
        return int(lib().or_state_hash(self.p))

    def neighbors(self, i, cap=4096):
        o = np.zeros(cap, np.uint32)
        n = lib().or_neighbors(self.p, i, _u(o), cap)
        return o[:min(n, cap)]

    # --- provider walls ------------------------------------------------------

    def set_halfspace_walls(self, providers):
        """Half-space providers queried every substep (the CPU path). Each is a
        dict: id, bounds ((lo), (hi)) or None, faces [(axis, plane, sign), ...]."""
        provs, faces = _wall_structs(providers)
        rc = lib().or_set_halfspace_walls(self.p, len(providers), provs, len(faces), faces)
        if rc != 0:
            raise ValueError("malformed half-space walls")

    def gather_wall_tables(self, side):
        """(ids, start, planes[np, 6]) of the current fluid positions (side 0
        mirror, 1 project; wall_gather.rs:52-100)."""
        npl = np.zeros(1, np.uint32)
        ne = lib().or_gather_wall_tables(self.p, side, 0, None, None, 0, None, _u(npl))
        ids = np.zeros(max(ne, 1), np.uint32)
        start = np.zeros(ne + 1, np.uint32)
        planes = np.zeros((max(int(npl[0]), 1), 6), np.float32)
        lib().or_gather_wall_tables(self.p, side, ne, _u(ids), _u(start), int(npl[0]), _f(planes), _u(npl))
        return ids[:ne], start, planes[:int(npl[0])]

    def set_wall_tables(self, side, ids, start, planes):
        """Plane tables held until replaced (the GPU path)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        start = np.ascontiguousarray(start, np.uint32)
        planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 6)
        ids_b = ids if len(ids) else np.zeros(1, np.uint32)
        pl_b = planes if len(planes) else np.zeros((1, 6), np.float32)
        rc = lib().or_set_wall_tables(self.p, side, len(ids), _u(ids_b), _u(start), len(planes), _f(pl_b))
        if rc != 0:
            raise ValueError("malformed wall table")

    def clear_walls(self):
        lib().or_clear_walls(self.p)

    def last_pushed(self):
        return int(lib().or_last_pushed(self.p))


def _wall_structs(providers):
    provs = (OrWallProvider * max(len(providers), 1))()
    flist = []
    for k, pv in enumerate(providers):
        provs[k].id = pv.get("id", k)
        b = pv.get("bounds")
        provs[k].has_bounds = 0 if b is None else 1
        if b is not None:
            provs[k].lo = (C.c_float * 3)(*b[0])
            provs[k].hi = (C.c_float * 3)(*b[1])
        for (axis, plane, sign) in pv["faces"]:
            flist.append((axis, plane, sign, k))
    faces = (OrWallFace * max(len(flist), 1))()
    for q, (axis, plane, sign, k) in enumerate(flist):
        faces[q].axis = axis
        faces[q].plane = plane
        faces[q].sign = sign
        faces[q].provider = k
    return provs, faces


# The reference's half-space test doubles (crates/vxl-phys-fluid/src/tests.rs):
# FloorY (:11-47): solid y < 0; Tank (:49-110): floor y = 0 and walls at
# x, z = +-0.2, faces in the order floor, left, right, back, front.
FLOOR_Y = dict(id=0, bounds=((-50.0, -1.0, -50.0), (50.0, 0.0, 50.0)), faces=[(1, 0.0, 1.0)])
TANK = dict(id=0, bounds=((-50.0, -50.0, -50.0), (50.0, 50.0, 50.0)),
            faces=[(1, 0.0, 1.0), (0, -0.2, 1.0), (0, 0.2, -1.0), (2, -0.2, 1.0), (2, 0.2, -1.0)])


# --- rigid-box leg (rigid_oracle.h) ------------------------------------------

def box_mass_props(half, density):
    """(inv_mass, inv_inertia[3]) of a box (mass.rs:35-48)."""
    half = np.asarray(half, np.float32)
    im = np.zeros(1, np.float32)
    ii = np.zeros(3, np.float32)
    lib().or_box_mass_props(_f(half), C.c_float(density), _f(im), _f(ii))
    return im[0], ii


def rigid_config(dt_tick, substeps=2, gravity=(0.0, -9.81, 0.0), max_linear=100.0, max_angular=50.0,
                 tank=(0.0, 0.0, 0.0)):
    c = OrRigidConfig()
    c.dt_tick = dt_tick
    c.substeps = substeps
    c.gravity = (C.c_float * 3)(*gravity)
    c.max_linear = max_linear
    c.max_angular = max_angular
    c.tank = (C.c_float * 3)(*tank)
    return c


class RigidBoxes:
    """Dynamic boxes advanced by or_rigid_tick; `bodies` are dicts with half,
    pos, rot, linvel, angvel, density."""

    def __init__(self, bodies):
        self.arr = (OrRigidBody * max(len(bodies), 1))()
        self.n = len(bodies)
        for k, b in enumerate(bodies):
            r = self.arr[k]
            r.half = (C.c_float * 3)(*b["half"])
            r.pos = (C.c_float * 3)(*b["pos"])
            r.rot = (C.c_float * 4)(*b.get("rot", (0.0, 0.0, 0.0, 1.0)))
            r.linvel = (C.c_float * 3)(*b.get("linvel", (0.0, 0.0, 0.0)))
            r.angvel = (C.c_float * 3)(*b.get("angvel", (0.0, 0.0, 0.0)))
            im, ii = box_mass_props(b["half"], b["density"])
            r.inv_mass = im
            r.inv_inertia = (C.c_float * 3)(*ii)

    def tick(self, cfg, react=None):
        """react: (n, 6) float32 (F, tau) per body, or None for zeros."""
        r = np.zeros((max(self.n, 1), 6), np.float32)
        if react is not None and self.n:
            r[:self.n] = np.asarray(react, np.float32).reshape(self.n, 6)
        lib().or_rigid_tick(C.byref(cfg), self.n, self.arr, _f(r))

    def state(self, k):
        b = self.arr[k]
        return dict(pos=np.array(b.pos[:], np.float32), rot=np.array(b.rot[:], np.float32),
                    linvel=np.array(b.linvel[:], np.float32), angvel=np.array(b.angvel[:], np.float32))
