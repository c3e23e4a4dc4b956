//! # vxl-phys-hip
//!
//! The WCSPH fluid step of `vxl_phys` on an MI355X (gfx950), behind the two
//! seams the wgpu backend uses:
//!
//! * [`HipBackend`] implements `PhysGpuBackend` (crates/vxl-phys-gpu/src/lib.rs:20-29);
//! * [`HipFluidStepper`] implements `FluidStepper`
//!   (crates/vxl-phys-core/src/interop.rs:293-316) with the semantics of
//!   `GpuFluidStepper` (crates/vxl-phys-gpu/src/pipeline/stepper.rs:70-147):
//!   the fluid state lives on the device after the first `step`, later steps
//!   upload only the boundary segment, `reactions` are the tick averages in
//!   span order, `bounds` is the AABB of the device state.
//!
//! Register it as the wgpu stepper is registered
//! (crates/vxl-phys/src/world_step/fluid_stepper.rs:46-58):
//!
//! ```ignore
//! let fs = &world.fluids()[fi].0;
//! let stepper = vxl_phys_hip::HipFluidStepper::new(fs, 0).expect("MI355X fluid stepper");
//! world.set_fluid_stepper(fi, Box::new(stepper));
//! ```
//!
//! Every call goes through libvxl_sph.so's C ABI ([`ffi`]). The step is bit
//! for bit `FluidSystem::step` (strict f32, the CPU path's summation order),
//! so a world stepped through this backend equals the CPU world exactly.
pub mod ffi;

use std::ffi::{CStr, CString};
use std::fmt;

use vxl_phys_core::interop::{FluidStepper, ProviderColliders};
use vxl_phys_core::Vec3;
use vxl_phys_fluid::{FluidFamily, FluidSystem};
use vxl_phys_gpu::PhysGpuBackend;

/// A status code of the C ABI and the library's message for it.
#[derive(Clone, Debug)]
pub struct HipError {
    pub code: ffi::Status,
    pub message: String,
}

impl fmt::Display for HipError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "vxl_sph status {}: {}", self.code, self.message)
    }
}

impl std::error::Error for HipError {}

/// `&[Vec3]` crosses the ABI as 3 floats per particle (Vec3 is #[repr(C)]).
const VEC3_IS_3_F32: () = assert!(std::mem::size_of::<Vec3>() == 12 && std::mem::align_of::<Vec3>() == 4);

fn check(code: ffi::Status) -> Result<(), HipError> {
    if code == ffi::OK {
        return Ok(());
    }
    // SAFETY: vxl_sph_last_error returns a NUL-terminated thread-local string
    // that stays valid until the next call on this thread.
    let message = unsafe { CStr::from_ptr(ffi::vxl_sph_last_error()) }.to_string_lossy().into_owned();
    Err(HipError { code, message })
}

/// GpuFluidStepper panics on a mismatched call (stepper.rs:80-96); this
/// backend keeps that contract for the trait methods, which return nothing.
fn expect(code: ffi::Status, what: &str) {
    if let Err(e) = check(code) {
        panic!("{what}: {e}");
    }
}

/// `PhysGpuBackend` for MI355X (gfx950): available when device 0 is a gfx950.
pub struct HipBackend;

impl PhysGpuBackend for HipBackend {
    fn name(&self) -> &'static str {
        "hip-gfx950"
    }

    fn available(&self) -> bool {
        // SAFETY: no arguments; the call only queries the HIP runtime.
        unsafe { ffi::vxl_sph_backend_available() != 0 }
    }

    /// Accepts the fluid batch ("fluid"); other tags stay on their own paths.
    fn submit(&self, tag: &'static str) -> bool {
        let Ok(t) = CString::new(tag) else { return false };
        // SAFETY: `t` is a valid NUL-terminated string for the duration of the call.
        unsafe { ffi::vxl_sph_backend_submit(t.as_ptr()) != 0 }
    }
}

/// `FluidStepper` on the device, one fluid system per stepper.
pub struct HipFluidStepper {
    h: *mut ffi::Stepper,
    reactions: Vec<(u32, Vec3, Vec3)>,
    /// Provider ids whose walls `gather_walls` collects (FluidSystem::boundaries).
    boundaries: Vec<u32>,
    /// Fluid positions at construction: what the first `gather_walls` reads
    /// (the device holds no state before the first step).
    first_fluid: Vec<Vec3>,
    /// Particle count of the last step (0 before the first).
    n: usize,
}

// SAFETY: the handle is used from one thread at a time (`&mut self` for
// every mutating call); the library keeps no thread affinity for it.
unsafe impl Send for HipFluidStepper {}

impl HipFluidStepper {
    /// A stepper for `fs` (its configuration, particle mass and spacing, its
    /// provider walls and current fluid positions) on HIP device `device`.
    /// Only the WCSPH family is stepped here.
    pub fn new(fs: &FluidSystem, device: i32) -> Result<Self, HipError> {
        let cfg = fs.config();
        if cfg.family != FluidFamily::CpuSph {
            return Err(HipError { code: ffi::UNSUPPORTED, message: "only the WCSPH family is stepped".into() });
        }
        // SAFETY: no arguments.
        let abi = unsafe { ffi::vxl_sph_abi_version() };
        if abi != ffi::ABI_VERSION {
            return Err(HipError {
                code: ffi::VERSION_MISMATCH,
                message: format!("libvxl_sph ABI {abi}, this crate {}", ffi::ABI_VERSION),
            });
        }
        let mut c = ffi::FluidConfig::default();
        // SAFETY: `c` is a valid, writable FluidConfig.
        check(unsafe { ffi::vxl_sph_default_config(&mut c) })?;
        c.rest_density = cfg.rest_density;
        c.smoothing_radius = cfg.smoothing_radius;
        c.xsph_viscosity = cfg.xsph_viscosity;
        c.sound_speed = cfg.sound_speed;
        c.gamma_tait = cfg.gamma_tait;
        c.substeps = cfg.substeps;
        c.tensile_clamp = cfg.tensile_instability_suppression as u32;
        c.boundary_layers = cfg.boundary_layers;
        c.gravity = [cfg.gravity.x, cfg.gravity.y, cfg.gravity.z];
        c.max_speed_frac = cfg.max_speed_frac;
        c.artificial_viscosity = cfg.artificial_viscosity;
        c.particle_mass = fs.particle_mass();
        c.spacing = fs.particle_spacing();
        c.max_bins = vxl_phys_core::grid::GRID_MAX_BINS as u64;
        let mut h = std::ptr::null_mut();
        // SAFETY: `c` outlives the call, `h` receives the handle.
        check(unsafe { ffi::vxl_sph_stepper_create(ffi::ABI_VERSION, &c, device, &mut h) })?;
        Ok(Self {
            h,
            reactions: Vec::new(),
            boundaries: fs.boundaries().to_vec(),
            first_fluid: fs.positions().to_vec(),
            n: 0,
        })
    }

    /// Positions and velocities of all particles on the device (fluid prefix
    /// first), after the last step.
    pub fn read_state(&self) -> Result<(Vec<Vec3>, Vec<Vec3>), HipError> {
        let mut p = vec![0f32; 3 * self.n];
        let mut v = vec![0f32; 3 * self.n];
        // SAFETY: both buffers hold 3 n floats.
        check(unsafe { ffi::vxl_sph_stepper_read_state(self.h, self.n as u32, p.as_mut_ptr(), v.as_mut_ptr()) })?;
        let to_vec3 = |a: &[f32]| -> Vec<Vec3> { a.chunks_exact(3).map(|c| Vec3::new(c[0], c[1], c[2])).collect() };
        Ok((to_vec3(&p), to_vec3(&v)))
    }

    fn upload_walls(&mut self, side: u32, table: (Vec<u32>, Vec<u32>, Vec<(Vec3, Vec3)>)) {
        let (ids, start, planes) = table;
        let flat: Vec<f32> = planes.iter().flat_map(|(p, n)| [p.x, p.y, p.z, n.x, n.y, n.z]).collect();
        // SAFETY: ids has ne entries, start ne + 1, flat 6 np floats.
        expect(
            unsafe {
                ffi::vxl_sph_stepper_set_wall_tables(self.h, side, ids.len() as u32, ids.as_ptr(), start.as_ptr(),
                                                     planes.len() as u32, flat.as_ptr())
            },
            "vxl_sph_stepper_set_wall_tables",
        );
    }
}

impl Drop for HipFluidStepper {
    fn drop(&mut self) {
        // SAFETY: `h` came from vxl_sph_stepper_create and is destroyed once.
        unsafe {
            ffi::vxl_sph_stepper_destroy(self.h);
        }
    }
}

impl FluidStepper for HipFluidStepper {
    fn step(&mut self, dt_tick: f32, n_fluid: usize, pos: &[Vec3], vel: &[Vec3], pmass: &[f32],
            spans: &[(u32, Vec3, u32, u32)]) {
        let () = VEC3_IS_3_F32;
        assert!(pos.len() == vel.len() && pos.len() == pmass.len(), "pos / vel / pmass lengths differ");
        let sp: Vec<ffi::Span> = spans
            .iter()
            .map(|&(body, o, start, end)| ffi::Span { body, origin: [o.x, o.y, o.z], start, end })
            .collect();
        // SAFETY: Vec3 is #[repr(C)] with three f32 fields (VEC3_IS_3_F32), so
        // pos and vel are 3 n contiguous floats; pmass holds n, sp nspans entries.
        expect(
            unsafe {
                ffi::vxl_sph_stepper_step(self.h, dt_tick, n_fluid as u32, pos.len() as u32,
                                          pos.as_ptr() as *const f32, vel.as_ptr() as *const f32, pmass.as_ptr(),
                                          sp.as_ptr(), sp.len() as u32)
            },
            "vxl_sph_stepper_step",
        );
        self.n = pos.len();
        let mut count = 0u32;
        // SAFETY: cap 0 with a null buffer queries the count.
        expect(unsafe { ffi::vxl_sph_stepper_reactions(self.h, 0, std::ptr::null_mut(), &mut count) },
               "vxl_sph_stepper_reactions");
        let mut out = vec![ffi::Reaction::default(); count as usize];
        // SAFETY: `out` holds `count` entries.
        expect(unsafe { ffi::vxl_sph_stepper_reactions(self.h, count, out.as_mut_ptr(), &mut count) },
               "vxl_sph_stepper_reactions");
        self.reactions = out
            .iter()
            .map(|r| (r.body, Vec3::new(r.force[0], r.force[1], r.force[2]),
                      Vec3::new(r.torque[0], r.torque[1], r.torque[2])))
            .collect();
    }

    fn reactions(&self) -> &[(u32, Vec3, Vec3)] {
        &self.reactions
    }

    /// The AABB of every particle on the device; `None` before the first step.
    fn bounds(&self) -> Option<(Vec3, Vec3)> {
        let (mut lo, mut hi) = ([0f32; 3], [0f32; 3]);
        // SAFETY: lo and hi hold 3 floats each.
        let st = unsafe { ffi::vxl_sph_stepper_bounds(self.h, lo.as_mut_ptr(), hi.as_mut_ptr()) };
        (st == ffi::OK).then(|| (Vec3::new(lo[0], lo[1], lo[2]), Vec3::new(hi[0], hi[1], hi[2])))
    }

    /// GpuFluidStepper::gather_walls (pipeline/stepper.rs:118-137): both plane
    /// tables gathered by the reference's own host code at the device's fluid
    /// positions (before the first step, the positions that step starts
    /// from), used from the next step on.
    fn gather_walls(&mut self, h: f32, providers: &dyn ProviderColliders) {
        if self.boundaries.is_empty() {
            return;
        }
        let ps: Vec<Vec3> = if self.n == 0 {
            self.first_fluid.clone()
        } else {
            let (p, _) = self.read_state().unwrap_or_else(|e| panic!("vxl_sph_stepper_read_state: {e}"));
            p[..self.first_fluid.len()].to_vec()
        };
        use vxl_phys_gpu::pipeline::{gather_wall_contacts, gather_wall_project_contacts};
        let mirror = gather_wall_contacts(&self.boundaries, h, &ps, providers);
        let project = gather_wall_project_contacts(&self.boundaries, h, &ps, providers);
        self.upload_walls(0, mirror);
        self.upload_walls(1, project);
    }
}
