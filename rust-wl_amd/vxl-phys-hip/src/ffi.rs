//! Raw declarations of include/vxl_sph.h: the part this crate binds (library,
//! backend and FluidStepper entry points). Layouts are `#[repr(C)]` twins of
//! the header's structures; status codes 0-5 keep the meaning of
//! vxl-phys-ffi's `FfiError` (crates/vxl-phys-ffi/src/lib.rs:24-34).
use std::os::raw::{c_char, c_int};

/// `VXL_SPH_ABI_VERSION`.
pub const ABI_VERSION: u32 = 1;

pub type Status = c_int;
pub const OK: Status = 0;
pub const NULL_POINTER: Status = 1;
pub const VERSION_MISMATCH: Status = 2;
pub const SIZE_MISMATCH: Status = 3;
pub const NOT_READY: Status = 4;
pub const PANIC_CAUGHT: Status = 5;
pub const INVALID_ARGUMENT: Status = 16;
pub const UNSUPPORTED: Status = 17;
pub const DEVICE_ERROR: Status = 18;
pub const OUT_OF_MEMORY: Status = 19;
pub const CAPACITY: Status = 20;
pub const NOT_FOUND: Status = 21;

/// `VxlSphHeader`: every configuration block starts with its size and the ABI version.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct Header {
    pub size: u32,
    pub version: u32,
}

/// `VxlSphFluidConfig`: FluidConfig (crates/vxl-phys-fluid/src/config.rs:6-69)
/// plus the values a scene needs (mass, spacing, grid budget, tick).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct FluidConfig {
    pub hdr: Header,
    pub rest_density: f32,
    pub smoothing_radius: f32,
    pub xsph_viscosity: f32,
    pub sound_speed: f32,
    pub gamma_tait: f32,
    pub substeps: u32,
    pub tensile_clamp: u32,
    pub boundary_layers: u32,
    pub gravity: [f32; 3],
    pub max_speed_frac: f32,
    pub artificial_viscosity: f32,
    pub particle_mass: f32,
    pub spacing: f32,
    pub max_bins: u64,
    pub dt_tick: f32,
    pub reserved: u32,
}

/// `VxlSphSpan`: one entry of FluidSystem::boundary_spans (fluid_boundary.rs:129-135).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct Span {
    pub body: u32,
    pub origin: [f32; 3],
    pub start: u32,
    pub end: u32,
}

/// `VxlSphReaction`: (body, force, torque about the body origin).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct Reaction {
    pub body: u32,
    pub force: [f32; 3],
    pub torque: [f32; 3],
}

/// Opaque `VxlSphStepper`.
#[repr(C)]
pub struct Stepper {
    _private: [u8; 0],
}

extern "C" {
    pub fn vxl_sph_abi_version() -> u32;
    pub fn vxl_sph_backend_name() -> *const c_char;
    pub fn vxl_sph_backend_available() -> i32;
    pub fn vxl_sph_backend_submit(tag: *const c_char) -> i32;
    pub fn vxl_sph_last_error() -> *const c_char;
    pub fn vxl_sph_default_config(out: *mut FluidConfig) -> Status;

    pub fn vxl_sph_stepper_create(abi_version: u32, cfg: *const FluidConfig, device: i32,
                                  out: *mut *mut Stepper) -> Status;
    pub fn vxl_sph_stepper_destroy(s: *mut Stepper) -> Status;
    pub fn vxl_sph_stepper_step(s: *mut Stepper, dt_tick: f32, n_fluid: u32, n: u32, pos: *const f32,
                                vel: *const f32, pmass: *const f32, spans: *const Span, nspans: u32) -> Status;
    pub fn vxl_sph_stepper_reactions(s: *mut Stepper, cap: u32, out: *mut Reaction, count: *mut u32) -> Status;
    pub fn vxl_sph_stepper_bounds(s: *mut Stepper, lo: *mut f32, hi: *mut f32) -> Status;
    pub fn vxl_sph_stepper_read_state(s: *mut Stepper, n: u32, pos: *mut f32, vel: *mut f32) -> Status;
    pub fn vxl_sph_stepper_set_wall_tables(s: *mut Stepper, side: u32, ne: u32, ids: *const u32,
                                           start: *const u32, np: u32, planes: *const f32) -> Status;
}

#[cfg(test)]
mod tests {
    use super::*;
    use std::mem::{align_of, size_of};

    /// The sizes and alignments of include/vxl_sph.h on x86_64.
    #[test]
    fn layouts_match_the_header() {
        assert_eq!(size_of::<Header>(), 8);
        assert_eq!(size_of::<FluidConfig>(), 88);
        assert_eq!(align_of::<FluidConfig>(), 8);
        assert_eq!(size_of::<Span>(), 24);
        assert_eq!(size_of::<Reaction>(), 28);
    }
}
