//! Facade check of the MI355X stepper: one scene stepped twice,
//! A = the whole CPU path (`World::step`), B = the same world with its fluid
//! stepped on the device (`World::set_fluid_stepper` with [`HipFluidStepper`]).
//! The scene is the one of the wgpu stepper's facade probe
//! (crates/vxl-phys-gpu/examples/gpu_facade_probe.rs): a static box container,
//! an 8^3 block of fluid at 0.05 m and a dynamic box lighter than water, so
//! reactions flow back into a moving body every tick.
//!
//! The device step is `FluidSystem::step` bit for bit, so the expectation is
//! stricter than the wgpu probe's tolerance: the floating box's position and
//! velocity are equal in A and B on every tick, and the stepper's `bounds()`
//! equals the AABB of A's particles.
//!
//! Run: `cargo run --release -p vxl-phys-hip --example hip_facade_probe -- [ticks] [--device K]`
use vxl_phys::{PhysConfig, Quat, Shape, Vec3, World};
use vxl_phys_core::interop::FluidStepper;
use vxl_phys_fluid::{FluidConfig, FluidSystem};
use vxl_phys_gpu::PhysGpuBackend;
use vxl_phys_hip::{HipBackend, HipFluidStepper};

fn scene() -> (World, usize) {
    let mut w = World::new(PhysConfig::default());
    let t = 0.05f32;
    w.add_static(Shape::Box { half: Vec3::new(0.3, t, 0.3) }, Vec3::new(0.0, 1.05 - t, 0.0), Quat::IDENTITY);
    for (x, z) in [(0.25f32, 0.0f32), (-0.25, 0.0), (0.0, 0.25), (0.0, -0.25)] {
        let half = if x != 0.0 { Vec3::new(t, 0.25, 0.3) } else { Vec3::new(0.3, 0.25, t) };
        w.add_static(Shape::Box { half }, Vec3::new(x, 1.3, z), Quat::IDENTITY);
    }
    let sys = FluidSystem::new(FluidConfig::default(), Vec3::new(-0.2, 1.05, -0.2), [8, 8, 8], 0.05);
    let fi = w.add_fluid_with_boundary_coupling(sys, &[]);
    w.add_dynamic(Shape::Box { half: Vec3::splat(0.1) }, Vec3::new(0.0, 1.55, 0.0), Quat::IDENTITY, 400.0);
    (w, fi)
}

/// (position, linear velocity) of the world's dynamic body, as raw bits.
fn floating_box(w: &World) -> [u32; 6] {
    let mut out = [0u32; 6];
    for i in 0..w.bodies.len() {
        if w.bodies.is_dynamic(i) {
            let (p, v) = (w.bodies.position[i], w.bodies.linvel[i]);
            out = [p.x.to_bits(), p.y.to_bits(), p.z.to_bits(), v.x.to_bits(), v.y.to_bits(), v.z.to_bits()];
        }
    }
    out
}

fn aabb(ps: &[Vec3]) -> (Vec3, Vec3) {
    let mut lo = Vec3::splat(f32::INFINITY);
    let mut hi = Vec3::splat(f32::NEG_INFINITY);
    for p in ps {
        lo = Vec3::new(lo.x.min(p.x), lo.y.min(p.y), lo.z.min(p.z));
        hi = Vec3::new(hi.x.max(p.x), hi.y.max(p.y), hi.z.max(p.z));
    }
    (lo, hi)
}

fn main() {
    let mut args = std::env::args().skip(1);
    let ticks: usize = args.next().and_then(|s| s.parse().ok()).unwrap_or(150);
    let rest: Vec<String> = args.collect();
    let device = rest
        .iter()
        .position(|a| a == "--device")
        .and_then(|k| rest.get(k + 1))
        .and_then(|v| v.parse().ok())
        .unwrap_or(0i32);
    if !HipBackend.available() {
        println!("no gfx950 device: this probe needs an MI355X");
        return;
    }
    let (mut wa, fi) = scene();
    let (mut wb, _) = scene();
    // One CPU tick in both worlds builds the boundary particles the stepper starts from.
    wa.step();
    wb.step();
    let stepper = HipFluidStepper::new(&wb.fluids()[fi].0, device).expect("MI355X fluid stepper");
    assert!(wb.set_fluid_stepper(fi, Box::new(stepper)), "fluid index out of range");

    for t in 1..=ticks {
        wa.step();
        wb.step();
        assert_eq!(floating_box(&wa), floating_box(&wb), "tick {t}: floating box differs between CPU and device");
        // After the tick A holds the stepped fluid and the boundary segment
        // rebuilt for this tick: the particles B's stepper holds on the device.
        let (alo, ahi) = aabb(wa.fluids()[fi].0.raw_particles().0);
        let st: &dyn FluidStepper = wb.fluids()[fi].2.as_deref().expect("registered stepper");
        let (lo, hi) = st.bounds().expect("bounds after a step");
        assert_eq!((lo, hi), (alo, ahi), "tick {t}: device bounds differ from the CPU particles' AABB");
        if t == 1 || t % 50 == 0 || t == ticks {
            let b = floating_box(&wa);
            println!("tick {t:>4}: box y {:.6} vy {:.6} (CPU = device, bit for bit)",
                     f32::from_bits(b[1]), f32::from_bits(b[4]));
        }
    }
    println!("{ticks} ticks: the device-stepped world equals the CPU world bit for bit");
}
