//! Links libvxl_sph.so (built by `make -C rust-wl_amd`) and records its
//! directory as the run path of test and example binaries.
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("VXL_SPH_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join(".."),
    };
    let lib = dir.join("libvxl_sph.so");
    if !lib.exists() {
        panic!("{} is missing: build it with `make -C rust-wl_amd` or set VXL_SPH_LIB_DIR", lib.display());
    }
    let dir = dir.canonicalize().unwrap();
    println!("cargo:rerun-if-env-changed=VXL_SPH_LIB_DIR");
    println!("cargo:rerun-if-changed={}", lib.display());
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=vxl_sph");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
