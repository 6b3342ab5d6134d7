// Links libalgebra_linear_hip.so, built by algebra_linear_amd/Makefile (hipcc --offload-arch=gfx950).
use std::env;
use std::path::PathBuf;

fn main() {
    let dir = match env::var("ALGEBRA_LINEAR_HIP_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join(".."),
    };
    let lib = dir.join("libalgebra_linear_hip.so");
    if !lib.exists() {
        panic!(
            "{} is missing: build it with `make -C {}` (hipcc, gfx950) or set ALGEBRA_LINEAR_HIP_DIR",
            lib.display(),
            dir.display()
        );
    }
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=algebra_linear_hip");
    // let test binaries and dependents find the shared library at run time
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    println!("cargo:rerun-if-env-changed=ALGEBRA_LINEAR_HIP_DIR");
    println!("cargo:rerun-if-changed={}", lib.display());
}
