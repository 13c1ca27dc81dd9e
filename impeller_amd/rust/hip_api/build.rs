// Links the crate against libimpeller_hip.so (the impeller_rt_* symbols) and
// the HIP runtime.  IMPELLER_LIB_DIR: the directory holding
// libimpeller_hip.so (default: the impeller_amd package, two levels up);
// ROCM_PATH: the ROCm installation (default /opt/rocm).
use std::env;
use std::path::PathBuf;

fn main() {
    let here = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let lib_dir = env::var("IMPELLER_LIB_DIR")
        .map(PathBuf::from)
        .unwrap_or_else(|_| here.join("..").join(".."));
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rerun-if-env-changed=IMPELLER_LIB_DIR");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-search=native={}/lib", rocm);
    println!("cargo:rustc-link-lib=dylib=impeller_hip");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
}
