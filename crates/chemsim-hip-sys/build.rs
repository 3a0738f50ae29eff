// Links libchemsim_hip.so.  CHEMSIM_HIP_LIB_DIR names the directory holding
// it (default: the library's in-tree build directory, chemsim_amd/lib, next
// to this crate's repository root); the directory is also put on the rpath.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=CHEMSIM_HIP_LIB_DIR");
    let dir = match env::var_os("CHEMSIM_HIP_LIB_DIR") {
        Some(d) => PathBuf::from(d),
        None => {
            let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR"));
            manifest.join("..").join("..").join("chemsim_amd").join("lib")
        }
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=chemsim_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    println!("cargo:root={}", dir.display());
}
