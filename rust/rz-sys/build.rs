// Links librz.so, built by rusterizer_amd/build.py (hipcc, gfx950) into
// rusterizer_amd/_lib/. RZ_LIB_DIR overrides that location.
use std::path::PathBuf;

fn main() {
    let dir = match std::env::var("RZ_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env!("CARGO_MANIFEST_DIR")).join("../../rusterizer_amd/_lib"),
    };
    println!("cargo:rerun-if-env-changed=RZ_LIB_DIR");
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=rz");
    // let binaries find librz.so next to where it was built
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
