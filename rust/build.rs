//! Links libvoronoids_hip, built by `python -c "import __graft_entry__ as g; g.build()"` (hipcc,
//! gfx950) into voronoids_amd/ next to this crate.
//!
//! VORONOIDS_HIP_LIB_DIR: the folder holding libvoronoids_hip.so / libvoronoids_hip.a (default:
//!   ../voronoids_amd relative to this crate).
//! VORONOIDS_HIP_STATIC=1: link the static archive; it then needs the HIP runtime (amdhip64 from
//!   $ROCM_PATH/lib, default /opt/rocm/lib) and the C++ runtime.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=VORONOIDS_HIP_LIB_DIR");
    println!("cargo:rerun-if-env-changed=VORONOIDS_HIP_STATIC");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
    let dir = match env::var("VORONOIDS_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR"))
            .join("..")
            .join("voronoids_amd"),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    let static_lib = env::var("VORONOIDS_HIP_STATIC").map(|v| v == "1").unwrap_or(false);
    if static_lib {
        let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
        println!("cargo:rustc-link-lib=static=voronoids_hip");
        println!("cargo:rustc-link-search=native={}/lib", rocm);
        println!("cargo:rustc-link-lib=dylib=amdhip64");
        println!("cargo:rustc-link-lib=dylib=stdc++");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{}/lib", rocm);
    } else {
        println!("cargo:rustc-link-lib=dylib=voronoids_hip");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
    }
}
