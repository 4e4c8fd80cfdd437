// Links libgzgpu.so.  GZGPU_LIB_DIR names the directory that holds it (the
// gzippy_amd package directory after `python -c "from gzippy_amd import build;
// build.build_all()"`); GZGPU_BUILD=1 runs that build first from
// GZGPU_SRC_DIR (the gzippy_amd repository root).
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    println!("cargo:rerun-if-env-changed=GZGPU_LIB_DIR");
    println!("cargo:rerun-if-env-changed=GZGPU_SRC_DIR");
    println!("cargo:rerun-if-env-changed=GZGPU_BUILD");
    let src = env::var("GZGPU_SRC_DIR").ok().map(PathBuf::from);
    if env::var("GZGPU_BUILD").ok().as_deref() == Some("1") {
        let root = src.clone().expect("GZGPU_BUILD=1 needs GZGPU_SRC_DIR");
        let ok = Command::new("python3")
            .args(["-c", "from gzippy_amd import build; build.build_all()"])
            .current_dir(&root)
            .status()
            .expect("python3 not found")
            .success();
        assert!(ok, "building libgzgpu.so failed");
    }
    let lib_dir = env::var("GZGPU_LIB_DIR")
        .map(PathBuf::from)
        .or_else(|_| src.map(|s| s.join("gzippy_amd")).ok_or(()))
        .expect("set GZGPU_LIB_DIR (directory of libgzgpu.so) or GZGPU_SRC_DIR");
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=gzgpu");
    // the HIP runtime libgzgpu.so was built against
    println!("cargo:rustc-link-search=native=/opt/rocm/lib");
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
}
