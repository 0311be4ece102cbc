// Builds liblbflow_hip.so from the HIP sources with hipcc (ROCm) for gfx950
// and links it. LBFLOW_HIP_SRC points at lbflow_amd/csrc (default: the copy
// in this repository), LBFLOW_HIP_ARCH overrides the offload target.
use std::{env, path::PathBuf, process::Command};

const UNITS: [&str; 7] = [
    "api.hip", "misc.hip", "serve.hip", "lat_d2q9.hip", "lat_d3q15.hip", "lat_d3q19.hip",
    "lat_d3q27.hip",
];

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let src = env::var("LBFLOW_HIP_SRC")
        .map(PathBuf::from)
        .unwrap_or_else(|_| manifest.join("../../lbflow_amd/csrc"));
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let arch = env::var("LBFLOW_HIP_ARCH").unwrap_or_else(|_| "gfx950".into());
    let mut objs = Vec::new();
    for unit in UNITS {
        let obj = out.join(unit.replace(".hip", ".o"));
        // contraction off: the FP64 path keeps the bits of a Rust build, which
        // never fuses a multiply-add; packed FP32 instructions stay out
        let status = Command::new("hipcc")
            .arg(format!("--offload-arch={arch}"))
            .args(["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off"])
            .args(["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-c"])
            .arg(src.join(unit))
            .arg("-o")
            .arg(&obj)
            .status()
            .expect("hipcc not found: install ROCm or put hipcc on PATH");
        assert!(status.success(), "hipcc failed on {unit}");
        println!("cargo:rerun-if-changed={}", src.join(unit).display());
        objs.push(obj);
    }
    let lib = out.join("liblbflow_hip.so");
    let status = Command::new("hipcc")
        .arg(format!("--offload-arch={arch}"))
        .args(["-shared", "-fPIC", "-o"])
        .arg(&lib)
        .args(&objs)
        .status()
        .expect("hipcc not found");
    assert!(status.success(), "linking liblbflow_hip.so failed");
    println!("cargo:rerun-if-changed={}", src.display());
    println!("cargo:rerun-if-env-changed=LBFLOW_HIP_SRC");
    println!("cargo:rerun-if-env-changed=LBFLOW_HIP_ARCH");
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=lbflow_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
}
