// Links libspicier_hip.so, built by `python -m spicier_amd.build`
// (hipcc --offload-arch=gfx950) into spicier_amd/ of this repository, or
// found in SPICIER_HIP_LIB_DIR.
fn main() {
    let dir = std::env::var("SPICIER_HIP_LIB_DIR").unwrap_or_else(|_| {
        let here = std::env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR");
        format!("{here}/../../spicier_amd")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=spicier_hip");
    println!("cargo:rerun-if-env-changed=SPICIER_HIP_LIB_DIR");
}
