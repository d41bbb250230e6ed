// Link libgeon_hip.so: from GEON_HIP_LIB_DIR when set, else from the
// package's build directory (geon_amd/lib beside this crate's parent).
fn main() {
    let dir = std::env::var("GEON_HIP_LIB_DIR").unwrap_or_else(|_| {
        let manifest = std::env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR");
        format!("{}/../../lib", manifest)
    });
    println!("cargo:rustc-link-search=native={}", dir);
    println!("cargo:rustc-link-lib=dylib=geon_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir);
    println!("cargo:rerun-if-env-changed=GEON_HIP_LIB_DIR");
}
