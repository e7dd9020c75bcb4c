// Links libhip_sys.so: from HIP_SYS_LIB_DIR where set, else from the library's in-tree build directory beside this
// crate (rustymath_amd/, where `python -m rustymath_amd.build` writes it).
fn main() {
    let dir = std::env::var("HIP_SYS_LIB_DIR").unwrap_or_else(|_| {
        let here = std::env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR");
        format!("{here}/../rustymath_amd")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=hip_sys");
    println!("cargo:rerun-if-env-changed=HIP_SYS_LIB_DIR");
}
