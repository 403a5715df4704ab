// Links libspecspars_hip.so from the directory named by SPEC_SPARS_HIP_LIB
// (spec_spars_amd/ of a built checkout), with an rpath so that the binary
// finds it at run time.
fn main() {
    let dir = std::env::var("SPEC_SPARS_HIP_LIB")
        .expect("set SPEC_SPARS_HIP_LIB to the directory holding libspecspars_hip.so");
    println!("cargo:rerun-if-env-changed=SPEC_SPARS_HIP_LIB");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=specspars_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
