fn main() {
    if std::env::var_os("CARGO_FEATURE_HIP").is_some() {
        let dir = std::env::var("LXV_HIP_LIB_DIR").unwrap_or_else(|_| "../lxmf-vanity_amd".into());
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-lib=dylib=lxmf_vanity_hip");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
}
