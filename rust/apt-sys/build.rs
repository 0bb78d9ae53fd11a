// Links libapt.so.  APT_LIB_DIR names the directory holding it (the
// appearance_amd package directory after `python -m appearance_amd.build`);
// without it the default search path is used.
fn main() {
    println!("cargo:rerun-if-env-changed=APT_LIB_DIR");
    if let Ok(dir) = std::env::var("APT_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    }
    println!("cargo:rustc-link-lib=dylib=apt");
}
