//! Compiles the library's HIP and C++ sources (../bonsai_amd/csrc) with hipcc for gfx950 into
//! libbonsai_gpu.so in OUT_DIR, with the flags bonsai_amd/build.py uses: every f32 operation rounds
//! once (-ffp-contract=off, no fast-math), denormals kept, packed-FP32 instructions off.
use std::env;
use std::path::{Path, PathBuf};
use std::process::Command;
use std::thread;

const ARCH: &str = "gfx950";

fn common(include: &Path) -> Vec<String> {
    let mut v: Vec<String> = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
                              "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument"]
        .iter()
        .map(|s| s.to_string())
        .collect();
    v.push(format!("-I{}", include.display()));
    v
}

fn device() -> Vec<String> {
    vec![
        format!("--offload-arch={ARCH}"),
        "-fno-gpu-flush-denormals-to-zero".into(),
        "-Xclang".into(), "-target-feature".into(), "-Xclang".into(), "-packed-fp32-ops".into(),
    ]
}

fn run(cmd: &mut Command) {
    let out = cmd.output().unwrap_or_else(|e| panic!("cannot run {:?}: {e}", cmd));
    if !out.status.success() {
        panic!("{:?} failed:\n{}\n{}", cmd, String::from_utf8_lossy(&out.stdout), String::from_utf8_lossy(&out.stderr));
    }
}

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let root = manifest.parent().unwrap().to_path_buf();
    let csrc = root.join("bonsai_amd").join("csrc");
    let include = root.join("include");
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".into());

    let mut sources: Vec<PathBuf> = std::fs::read_dir(&csrc)
        .expect("bonsai_amd/csrc")
        .map(|e| e.unwrap().path())
        .filter(|p| matches!(p.extension().and_then(|e| e.to_str()), Some("hip") | Some("cpp")))
        .collect();
    sources.sort();

    // one hipcc per source, in parallel
    let handles: Vec<_> = sources
        .iter()
        .map(|src| {
            let (src, out, hipcc, include) = (src.clone(), out.clone(), hipcc.clone(), include.clone());
            thread::spawn(move || {
                let obj = out.join(format!("{}.o", src.file_name().unwrap().to_str().unwrap()));
                let mut cmd = Command::new(&hipcc);
                cmd.args(common(&include));
                if src.extension().unwrap() == "hip" {
                    cmd.args(device()).args(["-x", "hip", "-c"]);
                } else {
                    cmd.args(["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", "-c"]);
                }
                cmd.arg(&src).arg("-o").arg(&obj);
                run(&mut cmd);
                obj
            })
        })
        .collect();
    let objs: Vec<PathBuf> = handles.into_iter().map(|h| h.join().expect("compile thread")).collect();

    let lib = out.join("libbonsai_gpu.so");
    let mut link = Command::new(&hipcc);
    link.arg(format!("--offload-arch={ARCH}")).args(["-shared", "-fPIC", "-o"]).arg(&lib).args(&objs).arg("-lpthread");
    run(&mut link);

    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=bonsai_gpu");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", out.display());
    println!("cargo:rerun-if-changed={}", csrc.display());
    println!("cargo:rerun-if-changed={}", include.display());
    println!("cargo:rerun-if-env-changed=HIPCC");
}
