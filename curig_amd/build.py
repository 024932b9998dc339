"""Builds libcurafft.so (HIP for gfx950) in-tree, next to this file.

    python -m curig_amd.build
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libcurafft.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["kern_lds_spread_f32.hip", "kern_lds_interp_f32.hip", "kern_lds_cols_f32.hip", "kern_lds_spread_f64.hip",
           "kern_lds_interp_f64.hip", "kern_lds_cols_f64.hip", "kern_plane.hip", "kern_prep.hip", "gridder.hip",
           "plan_math.cpp", "fft.cpp", "capi.cpp"]
ARCH = os.environ.get("CURIG_OFFLOAD_ARCH", "gfx950")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-x", "hip",
        # no packed-FP32 instructions in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-result",
        "-Wno-unused-command-line-argument",
    ]


def build(verbose=False, jobs=None, defines=(), out=None, objdir=None):
    """Compile and link the library.  defines: extra -D macros (e.g. the
    layout constants of curig_internal.h); out / objdir: another library path
    and object folder (default: the package's libcurafft.so)."""
    objdir = objdir or OBJDIR
    out = out or LIB
    os.makedirs(objdir, exist_ok=True)
    hipcc = _hipcc()
    dflags = [f"-D{d}" for d in defines]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + _flags() + dflags + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr)
        return obj

    jobs = jobs or min(len(SOURCES), max(1, min(os.cpu_count() or 4, 16)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = out + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocfft", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    return out


def build_test_driver(lib=None):
    """The tests' driver of the C++ plan APIs (tests/plan_driver.cpp): host C++
    compiled by hipcc into tests/_bin/libcurig_plan_driver.so, linked against
    the library it drives (found next to it through the rpath)."""
    lib = lib or LIB
    src = os.path.join(ROOT, "tests", "plan_driver.cpp")
    outdir = os.path.join(ROOT, "tests", "_bin")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libcurig_plan_driver.so")
    rel = os.path.relpath(os.path.dirname(lib), outdir)
    cmd = [_hipcc(), "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I", INCLUDE, src, "-o", out + ".tmp",
           "-L", os.path.dirname(lib), "-l:" + os.path.basename(lib), f"-Wl,-rpath,$ORIGIN/{rel}",
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_test_driver())
