"""Build recipe of the native libraries of this package.

libbm.so        HIP kernels for gfx950 + the C++ host layer + the C ABI
                (include/bm.h), built with hipcc.
libbm_synth.so  the synthetic-volume generator (host C++, OpenMP).

Every HIP translation unit is compiled with FMA contraction disabled (the
f32/f64 arithmetic that must match the reference stays unfused) and with
packed-FP32 instruction selection disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["bm_kernels.hip", "bm_tiles.hip", "bm_engine.hip", "bm_quantiles.hip", "bm_smooth.hip",
               "bm_topology.hip", "bm_noalloc.hip",
               "bm_tune.cpp", "bm_serve.cpp", "bm_tiff.cpp", "bm_api.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_libbm(jobs=8):
    """Compile every HIP source to an object (in parallel) and link libbm.so."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [hipcc] + HIPCC_FLAGS + lang + ["-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(cmd[-3])
    if failed:
        raise RuntimeError("hipcc failed on: " + ", ".join(failed))
    out = os.path.join(HERE, "libbm.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"], CSRC)
    os.replace(tmp, out)
    return out


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build of the whole
# library (device code is built as usual: each -fsanitize= applies to the host
# pass only).  Loaded through BM_LIBRARY with the sanitizer runtime preloaded,
# it runs the CPU tests of the host code (the TIFF decoder, argument checks).
SANITIZE_FLAGS = ("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
                  "-Xarch_host -fno-sanitize-recover=undefined "
                  "-Xarch_host -fno-omit-frame-pointer -Xarch_host -shared-libasan").split()
SANITIZED_LIB = os.path.join(HERE, "_build", "sanitized", "libbm.so")


def sanitizer_runtime():
    """The clang ASan runtime the sanitized library links against."""
    hipcc = _hipcc()
    r = subprocess.run([os.path.join(os.path.dirname(os.path.realpath(hipcc)), "..", "llvm",
                                     "bin", "clang++"), "--print-file-name=libclang_rt.asan-x86_64.so"],
                       stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    path = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        import glob
        hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        path = hits[0] if hits else ""
    return path


def build_sanitized(jobs=8):
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build", "sanitized")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [hipcc] + HIPCC_FLAGS + SANITIZE_FLAGS + ["-O1"] + lang + ["-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(cmd[-3])
    if failed:
        raise RuntimeError("hipcc (sanitized) failed on: " + ", ".join(failed))
    tmp = SANITIZED_LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + SANITIZE_FLAGS + objs +
         ["-lz"], CSRC)
    os.replace(tmp, SANITIZED_LIB)
    return SANITIZED_LIB


def build_synth():
    out = os.path.join(HERE, "libbm_synth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
          "-std=c++17", "-o", tmp, os.path.join(CSRC, "synth.cpp")], CSRC)
    os.replace(tmp, out)
    return out


def build_all():
    build_synth()
    return build_libbm()


if __name__ == "__main__":
    print(build_all())
