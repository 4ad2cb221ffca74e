"""Builds the host-side code of the library and the oracle with
AddressSanitizer and UndefinedBehaviorSanitizer, plus two drivers that run
their CPU-only paths (no device needed):

- oracle_driver: every oracle kernel on small states (qubits straddling the
  4096-amplitude block, edge draws for sampling);
- lib_driver: circuit text parsing and printing (valid and invalid texts),
  circuit building, the fused-schedule planner and its host emulation, and
  the batched form's planner, FP64 emulator and host run of its kernel code.

The sanitizers instrument host code only (the HIP kernels are compiled as
usual).  Outputs go to tests/sanitize/_build/.  Usage:

    python tests/sanitize/build_san.py      # build and run both drivers
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
CSRC = os.path.join(ROOT, "ncsim_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANGXX = os.environ.get("SAN_CXX", "/opt/rocm/llvm/bin/clang++")
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
COMMON = ["-O1", "-g", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
          "-Wno-unused-function", "-Wno-unused-variable", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CPP = ["lower.cpp", "circuit.cpp", "api.cpp", "fuse.cpp", "jit.cpp", "batch.cpp", "batch_sim.cpp"]
HIP = ["kernels.hip", "tile.hip", "batch.hip", "frames.hip"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(" ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def compile_one(src):
    obj = os.path.join(OUT, src + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        run([HIPCC, "-x", "hip", "-c", path, "-o", obj] + COMMON)
    else:
        run([HIPCC, "-c", path, "-o", obj] + COMMON + HOST_SAN)
    return obj


def san_runtime_dir():
    """The directory of the compiler's shared sanitizer runtime."""
    r = run([CLANGXX, "-print-file-name=libclang_rt.asan-x86_64.so"])
    path = r.stdout.strip()
    if not os.path.isabs(path):
        r = run([CLANGXX, "--print-runtime-dir"])
        return r.stdout.strip()
    return os.path.dirname(path)


def build():
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(compile_one, CPP + HIP))
    lib = os.path.join(OUT, "libncsv_san.so")
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", lib] + objs +
        ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan",
         "-lhiprtc", "-lpthread"])
    drv = os.path.join(OUT, "lib_driver")
    run([CLANGXX, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         "-shared-libsan", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "lib_driver.cpp"), lib,
         "-Wl,-rpath," + OUT, "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath," + san_runtime_dir(), "-o", drv])
    odrv = os.path.join(OUT, "oracle_driver")
    run(["gcc", "-O1", "-g", "-std=gnu11", "-fopenmp", "-ffp-contract=off", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", os.path.join(HERE, "oracle_driver.c"),
         os.path.join(ROOT, "oracle", "sv_oracle.c"), "-lm", "-o", odrv])
    return drv, odrv


def main():
    drv, odrv = build()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    for d in (odrv, drv):
        r = subprocess.run([d], capture_output=True, text=True, env=env)
        sys.stdout.write(r.stdout[-2000:])
        sys.stderr.write(r.stderr[-6000:])
        if r.returncode:
            return r.returncode
    return 0


if __name__ == "__main__":
    sys.exit(main())
