"""Builds tests/host/host_harness.cpp (g++, against the in-tree libpsvfwi.so).

With sanitize=True the host layer's sources (INI reader, FD* classes, FWI driver, fd_kernel
entry points) are compiled into the harness itself with AddressSanitizer and
UndefinedBehaviorSanitizer, so its calls reach the instrumented copies; the HIP library
stays uninstrumented (sanitizers run on host code only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cuda_fwi_amd", "lib")
HOST_SOURCES = ("ini_reader.cpp", "fd_host.cpp", "fd_wave_model.cpp", "fd_kernel_dropin.cpp")
SANITIZER_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
                 "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def build_harness(outdir, sanitize=False):
    exe = os.path.join(str(outdir), "host_harness_san" if sanitize else "host_harness")
    if not os.path.exists(os.path.join(LIBDIR, "libpsvfwi.so")):
        raise RuntimeError("libpsvfwi.so is not built (make -C cuda_fwi_amd/csrc)")
    cmd = ["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host", "host_harness.cpp")]
    if sanitize:
        cmd[1:2] = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined"]
        cmd += ["-I" + os.path.join(ROOT, "cuda_fwi_amd", "csrc", "host")]
        cmd += [os.path.join(ROOT, "cuda_fwi_amd", "csrc", "host", f) for f in HOST_SOURCES]
    cmd += ["-o", exe, "-L" + LIBDIR, "-lpsvfwi", "-Wl,-rpath," + LIBDIR]
    subprocess.run(cmd, check=True)
    return exe


def run(exe, *args, cwd=None, timeout=300):
    env = dict(os.environ, **SANITIZER_ENV) if exe.endswith("_san") else None
    r = subprocess.run([exe, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"host_harness {args[0]} failed ({r.returncode}): {r.stderr[-2000:]}")
    return r
