"""Builds tests/host/host_harness.cpp (g++, against the in-tree libpsvfwi.so)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cuda_fwi_amd", "lib")


def build_harness(outdir):
    exe = os.path.join(str(outdir), "host_harness")
    if not os.path.exists(os.path.join(LIBDIR, "libpsvfwi.so")):
        raise RuntimeError("libpsvfwi.so is not built (make -C cuda_fwi_amd/csrc)")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host", "host_harness.cpp"), "-o", exe,
                    "-L" + LIBDIR, "-lpsvfwi", "-Wl,-rpath," + LIBDIR], check=True)
    return exe


def run(exe, *args, cwd=None, timeout=300):
    r = subprocess.run([exe, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"host_harness {args[0]} failed ({r.returncode}): {r.stderr[-2000:]}")
    return r
