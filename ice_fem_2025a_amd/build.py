"""Builds libife.so (HIP kernels + C-ABI) for gfx950 with hipcc, in-tree.

Each source compiles to its own object in parallel, then one link.  build_engine(force=True)
(what __graft_entry__.build() calls) recompiles everything; force=False recompiles only the
objects older than their source or the headers."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_build")
OUT = os.path.join(OUT_DIR, "libife.so")
SOURCES = ["ife_host.cpp", "ife_mesh_hdf5.cpp", "ife_krylov.hip", "ife_linesearch.hip", "ife_kernels.hip"]
HEADERS = [os.path.join(CSRC, "ife_internal.h"), os.path.join(ROOT, "include", "ife.h")]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"),
]
LINK = ["--offload-arch=gfx950", "-shared", "-fPIC", "-ldl", "-L/opt/rocm/lib", "-lrocsparse",
        "-Wl,-rpath,/opt/rocm/lib"]


def _compile(src, force, verbose):
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    deps = [src] + HEADERS
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj
    cmd = [HIPCC] + FLAGS + ["-c", "-o", obj + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed compiling {os.path.basename(src)}")
    os.replace(obj + ".tmp", obj)
    return obj


def build_engine(verbose=False, force=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    with ThreadPoolExecutor(len(srcs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, force, verbose), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + ["-o", tmp] + objs + LINK
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libife.so")
    os.replace(tmp, OUT)
    return OUT


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "ife_driver.cpp")
DRIVER = os.path.join(OUT_DIR, "ife_driver")


def build_driver():
    """The C++ host driver (plain g++ over include/ife_model.hpp and the C-ABI)."""
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I/opt/rocm/include", DRIVER_SRC, "-o", DRIVER, "-L" + OUT_DIR, "-life", "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,-rpath," + OUT_DIR, "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the C++ host driver")
    return DRIVER


if __name__ == "__main__":
    inc = "--incremental" in sys.argv
    print(build_engine(verbose=True, force=not inc))
    print(build_driver())
