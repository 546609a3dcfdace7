"""Builds libsuperconga_amd.so in-tree for gfx950.

Host tables (sc_setup.cpp) are compiled with g++ and -ffp-contract=off so
that geometric decisions are taken on plainly rounded FP64 arithmetic; the
HIP translation units are compiled with hipcc for gfx950 with packed FP32
instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libsuperconga_amd.so")
SIMULATION = os.path.join(LIBDIR, "simulation_amd")
ARCH = os.environ.get("SUPERCONGA_AMD_ARCH", "gfx950")


def _hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"),
              shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              "-Wextra", "-Wno-unused-parameter", "-pthread"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = _hipcc()
    jobs = [
        ["g++"] + HOST_FLAGS + ["-c", os.path.join(SRC, "sc_setup.cpp"),
                                "-o", os.path.join(OBJ, "sc_setup.o")],
    ]
    for name in ("sc_engine", "sc_capi", "sc_fft", "sc_newton", "sc_ldos", "sc_adjoint",
                 "sc_noalloc"):
        src = os.path.join(SRC, name + ".hip")
        if not os.path.exists(src):
            continue
        jobs.append([hipcc] + HIP_FLAGS + ["-c", src, "-o",
                                           os.path.join(OBJ, name + ".o")])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
         + ["-pthread", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    # the drop-in for the reference's `simulation` binary, host code on the
    # C++ adaptors (include/superconga_amd.hpp) linked against the library
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    tmp = SIMULATION + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra",
          "-I" + os.path.join(os.path.dirname(HERE), "include"),
          os.path.join(SRC, "simulation_amd.cpp"), "-o", tmp, "-L" + LIBDIR,
          "-lsuperconga_amd", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + rocm_lib])
    os.replace(tmp, SIMULATION)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
