"""Build the gfx950 shared library ``libanelastic_mg.so`` in-tree with hipcc.

Usage: ``python -m anelastic_solve_amd.build`` (or ``build()`` from Python).
The library is the product: the HIP kernels, the host orchestrator and the
C ABI of ``include/amg.h``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libanelastic_mg.so")

SOURCES = ["amg_kernels.hip", "amg_flow.hip", "amg_hierarchy.cpp", "amg_direct.cpp",
           "amg_solver.cpp", "amg_server.cpp", "amg_capi.cpp", "amg_noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

# Fortran drop-in shim (ISO_C_BINDING) and its test driver, in link order
FORTRAN = ["amg_cabi.F90", "amg_host_modules.F90", "amg_mg_shim.F90", "amg_shim_driver.F90"]
FDIR = os.path.join(HERE, "fortran")
FDRIVER = os.path.join(FDIR, "bin", "amg_shim_driver")
# restart-dump reader/writer in Fortran (round-trip checks of restart.py)
FRESTART = os.path.join(FDIR, "bin", "amg_restart_io")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-function", "-Wno-unused-result",
    # packed-FP32 instructions stay off (see the design notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build(verbose=False, jobs=4):
    """Compile every source for gfx950 and link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def build_fortran(verbose=False):
    """Compile the Fortran shim and its driver against the library (skipped,
    with a note, when the image has no flang)."""
    if not os.path.exists(FLANG):
        sys.stderr.write("flang not found: Fortran shim not built\n")
        return None
    moddir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(moddir, exist_ok=True)
    os.makedirs(os.path.dirname(FDRIVER), exist_ok=True)
    objs = []
    for src in FORTRAN:
        obj = os.path.join(moddir, os.path.splitext(src)[0] + ".o")
        _run([FLANG, "-O2", "-fPIC", "-module-dir", moddir, "-c", os.path.join(FDIR, src),
              "-o", obj], verbose)
        objs.append(obj)
    tmp = FDRIVER + ".tmp"
    _run([FLANG, *objs, "-L" + HERE, "-lanelastic_mg", "-Wl,-rpath,$ORIGIN/../..", "-o", tmp],
         verbose)
    os.replace(tmp, FDRIVER)
    tmp = FRESTART + ".tmp"
    _run([FLANG, "-O2", "-module-dir", moddir, os.path.join(FDIR, "amg_restart_io.F90"), "-o", tmp],
         verbose)
    os.replace(tmp, FRESTART)
    return FDRIVER


def _wait(item):
    cmd, p = item
    out, err = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out + err)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    msg = (out + err).strip()
    if msg:
        # keep real warnings visible, drop the host pass's note on the packed-fp32 feature
        lines = [l for l in msg.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
