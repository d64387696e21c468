"""Builds the gfx950 engine library in-tree.

    python -m vlsi_layout_amd.build          # libvlayout.so (+ the Fortran layer)

Every HIP source is compiled with hipcc for gfx950 only, then linked into
vlsi_layout_amd/libvlayout.so.  The Fortran host layer (vl_layer_ops module
and the MAGPARSER-style `exec run` driver) is compiled with flang against it.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libvlayout.so")
FLIB = os.path.join(PKG, "libvlayout_f.so")
FEXE = os.path.join(PKG, "vl_magparser")
CHECK_EXE = os.path.join(ROOT, "tests", "fortran", "check_layer_ops")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")
FLANG = os.environ.get("FLANG", shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off in every gfx950 build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]
HOST_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include")]

HIP_SOURCES = ["vl_prims.hip", "vl_sort.hip", "vl_tile.hip", "vl_fused.hip", "vl_stitch.hip", "vl_dense.hip", "vl_window.hip", "vl_pnum.hip", "vl_snap.hip", "vl_io.hip", "vl_noalloc.hip", "vl_engine.hip"]
CPP_SOURCES = ["vl_synth.cpp"]
F_SOURCES = ["vl_layer_ops.f90", "vl_rect_boolean.f90", "vl_control.f90"]
F_MAIN = "vl_magparser.f90"


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", os.path.join(BUILD, s + ".o")])
    for s in CPP_SOURCES:   # host-only C++: no device code
        jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", os.path.join(BUILD, s + ".o")])
    workers = min(len(jobs), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(BUILD, s + ".o") for s in HIP_SOURCES + CPP_SOURCES]
    tmp = LIB + ".tmp"
    # zlib for SNAP files (zstd is opened at run time when a file needs it)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz", "-ldl"])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    if not os.path.isdir(FSRC) or not all(os.path.exists(os.path.join(FSRC, f)) for f in F_SOURCES):
        return None
    os.makedirs(BUILD, exist_ok=True)
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for f in F_SOURCES:
        o = os.path.join(BUILD, f + ".o")
        _run([FLANG, "-O2", "-fPIC", "-module-dir", mod, "-I", mod, "-c", os.path.join(FSRC, f), "-o", o])
        objs.append(o)
    tmp = FLIB + ".tmp"
    _run([FLANG, "-shared", "-fPIC", "-o", tmp] + objs + ["-L", PKG, "-lvlayout", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    main_o = os.path.join(BUILD, F_MAIN + ".o")
    _run([FLANG, "-O2", "-module-dir", mod, "-I", mod, "-c", os.path.join(FSRC, F_MAIN), "-o", main_o])
    tmp = FEXE + ".tmp"
    _run([FLANG, "-o", tmp, main_o] + objs + ["-L", PKG, "-lvlayout", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FEXE)
    # the test helper that calls the module procedures directly
    check_src = os.path.join(ROOT, "tests", "fortran", "check_layer_ops.f90")
    if os.path.exists(check_src):
        check_o = os.path.join(BUILD, "check_layer_ops.o")
        _run([FLANG, "-O2", "-module-dir", mod, "-I", mod, "-c", check_src, "-o", check_o])
        tmp = CHECK_EXE + ".tmp"
        _run([FLANG, "-o", tmp, check_o] + objs + ["-L", PKG, "-lvlayout", "-Wl,-rpath," + PKG])
        os.replace(tmp, CHECK_EXE)
    return FLIB


def main(argv=None):
    build_hip(verbose="-v" in (argv or sys.argv[1:]))
    build_fortran()
    print(LIB)


if __name__ == "__main__":
    main()
