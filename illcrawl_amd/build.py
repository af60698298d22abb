"""Build the gfx950 shared library libilcpsg.so in-tree with hipcc.

    python -m illcrawl_amd.build

compiles every source under illcrawl_amd/csrc into illcrawl_amd/lib/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libilcpsg.so")
# C++ program driving the library through include/illcrawl_backend.hpp (GPU tests run it)
DEMO_SRC = os.path.join(ROOT, "tests", "cpp", "backend_demo.cpp")
DEMO = os.path.join(LIBDIR, "backend_demo")
# C++ data crawler over an HDF5 snapshot (include/illcrawl_crawler.hpp)
CRAWL_SRC = os.path.join(ROOT, "tests", "cpp", "crawl_demo.cpp")
CRAWL = os.path.join(LIBDIR, "crawl_demo")
# the render tool (illcrawl_render: image, animation and tomography targets)
RENDER_SRC = os.path.join(HERE, "app", "illcrawl_render.cpp")
RENDER = os.path.join(LIBDIR, "illcrawl_render")
IO_PREFIX = os.environ.get("ILC_IO_PREFIX", "/opt/conda")  # HDF5 C++ and cfitsio

SOURCES = ["host.cpp", "gridbuild.cpp", "capi.cpp", "batch.cpp", "crawl.hip", "grid.hip",
           "gather.hip", "xray.hip", "smooth3d.hip",
           "project.hip", "batch.hip"]

ARCH = "gfx950"

# Packed FP32 (v_pk_*_f32) is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_io_program(src, out):
    """A host program over libilcpsg, HDF5's C++ API and cfitsio."""
    inc, lib = os.path.join(IO_PREFIX, "include"), os.path.join(IO_PREFIX, "lib")
    # host code only (HIP runtime API): g++, whose libstdc++ matches the one
    # libhdf5_cpp was built against
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    # The prefix ships an older libstdc++ of its own: its libraries are linked
    # by path, and the system library directory precedes it in the run path.
    syslib = "/usr/lib/x86_64-linux-gnu"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-isystem", os.path.join(rocm, "include"),
           "-isystem", inc, src, "-L", LIBDIR, "-lilcpsg",
           "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           os.path.join(lib, "libhdf5_cpp.so"), os.path.join(lib, "libhdf5.so"),
           os.path.join(lib, "libcfitsio.so"), f"-Wl,-rpath-link,{lib}",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib"),
           f"-Wl,-rpath,{syslib}", f"-Wl,-rpath,{lib}", "-o", out + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    cc = hipcc()
    common = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I",
              os.path.join(ROOT, "include"), "-I", CSRC, "-Wall", "-Wno-unused-function",
              "-Wno-unused-command-line-argument"] + DEVICE_FLAGS
    tmp = tempfile.mkdtemp(prefix="ilcbuild_")
    objs = []
    procs = []
    try:
        for src in SOURCES:
            obj = os.path.join(tmp, src + ".o")
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            cmd = [cc] + common + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT)))
            objs.append(obj)
        failed = []
        for src, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                failed.append((src, out.decode(errors="replace")))
            elif verbose and out:
                print(out.decode(errors="replace"))
        if failed:
            msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
            raise RuntimeError("hipcc failed:\n" + msg)
        tmp_lib = LIB + ".tmp"
        cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp_lib] + objs
        subprocess.run(cmd, check=True)
        os.replace(tmp_lib, LIB)
        if os.path.exists(DEMO_SRC):
            cmd = [cc, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
                   DEMO_SRC, "-L", LIBDIR, "-lilcpsg", "-Wl,-rpath,$ORIGIN", "-o", DEMO + ".tmp"]
            subprocess.run(cmd, check=True)
            os.replace(DEMO + ".tmp", DEMO)
        # the data crawler and the render tool (HDF5 in, FITS out) need
        # HDF5's C++ API and cfitsio
        if os.path.exists(os.path.join(IO_PREFIX, "include", "H5Cpp.h")):
            for src, out in ((CRAWL_SRC, CRAWL), (RENDER_SRC, RENDER)):
                if os.path.exists(src):
                    build_io_program(src, out)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
