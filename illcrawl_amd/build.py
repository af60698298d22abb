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

SOURCES = ["host.cpp", "gridbuild.cpp", "capi.cpp", "batch.cpp", "crawl.hip", "grid.hip",
           "gather.hip", "xray.hip",
           "project.hip", "batch.hip"]

ARCH = "gfx950"

# Packed FP32 (v_pk_*_f32) is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


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
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
