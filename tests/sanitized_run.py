"""Host-side AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU
tests of the C ABI and the oracle.

    python tests/sanitized_run.py [pytest args]

builds, into a temporary directory,
  - libilcpsg.so from illcrawl_amd/csrc with the host code instrumented
    (hipcc with -Xarch_host before every -fsanitize= flag: the gfx950 device
    code is compiled as usual and not instrumented), and
  - the oracle (oracle/oracle.c) with clang and the same sanitizers,
then runs tests/test_capi_cpu.py and tests/test_oracle.py against those
builds (ILC_LIB_PATH, ILC_ORACLE_LIB) with the sanitizer runtime preloaded
into the Python process.  Any report aborts the run (halt_on_error); leak
checking is off, since the interpreter itself does not free everything at
exit.  No GPU is needed or touched.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from illcrawl_amd import build as B  # noqa: E402

HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host",
            "-fno-omit-frame-pointer"]


def llvm_dir():
    return os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm")


def asan_runtime():
    hits = sorted(glob.glob(os.path.join(llvm_dir(), "lib", "clang", "*", "lib", "linux",
                                         "libclang_rt.asan-x86_64.so")))
    if not hits:
        raise SystemExit("clang's ASan runtime (libclang_rt.asan-x86_64.so) not found")
    return hits[-1]


def build_library(out_dir):
    cc = B.hipcc()
    common = ["-O1", "-g", "-std=c++17", "-fPIC", f"--offload-arch={B.ARCH}", "-I",
              os.path.join(ROOT, "include"), "-I", B.CSRC] + B.DEVICE_FLAGS + HOST_SAN
    objs, procs = [], []
    for src in B.SOURCES:
        obj = os.path.join(out_dir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        procs.append((src, subprocess.Popen([cc] + common + lang +
                                            ["-c", os.path.join(B.CSRC, src), "-o", obj],
                                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode:
            raise SystemExit(f"hipcc failed on {src}:\n{out.decode(errors='replace')[-4000:]}")
    lib = os.path.join(out_dir, "libilcpsg.so")
    subprocess.run([cc, "-shared", f"--offload-arch={B.ARCH}"] + HOST_SAN + ["-o", lib] + objs,
                   check=True)
    return lib


def build_oracle(out_dir):
    lib = os.path.join(out_dir, "liboracle.so")
    clang = os.path.join(llvm_dir(), "bin", "clang")
    subprocess.run([clang, "-O1", "-g", "-std=c11", "-D_GNU_SOURCE", "-ffp-contract=off",
                    "-fno-fast-math", "-fopenmp", "-fPIC", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-shared",
                    "-o", lib, os.path.join(ROOT, "oracle", "oracle.c"), "-lm",
                    "-Wl,-rpath," + os.path.join(llvm_dir(), "lib")], check=True)
    return lib


def instrumented(lib):
    """Number of sanitizer report hooks the library imports (0: the flags did
    not take)."""
    out = subprocess.run(["nm", "-D", "--undefined-only", lib], capture_output=True,
                         text=True, check=True).stdout
    return sum(1 for line in out.splitlines()
               if "__asan_report" in line or "__ubsan_handle" in line)


def main(argv):
    tmp = tempfile.mkdtemp(prefix="ilc_sanitized_")
    try:
        print("[sanitized] building the library (host code instrumented) ...", flush=True)
        lib = build_library(tmp)
        print("[sanitized] building the oracle ...", flush=True)
        olib = build_oracle(tmp)
        for name, path in (("library", lib), ("oracle", olib)):
            k = instrumented(path)
            print(f"[sanitized] {name}: {k} sanitizer hooks", flush=True)
            if k == 0:
                raise SystemExit(f"{path} is not instrumented")
        env = dict(os.environ)
        env.update(ILC_LIB_PATH=lib, ILC_ORACLE_LIB=olib, LD_PRELOAD=asan_runtime(),
                   ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
                   UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
        cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
               os.path.join(ROOT, "tests", "test_capi_cpu.py"),
               os.path.join(ROOT, "tests", "test_oracle.py")] + list(argv)
        print("[sanitized] " + " ".join(cmd[1:]), flush=True)
        return subprocess.run(cmd, env=env, cwd=ROOT).returncode
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
