"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, validates arguments without a GPU, and its host camera math
equals the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

from illcrawl_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            syms |= set(re.findall(r"\b(ilc_[a-z0-9_]+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    L = N.lib()
    syms = header_symbols()
    assert len(syms) >= 19
    for s in sorted(syms):
        assert hasattr(L, s), s
    # and the binding declares them all
    assert syms <= set(N.SIGNATURES)


def test_abi_version_and_error_paths():
    L = N.lib()
    assert L.ilc_abi_version() == 1
    # null handles / arguments are rejected without touching a GPU
    assert L.ilc_psg_run(None, None) == N.ILC_EINVAL
    assert b"null" in L.ilc_last_error()
    assert L.ilc_psg_init_backend(None, 10) == N.ILC_EINVAL
    assert L.ilc_psg_create(None, None) == N.ILC_EINVAL
    assert L.ilc_psg_destroy(None) == N.ILC_OK
    with pytest.raises(N.IllcrawlError):
        N.call("ilc_psg_run", None, None)


def test_create_rejects_bad_config():
    cfg = N.PsgConfig()
    cfg.mode = 7
    h = ctypes.c_void_p()
    assert N.lib().ilc_psg_create(ctypes.byref(cfg), ctypes.byref(h)) == N.ILC_EINVAL
    cfg.mode = N.MODE_REFERENCE
    cfg.kernel = N.KERNEL_CUBIC
    assert N.lib().ilc_psg_create(ctypes.byref(cfg), ctypes.byref(h)) == N.ILC_EINVAL


def test_camera_basis_matches_oracle(orc):
    from illcrawl_amd.backend import camera_basis
    rng = np.random.default_rng(11)
    cams = [dict(position=[5000.0, 5000.0, -3660.0], look_at=[0.0, 0.0, 1.0], roll=0.0,
                 width=2000.0, nx=512, ny=512)]
    for _ in range(8):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cams.append(dict(position=rng.uniform(0, 1e4, 3), look_at=v, roll=rng.uniform(0, 6),
                         width=rng.uniform(100, 5000), nx=int(rng.integers(1, 3000)),
                         ny=int(rng.integers(1, 3000))))
    for cam in cams:
        b0, b1, mp, ps = camera_basis(cam)
        o = orc.Camera(cam["position"], cam["look_at"], cam["roll"], cam["width"], cam["nx"],
                       cam["ny"])
        assert np.array_equal(b0, o.b0) and np.array_equal(b1, o.b1)
        assert np.array_equal(mp, o.min_position) and ps == o.pixel_size


def test_camera_rejects_degenerate():
    from illcrawl_amd.backend import camera_basis
    with pytest.raises(N.IllcrawlError):
        camera_basis(dict(position=[0, 0, 0], look_at=[0, 0, 1], width=0.0, nx=4, ny=4))


def test_cpp_header_compiles_standalone(tmp_path):
    """include/illcrawl_backend.hpp is plain C++17 over the C ABI and the HIP
    runtime API: it compiles with the host compiler alone."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "illcrawl_backend.hpp"\n'
                   'int main() { namespace dm = illcrawl::reconstruction_backends::dm;\n'
                   '  illcrawl::reconstruction_backend* b = nullptr; (void)b;\n'
                   '  return sizeof(dm::projective_smoothing_grid) > 0 ? 0 : 1; }\n')
    inc = os.path.join(ROOT, "include")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra",
                        "-D__HIP_PLATFORM_AMD__", "-I", inc, "-I", "/opt/rocm/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_demo_built_and_reports_usage():
    import subprocess
    demo = os.path.join(ROOT, "illcrawl_amd", "lib", "backend_demo")
    if not os.path.exists(demo):
        pytest.skip("backend_demo not built")
    r = subprocess.run([demo], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_gaunt_table_loader(tmp_path):
    from illcrawl_amd import backend
    f = tmp_path / "g.dat"
    f.write_text("#\n# comment\n#\n  1.0e+00  2.0e+00 3.5\n  4.0 5.0 6.0\n")
    t = backend.load_gaunt_table(str(f))
    assert t.shape == (2, 3) and t.dtype == np.float32 and t[1, 2] == 6.0
    f.write_text("1 2 3\n4 5\n")
    with pytest.raises(ValueError):
        backend.load_gaunt_table(str(f))


def test_static_crawl_block_refuses_oversized_blocks_without_a_gpu():
    """ilc_crawl_block_static checks its size limit before any HIP call."""
    vol = N.VolumeC((ctypes.c_double * 3)(0, 0, 0), (ctypes.c_double * 3)(1, 1, 1),
                    (ctypes.c_double * 3)(10, 10, 10), 0.0)
    m = ctypes.c_uint64(7)
    st = N.lib().ilc_crawl_block_static(ctypes.byref(vol), None, (1 << 29) + 1, None, 0, None,
                                        None, None, ctypes.byref(m), None)
    assert st == N.ILC_ERANGE
    assert b"allocation-free maximum" in N.lib().ilc_last_error()
    assert m.value == 7
