"""The C++ backend classes (include/illcrawl_backend.hpp) on the GPU.

tests/cpp/backend_demo.cpp renders through dm::projective_smoothing over
dm::projective_smoothing_grid (the reference's class stack); its image must
be bit-identical to the Python mirror's on the same inputs and within the
parity bars of the oracle's reference-semantics gather; in exact mode its
backward must match the FP64 oracle adjoint.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "illcrawl_amd", "lib", "backend_demo")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode,kernel", [(0, 0), (1, 1)])
def test_cpp_backend_stack(orc, tmp_path, mode, kernel):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(DEMO):
        pytest.fail("backend_demo is not built (run __graft_entry__.build())")
    from illcrawl_amd import backend as B, synth
    say(f"case: C++ backend stack mode={mode} kernel={kernel}")
    snap = synth.make_snapshot(20000, 4, 10000.0, n_halos=1, rs_range=(300.0, 300.0))
    v = synth.centred_view(snap.box, look_at=(0.2, 0.1, 0.97), width=3000.0, nx=80, ny=64,
                           roll=0.3)
    p4 = snap.particles4().astype(np.float32)
    hs = snap.hsml.astype(np.float32)
    p4.tofile(tmp_path / "particles.f32")
    hs.tofile(tmp_path / "hsml.f32")
    args = [DEMO, str(tmp_path), str(v["nx"]), str(v["ny"]), repr(v["width"]), repr(v["depth"]),
            *[repr(float(x)) for x in v["position"]], *[repr(float(x)) for x in v["look_at"]],
            repr(v["roll"]), str(mode), str(kernel)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    say(r.stdout + r.stderr)
    assert r.returncode == 0
    assert "error-path ok" in r.stdout
    img = np.fromfile(tmp_path / "image.f32", dtype=np.float32)

    # the Python mirror on the same inputs
    names = ("reference", "exact")
    kernels = ("quartic_ref", "quartic", "cubic")
    ps = B.ProjectiveSmoothing(v, v["depth"], 1.0,
                               B.ProjectiveSmoothingGrid(mode=names[mode], kernel=kernels[kernel]))
    ps.init_backend(p4.shape[0])
    ps.setup_particles(p4, [hs])
    ps.setup_evaluation_points()
    ps.run()
    py_img = ps.retrieve_results().cpu().numpy().ravel()
    assert np.array_equal(img, py_img)

    # the oracle
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    pp = cam.project_particles(p4, v["depth"])
    pts = cam.pixel_points()
    if mode == 0:
        ref = orc.Grid(pp, hs).gather_ref(pts)
    else:
        ref = orc.gather_exact(pp, hs, pts, kind=kernel)
    m = ref > 1e-6 * ref.max()
    linf = float(np.max(np.abs(img[m] - ref[m]) / ref[m]))
    ssum = abs(float(img.astype(np.float64).sum()) - ref.sum()) / ref.sum()
    say(f"  linf {linf:.2e} sum {ssum:.2e}")
    assert linf < 1e-5 and ssum < 1e-6

    if mode == 1:
        g = np.fromfile(tmp_path / "grad.f32", dtype=np.float32).reshape(4, -1)
        refg = orc.backward_exact(pp, hs, pts, np.ones(pts.shape[0]), kernel)
        for k in range(4):
            e = np.max(np.abs(g[k] - refg[k])) / np.max(np.abs(refg[k]))
            say(f"  grad {k}: rel-to-max {e:.2e}")
            assert e < 1e-4
