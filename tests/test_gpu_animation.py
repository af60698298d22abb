"""Animations through the batched launch: the frames of a camera stepper
(illcrawl_amd/animation.py) rendered as one batch equal the single-view
renders of the same frames (to the grouping of the FP32 partial sums: a
batched view's grid keeps only the particles that reach its screen), and a
half turn about the screen's up axis mirrors the image."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from illcrawl_amd import backend
    return backend


@pytest.fixture(scope="module")
def snap():
    from illcrawl_amd import synth
    return synth.make_snapshot(1 << 15, 9, 10000.0, n_halos=2, rs_range=(200.0, 400.0),
                               gas=True)


def single_view(B, snap, v, mode):
    p4 = torch.empty((snap.n, 4), dtype=torch.float32, device="cuda")
    p4[:, :3] = torch.from_numpy(snap.pos).cuda()
    p4[:, 3] = B.quantity_scaled(snap.mass, 1.0)
    proj = B.project_particles(v, p4)
    b = B.ProjectiveSmoothingGrid(mode=mode)
    b.init_backend(snap.n)
    b.setup_projected_particles(proj, snap.hsml)
    b.setup_evaluation_points(B.camera_pixel_points(v), layout_nx=v["nx"])
    b.run()
    return b.retrieve_results_f64().cpu().numpy().reshape(v["ny"], v["nx"])


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_animation_frames_equal_single_views(B, snap, mode):
    from illcrawl_amd import animation as A, synth
    v0 = synth.centred_view(snap.box, look_at=(0.0, 0.6, 0.8), width=3000.0, nx=64)
    v0["quantity"] = 0
    st = A.illcrawl_render_animation_stepper(np.full(3, 0.5 * snap.box), v0, 1.5)
    views = A.animation_views(st, 6)
    say(f"case: 6 animation frames, {mode}")
    r = B.BatchRenderer(mode=mode)
    r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
    frames = A.render_animation(r, views, f64=True)
    assert frames.shape == (6, 64, 64)
    for k, v in enumerate(views):
        ref = single_view(B, snap, v, mode)
        assert np.array_equal(frames[k] == 0, ref == 0), f"frame {k}"
        m = ref != 0
        assert np.max(np.abs(frames[k][m] - ref[m]) / ref[m]) < 1e-6, f"frame {k}"
    assert frames.max() > 0


def test_half_turn_mirrors_the_image(B, snap):
    from illcrawl_amd import animation as A, synth
    v0 = synth.centred_view(snap.box, look_at=(0.0, 0.6, 0.8), width=3000.0, nx=64)
    v0["quantity"] = 0
    _, b1, _, _ = B.camera_basis(v0)
    st = A.RotationAroundPoint(np.full(3, 0.5 * snap.box), b1, v0, 360.0)
    views = A.animation_views(st, 2)          # 0 and 180 degrees about the up axis
    r = B.BatchRenderer(mode="exact")
    r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
    f = A.render_animation(r, views, f64=True)
    # pixel corner x of the turned frame sits where corner nx - x of frame 0 does
    a, b = f[1][:, 1:], f[0][:, :0:-1]
    m = b > 1e-4 * b.max()
    assert m.sum() > 1000
    rel = np.max(np.abs(a[m] - b[m]) / b[m])
    say(f"case: half turn, rel L-inf {rel:.2e}")
    assert rel < 1e-4
