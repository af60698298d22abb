"""The line-of-sight kernel clip of the batch (ilc_batch_set_depth_clip)
against the oracle's FP64 quadrature of the clipped 3-D quartic
(orc_gather_clip): slabs of a tomography stack, a rotated view with a depth
cut, and the stack's slabs adding up to one range over the whole stack.
Bars as exact mode's: relative L-inf 1e-5 over pixels above 1e-6 of the
maximum, 1e-6 on the image sum."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

LINF_TOL = 1e-5
SUM_TOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from illcrawl_amd import backend
    return backend


@pytest.fixture(scope="module")
def snap():
    from illcrawl_amd import synth
    return synth.make_snapshot(1 << 15, 17, 10000.0, n_halos=2, rs_range=(200.0, 400.0),
                               gas=True)


def errors(got, ref):
    m = ref > 1e-6 * ref.max()
    linf = float(np.max(np.abs(got[m] - ref[m]) / ref[m]))
    return linf, float(abs(got.sum() - ref.sum()) / ref.sum())


def oracle_clip(orc, snap, v, kind):
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    pz = cam.project_particles_depth(snap.particles4())
    return orc.gather_clip(pz, snap.hsml, cam.pixel_points(), v["depth"], kind)


def clip_batch(B, snap, views, kernel="quartic_ref"):
    r = B.BatchRenderer(mode="exact", kernel=kernel)
    r.set_particles(snap.pos, snap.mass, snap.hsml)
    r.set_depth_clip(True)
    r.set_views(views)
    r.render()
    return r


@pytest.mark.parametrize("kernel,kind", [("quartic_ref", 0), ("quartic", 1)])
def test_clipped_slabs_vs_oracle(B, orc, snap, kernel, kind):
    from illcrawl_amd import synth
    say(f"case: clipped slabs, {kernel}")
    views = synth.slab_views(snap.box, n_slabs=6, thickness=500.0, nx=80, width=4000.0,
                             quantity=0)
    r = clip_batch(B, snap, views, kernel)
    for k, v in enumerate(views):
        got = r.image_f64(k).cpu().numpy().reshape(-1)
        ref = oracle_clip(orc, snap, v, kind)
        linf, s = errors(got, ref)
        say(f"  slab {k}: linf {linf:.2e} sum {s:.2e}")
        assert linf < LINF_TOL and s < SUM_TOL, k


def test_clipped_slabs_add_up_to_the_stack(B, orc, snap):
    """By linearity the slabs' clipped images sum to one range spanning the
    whole stack (the oracle's)."""
    from illcrawl_amd import synth
    say("case: clipped slabs sum to the stack")
    n, th = 6, 500.0
    views = synth.slab_views(snap.box, n_slabs=n, thickness=th, nx=80, width=4000.0,
                             quantity=0)
    r = clip_batch(B, snap, views)
    tot = sum(r.image_f64(k).cpu().numpy().reshape(-1) for k in range(n))
    whole = dict(views[0], depth=n * th)
    ref = oracle_clip(orc, snap, whole, 0)
    linf, s = errors(tot, ref)
    say(f"  sum of slabs vs whole stack: linf {linf:.2e} sum {s:.2e}")
    assert linf < LINF_TOL and s < SUM_TOL
    # the centre mask puts each particle in one slab whole: a different image
    rm = B.BatchRenderer(mode="exact")
    rm.set_particles(snap.pos, snap.mass, snap.hsml)
    rm.set_views(views)
    rm.render()
    mask_img = rm.image_f64(2).cpu().numpy().reshape(-1)
    clip_img = r.image_f64(2).cpu().numpy().reshape(-1)
    assert float(np.max(np.abs(mask_img - clip_img))) > 1e-3 * float(clip_img.max())


def test_clipped_rotated_view_vs_oracle(B, orc, snap):
    """A rotated, windowed view whose depth range ends inside the box, and an
    X-ray view of the same camera."""
    from illcrawl_amd import synth
    say("case: clipped rotated view")
    v = synth.centred_view(snap.box, look_at=(0.4, -0.3, 0.85), width=3000.0, nx=72, roll=0.7)
    v = dict(v, depth=0.55 * v["depth"], quantity=0)
    r = clip_batch(B, snap, [v])
    got = r.image_f64(0).cpu().numpy().reshape(-1)
    ref = oracle_clip(orc, snap, v, 0)
    linf, s = errors(got, ref)
    say(f"  rotated view: linf {linf:.2e} sum {s:.2e}")
    assert linf < LINF_TOL and s < SUM_TOL


def test_clip_error_paths(B, snap):
    from illcrawl_amd import IllcrawlError, synth
    say("case: clip error paths")
    r = B.BatchRenderer(mode="reference")
    with pytest.raises(IllcrawlError):
        r.set_depth_clip(True)
    r = B.BatchRenderer(mode="exact", kernel="cubic")
    with pytest.raises(IllcrawlError):
        r.set_depth_clip(True)
    views = synth.slab_views(snap.box, n_slabs=2, thickness=500.0, nx=32, quantity=0)
    r = clip_batch(B, snap, views)
    with pytest.raises(IllcrawlError):
        r.render(plain=True)
    g = [torch.zeros(32 * 32, dtype=torch.float32, device="cuda") for _ in views]
    with pytest.raises(IllcrawlError):
        r.backward(g)
    # the centre mask again after switching the clip off
    r.set_depth_clip(False)
    r.render(plain=True)
