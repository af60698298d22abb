"""GPU adjoint (ilc_psg_backward and the torch.autograd binding) against FP64.

Bars: every gradient component within 1e-4 of the FP64 value, relative to
that component's largest magnitude over the particles checked (the gradients
are signed sums, so a per-particle relative bar is meaningless where a sum
cancels); the same against FP64 central differences of an FP64 forward on
256 random particles.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GRAD_TOL = 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from illcrawl_amd import backend
    return backend


def rel_to_max(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.max(np.abs(ref))
    return float(np.max(np.abs(a - ref)) / scale) if scale > 0 else float(np.max(np.abs(a)))


def projected_scene(n, seed, nx, ny, width=1000.0):
    """Projected-frame particles (x, y, 0, w) and h around a lattice of
    nx x ny points (centres of pixels of size width/nx), with halo-like
    clustering, some w = 0 and some particles far outside the image."""
    rng = np.random.default_rng(seed)
    ps = width / nx
    x = rng.uniform(-0.2 * width, 1.2 * width, n)
    y = rng.uniform(-0.2 * width * ny / nx, 1.2 * width * ny / nx, n)
    k = n // 3
    x[:k] = rng.normal(0.45 * width, 0.08 * width, k)
    y[:k] = rng.normal(0.5 * width * ny / nx, 0.08 * width, k)
    h = np.exp(rng.uniform(np.log(0.7 * ps), np.log(40 * ps), n))
    w = rng.lognormal(0.0, 0.3, n)
    w[rng.random(n) < 0.05] = 0.0
    p4 = np.stack([x, y, np.zeros(n), w], 1).astype(np.float32)
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([(ix.ravel() + 0.5) * ps, (iy.ravel() + 0.5) * ps,
                    np.zeros(nx * ny), np.zeros(nx * ny)], 1).astype(np.float32)
    return p4, h.astype(np.float32), pts


@pytest.mark.parametrize("kernel,kind,seed", [("quartic", 1, 0), ("quartic_ref", 0, 1),
                                              ("cubic", 2, 2)])
def test_backward_vs_fp64_oracle(B, orc, kernel, kind, seed):
    nx, ny = 96, 80
    say(f"case: backward vs oracle kernel={kernel} seed={seed} n=3000 points={nx}x{ny}")
    p4, h, pts = projected_scene(3000, seed, nx, ny)
    rng = np.random.default_rng(seed + 10)
    G = rng.normal(size=nx * ny).astype(np.float32)
    scale = 1.7
    b = B.ProjectiveSmoothingGrid(mode="exact", kernel=kernel, target_particles_per_cell=50)
    b.init_backend(p4.shape[0])
    b.setup_projected_particles(p4, h)
    b.setup_evaluation_points(pts, layout_nx=nx)
    gw, gh, gx, gy = (t.cpu().numpy() for t in b.backward(torch.from_numpy(G).cuda(), scale))
    torch.cuda.synchronize()
    ref = orc.backward_exact(p4, h, pts, G.astype(np.float64) * np.float64(np.float32(scale)),
                             kind)
    # the cubic's shape and derivative are tables (4096 nodes, linear): its
    # gradients carry their interpolation error, like its forward (2e-5)
    tol = 3e-4 if kernel == "cubic" else GRAD_TOL
    for name, got, r in zip(("w", "h", "x", "y"), (gw, gh, gx, gy), ref):
        e = rel_to_max(got, r)
        say(f"  dL/d{name}: rel-to-max err {e:.2e}")
        assert e < tol, name
    # w = 0 particles still get dL/dw, and nothing else
    z = p4[:, 3] == 0
    assert np.all(gh[z] == 0) and np.all(gx[z] == 0) and np.any(gw[z] != 0)


def test_backward_repeats_and_sparse_gradient(B, orc):
    nx, ny = 64, 64
    say("case: backward twice, gradient nonzero on one 8x8 block only")
    p4, h, pts = projected_scene(2000, 5, nx, ny)
    G = np.zeros((ny, nx), np.float32)
    G[40:48, 8:16] = 1.0
    b = B.ProjectiveSmoothingGrid(mode="exact", kernel="quartic")
    b.init_backend(p4.shape[0])
    b.setup_projected_particles(p4, h)
    b.setup_evaluation_points(pts, layout_nx=nx)
    g1 = [t.cpu().numpy() for t in b.backward(torch.from_numpy(G.ravel()).cuda())]
    g2 = [t.cpu().numpy() for t in b.backward(torch.from_numpy(G.ravel()).cuda())]
    ref = orc.backward_exact(p4, h, pts, G.ravel().astype(np.float64), 1)
    for a, c, r in zip(g1, g2, ref):
        assert rel_to_max(a, r) < GRAD_TOL
        assert rel_to_max(c, r) < GRAD_TOL
    # particles whose support misses the block get exactly 0
    far = ref[0] == 0
    assert far.any() and np.all(g1[0][far] == 0)


def test_backward_errors(B):
    say("case: backward error paths")
    from illcrawl_amd import IllcrawlError
    p4, h, pts = projected_scene(100, 2, 32, 32)
    G = torch.zeros(32 * 32, device="cuda")
    for mode, kernel in (("reference", "quartic_ref"),):
        b = B.ProjectiveSmoothingGrid(mode=mode, kernel=kernel)
        b.init_backend(p4.shape[0])
        b.setup_projected_particles(p4, h)
        b.setup_evaluation_points(pts, layout_nx=32)
        with pytest.raises(IllcrawlError):
            b.backward(G)
    b = B.ProjectiveSmoothingGrid(mode="exact", kernel="quartic")
    with pytest.raises(IllcrawlError):
        b.backward(torch.zeros(0, device="cuda"))
    b.init_backend(p4.shape[0])
    b.setup_projected_particles(p4, h)
    b.setup_evaluation_points(pts, layout_nx=32)
    with pytest.raises(ValueError):
        b.backward(torch.zeros(5, device="cuda"))


def _single_image(xy, w, h, px, py, K, scale):
    """FP64 image of one projected particle on the points (px, py)."""
    r2 = (px - xy[0]) ** 2 + (py - xy[1]) ** 2
    s = np.maximum(1.0 - r2 / (h * h), 0.0)
    return scale * w * K * s ** 2.5 / (h * h)


def test_autograd_vs_fp64_central_differences(B, orc):
    """L = 1/2 |I(m, h, x) - I(m', h, x)|^2 with m' = m * lognormal(0.05); the
    binding's gradients w.r.t. m, h and world x, y, z against FP64 central
    differences (relative step 1e-6) on 256 random particles."""
    from illcrawl_amd import synth
    from illcrawl_amd.autograd import projected_image
    nx = 64
    say(f"case: autograd vs central differences n=4096 {nx}x{nx}")
    snap = synth.make_snapshot(4096, 3, 10000.0, n_halos=1, rs_range=(300.0, 300.0))
    v = synth.centred_view(snap.box, look_at=(0.3, -0.2, 0.93), width=3000.0, nx=nx, roll=0.4)
    pos64 = snap.pos.astype(np.float32)
    m = snap.mass.astype(np.float32)
    hs = snap.hsml.astype(np.float32)
    rng = np.random.default_rng(7)
    m_t = (m * rng.lognormal(0.0, 0.05, m.shape[0])).astype(np.float32)
    dev = torch.device("cuda", 0)
    pos = torch.from_numpy(pos64).to(dev).requires_grad_(True)
    w = torch.from_numpy(m).to(dev).requires_grad_(True)
    h = torch.from_numpy(hs).to(dev).requires_grad_(True)
    scale = 1.0e4
    with torch.no_grad():
        target = projected_image(pos, torch.from_numpy(m_t).to(dev), h, v, scale=scale)
    img = projected_image(pos, w, h, v, scale=scale)
    loss = 0.5 * ((img.double() - target.double()) ** 2).sum()
    loss.backward()
    g_pos = pos.grad.cpu().numpy().astype(np.float64)
    g_w = w.grad.cpu().numpy().astype(np.float64)
    g_h = h.grad.cpu().numpy().astype(np.float64)

    # FP64 forward: the projection the library applies (FP32 matrix, FP64 math)
    minv, ctr = B.camera_projection(v)
    minv = minv.astype(np.float64)
    ctr = ctr.astype(np.float64)
    pts = B.camera_pixel_points(v).cpu().numpy()
    px, py = pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64)
    K = 3.5 / np.pi
    proj = ((pos64.astype(np.float64) - ctr) @ minv.T)
    depth_ok = (proj[:, 2] >= 0) & (proj[:, 2] <= v["depth"])
    assert depth_ok.all()
    # the FP64 single-particle image agrees with the oracle's FP64 gather
    p4 = np.concatenate([proj[:, :2], np.zeros((len(m), 1)), m[:, None]], 1).astype(np.float32)
    I64 = orc.gather_exact(p4, hs, pts, kind=1) * scale
    j0 = int(np.argmax(m))
    one = _single_image(p4[j0, :2].astype(np.float64), float(m[j0]), float(hs[j0]), px, py, K,
                        scale)
    ref_one = orc.gather_exact(p4[j0:j0 + 1], hs[j0:j0 + 1], pts, kind=1) * scale
    assert np.allclose(one, ref_one, rtol=1e-12, atol=1e-300)
    # the residual the backward receives (the GPU image minus the target);
    # the central differences then test dI/dtheta alone
    R = img.detach().cpu().numpy().ravel().astype(np.float64) - \
        target.cpu().numpy().ravel().astype(np.float64)
    assert np.allclose(img.detach().cpu().numpy().ravel(), I64, rtol=1e-5,
                       atol=1e-6 * float(np.abs(I64).max()))

    # 256 particles whose support reaches the image
    reach = np.array([np.any(_single_image(p4[j, :2].astype(np.float64), 1.0, float(hs[j]), px,
                                           py, K, 1.0) > 0) for j in range(len(m))])
    cand = np.flatnonzero(reach)
    sel = rng.choice(cand, size=min(256, cand.size), replace=False)
    fd = np.zeros((sel.size, 5))
    for a, j in enumerate(sel):
        base = _single_image(proj[j, :2], float(m[j]), float(hs[j]), px, py, K, scale)
        theta = [float(m[j]), float(hs[j]), *pos64[j].astype(np.float64)]
        for q in range(5):
            eps = 1e-6 * max(abs(theta[q]), 1.0 if q >= 2 else 0.0)
            vals = []
            for sgn in (1.0, -1.0):
                t = list(theta)
                t[q] += sgn * eps
                xy = ((np.array(t[2:]) - ctr) @ minv.T)[:2] if q >= 2 else proj[j, :2]
                img_j = _single_image(xy, t[0], t[1], px, py, K, scale)
                vals.append(0.5 * np.sum((R + img_j - base) ** 2))
            fd[a, q] = (vals[0] - vals[1]) / (2 * eps)
    got = np.stack([g_w[sel], g_h[sel], g_pos[sel, 0], g_pos[sel, 1], g_pos[sel, 2]], 1)
    for q, name in enumerate(("m", "h", "x", "y", "z")):
        e = rel_to_max(got[:, q], fd[:, q])
        say(f"  dL/d{name}: rel-to-max err vs central differences {e:.2e}")
        assert e < GRAD_TOL, name


def test_autograd_depth_mask(B):
    """Particles the depth mask removes get dL/dm = 0 and no other gradient."""
    from illcrawl_amd.autograd import projected_image
    say("case: autograd depth mask")
    rng = np.random.default_rng(11)
    n = 512
    pos = np.stack([rng.uniform(0, 100, n), rng.uniform(0, 100, n),
                    rng.uniform(-50, 150, n)], 1).astype(np.float32)
    cam = dict(position=(50.0, 50.0, 0.0), look_at=(0.0, 0.0, 1.0), roll=0.0, width=100.0,
               nx=32, ny=32, depth=100.0)
    dev = torch.device("cuda", 0)
    P = torch.from_numpy(pos).to(dev).requires_grad_(True)
    W = torch.ones(n, device=dev, requires_grad=True)
    H = torch.full((n,), 8.0, device=dev, requires_grad=True)
    img = projected_image(P, W, H, cam)
    img.sum().backward()
    z = pos[:, 2]
    out = (z < 0) | (z > 100)
    gw = W.grad.cpu().numpy()
    assert out.any() and np.all(gw[out] == 0) and np.all(gw[~out] > 0)
    assert np.all(H.grad.cpu().numpy()[out] == 0)
    assert np.all(P.grad.cpu().numpy()[out] == 0)
