"""Pins for the CPU oracle (oracle/oracle.c).

Each test fixes an oracle function with a check that a plausible mistake
(dropped term, wrong sign or index, transposed operand, wrong factor) fails:
closed forms, known answers computed by hand, brute force in numpy, and
invariants.  The reference holds no tests or golden vectors for this path
(its only data file on the path family is the van Hoof Gaunt table), so these
are the pins.
"""
import math

import numpy as np
import pytest


def w2d_ref_closed(r, h):
    """The reference's projected quartic in closed form: 7/(pi h^2)(1-q^2)^2.5
    (its float constants 2*1.04445*1.066667 = 2.228161 ~ 7/pi)."""
    q2 = (r / h) ** 2
    return 2.0 * 1.04445 * 1.066667 / (h * h) * max(1.0 - q2, 0.0) ** 2.5


# ---------------------------------------------------------------- camera

def test_camera_look_z_closed_form(orc):
    cam = orc.Camera([10.0, 20.0, -5.0], [0.0, 0.0, 1.0], 0.0, 100.0, 8, 4)
    assert np.allclose(cam.b0, [1, 0, 0]) and np.allclose(cam.b1, [0, 1, 0])
    assert cam.pixel_size == pytest.approx(12.5)
    # lower-left corner: pos - (nx/2) ps b0 - (ny/2) ps b1
    assert np.allclose(cam.min_position, [10 - 4 * 12.5, 20 - 2 * 12.5, -5.0])


def test_camera_random_orthonormal_and_roll(orc):
    rng = np.random.default_rng(5)
    for _ in range(20):
        la = rng.normal(size=3)
        la /= np.linalg.norm(la)
        c0 = orc.Camera([0, 0, 0], la, 0.0, 1.0, 2, 2)
        a = rng.uniform(0, 2 * math.pi)
        c1 = orc.Camera([0, 0, 0], la, a, 1.0, 2, 2)
        for c in (c0, c1):
            assert abs(np.dot(c.b0, c.b1)) < 1e-12
            assert abs(np.linalg.norm(c.b0) - 1) < 1e-12
            assert np.allclose(np.cross(c.b0, c.b1), la, atol=1e-12)
        # roll by a about look_at: b0(a) = cos a b0 + sin a b1
        assert np.allclose(c1.b0, math.cos(a) * c0.b0 + math.sin(a) * c0.b1, atol=1e-12)


def test_projection_matrix_is_transpose_for_orthonormal(orc):
    rng = np.random.default_rng(6)
    for _ in range(10):
        la = rng.normal(size=3)
        la /= np.linalg.norm(la)
        cam = orc.Camera([0, 0, 0], la, rng.uniform(0, 6), 1.0, 2, 2)
        m = cam.projection_matrix().reshape(3, 3).astype(np.float64)
        M = np.stack([cam.b0, cam.b1, la], axis=1)
        assert np.allclose(m, M.T, atol=2e-6)


def test_project_particles_depth_mask(orc):
    cam = orc.Camera([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 0.0, 10.0, 2, 2)
    p = np.array([[1, 2, -1, 5], [1, 2, 0, 5], [1, 2, 50, 5], [1, 2, 100, 5],
                  [1, 2, 100.5, 5], [3, -4, 20, 7]], np.float32)
    out = cam.project_particles(p, 100.0)
    assert np.array_equal(out[:, 3], [0, 5, 5, 5, 0, 7])
    assert np.all(out[:, 2] == 0)
    assert np.array_equal(out[:, :2], p[:, :2])
    # rotated camera: camera-frame coords are (b0, b1) components of x - pos
    la = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    cam = orc.Camera([0.0, 0.0, 0.0], la, 0.3, 10.0, 2, 2)
    x = np.array([[3.0, 1.0, 2.0, 1.0]], np.float32)
    out = cam.project_particles(x, 100.0)
    exp = [np.dot(x[0, :3], cam.b0), np.dot(x[0, :3], cam.b1)]
    assert np.allclose(out[0, :2], exp, atol=1e-5)


def test_pixel_points_are_corners(orc):
    cam = orc.Camera([100.0, 200.0, 0.0], [0.0, 0.0, 1.0], 0.0, 64.0, 8, 4)
    pts = cam.pixel_points()
    ps = 8.0
    ix, iy = np.meshgrid(np.arange(8), np.arange(4))
    # corner convention: (ix, iy) -> min + ix ps b0 + iy ps b1; camera frame
    # origin at the screen centre -> x = -W/2 + ix ps, y = -(ny/2) ps + iy ps
    assert np.array_equal(pts[:, 0], (-32.0 + ps * ix).reshape(-1).astype(np.float32))
    assert np.array_equal(pts[:, 1], (-16.0 + ps * iy).reshape(-1).astype(np.float32))


# ---------------------------------------------------------------- grid

def test_grid_params_hand_computed(orc):
    p = np.array([[0, 0, 0, 1], [10, 20, 0, 1]], np.float32)
    g = orc.Grid(p, np.ones(2, np.float32), target=1)
    ext = np.array([10.2, 20.2, 0.2])
    w = (1.0 / 2.0 * (ext[0] * ext[1] * ext[2])) ** (1.0 / 3.0)
    assert list(g.nc) == [math.ceil(10.2 / w), math.ceil(20.2 / w), 1] == [4, 8, 1]
    assert g.cs[0] == np.float32(w)
    assert np.array_equal(g.minc, np.float32([-0.1, -0.1, -0.1]))


def _random_scene(n, seed, box=100.0, z0=True):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), np.float32)
    p[:, :2] = rng.uniform(-box / 2, box / 2, (n, 2))
    if not z0:
        p[:, 2] = rng.uniform(-box / 2, box / 2, n)
    p[:, 3] = rng.lognormal(0, 0.5, n)
    # clustered half with duplicates to create multi-particle cells
    p[: n // 4, :2] = rng.normal(0, 2.0, (n // 4, 2))
    p[n // 4: n // 4 + 8, :2] = p[0, :2]
    h = rng.uniform(0.5, 6.0, n).astype(np.float32)
    return p, h


@pytest.mark.parametrize("z0", [True, False])
def test_grid_build_against_numpy(orc, z0):
    p, h = _random_scene(3000, 1, z0=z0)
    g = orc.Grid(p, h, target=20)
    # keys: float32 floor((p - min)/w), key = x + y nx + z nx ny
    idx = np.floor((p[:, :3] - g.minc) / g.cs).astype(np.int64)
    nx, ny, nz = (int(v) for v in g.nc)
    keys = idx[:, 0] + idx[:, 1] * nx + idx[:, 2] * nx * ny
    assert np.array_equal(g.keys(p).astype(np.int64), keys)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(g.perm.astype(np.int64), order)
    assert np.array_equal(g.sorted_p4, p[order])
    assert np.array_equal(g.sorted_h, h[order])
    cnt = np.bincount(keys, minlength=nx * ny * nz)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    exp_cells = np.where(cnt[:, None] > 0, np.stack([start, start + cnt], 1), 0)
    assert np.array_equal(g.cells, exp_cells)
    mx = np.zeros(nx * ny * nz, np.float32)
    np.maximum.at(mx, keys, h)
    assert np.array_equal(g.maxh_cell, mx)
    assert g.max_h == h.max()


def test_grid_key_out_of_range_is_an_error(orc):
    p = np.array([[0, 0, 0, 1], [1, 1, 0, 1]], np.float32)
    with pytest.raises(ValueError):
        orc.Grid(p, np.ones(2, np.float32),
                 params=(np.array([1, 1, 1], np.int32), np.zeros(3, np.float32),
                         np.full(3, 0.5, np.float32)))


# ---------------------------------------------------------------- kernel

def test_quartic_closed_forms(orc):
    for h in (0.3, 1.0, 7.5, 2500.0):
        # 7/(pi h^2) up to the reference's rounded constants (1.04445, 1.066667)
        assert orc.quartic_projection(0.0, h) == pytest.approx(7 / (math.pi * h * h), rel=1e-5)
        assert orc.quartic_projection(0.0, h) == pytest.approx(
            2 * 1.04445 * 1.066667 / (h * h), rel=1e-6)
        assert orc.quartic_projection(h, h) == 0.0
        assert orc.quartic_projection(1.5 * h, h) == 0.0
        for q in (0.1, 0.5, 0.9):
            assert orc.quartic_projection(q * h, h) == pytest.approx(w2d_ref_closed(q * h, h),
                                                                     rel=2e-5)


@pytest.mark.parametrize("kind,integral", [(0, 2.0), (1, 1.0), (2, 1.0)])
def test_kernel_plane_integral(orc, kind, integral):
    h = 3.0
    n = 241
    xs = np.linspace(-7.2, 7.2, n)
    X, Y = np.meshgrid(xs, xs)
    pts = np.zeros((n * n, 4), np.float32)
    pts[:, 0] = X.reshape(-1)
    pts[:, 1] = Y.reshape(-1)
    p = np.array([[0.013, -0.021, 0, 1.0]], np.float32)
    v = orc.gather_exact(p, np.array([h], np.float32), pts, kind)
    dA = (xs[1] - xs[0]) ** 2
    assert v.sum() * dA == pytest.approx(integral, rel=2e-4)


def test_cubic_projected_closed_forms(orc):
    assert orc.cubic_projected_unit(0.0) == pytest.approx(1.5 / math.pi, rel=1e-12)
    assert orc.cubic_projected_unit(2.0) == 0.0
    # b in (1,2): only the outer piece, closed form 1/(4pi) * 2 int (2-q)^3 dz
    b = 1.5
    zs = np.linspace(0, math.sqrt(4 - b * b), 200001)
    q = np.sqrt(b * b + zs * zs)
    f = 0.25 * (2 - q) ** 3 / math.pi
    ref = 2 * np.trapezoid(f, zs)
    assert orc.cubic_projected_unit(b) == pytest.approx(ref, rel=1e-8)


# ---------------------------------------------------------------- gather

def test_gather_ref_uniform_h_equals_exact(orc):
    """With one smoothing length everywhere the F3 radius (own cell's max h
    + cell radius) is conservative for points in non-empty cells, so the
    reference gather includes every particle within h there.  In an empty
    cell the radius is the cell radius alone: particles between it and h
    are dropped."""
    rng = np.random.default_rng(2)
    n = 2000
    p = np.zeros((n, 4), np.float32)
    p[:, :2] = rng.uniform(-50, 50, (n, 2))
    p[:, 3] = rng.lognormal(0, 0.5, n)
    p[::7, 3] = 0.0  # depth-masked particles contribute nothing
    h = np.full(n, 6.0, np.float32)
    g = orc.Grid(p, h, target=100)
    pts = np.zeros((800, 4), np.float32)
    pts[:, :2] = rng.uniform(-48, 48, (800, 2))
    ref = g.gather_ref(pts, acc_mode=1)
    ex = orc.gather_exact(p, h, pts, 0)
    own = g.maxh_cell[g.keys(pts).astype(np.int64)]
    full = own > 0
    assert full.sum() > 400 and (~full).sum() >= 5
    m = full & (ex > 1e-6 * ex.max())
    assert np.max(np.abs(ref[m] - ex[m]) / ex[m]) < 5e-6
    assert np.all(ref[ex == 0] == 0)
    # empty own cell: radius = cell radius < h -> never more than exact
    radius = 0.5 * math.sqrt(3) * float(g.cs[0])
    assert radius < 6.0
    e = ~full
    assert np.all(ref[e] <= ex[e] * (1 + 1e-6))
    assert np.any(ref[e] < ex[e] * (1 - 1e-3))


def test_gather_ref_f3_drop_known_answer(orc):
    """F3: the search radius comes from the max h of the point's own cell.
    A: small-h particle in the point's cell; B: big-h particle 6 units away
    (inside h_B) but beyond max_h[cell(P)] + cell_radius -> dropped for P.
    For Q, in B's cell, the radius is h_B-based, so both A and B count."""
    p = np.array([[0.0, 0.0, 0, 1.0],    # A
                  [6.0, 0.0, 0, 2.0],    # B
                  [20.0, 20.0, 0, 1.0],  # C (bbox)
                  [-20.0, -20.0, 0, 1.0]], np.float32)  # D (bbox)
    h = np.array([0.5, 8.0, 0.5, 0.5], np.float32)
    g = orc.Grid(p, h, params=(np.array([41, 41, 1], np.int32),
                               np.array([-20.1, -20.1, -0.1], np.float32),
                               np.ones(3, np.float32)))
    # P in A's cell (20,20): R = 0.5 + sqrt(3)/2 -> x cells [18, 21]; B is in 26
    # Q in B's cell (26,20): R = min(8 + 0.866, 8) = 8 -> A's cell included
    P = np.array([[0.2, 0.1, 0, 0], [6.2, 0.2, 0, 0]], np.float32)
    got = g.gather_ref(P, acc_mode=1)
    rA = math.hypot(0.2, 0.1)
    assert got[0] == pytest.approx(1.0 * w2d_ref_closed(rA, 0.5), rel=1e-5)
    rBQ = math.hypot(0.2, 0.2)
    assert got[1] == pytest.approx(2.0 * w2d_ref_closed(rBQ, 8.0), rel=1e-5)
    # the exact oracle keeps B for P
    ex = orc.gather_exact(p, h, P, 0)
    rBP = math.hypot(5.8, 0.1)
    assert ex[0] == pytest.approx(w2d_ref_closed(rA, 0.5) + 2.0 * w2d_ref_closed(rBP, 8.0),
                                  rel=1e-6)


def test_gather_ref_f4_alias_known_answer(orc):
    """F4: a point right of the grid (x index = nx) has an unclamped key
    that aliases cell (0, y+1); its radius then comes from that cell's max h
    instead of the overall max."""
    # grid with nx = ny = 2 cells of width ~10 (target chosen accordingly)
    p = np.array([[1.0, 1.0, 0, 1.0],      # E in cell (0,0), small h
                  [1.0, 11.0, 0, 1.0],     # F in cell (0,1), h 1 -> aliased radius
                  [19.0, 19.0, 0, 1.0],    # G cell (1,1), big h (overall max)
                  [19.0, 1.0, 0, 1.0]], np.float32)  # H cell (1,0)
    h = np.array([0.5, 1.0, 30.0, 0.5], np.float32)
    nc = np.array([2, 2, 1], np.int32)
    minc = np.array([0.0, 0.0, -0.1], np.float32)
    cs = np.array([10.0, 10.0, 10.0], np.float32)
    g = orc.Grid(p, h, params=(nc, minc, cs))
    # Q at x index 2 (= nx), y index 0 -> key 2 = cell (0,1) -> R = 1 + 8.66
    Q = np.array([[21.0, 5.0, 0, 0]], np.float32)
    got = g.gather_ref(Q, acc_mode=1)
    # index box [floor((21-9.66)/10), floor((21+9.66)/10)] = [1, 3] in x,
    # [floor((5-9.66)/10), floor(14.66/10)] = [-1, 1] in y -> cells (1,0),(1,1)
    # both within the cull; G (r = hypot(2, 14) < 30) counts, H (r > 0.5) 0
    rG = math.hypot(2.0, 14.0)
    assert got[0] == pytest.approx(w2d_ref_closed(rG, 30.0), rel=1e-5)
    # a point whose key is negative (below the grid) uses the overall max:
    # index box covers every cell -> equals the exact sum
    Qn = np.array([[5.0, -3.0, 0, 0]], np.float32)
    got = g.gather_ref(Qn, acc_mode=1)
    ex = orc.gather_exact(p, h, Qn, 0)
    assert got[0] == pytest.approx(ex[0], rel=1e-5)


def test_gather_ref_fp32_accumulator_is_reference_order(orc):
    """acc_mode 0 accumulates in FP32: its difference from acc_mode 1 is
    bounded by n * 2^-24 relative and is nonzero for long sums."""
    p, h = _random_scene(4000, 3)
    g = orc.Grid(p, h, target=40)
    pts = np.zeros((300, 4), np.float32)
    pts[:, :2] = np.random.default_rng(4).normal(0, 3, (300, 2))
    r32, pairs = g.gather_ref(pts, acc_mode=0, with_pairs=True)
    r64 = g.gather_ref(pts, acc_mode=1)
    m = r64 > 0
    rel = np.abs(r32 - r64)[m] / r64[m]
    assert np.all(rel <= pairs[m] * 2.0 ** -24 + 1e-12)
    assert np.any(rel > 0)
    assert np.all(r32.astype(np.float32) == r32)


def test_exact_counts_and_lattice_count(orc):
    p, h = _random_scene(1500, 7)
    p[::5, 3] = 0
    cam = orc.Camera([0.0, 0.0, -10.0], [0.0, 0.0, 1.0], 0.0, 60.0, 48, 40)
    pts = cam.pixel_points()
    _, cnt = orc.gather_exact(p, h, pts, 0, with_counts=True)
    # brute force in numpy
    d2 = ((pts[:, None, 0].astype(np.float64) - p[None, :, 0]) ** 2
          + (pts[:, None, 1].astype(np.float64) - p[None, :, 1]) ** 2)
    bf = ((d2 < (h.astype(np.float64) ** 2)[None, :]) & (p[None, :, 3] != 0)).sum()
    assert cnt.sum() == bf
    assert orc.count_evals_lattice(p, h, pts, 48, 40) == bf
    _, cnt2 = orc.gather_exact(p, h, pts, 2, with_counts=True)
    assert orc.count_evals_lattice(p, h, pts, 48, 40, 2.0) == cnt2.sum()


# ---------------------------------------------------------------- quantities

def test_tabulated_1d_half_texel_and_border(orc):
    y = [1.0, 2.0, 4.0]
    cases = {0.0: 0.5, 0.5: 1.0, 1.0: 1.5, 1.5: 2.0, 2.5: 4.0, 3.0: 2.0, 3.5: 0.0,
             -0.5: 0.0, 0.25: 0.75, 1.75: 2.5}
    for x, v in cases.items():
        assert orc.tabulated_1d(y, 0.0, 1.0, x) == pytest.approx(v, abs=1e-7), x
    # x_min / dx scaling
    assert orc.tabulated_1d(y, 10.0, 0.5, 10.25) == pytest.approx(1.0)


def test_tabulated_2d_half_texel(orc):
    data = np.array([[1.0, 2.0], [3.0, 5.0]], np.float32)   # (ny, nx)
    assert orc.tabulated_2d(data, 0, 0, 1, 1, 0.5, 0.5) == pytest.approx(1.0)
    assert orc.tabulated_2d(data, 0, 0, 1, 1, 1.5, 1.5) == pytest.approx(5.0)
    assert orc.tabulated_2d(data, 0, 0, 1, 1, 1.0, 0.5) == pytest.approx(1.5)
    assert orc.tabulated_2d(data, 0, 0, 1, 1, 0.5, 1.0) == pytest.approx(2.0)
    assert orc.tabulated_2d(data, 0, 0, 1, 1, 1.0, 1.0) == pytest.approx(11 / 4)
    assert orc.tabulated_2d(data, 0, 0, 1, 1, 0.0, 0.5) == pytest.approx(0.5)


def test_temperature_closed_form(orc):
    for u, xe in ((1e4, 1.1), (350.0, 1.0), (2e6, 1.2)):
        mu = 4.0 / (1 + 3 * 0.76 + 4 * 0.76 * xe)
        assert orc.temperature(u, xe) == pytest.approx(2.0 / 3.0 * u * mu * 121.1475502,
                                                       rel=1e-6)


def test_xray_bolometric_closed_form(orc):
    gy = np.linspace(1.0, 2.0, 11).astype(np.float32)
    x0, dx = np.float32(math.log(1e4)), np.float32(0.5)
    # T at a texel centre (u = i + 0.5) reads gy[i] exactly
    i = 4
    T = np.float32(math.exp(float(x0) + (i + 0.5) * float(dx)))
    rho = np.float32(2e-6)
    xe = np.float32(1.1)
    w = orc.xray_bolometric([rho], [T], [xe], gy, x0, dx)[0]
    exp = (4.046263 * 2e-6 * 1.1) ** 2 * math.sqrt(float(T)) * float(gy[i])
    assert w == pytest.approx(exp, rel=3e-5)


# ---------------------------------------------------------------- adjoint

@pytest.mark.parametrize("kind", [0, 1, 2])
def test_backward_exact_vs_central_differences(orc, kind):
    rng = np.random.default_rng(17)
    n = 40
    p = np.zeros((n, 4), np.float32)
    p[:, :2] = rng.uniform(-3, 3, (n, 2))
    p[:, 3] = rng.lognormal(0, 0.3, n)
    h = rng.uniform(0.8, 2.5, n).astype(np.float32)
    pts = np.zeros((30 * 30, 4), np.float32)
    xs = np.linspace(-4, 4, 30, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs)
    pts[:, 0], pts[:, 1] = X.reshape(-1), Y.reshape(-1)
    G = rng.normal(size=pts.shape[0])
    gw, gh, gx, gy = orc.backward_exact(p, h, pts, G, kind)

    def L(pp, hh):
        return float(np.dot(G, orc.gather_exact(pp, hh, pts, kind)))

    for j in (0, 7, 19, 33):
        # central differences in float32-representable steps (inputs are float32)
        for comp, g in ((3, gw), (0, gx), (1, gy)):
            eps = 1e-3 * max(abs(float(p[j, comp])), 1.0)
            a, b = p.copy(), p.copy()
            a[j, comp] += eps
            b[j, comp] -= eps
            fd = (L(a, h) - L(b, h)) / (float(a[j, comp]) - float(b[j, comp]))
            assert g[j] == pytest.approx(fd, rel=2e-3, abs=1e-6 * abs(g).max())
        eps = 1e-3 * float(h[j])
        ha, hb = h.copy(), h.copy()
        ha[j] += eps
        hb[j] -= eps
        fd = (L(p, ha) - L(p, hb)) / (float(ha[j]) - float(hb[j]))
        assert gh[j] == pytest.approx(fd, rel=2e-3, abs=1e-6 * abs(gh).max())


def test_crawl_block_known_rows(orc):
    """Periodic images moved in by one period per axis and sign, the border
    tolerance, inclusive faces, and input order (reconstructing_data_crawler
    block body)."""
    c, e, L = (50.0, 50.0, 50.0), (20.0, 20.0, 20.0), (100.0, 100.0, 100.0)
    rows = np.array([
        [50, 50, 50],      # inside
        [150, 50, 50],     # x - L inside
        [50, -55, 50],     # y + L = 45 inside
        [50, 50, 141],     # z - L = 41 inside
        [40, 50, 50],      # inside, on the min face
        [60, 60, 60],      # inside, on the max faces
        [39.5, 50, 50],    # just outside
        [50, 50, 61],      # outside (z)
        [161, 50, 50],     # x - L = 61: outside
    ], np.float32)
    h = np.arange(len(rows), dtype=np.float32) + 1
    p4, (hk,), idx = orc.crawl_block(c, e, L, 0.0, rows, [h])
    assert list(idx) == [0, 1, 2, 3, 4, 5]
    assert np.array_equal(p4[:, :3], np.array([[50, 50, 50], [50, 50, 50], [50, 45, 50],
                                               [50, 50, 41], [40, 50, 50], [60, 60, 60]],
                                              np.float32))
    assert np.all(p4[:, 3] == 0) and np.array_equal(hk, h[:6])
    # a border lets a row move into the grown box, but the volume test has no border:
    # 138.5 - 100 = 38.5 is within [40 - 2, 60 + 2] -> moved, then outside the volume
    p4b, _, idxb = orc.crawl_block(c, e, L, 2.0, np.array([[138.5, 50, 50]], np.float32))
    assert len(idxb) == 0
    p4c, _, idxc = orc.crawl_block(c, e, L, 2.0, np.array([[140.0, 50, 50]], np.float32))
    assert list(idxc) == [0] and p4c[0, 0] == 40.0


def test_crawl_block_vs_numpy(orc):
    rng = np.random.default_rng(3)
    n = 20000
    xyz = rng.uniform(-120, 220, (n, 3)).astype(np.float32)
    col = rng.random(n).astype(np.float32)
    c, e, L, bd = (37.5, 61.25, 80.0), (30.0, 45.0, 70.0), (100.0, 100.0, 120.0), 3.0
    p4, (ck,), idx = orc.crawl_block(c, e, L, bd, xyz, [col])
    x = xyz.astype(np.float64)
    c, e, L = np.array(c), np.array(e), np.array(L)
    bmin, bmax = c - 0.5 * e, c + 0.5 * e
    lo, hi = x - L, x + L
    in_lo = (bmin - bd <= lo) & (lo <= bmax + bd)
    in_hi = (bmin - bd <= hi) & (hi <= bmax + bd)
    xc = np.where(in_lo, lo, np.where(in_hi, hi, x))
    keep = np.all((xc >= c - 0.5 * e) & (xc <= c + 0.5 * e), axis=1)
    assert np.array_equal(idx, np.flatnonzero(keep))
    assert np.array_equal(p4[:, :3], xc[keep].astype(np.float32))
    assert np.array_equal(ck, col[keep])


K_B, RYD = 8.61733e-5, 13.605


def _temperature64(u, xe):
    mu = 4.0 / (1.0 + 3 * 0.76 + 4 * 0.76 * xe)
    return (5.0 / 3.0 - 1.0) * u * mu * 121.1475502


def test_xray_spectral_flux_closed_form(orc):
    """One energy bin with a constant Gaunt table g0: 0.002535 dE ion^2 g0
    T^-1/2 exp(-E/kT) / r^2, ion = 4.046263 rho xe (quantities.cl:97-136)."""
    g0 = 1.37
    tab = np.full((146, 81), g0, np.float32)
    rho = np.array([1e4, 3e5, 2e7], np.float32)
    u = np.array([2e4, 8e4, 6e5], np.float32)   # T inside the table's gamma^2 range
    xe = np.array([1.05, 1.16, 1.2], np.float32)
    z, dl, E, dE = 0.05, 1.2e5, 1.3, 0.25
    got = orc.xray_spectral_flux(rho, u, xe, z, dl, tab, E, dE)
    T = _temperature64(u.astype(np.float64), xe.astype(np.float64))
    ion = 4.046263 * rho.astype(np.float64) * xe
    Ez, dEz = (1 + z) * E, (1 + z) * dE
    want = 0.002535 * dEz * ion ** 2 * g0 / np.sqrt(T) * np.exp(-Ez * 1000.0 / (K_B * T)) / dl ** 2
    assert np.allclose(got, want, rtol=2e-5)


def test_xray_gaunt_coordinates(orc):
    """A table linear in its indices, t[j][i] = i + 1000 j, is reproduced
    exactly by the linear sampler inside the table: the flux then reveals the
    sampling position (ln(gamma^2) + 6) / 0.2 + 1000 (ln(u) + 16) / 0.2 - 500.5."""
    j, i = np.meshgrid(np.arange(146), np.arange(81), indexing="ij")
    tab = (i + 1000.0 * j).astype(np.float32)
    rho, u, xe = (np.array([v], np.float32) for v in (2e6, 3e5, 1.1))
    z, dl, E, dE = 0.0, 1.0, 2.0, 1.0
    got = float(orc.xray_spectral_flux(rho, u, xe, z, dl, tab, E, dE)[0])
    T = float(_temperature64(3e5, np.float64(np.float32(1.1))))
    kT = K_B * T
    gx = (np.log(RYD / kT) + 6.0) / 0.2 - 0.5
    gy = (np.log(1000.0 * E / kT) + 16.0) / 0.2 - 0.5
    assert 1 < gx < 79 and 1 < gy < 144
    pref = 0.002535 * (4.046263 * 2e6 * 1.1) ** 2 / np.sqrt(T) * np.exp(-E * 1000.0 / kT)
    assert abs(got / pref - (gx + 1000.0 * gy)) < 2e-3 * (gx + 1000.0 * gy) * 1e-2


def test_xray_flux_band_integral(orc):
    """Constant table: the RKF band integral equals
    C kT (exp(-E0/kT) - exp(-E1/kT)) over the redshifted band.  The
    reference's step control compares step errors with an absolute
    tolerance of 1/num_samples and a relative one scaled by the running
    integral (integration.cl:86-96 as called at quantities.cl:226-237): it
    is accurate for integrands of order one or more that vary slowly over
    the band (kT above about 1.3 keV here), and the densities, distance and
    temperatures are chosen for that (steep integrands take steps of
    several keV and miss by up to ~10%, as the reference does).  Its
    4th-order weights (2197/4101 where Fehlberg has 2197/4104,
    integration.cl:69) add a bias of about 4e-4 per step."""
    g0 = 1.21
    tab = np.full((146, 81), g0, np.float32)
    rng = np.random.default_rng(4)
    n = 64
    rho = (10.0 ** rng.uniform(5.5, 6.5, n)).astype(np.float32)
    u = (10.0 ** rng.uniform(5.5, 5.95, n)).astype(np.float32)
    xe = rng.uniform(1.0, 1.2, n).astype(np.float32)
    z, dl, e0, e1 = 0.02, 450.0, 0.5, 7.0
    T = _temperature64(u.astype(np.float64), xe.astype(np.float64))
    t_inv = 1000.0 / (K_B * T)
    ion = 4.046263 * rho.astype(np.float64) * xe
    C = 0.002535 * ion ** 2 * g0 / np.sqrt(T) / dl ** 2
    a, b = (1 + z) * e0, (1 + z) * e1
    want = C * (np.exp(-a * t_inv) - np.exp(-b * t_inv)) / t_inv
    errs, steps_mean = [], []
    for ns in (100, 1000):
        got, steps = orc.xray_flux(rho, u, xe, z, dl, tab, e0, e1, ns)
        errs.append(float(np.max(np.abs(got - want) / want)))
        steps_mean.append(float(steps.mean()))
    assert errs[0] < 5e-3 and errs[1] < 5e-3


# ------------------------------------------------- volumetric smoothing (dm::grid)

def test_quartic3d_closed_forms(orc):
    """quartic_polynomial3d (smoothing_kernel.cl:27-37): W(0) = 1.04445/h^3,
    W = 0 from r = h on, and 4 pi int_0^h r^2 W dr = 1.04445 * 32 pi / 105
    (the shape (1 - q^2)^2 integrates to 32 pi / 105 h^3)."""
    h = 1.7
    assert orc.quartic3d(0.0, h) == pytest.approx(1.04445 / h ** 3, rel=1e-6)
    assert orc.quartic3d(h, h) == 0.0 and orc.quartic3d(1.3 * h, h) == 0.0
    r = (np.arange(200000) + 0.5) * (h / 200000)
    w = np.array([orc.quartic3d(x, h) for x in r[::100]])
    integral = float(np.sum(4.0 * np.pi * r[::100] ** 2 * w) * (h / 2000))
    assert integral == pytest.approx(1.04445 * 32.0 * np.pi / 105.0, rel=2e-5)
    # a dropped q^4 or a wrong sign on 2 q^2 breaks the half-radius value
    q = 0.5
    assert orc.quartic3d(q * h, h) == pytest.approx(1.04445 / h ** 3 * (1 - q * q) ** 2, rel=1e-6)


def test_smooth3d_exact_single_particle_and_volume_integral(orc):
    p4 = np.array([[1.0, 2.0, 3.0, 2.5]], np.float32)
    h = np.array([0.8], np.float32)
    pts = np.array([[1.0, 2.0, 3.0, 0], [1.4, 2.0, 3.0, 0], [1.0, 2.0, 3.9, 0]], np.float32)
    got = orc.smooth3d_exact(p4, h, pts)
    K = 105.0 / (32.0 * np.pi)
    h8 = float(np.float32(0.8))
    assert got[0] == pytest.approx(2.5 * K / h8 ** 3, rel=1e-12)
    s = 1.0 - (float(np.float32(1.4)) - 1.0) ** 2 / h8 ** 2
    assert got[1] == pytest.approx(2.5 * K / h8 ** 3 * s * s, rel=1e-6)
    assert got[2] == 0.0
    # sum over a fine lattice times the cell volume is the particle's weight
    d = 0.04
    ax = np.arange(-1.0, 1.0 + 1e-9, d)
    X, Y, Z = np.meshgrid(ax + 1.0, ax + 2.0, ax + 3.0, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.zeros(X.size)], 1).astype(np.float32)
    vol = orc.smooth3d_exact(p4, h, lat).sum() * d ** 3
    assert vol == pytest.approx(2.5, rel=2e-3)


def test_smooth3d_ref_equals_exact_when_nothing_is_dropped(orc):
    """Equal h everywhere: every point's search radius (max h of its cell +
    cell half-diagonal) covers every particle in reach, so the reference sum
    is the exact sum with the 1.04445f constant.  The reference's FP32 form
    1 + q^4 - 2 q^2 cancels near q = 1 (up to ~5e-5 of a point's value
    here); a dropped or extra particle is an error of 1e-3 or more."""
    rng = np.random.default_rng(8)
    n = 3000
    p4 = np.concatenate([rng.uniform(0, 20, (n, 3)), rng.uniform(0.5, 1.5, (n, 1))],
                        1).astype(np.float32)
    h = np.full(n, 1.3, np.float32)
    g = orc.Grid(p4, h, target=20)
    assert int(g.nc.min()) > 1
    pts = np.concatenate([rng.uniform(1, 19, (400, 3)), np.zeros((400, 1))], 1).astype(np.float32)
    ref = g.smooth3d_ref(pts)
    ex = orc.smooth3d_exact(p4, h, pts) * (float(np.float32(1.04445)) / (105.0 / (32.0 * np.pi)))
    m = ex > 0
    assert m.sum() > 300
    assert np.max(np.abs(ref[m] - ex[m]) / ex[m]) < 1e-4
    assert np.all(ref[~m] == 0.0)
    # the stable FP64 form of the same search agrees with the exact sum to rounding
    ref1 = g.smooth3d_ref(pts, form=1)
    assert np.max(np.abs(ref1[m] - ex[m]) / ex[m]) < 1e-12


def test_smooth3d_ref_drops_what_the_search_box_excludes(orc):
    """The search radius comes from the point's own cell: a big particle in
    another cell whose support reaches the point is dropped when it lies
    beyond max h (point's cell) + cell half-diagonal (the F3 heuristic)."""
    cs = np.float32(1.0)
    params = (np.array([10, 10, 10], np.int32), np.zeros(3, np.float32), np.full(3, cs))
    # small particles at the point's cell, one big one 3.2 away (h = 4)
    p4 = np.array([[2.5, 2.5, 2.5, 1.0], [2.6, 2.4, 2.5, 1.0], [5.7, 2.5, 2.5, 1.0]], np.float32)
    h = np.array([0.6, 0.6, 4.0], np.float32)
    g = orc.Grid(p4, h, params=params)
    pts = np.array([[2.5, 2.5, 2.5, 0.0]], np.float32)
    ref = g.smooth3d_ref(pts)[0]
    near = sum(orc.quartic3d(np.linalg.norm(p4[k, :3] - pts[0, :3]), h[k]) for k in range(2))
    big = orc.quartic3d(np.linalg.norm(p4[2, :3] - pts[0, :3]), h[2])
    assert big > 0                         # the big particle's support reaches the point
    assert ref == pytest.approx(near, rel=1e-6)   # but the search box drops it


# ------------------------------------------- RKF band integral and Gaunt data pins

_W_BIAS = 2197.0 * 3.0 / (4101.0 * 4104.0)  # 2197/4101 - 2197/4104


def _brems_closed(rho, u, xe, g0):
    """C and kT (keV) of the constant-table integrand f(E) = C exp(-E/kT)."""
    T = _temperature64(np.float64(np.float32(u)), np.float64(np.float32(xe)))
    ion = 4.046263 * np.float64(np.float32(rho)) * np.float64(np.float32(xe))
    return 0.002535 * ion ** 2 * g0 / np.sqrt(T), K_B * T / 1000.0


@pytest.mark.parametrize("ns,e0,e1", [(1, 2.0, 2.5), (2, 2.0, 2.5), (1, 2.0, 3.0), (4, 1.0, 2.5)])
def test_xray_flux_rkf_weights_closed_form(orc, ns, e0, e1):
    """With num_samples = ns the first step is h0 = (E1 - E0) / ns
    (quantities.cl:201-202).  For an integrand of order one the step control
    accepts it and grows the next step about sevenfold, past the band's end,
    which trims it to the end (integration.cl:86-130): the band is two steps,
    h0 and E1 - E0 - h0 (one when ns = 1).  For f = C exp(-E/kT) with steps
    below kT/2, Fehlberg's 4th-order rule is exact to 1e-7, so the
    reference's weight 2197/4101 (integration.cl:69, where Fehlberg has
    2197/4104) adds exactly h * (2197/4101 - 2197/4104) * f(E + 12h/13) per
    step to the closed form C kT (exp(-E0/kT) - exp(-E1/kT)).  A 4104
    transcription misses by 3.9e-4."""
    g0 = 1.21
    tab = np.full((146, 81), g0, np.float32)
    xe = 1.1
    mu = 4.0 / (1.0 + 3 * 0.76 + 4 * 0.76 * xe)
    u = (4000.0 / K_B) / ((2.0 / 3.0) * mu * 121.1475502)  # kT = 4 keV
    rho = 450.0                                            # f(2 keV) ~ 1
    got, steps = orc.xray_flux(np.array([rho]), np.array([u]), np.array([xe]), 0.0, 1.0, tab,
                               e0, e1, ns)
    C, kT = _brems_closed(rho, u, xe, g0)
    f = lambda E: C * np.exp(-E / kT)  # noqa: E731
    exact = C * kT * (np.exp(-e0 / kT) - np.exp(-e1 / kT))
    h0 = (e1 - e0) / ns
    hs = [h0] if ns == 1 else [h0, e1 - e0 - h0]
    starts = [e0, e0 + h0]
    bias = sum(_W_BIAS * h * f(a + 12.0 / 13.0 * h) for a, h in zip(starts, hs))
    assert int(steps[0]) == len(hs)
    assert float(got[0]) / (exact + bias) == pytest.approx(1.0, abs=1e-6)
    assert float(got[0]) / exact - 1.0 > 0.8 * _W_BIAS


def test_xray_flux_relative_tolerance_is_scale_free(orc):
    """With a negligible absolute tolerance (num_samples = 1e6) the step
    control compares each step's error with |running integral / position|
    (integration.cl:88-96): scaling the integrand by lambda scales both, so the
    step sequence is the same and the band integral scales by lambda.  Here
    lambda = (rho / rho0)^2 through the ion density."""
    tab = np.full((146, 81), 1.21, np.float32)
    xe = 1.1
    mu = 4.0 / (1.0 + 3 * 0.76 + 4 * 0.76 * xe)
    u = (4000.0 / K_B) / ((2.0 / 3.0) * mu * 121.1475502)
    rho = np.array([450.0, 900.0, 1800.0, 3600.0], np.float32)
    got, steps = orc.xray_flux(rho, np.full(4, u), np.full(4, xe), 0.0, 1.0, tab, 0.5, 7.0,
                               10 ** 6)
    assert len(set(steps.tolist())) == 1 and steps[0] > 4
    lam = (rho.astype(np.float64) / 450.0) ** 2
    assert np.allclose(got / got[0], lam, rtol=1e-6)


def _gaunt_fixture():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "gaunt_factor_vanhoof.npy"))


def _g_at(orc, tab, cg, cu):
    """g_ff (quantities.cl:69-83) read back from one bin of
    xray_spectral_flux at the texel coordinates (cg, cu): ln(gamma^2) =
    -6 + 0.2 cg, ln(u) = -16 + 0.2 cu (the table's first node at 0)."""
    kT_eV = RYD / np.exp(-6.0 + 0.2 * cg)
    T = kT_eV / K_B
    xe = 1.1
    mu = 4.0 / (1.0 + 3 * 0.76 + 4 * 0.76 * xe)
    u_int = T / ((2.0 / 3.0) * mu * 121.1475502)
    E = np.exp(-16.0 + 0.2 * cu) * kT_eV / 1000.0
    rho = 1e3
    got = float(orc.xray_spectral_flux(np.array([rho]), np.array([u_int]), np.array([xe]), 0.0,
                                       1.0, tab, E, 1.0)[0])
    Tf = _temperature64(np.float64(np.float32(u_int)), np.float64(np.float32(xe)))
    ion = 4.046263 * np.float64(np.float32(rho)) * np.float64(np.float32(xe))
    pref = 0.002535 * ion ** 2 / np.sqrt(Tf) * np.exp(-float(np.float32(E)) * 1000.0 / (K_B * Tf))
    return got / pref


def test_gaunt_table_fixture_shape_and_values():
    """The reference's table (tests/golden, from gaunt_factor.dat): 146 rows of
    ln u by 81 columns of ln gamma^2; g_ff is near 20.76 at the first entry,
    falls towards large u, and is positive everywhere."""
    tab = _gaunt_fixture()
    assert tab.shape == (146, 81) and tab.dtype == np.float32
    assert tab[0, 0] == np.float32(2.07575626e+01) and tab[0, 1] == np.float32(2.07575534e+01)
    assert tab[1, 0] == np.float32(2.05036528e+01)
    assert (tab > 0).all() and tab[-1].max() < tab[0].min()


def test_gff_at_texel_centres_and_nodes_of_the_reference_table(orc):
    """read_imagef with a linear filter on unnormalised coordinates
    (tabulated_function.cl:46-59): at a texel centre (coordinate i + 0.5) it
    returns that table entry; at a grid node (coordinate i) the mean of the
    four entries around it.  Checked on the reference's own table at nodes
    where the mean differs from every single entry by more than the
    tolerance, so a missing half-texel shift, swapped axes or log10 axes
    fail."""
    tab = _gaunt_fixture()
    checked = 0
    for cg in (3, 11, 20, 33, 47, 60, 71, 79):
        for cu in (4, 17, 30, 45, 58, 70, 84):
            centre = _g_at(orc, tab, cg + 0.5, cu + 0.5)
            assert centre == pytest.approx(float(tab[cu, cg]), rel=2e-5), (cg, cu)
            quad = tab[cu - 1:cu + 1, cg - 1:cg + 1].astype(np.float64)
            mean = float(quad.mean())
            node = _g_at(orc, tab, cg, cu)
            assert node == pytest.approx(mean, rel=2e-5), (cg, cu)
            if np.min(np.abs(quad - mean)) > 1e-4 * mean:
                checked += 1
    assert checked >= 20


# ------------------------------------------- line-of-sight clipped projection

def _clip_scene(n=300, seed=9):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), np.float32)
    p[:, 0] = rng.uniform(-50, 50, n)
    p[:, 1] = rng.uniform(-50, 50, n)
    p[:, 2] = rng.uniform(100, 400, n)     # camera depth z'
    p[:, 3] = rng.lognormal(0.0, 0.3, n)
    h = rng.uniform(5.0, 60.0, n).astype(np.float32)
    g = np.stack(np.meshgrid(np.linspace(-60, 60, 17), np.linspace(-60, 60, 17)), -1)
    pts = np.zeros((17 * 17, 4), np.float32)
    pts[:, :2] = g.reshape(-1, 2)
    return p, h, pts


@pytest.mark.parametrize("kind", [0, 1])
def test_clip_whole_support_equals_unclipped(orc, kind):
    """A depth range holding every particle's whole support gives the plain
    projected kernel (orc_gather_exact)."""
    p, h, pts = _clip_scene()
    got = orc.gather_clip(p, h, pts, 1000.0, kind)
    flat = p.copy()
    flat[:, 2] = 0.0
    ref = orc.gather_exact(flat, h, pts, kind)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_clip_ranges_add_and_halve(orc):
    """Complementary depth ranges add up to the whole (the depth range of a
    view starts at 0, so [0, d1] + [d1, d] is checked as the whole minus the
    first part shifted); a range ending at a particle's centre takes half of
    its column; a thin slab of thickness e at depth offset z takes e times
    the 3-D quartic K (15/16) (1 - q^2)^2 / h^3 at q^2 = (r^2 + z^2) / h^2."""
    p, h, pts = _clip_scene(n=80, seed=4)
    whole = orc.gather_clip(p, h, pts, 1000.0)
    first = orc.gather_clip(p, h, pts, 250.0)
    q = p.copy()
    q[:, 2] -= 250.0                        # the range [250, 1000] seen from 250
    second = orc.gather_clip(q, h, pts, 750.0)
    assert np.allclose(first + second, whole, rtol=1e-12, atol=1e-300)
    # one particle: a range ending at its centre
    one = np.array([[3.0, -4.0, 200.0, 1.7]], np.float32)
    h1 = np.array([20.0], np.float32)
    half = orc.gather_clip(one, h1, pts, 200.0)
    full = orc.gather_clip(one, h1, pts, 1000.0)
    m = full > 0
    assert m.sum() > 4 and np.allclose(half[m], 0.5 * full[m], rtol=1e-13)
    # a thin slab [z0, z0 + e]: the midpoint rule to O(e^2)
    K = 2.0 * float(np.float32(1.04445)) * float(np.float32(1.066667))
    e, z0 = 1e-3, 7.0
    sl = one.copy()
    sl[0, 2] = -z0                           # the particle z0 before the range's start
    thin = orc.gather_clip(sl, h1, pts, e)
    r2 = (pts[:, 0].astype(np.float64) - 3.0) ** 2 + (pts[:, 1].astype(np.float64) + 4.0) ** 2
    zc = z0 + 0.5 * e
    q2 = (r2 + zc * zc) / 400.0
    w3 = np.where(q2 < 1.0, (1.0 - q2) ** 2, 0.0) * K * 15.0 / 16.0 / 8000.0
    m = w3 > 1e-3 * w3.max()
    assert np.allclose(thin[m], 1.7 * e * w3[m], rtol=1e-6)


def test_project_particles_depth_matches_masked_projection(orc):
    """x', y' equal the masked projection's; the masked one zeroes w exactly
    where z' lies outside [0, depth]."""
    from illcrawl_amd import synth
    snap = synth.make_snapshot(1 << 12, 3, 10000.0)
    v = synth.centred_view(snap.box, look_at=(0.3, -0.5, 0.8), width=2000.0, nx=16, roll=0.4)
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    depth = 0.6 * v["depth"]
    pz = cam.project_particles_depth(snap.particles4())
    pm = cam.project_particles(snap.particles4(), depth)
    assert np.array_equal(pz[:, :2], pm[:, :2])
    inside = (pz[:, 2] >= 0) & (pz[:, 2] <= np.float32(depth))
    assert np.array_equal(pm[:, 3] != 0, inside & (pz[:, 3] != 0))
    assert 0 < inside.sum() < snap.n
