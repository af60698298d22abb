"""GPU tests of the batched multi-view path (ilc_batch_* through the C ABI):
every view of a persistent, graph-replayed or plain batch equals the
single-view backend on the same view, which is itself pinned to the oracle;
one view per batch is checked against the oracle directly."""
import math
import sys

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
def scene():
    from illcrawl_amd import synth
    snap = synth.make_snapshot(1 << 15, 7, 10000.0, n_halos=2, rs_range=(200.0, 400.0),
                               gas=True)
    views = synth.random_views(snap.box, n_views=8, seed=2, resolutions=(48, 64, 96, 128),
                               widths=(1500.0, 3000.0))
    for k, v in enumerate(views):
        v["scale"] = 1.0 if k % 3 else 0.5
    views[0]["depth"] *= 0.5   # a depth-masked view
    return snap, views


# A batched view's grid keeps only the particles whose support reaches the
# screen (the others only raise their cell's max h), so its tiles stream fewer
# candidates than the single-view grid's and group the same list entries into
# FP32 partial sums at other places: the images agree to the FP32 rounding
# of those sums.  The plain path builds each view's grid alone and stays bit
# for bit equal to the single-view backend.
GROUPING_TOL = 1e-6


def assert_same_sum(got, ref, what):
    assert np.array_equal(got == 0, ref == 0), what
    m = ref != 0
    rel = float(np.max(np.abs(got[m] - ref[m]) / np.abs(ref[m]))) if m.any() else 0.0
    assert rel < GROUPING_TOL, (what, rel)


def single_view(B, snap, v, mode, tab, kernel="quartic_ref"):
    if v["quantity"] == 1:
        w = B.quantity_xray_bolometric(snap.rho, snap.temp, snap.xe, *tab)
    else:
        w = B.quantity_scaled(snap.mass, 1.0)
    p4 = torch.empty((snap.n, 4), dtype=torch.float32, device="cuda")
    p4[:, :3] = torch.from_numpy(snap.pos).cuda()
    p4[:, 3] = w
    proj = B.project_particles(v, p4)
    b = B.ProjectiveSmoothingGrid(mode=mode, kernel=kernel)
    b.init_backend(snap.n)
    b.setup_projected_particles(proj, snap.hsml)
    b.setup_evaluation_points(B.camera_pixel_points(v), layout_nx=v["nx"])
    b.run()
    return b.retrieve_results_f64().cpu().numpy().reshape(v["ny"], v["nx"]), proj


def make_batch(B, snap, views, mode, tab, tile=None, kernel="quartic_ref"):
    r = B.BatchRenderer(mode=mode, tile=tile, kernel=kernel)
    r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
    r.set_gaunt_table(*tab)
    r.set_views(views)
    return r


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_batch_equals_single_view(B, scene, mode):
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say(f"case: batch vs single, mode={mode}")
    r = make_batch(B, snap, views, mode, tab)
    r.render()
    for k, v in enumerate(views):
        ref, _ = single_view(B, snap, v, mode, tab)
        got = r.image_f64(k).cpu().numpy()
        assert_same_sum(got, ref, f"view {k}")
        img = r.image(k).cpu().numpy()
        assert np.array_equal(img, (got * np.float64(np.float32(v["scale"]))).astype(np.float32))
    # the plain path: each view's grid alone, bit for bit the single view's
    r.render(plain=True)
    for k, v in enumerate(views):
        ref, _ = single_view(B, snap, v, mode, tab)
        assert np.array_equal(r.image_f64(k).cpu().numpy(), ref), f"plain view {k}"


@pytest.mark.parametrize("tile", [None, (16, 16, 64), (8, 8, 64)])
def test_batch_view_vs_oracle(B, scene, orc, tile):
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say(f"case: batch view 1 vs oracle, tile {tile}")
    r = make_batch(B, snap, views, "reference", tab, tile=tile)
    r.render()
    for k in (0, 1):
        v = views[k]
        cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
        w = orc.xray_bolometric(snap.rho, snap.temp, snap.xe, *tab) if v["quantity"] == 1 \
            else snap.mass
        pp = cam.project_particles(snap.particles4(w), v["depth"])
        ref = orc.Grid(pp, snap.hsml).gather_ref(cam.pixel_points(), 1)
        got = r.image_f64(k).cpu().numpy().reshape(-1)
        m = ref > 1e-6 * ref.max()
        assert np.max(np.abs(got[m] - ref[m]) / ref[m]) < 1e-5
        assert abs(got.sum() - ref.sum()) / ref.sum() < 1e-6


def test_plain_and_tiles_agree(B, scene):
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say("case: plain path and tile shapes")
    r = make_batch(B, snap, views, "reference", tab)
    r.render()
    base = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    r.render(plain=True)
    for k in range(len(views)):
        assert_same_sum(r.image_f64(k).cpu().numpy(), base[k], f"plain view {k}")
    # (16, 16, 64): one-wave workgroups, four points per lane
    for tile in ((16, 16), (32, 16), (64, 16), (8, 8, 64), (16, 16, 64)):
        r.set_tile(*tile)
        r.render()
        for k in range(len(views)):
            a = r.image_f64(k).cpu().numpy()
            rel = np.abs(a - base[k]).max() / base[k].max()
            assert rel < 1e-6, (tile, k, rel)
    assert r.tile() == (16, 16, 64)
    r.render(plain=True)
    for k in range(len(views)):
        a = r.image_f64(k).cpu().numpy()
        rel = np.abs(a - base[k]).max() / base[k].max()
        assert rel < 1e-6, ("plain, one-wave tiles", k, rel)


def test_graph_replays_match_eager(B, scene):
    """Replays of the captured device work equal the eager render, with eager
    work and synchronisations between the replays."""
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say("case: graph replays")
    r = make_batch(B, snap, views, "reference", tab)
    r.render()
    eager = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    for rep in range(4):
        r.render(graph=True)
        torch.cuda.synchronize()
        for k in range(len(views)):
            assert np.array_equal(r.image_f64(k).cpu().numpy(), eager[k]), (rep, k)
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        r.render()
        for k in range(len(views)):
            assert np.array_equal(r.image_f64(k).cpu().numpy(), eager[k])


def test_graph_after_plain_render(B, scene):
    """A plain render between graph renders does not leak its per-view grids
    into the replayed persistent launch: graph, plain, graph all give the
    eager batched images bit for bit (the plain ones, the single views')."""
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say("case: graph / plain / graph")
    r = make_batch(B, snap, views, "reference", tab)
    r.render()
    eager = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    r.render(graph=True)
    for k in range(len(views)):
        assert np.array_equal(r.image_f64(k).cpu().numpy(), eager[k]), ("graph 1", k)
    r.render(plain=True)
    plain = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    for k in range(len(views)):
        assert_same_sum(plain[k], eager[k], f"plain view {k}")
    for rep in range(2):
        r.render(graph=True)
        torch.cuda.synchronize()
        for k in range(len(views)):
            assert np.array_equal(r.image_f64(k).cpu().numpy(), eager[k]), ("graph 2", rep, k)
        # the views' grids are the shared build's again
        assert r.view_info(1)["num_particles"] <= snap.n


def test_graph_after_larger_snapshot(B, scene):
    """set_particles with a larger snapshot after graph renders: the next
    graph render sizes its buffers eagerly before it captures, and its images
    equal an eager render of the new snapshot on a fresh batch."""
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say("case: graph render after a larger snapshot")
    half = snap.n // 2
    r = B.BatchRenderer(mode="reference")
    r.set_particles(snap.pos[:half], snap.mass[:half], snap.hsml[:half], snap.rho[:half],
                    snap.temp[:half], snap.xe[:half])
    r.set_gaunt_table(*tab)
    r.set_views(views)
    r.render()
    r.render(graph=True)
    r.render(graph=True)
    r.set_particles(snap.pos, snap.mass, snap.hsml, snap.rho, snap.temp, snap.xe)
    r.render(graph=True)
    got = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    r.render(graph=True)
    again = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    fresh = make_batch(B, snap, views, "reference", tab)
    fresh.render()
    for k in range(len(views)):
        ref = fresh.image_f64(k).cpu().numpy()
        assert np.array_equal(got[k], ref), ("first graph render", k)
        assert np.array_equal(again[k], ref), ("replay", k)


def test_batch_counts_and_autotune(B, scene, orc):
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say("case: counts / autotune")
    r = make_batch(B, snap, views, "exact", tab)
    r.render()
    counts = r.count_evals()
    v = views[2]
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    w = orc.xray_bolometric(snap.rho, snap.temp, snap.xe, *tab) if v["quantity"] == 1 \
        else snap.mass
    pp = cam.project_particles(snap.particles4(w), v["depth"])
    assert counts[2] == orc.count_evals_lattice(pp, snap.hsml, cam.pixel_points(), v["nx"],
                                                v["ny"])
    (tw, th), ms = r.autotune()
    w, h, wg = r.tile()
    assert (w, h) == (tw, th) and wg in (64, 256) and tw * th % wg == 0 and ms > 0
    r.render()
    assert r.count_evals() == counts


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_windowed_grids_equal_single_view(B, mode):
    """Narrow views deep inside the snapshot build their grids from the
    particles around the screen only; their images equal the single-view
    backend's full grid to the grouping of the FP32 partial sums.  A view
    reaching past the particle bounding box keeps every particle (reference
    mode: a point outside the grid can alias any cell)."""
    from illcrawl_amd import synth
    say(f"case: windowed batch grids, mode={mode}")
    snap = synth.make_snapshot(1 << 16, 9, 10000.0, n_halos=2, rs_range=(200.0, 400.0),
                               gas=True)
    tab = synth.gaunt_table_1d()
    centre = snap.halos[0][0]
    views = []
    for k, (w, nx) in enumerate(((300.0, 64), (600.0, 96), (900.0, 128), (450.0, 80))):
        c = np.asarray(centre, np.float64) + 150.0 * k
        v = synth.centred_view(snap.box, look_at=(0.1 * k, -0.05 * k, 1.0), width=w, nx=nx,
                               roll=0.2 * k, centre=c)
        v["quantity"] = k % 2
        v["scale"] = 1.0
        views.append(v)
    edge = synth.centred_view(snap.box, width=6000.0, nx=64, centre=(0.0, 0.0, 5000.0))
    edge["quantity"], edge["scale"] = 0, 1.0
    views.append(edge)
    r = make_batch(B, snap, views, mode, tab)
    r.render()
    n_windowed = 0
    for k, v in enumerate(views):
        ref, _ = single_view(B, snap, v, mode, tab)
        got = r.image_f64(k).cpu().numpy()
        info = r.view_info(k)
        say(f"  view {k}: {info['num_particles']} of {snap.n} particles, "
            f"max {float(ref.max()):.3e}")
        assert_same_sum(got, ref, f"view {k}")
        assert ref.max() > 0
        n_windowed += info["num_particles"] < snap.n
    assert n_windowed >= 3
    if mode == "reference":
        assert r.view_info(len(views) - 1)["num_particles"] == snap.n
    # graph replays give the eager images
    eager = [r.image_f64(k).cpu().numpy() for k in range(len(views))]
    r.render(graph=True)
    r.render(graph=True)
    for k in range(len(views)):
        assert np.array_equal(r.image_f64(k).cpu().numpy(), eager[k]), f"graph view {k}"


@pytest.mark.parametrize("mode,tile", [("reference", None), ("exact", None),
                                       ("reference", (16, 16, 64)), ("reference", (8, 8, 64))])
def test_slab_stack_shares_one_grid(B, orc, mode, tile):
    """Axis-aligned slabs project every particle to the same (x', y'): the
    batch builds one full grid for the stack (cells, per-cell max h) and each
    slab's lists from its depth cut only.  Particles outside the cut have
    w = 0 and add nothing, so the images equal the single-view ones up to the
    FP32 grouping of the partial sums, and match the oracle."""
    from illcrawl_amd import synth
    say(f"case: slab stack, mode={mode}, tile {tile}")
    snap = synth.make_snapshot(1 << 16, 11, 10000.0, n_halos=3, rs_range=(200.0, 500.0),
                               gas=True)
    tab = synth.gaunt_table_1d()
    views = synth.slab_views(snap.box, n_slabs=6, thickness=800.0, nx=96, width=8000.0,
                             quantity=1)
    for v in views:
        v["scale"] = 1.0
    r = make_batch(B, snap, views, mode, tab, tile=tile)
    r.render()
    for k, v in enumerate(views):
        ref, proj = single_view(B, snap, v, mode, tab)
        got = r.image_f64(k).cpu().numpy()
        info = r.view_info(k)
        kept = int((proj[:, 3] != 0).sum().item())
        m = ref > 1e-6 * ref.max()
        e = float(np.max(np.abs(got[m] - ref[m]) / ref[m]))
        say(f"  slab {k}: grid lists {info['num_particles']} (depth cut {kept}) of {snap.n}, "
            f"rel diff vs single view {e:.2e}")
        assert info["num_particles"] < snap.n // 2
        assert e < 1e-6
    # one slab against the oracle
    v = views[2]
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    w = orc.xray_bolometric(snap.rho, snap.temp, snap.xe, *tab)
    pp = cam.project_particles(snap.particles4(w), v["depth"])
    pts = cam.pixel_points()
    if mode == "reference":
        ref = orc.Grid(pp, snap.hsml).gather_ref(pts)
    else:
        ref = orc.gather_exact(pp, snap.hsml, pts, kind=0)
    got = r.image_f64(2).cpu().numpy().ravel()
    m = ref > 1e-6 * ref.max()
    linf = float(np.max(np.abs(got[m] - ref[m]) / ref[m]))
    ssum = abs(got.sum() - ref.sum()) / ref.sum()
    say(f"  slab 2 vs oracle: linf {linf:.2e} sum {ssum:.2e}")
    assert linf < 1e-5 and ssum < 1e-6
    r.render(graph=True)
    r.render(graph=True)
    assert np.array_equal(r.image_f64(2).cpu().numpy().ravel(), got)


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_slab_stack_pass_equals_per_view_pass(B, mode):
    """The multi-view passes of a slab stack test two slabs per particle (the
    last slab starting at or before its depth, and the one before); beside a
    view that is no slab they test every view.  The slabs' kept rows, their
    order, grids and images are the same either way, bit for bit."""
    from illcrawl_amd import synth
    say(f"case: slab stack pass vs per-view pass, mode={mode}")
    snap = synth.make_snapshot(1 << 16, 12, 10000.0, n_halos=3, rs_range=(200.0, 500.0),
                               gas=True)
    tab = synth.gaunt_table_1d()
    slabs = synth.slab_views(snap.box, n_slabs=6, thickness=800.0, nx=64, width=8000.0,
                             quantity=1)
    for v in slabs:
        v["scale"] = 1.0
    other = synth.centred_view(snap.box, look_at=(0.3, 0.2, 0.9), width=6000.0, nx=64)
    other.update(quantity=1, scale=1.0)
    r = make_batch(B, snap, slabs, mode, tab)
    r.render()
    a = [r.image_f64(k).cpu().numpy() for k in range(len(slabs))]
    kept = [r.view_info(k)["num_particles"] for k in range(len(slabs))]
    r2 = make_batch(B, snap, slabs + [other], mode, tab)
    r2.render()
    for k in range(len(slabs)):
        assert r2.view_info(k)["num_particles"] == kept[k], k
        assert np.array_equal(r2.image_f64(k).cpu().numpy(), a[k]), k
    assert sum(kept) < snap.n and min(kept) > 0


def test_batch_backward_equals_sum_of_single_views(B):
    """ilc_batch_backward (exact mode) equals, per particle, the sum over
    views of the single-view adjoint (itself pinned to the FP64 oracle)
    carried to world positions, with dL/dw taken only from views whose depth
    cut keeps the particle.  Covers full, windowed and slab-group grids."""
    from illcrawl_amd import synth
    say("case: batch backward vs single views")
    snap = synth.make_snapshot(1 << 15, 21, 10000.0, n_halos=2, rs_range=(200.0, 400.0),
                               gas=True)
    tab = synth.gaunt_table_1d()
    views = synth.random_views(snap.box, n_views=3, seed=5, resolutions=(48, 64, 80),
                               widths=(1500.0, 3000.0))
    views[0]["depth"] *= 0.5
    c = np.asarray(snap.halos[0][0], np.float64)
    views.append(dict(synth.centred_view(snap.box, width=500.0, nx=64, centre=c), quantity=0))
    views += synth.slab_views(snap.box, n_slabs=3, thickness=1500.0, nx=64, width=8000.0,
                              quantity=0)
    for v in views:
        v["scale"] = 0.8
    r = make_batch(B, snap, views, "exact", tab)
    r.render()
    rng = np.random.default_rng(9)
    grads = [torch.from_numpy(rng.normal(size=(v["ny"], v["nx"])).astype(np.float32)).cuda()
             for v in views]
    gw, gh, gp = (t.cpu().numpy().astype(np.float64) for t in r.backward(grads))
    info = [r.view_info(k)["num_particles"] for k in range(len(views))]
    say(f"  grid sizes {info} of {snap.n}")
    assert min(info) < snap.n
    ew, eh, ep = np.zeros(snap.n), np.zeros(snap.n), np.zeros((snap.n, 3))
    for v, g in zip(views, grads):
        if v["quantity"] == 1:
            w = B.quantity_xray_bolometric(snap.rho, snap.temp, snap.xe, *tab)
        else:
            w = B.quantity_scaled(snap.mass, 1.0)
        p4 = torch.empty((snap.n, 4), dtype=torch.float32, device="cuda")
        p4[:, :3] = torch.from_numpy(snap.pos).cuda()
        p4[:, 3] = w
        proj = B.project_particles(v, p4)
        ones = p4.clone()
        ones[:, 3] = 1.0
        kept = (B.project_particles(v, ones)[:, 3] != 0).cpu().numpy()
        b = B.ProjectiveSmoothingGrid(mode="exact")
        b.init_backend(snap.n)
        b.setup_projected_particles(proj, snap.hsml)
        b.setup_evaluation_points(B.camera_pixel_points(v), layout_nx=v["nx"])
        sw, sh, sx, sy = (t.cpu().numpy().astype(np.float64)
                          for t in b.backward(g.reshape(-1), float(np.float32(v["scale"]))))
        minv, _ = B.camera_projection(v)
        ew += np.where(kept, sw, 0.0)
        eh += sh
        ep += sx[:, None] * minv[0][None, :] + sy[:, None] * minv[1][None, :]
    for name, got, ref in (("w", gw, ew), ("h", gh, eh), ("pos", gp, ep)):
        e = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        say(f"  dL/d{name}: rel-to-max diff {e:.2e}")
        assert e < 1e-5, name


def test_cubic_batch_autotunes_before_its_first_render(B, scene, orc):
    """The cubic spline's table is set up by every entry point that launches
    the gather, the autotuner included (a regression test: the autotuner once
    ran before the first render without it)."""
    from illcrawl_amd import synth
    snap, views = scene
    tab = synth.gaunt_table_1d()
    say("case: cubic batch, autotune first")
    r = make_batch(B, snap, views[:3], "exact", tab, kernel="cubic")
    (tw, th), ms = r.autotune()
    assert ms > 0
    r.render()
    for k, v in enumerate(views[:3]):
        ref, _ = single_view(B, snap, v, "exact", tab, kernel="cubic")
        assert_same_sum(r.image_f64(k).cpu().numpy(), ref, f"cubic view {k}")
    v = views[1]
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    w = orc.xray_bolometric(snap.rho, snap.temp, snap.xe, *tab) if v["quantity"] == 1 \
        else snap.mass
    pp = cam.project_particles(snap.particles4(w), v["depth"])
    ref = orc.gather_exact(pp, snap.hsml, cam.pixel_points(), kind=2)
    got = r.image_f64(1).cpu().numpy().reshape(-1)
    m = ref > 1e-6 * ref.max()
    assert np.max(np.abs(got[m] - ref[m]) / ref[m]) < 2e-5
    assert abs(got.sum() - ref.sum()) / ref.sum() < 2e-6
