"""The reference's spectral X-ray quantities on the GPU (ilc_quantity_xray_flux,
ilc_quantity_xray_spectral_flux) against the oracle.

Both sides evaluate the same FP32 operations without contraction and use
correctly rounded FP32 exp/log/pow/rsqrt (computed in FP64 and rounded), so
results and RKF step counts agree bit for bit except where an FP64 library
result lands within one FP64 ulp of an FP32 rounding boundary on one side
only (a rare event): the bar is bit-exact for at least 99.9% of the
particles and 1e-3 relative (one RKF tolerance step) for the rest.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from illcrawl_amd import backend
    return backend


def gaunt_like_table():
    """A smooth table of the van Hoof layout (146 rows of ln u, 81 columns of
    ln gamma^2) with values in the Gaunt factor's range."""
    j, i = np.meshgrid(np.arange(146), np.arange(81), indexing="ij")
    return (1.0 + 0.6 * np.exp(-((j - 90.0) / 30.0) ** 2) + 0.2 * np.sin(i / 9.0)
            + 0.002 * j).astype(np.float32)


def gas(n, seed):
    rng = np.random.default_rng(seed)
    rho = (10.0 ** rng.uniform(3.0, 8.0, n)).astype(np.float32)
    u = (10.0 ** rng.uniform(3.0, 7.0, n)).astype(np.float32)   # T ~ 5e4 .. 5e8 K
    xe = rng.uniform(1.0, 1.2, n).astype(np.float32)
    return rho, u, xe


def agree(got, ref, what):
    got = np.asarray(got)
    ref = np.asarray(ref)
    same = (got.view(np.uint32) == ref.view(np.uint32)) if got.dtype == np.float32 \
        else got == ref
    frac = float(same.mean())
    # equal infinities (same bits) give inf - inf; only unequal values are read
    with np.errstate(invalid="ignore"):
        rel = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-300)
    worst = float(rel[~same].max()) if (~same).any() else 0.0
    say(f"  {what}: bit-identical {frac:.6f}, worst other rel diff {worst:.2e}")
    assert frac >= 0.999
    assert worst < 1e-3


@pytest.mark.parametrize("num_samples,band", [(10, (0.5, 7.0)), (100, (0.3, 2.0))])
def test_xray_flux_vs_oracle(B, orc, num_samples, band):
    say(f"case: xray_flux num_samples={num_samples} band={band}")
    rho, u, xe = gas(50000, num_samples)
    tab = gaunt_like_table()
    z, dl = 0.03, 1.5e5
    got, steps = B.quantity_xray_flux(rho, u, xe, tab, z, dl, band[0], band[1], num_samples,
                                      with_steps=True)
    ref, rsteps = orc.xray_flux(rho, u, xe, z, dl, tab, band[0], band[1], num_samples)
    agree(got.cpu().numpy(), ref, "flux")
    st = steps.cpu().numpy()
    assert float((st == rsteps).mean()) >= 0.999
    say(f"  steps: mean {st.mean():.1f}, max {st.max()}")
    assert st.max() < 65536


def reference_table():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "gaunt_factor_vanhoof.npy"))


@pytest.mark.parametrize("num_samples", [10, 1000])
def test_xray_flux_with_the_reference_gaunt_table(B, orc, num_samples):
    """The reference's own van Hoof table (gaunt_factor.dat, tests/golden)."""
    say(f"case: xray_flux, reference Gaunt table, num_samples={num_samples}")
    rho, u, xe = gas(50000, 100 + num_samples)
    tab = reference_table()
    got, steps = B.quantity_xray_flux(rho, u, xe, tab, 0.02, 1e5, 0.5, 7.0, num_samples,
                                      with_steps=True)
    ref, rsteps = orc.xray_flux(rho, u, xe, 0.02, 1e5, tab, 0.5, 7.0, num_samples)
    agree(got.cpu().numpy(), ref, "flux")
    assert float((steps.cpu().numpy() == rsteps).mean()) >= 0.999
    for E, dE in ((0.7, 0.1), (5.0, 1.0)):
        got = B.quantity_xray_spectral_flux(rho, u, xe, tab, 0.0, 1e5, E, dE).cpu().numpy()
        ref = orc.xray_spectral_flux(rho, u, xe, 0.0, 1e5, tab, E, dE)
        agree(got, ref, f"spectral E={E}")


def test_xray_spectral_flux_vs_oracle(B, orc):
    say("case: xray_spectral_flux")
    rho, u, xe = gas(100000, 7)
    tab = gaunt_like_table()
    for E, dE in ((0.7, 0.1), (3.0, 0.5)):
        got = B.quantity_xray_spectral_flux(rho, u, xe, tab, 0.1, 2e5, E, dE).cpu().numpy()
        ref = orc.xray_spectral_flux(rho, u, xe, 0.1, 2e5, tab, E, dE)
        agree(got, ref, f"E={E}")


def test_xray_flux_errors(B):
    from illcrawl_amd import IllcrawlError
    rho, u, xe = gas(10, 1)
    with pytest.raises(IllcrawlError):
        B.quantity_xray_flux(rho, u, xe, gaunt_like_table(), 0.0, 1e5, 0.5, 7.0, 0)
    with pytest.raises(ValueError):
        B.quantity_xray_flux(rho, u, xe, np.ones(10, np.float32), 0.0, 1e5, 0.5, 7.0, 10)


def test_xray_flux_projection_like_the_oracle(B, orc):
    """The crawler's composition (F5): the quantity writes w, the grid path
    sums w W over the image (reference mode)."""
    from illcrawl_amd import synth
    say("case: xray_flux -> projection")
    snap = synth.make_snapshot(1 << 15, 13, 10000.0, n_halos=1, rs_range=(300.0, 300.0),
                               gas=True)
    rng = np.random.default_rng(3)
    u = (10.0 ** rng.uniform(4.5, 6.0, snap.n)).astype(np.float32)
    tab = gaunt_like_table()
    w = B.quantity_xray_flux(snap.rho, u, snap.xe, tab, 0.02, 1e5, 0.5, 7.0, 20)
    wr, _ = orc.xray_flux(snap.rho, u, snap.xe, 0.02, 1e5, tab, 0.5, 7.0, 20)
    v = synth.centred_view(snap.box, width=3000.0, nx=64)
    p4 = torch.empty((snap.n, 4), dtype=torch.float32, device="cuda")
    p4[:, :3] = torch.from_numpy(snap.pos).cuda()
    p4[:, 3] = w
    ps = B.ProjectiveSmoothing(v, v["depth"], 1.0, B.ProjectiveSmoothingGrid())
    ps.init_backend(snap.n)
    ps.setup_particles(p4, [snap.hsml])
    ps.setup_evaluation_points()
    ps.run()
    img = ps.retrieve_results().cpu().numpy().ravel().astype(np.float64)
    cam = orc.Camera(v["position"], v["look_at"], v["roll"], v["width"], v["nx"], v["ny"])
    pp = cam.project_particles(snap.particles4(wr), v["depth"])
    ref = orc.Grid(pp, snap.hsml).gather_ref(cam.pixel_points())
    m = ref > 1e-6 * ref.max()
    linf = float(np.max(np.abs(img[m] - ref[m]) / ref[m]))
    say(f"  image linf {linf:.2e}")
    assert linf < 1e-5
