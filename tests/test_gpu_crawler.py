"""The data crawler end to end on the GPU: tests/cpp/crawl_demo.cpp streams an
Illustris-layout HDF5 snapshot through include/illcrawl_crawler.hpp in row
blocks: periodic correction, cutout and compaction on the device
(ilc_crawl_block), the quantity transformation, and a projective smoothing
reconstruction per block.  It writes a FITS image (cfitsio).

The same blocks are replayed through the Python mirror of the ABI
(crawl_block, the quantity functions, ProjectiveSmoothing over the grid
backend).  Every block builds its own grid, as in the reference, whose
reference-mode search radius depends on the block (F13), so the replay uses
the same block boundaries.  The two images may differ only by the FP32
rounding of the evaluation points (world pixel coordinates projected on the
C++ side, camera-frame pixels on the Python side) and of the dA scaling: the
bar is 1e-5 relative L-inf over pixels above 1e-6 of the maximum and 1e-6
on the sum."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "illcrawl_amd", "lib", "crawl_demo")
BOX, N, BLOCK, NX, WIDTH = 100.0, 6000, 1000, 96, 100.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(DEMO):
        pytest.skip("crawl_demo not built (HDF5 C++ / cfitsio absent)")
    from illcrawl_amd import backend
    return backend


def gaunt_like_table():
    j, i = np.meshgrid(np.arange(146), np.arange(81), indexing="ij")
    return (1.0 + 0.6 * np.exp(-((j - 90.0) / 30.0) ** 2) + 0.2 * np.sin(i / 9.0)
            + 0.002 * j).astype(np.float32)


@pytest.fixture(scope="module")
def snap(tmp_path_factory, B):
    d = str(tmp_path_factory.mktemp("crawl"))
    rng = np.random.default_rng(11)
    # positions partly outside the box: the crawler wraps them back
    c = dict(pos=rng.uniform(-10.0, BOX + 10.0, (N, 3)),
             rho=(10.0 ** rng.uniform(4.0, 7.0, N)).astype(np.float32),
             u=(10.0 ** rng.uniform(4.0, 6.5, N)).astype(np.float32),
             xe=rng.uniform(1.0, 1.2, N).astype(np.float32),
             mass=rng.lognormal(-2.0, 0.3, N).astype(np.float32),
             hsml=rng.uniform(2.0, 6.0, N).astype(np.float32))
    c["pos"].astype(np.float64).tofile(os.path.join(d, "pos.f64"))
    for k in ("rho", "u", "xe", "mass", "hsml"):
        c[k].tofile(os.path.join(d, k + ".f32"))
    tab = gaunt_like_table()
    tab.tofile(os.path.join(d, "gaunt.f32"))
    with open(os.path.join(d, "gaunt.dims"), "w") as f:
        f.write(f"{tab.shape[1]} {tab.shape[0]}\n")
    r = subprocess.run([DEMO, "write-snapshot", d, str(N)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    c["dir"], c["tab"] = d, tab
    return c


def replay(B, s, quantity):
    """The crawler's block loop through the Python mirror."""
    c = 0.5 * BOX
    cam = dict(position=(c, c, -1.0), look_at=(0.0, 0.0, 1.0), roll=0.0, width=WIDTH, nx=NX,
               ny=NX)
    dA = {"density": 1.0, "mass": None, "xray": None}[quantity]
    ps = B.ProjectiveSmoothing(cam, BOX + 2.0, (lambda a, depth: a) if dA is None else dA,
                               B.ProjectiveSmoothingGrid(mode="reference"))
    ps.init_backend(BLOCK)
    ps.setup_evaluation_points()
    for first in range(0, N, BLOCK):
        sl = slice(first, first + BLOCK)
        xyz = s["pos"][sl].astype(np.float32)   # the reference casts to the device scalar
        cols = [s["hsml"][sl], s["rho"][sl], s["u"][sl], s["xe"][sl]]
        p4, (h, rho, u, xe) = B.crawl_block((c, c, c), (BOX,) * 3, (BOX,) * 3, 0.0, xyz, cols)
        if p4.shape[0] == 0:
            continue
        if quantity == "xray":
            w = B.quantity_xray_flux(rho, u, xe, s["tab"], 0.0, 1.0e5, 0.5, 5.0, 100)
        else:
            w = B.quantity_scaled(rho, 1.0)
        p4[:, 3] = w
        ps.setup_particles(p4, [h])
        ps.run()
    return ps.retrieve_results().cpu().numpy()


@pytest.mark.parametrize("quantity", ["density", "mass", "xray"])
def test_crawler_from_hdf5_equals_block_replay(B, snap, quantity):
    from illcrawl_amd import io as I
    say(f"case: C++ crawler over HDF5, quantity={quantity}, n={N}, blocks of {BLOCK}")
    r = subprocess.run([DEMO, "render", snap["dir"], str(BLOCK), quantity, str(NX), str(WIDTH),
                        str(BOX)], capture_output=True, text=True, timeout=120)
    say("  " + (r.stdout.strip() or r.stderr.strip()))
    assert r.returncode == 0, r.stderr
    got = I.read_fits(os.path.join(snap["dir"], f"image_{quantity}.fits")).astype(np.float64)
    ref = replay(B, snap, quantity).astype(np.float64)
    assert got.shape == ref.shape == (NX, NX)
    m = ref > 1e-6 * ref.max()
    assert m.sum() > NX * NX // 2
    linf = float(np.max(np.abs(got[m] - ref[m]) / ref[m]))
    ssum = abs(got.sum() - ref.sum()) / ref.sum()
    say(f"  rel L-inf {linf:.2e}, rel sum {ssum:.2e}, max {ref.max():.4e}")
    assert linf < 1e-5 and ssum < 1e-6


def volume_replay(B, s, nx, nz):
    """The crawler's blocks as (particles4, h) for the volumetric mirror."""
    c = 0.5 * BOX
    out = []
    for first in range(0, N, BLOCK):
        sl = slice(first, first + BLOCK)
        xyz = s["pos"][sl].astype(np.float32)
        p4, (h, rho) = B.crawl_block((c, c, c), (BOX,) * 3, (BOX,) * 3, 0.0, xyz,
                                     [s["hsml"][sl], s["rho"][sl]])
        if p4.shape[0]:
            p4[:, 3] = B.quantity_scaled(rho, 1.0)
            out.append((p4, h))
    return out


@pytest.mark.parametrize("quantity,nz", [("density", 0), ("mass", 0), ("density", 5)])
def test_volumetric_crawler_from_hdf5_equals_block_replay(B, snap, quantity, nz):
    """volumetric_slice / volumetric_tomography (include/illcrawl_volumetric.hpp)
    over the crawler with dm::grid, against the same blocks through the
    Python mirror: the points are the same doubles rounded to float on both
    sides and every block builds the same grid, so the values agree to the
    FP32 rounding of the dV scaling."""
    from illcrawl_amd import io as I
    nx = 48
    say(f"case: C++ volumetric crawler over HDF5, quantity={quantity}, nz={nz}")
    r = subprocess.run([DEMO, "volume", snap["dir"], str(BLOCK), quantity, str(nx), str(WIDTH),
                        str(BOX), str(nz), "0"], capture_output=True, text=True, timeout=120)
    say("  " + (r.stdout.strip() or r.stderr.strip()))
    assert r.returncode == 0, r.stderr
    name = ("slice_" if nz == 0 else "cube_") + quantity + ".fits"
    got = I.read_fits(os.path.join(snap["dir"], name)).astype(np.float64)
    c = 0.5 * BOX
    cam = dict(position=(c, c, c), look_at=(0.0, 0.0, 1.0), roll=0.0, width=WIDTH, nx=nx, ny=nx)
    ps = WIDTH / nx
    dV = 1.0 if quantity == "density" else ps ** 3
    blocks = volume_replay(B, snap, nx, nz)
    if nz == 0:
        ref = B.volumetric_slice(cam, blocks, dV=dV).astype(np.float64)
    else:
        ref = B.volumetric_tomography(cam, blocks, num_steps=nz, dV=dV).astype(np.float64)
    assert got.shape == ref.shape
    m = ref > 1e-6 * ref.max()
    assert m.sum() > ref.size // 2
    linf = float(np.max(np.abs(got[m] - ref[m]) / ref[m]))
    ssum = abs(got.sum() - ref.sum()) / ref.sum()
    say(f"  rel L-inf {linf:.2e}, rel sum {ssum:.2e}")
    assert linf < 1e-6 and ssum < 1e-6


def test_volumetric_slice_with_jittered_samples(B, snap):
    """volumetric_slice with additional samples per pixel (seeded jitter of
    up to half a pixel, averaged): the same seed gives the same slice, and on
    this smooth field the average stays within a few per cent of the pixel
    corner values in total."""
    from illcrawl_amd import io as I
    nx = 48
    outs = []
    for samples in (0, 3, 3):
        r = subprocess.run([DEMO, "volume", snap["dir"], str(BLOCK), "density", str(nx),
                            str(WIDTH), str(BOX), "0", str(samples)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(I.read_fits(os.path.join(snap["dir"], "slice_density.fits"))
                    .astype(np.float64))
    base, jit, again = outs
    say(f"  jittered/base sum {jit.sum() / base.sum():.4f}")
    assert np.all(np.isfinite(jit))
    assert np.array_equal(jit, again)
    assert not np.array_equal(jit, base)
    assert abs(jit.sum() / base.sum() - 1.0) < 0.05
