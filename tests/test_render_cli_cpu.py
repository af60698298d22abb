"""The render tool's option and quantity layer on the CPU
(include/illcrawl_app.hpp, illcrawl_amd/app/illcrawl_render.cpp): every case
here fails before any device work, with exit status 1 and the reference's
messages (illcrawl_app.cpp:594-835, illcrawl_render.cpp:336-343)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "illcrawl_amd", "lib")
RENDER = os.path.join(LIB, "illcrawl_render")
DEMO = os.path.join(LIB, "crawl_demo")


@pytest.fixture(scope="module")
def snap(tmp_path_factory):
    if not (os.path.exists(RENDER) and os.path.exists(DEMO)):
        pytest.skip("illcrawl_render not built (HDF5 C++ / cfitsio absent)")
    d = str(tmp_path_factory.mktemp("cli"))
    rng = np.random.default_rng(3)
    n = 200
    rng.uniform(0.0, 10.0, (n, 3)).tofile(os.path.join(d, "pos.f64"))
    for k in ("rho", "u", "xe", "mass", "hsml"):
        rng.uniform(1.0, 2.0, n).astype(np.float32).tofile(os.path.join(d, k + ".f32"))
    r = subprocess.run([DEMO, "write-snapshot", d, str(n)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("crawl_demo cannot run here: " + r.stderr[-200:])
    return os.path.join(d, "snap.hdf5")


def run(*args):
    r = subprocess.run([RENDER] + list(args), capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


def test_missing_required_options(snap):
    rc, out = run()
    assert rc == 1 and "'--input' is required" in out and "--render_target" in out
    rc, out = run("-i", snap)
    assert rc == 1 and "'--render_target' is required" in out


def test_unknown_option_and_bad_values(snap):
    rc, out = run("-i", snap, "-t", "image", "--no_such_option", "1")
    assert rc == 1 and "unrecognised option '--no_such_option'" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "dm_mass", "--output.resolution.x", "12x")
    assert rc == 1 and "cannot convert" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "dm_mass", "--quantity.dm.particle_mass", "0")
    assert rc == 1 and "Dark matter particle mass must be > 0" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "dm_mass", "--integration.relative_tolerance",
                  "0")
    assert rc == 1 and "integration tolerance" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "xray_flux", "--quantity.xray_flux.min_energy",
                  "5", "--quantity.xray_flux.max_energy", "1")
    assert rc == 1 and "End of integration range must be >" in out


def test_quantities_and_targets_outside_this_build(snap):
    rc, out = run("-i", snap, "-t", "image")  # the reference's default quantity
    assert rc == 1 and "chandra_count_rate is not reconstructed by this build" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "no_such_quantity")
    assert rc == 1 and "Unknwon quantity: no_such_quantity" in out
    rc, out = run("-i", snap, "-t", "xray_spectrum", "-q", "dm_mass")
    assert rc == 1 and "needs the Voronoi reconstruction" in out
    rc, out = run("-i", snap, "-t", "movie", "-q", "dm_mass")
    assert rc == 1 and "Invalid render target: movie" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "dm_mass", "--reconstructor.voronoi", "nn8")
    assert rc == 1 and "Unknown voronoi reconstructor: nn8" in out
    rc, out = run("-i", snap, "-t", "image", "-q", "xray_flux",
                  "--quantity.xray_flux.gaunt_table", "/no/such/gaunt_factor.dat")
    assert rc == 1 and "Could not load data table" in out


def test_missing_snapshot(snap):
    rc, out = run("-i", snap + ".absent", "-t", "image", "-q", "dm_mass")
    assert rc == 1 and "Error: HDF5" in out
