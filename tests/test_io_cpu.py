"""Snapshot input and image output of the data crawler, on the CPU:

- FITS: the numpy writer and reader of illcrawl_amd.io against themselves and
  against cfitsio (the reference's FITS library, fits.hpp:210-300) through
  the C++ program tests/cpp/crawl_demo.cpp;
- HDF5: an Illustris-layout snapshot written by the C++ program (HDF5's C++
  API, the reference's) read back by the ctypes loader of illcrawl_amd.io,
  whole and in row blocks (the crawler's streaming, hdf5_io.hpp:39-107)."""
import os
import subprocess

import numpy as np
import pytest

from illcrawl_amd import io as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "illcrawl_amd", "lib", "crawl_demo")


def need_demo():
    if not os.path.exists(DEMO):
        pytest.skip("crawl_demo not built (HDF5 C++ / cfitsio absent)")


def test_fits_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    for dt in (np.float32, np.float64):
        a = rng.normal(size=(13, 29)).astype(dt)
        p = str(tmp_path / f"a_{np.dtype(dt).name}.fits")
        I.write_fits(p, a)
        raw = open(p, "rb").read()
        assert len(raw) % 2880 == 0
        head = raw[:2880].decode("ascii")
        assert head.startswith("SIMPLE  =                    T")
        cards = [head[i:i + 80] for i in range(0, 2880, 80)]
        keys = {c[:8].strip(): c[10:30].strip() for c in cards if c[8:10] == "= "}
        assert keys["BITPIX"] == ("-32" if dt == np.float32 else "-64")
        assert keys["NAXIS1"] == "29" and keys["NAXIS2"] == "13"  # NAXIS1: the fastest index
        b = I.read_fits(p)
        assert b.dtype == dt and b.shape == a.shape and np.array_equal(a, b)


def test_fits_interoperates_with_cfitsio(tmp_path):
    need_demo()
    rng = np.random.default_rng(4)
    img = rng.uniform(-5, 5, size=(7, 11)).astype(np.float32)
    raw, cf, py, back = (str(tmp_path / n) for n in ("img.f32", "cf.fits", "py.fits", "back.f32"))
    img.tofile(raw)
    # cfitsio writes (and checks its own read back); numpy reads
    r = subprocess.run([DEMO, "write-fits", raw, "11", "7", cf], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(I.read_fits(cf), img)
    # numpy writes; cfitsio reads
    I.write_fits(py, img)
    r = subprocess.run([DEMO, "fits-to-raw", py, back], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "extents 11 7" in r.stdout
    assert np.array_equal(np.fromfile(back, dtype=np.float32).reshape(7, 11), img)


def write_snapshot(d, n, seed=5):
    rng = np.random.default_rng(seed)
    cols = dict(pos=rng.uniform(-10.0, 110.0, (n, 3)),
                rho=rng.lognormal(0.0, 1.0, n).astype(np.float32),
                u=rng.uniform(1e3, 1e5, n).astype(np.float32),
                xe=rng.uniform(1.0, 1.2, n).astype(np.float32),
                mass=rng.lognormal(-2.0, 0.3, n).astype(np.float32),
                hsml=rng.uniform(2.0, 6.0, n).astype(np.float32))
    cols["pos"].astype(np.float64).tofile(os.path.join(d, "pos.f64"))
    for k in ("rho", "u", "xe", "mass", "hsml"):
        cols[k].tofile(os.path.join(d, k + ".f32"))
    r = subprocess.run([DEMO, "write-snapshot", d, str(n)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return cols


def test_hdf5_snapshot_read_whole_and_in_blocks(tmp_path):
    need_demo()
    try:
        I._h5()
    except I.HDF5Missing as e:
        pytest.skip(str(e))
    d = str(tmp_path)
    n = 2345
    cols = write_snapshot(d, n)
    with I.IllustrisDataLoader(os.path.join(d, "snap.hdf5"), 0) as L:
        assert L.current_group_name == "PartType0"
        assert L.has_dataset("Coordinates") and not L.has_dataset("SubfindHsml")
        c = L.get_dataset(I.IllustrisDataLoader.COORDINATES)
        assert c.shape == (n, 3)
        assert np.array_equal(c.read(), cols["pos"])           # float64 kept exactly
        for name, key in (("Density", "rho"), ("InternalEnergy", "u"),
                          ("ElectronAbundance", "xe"), ("Masses", "mass"),
                          ("SmoothingLength", "hsml")):
            ds = L.get_dataset(name)
            assert ds.shape == (n,)
            assert np.array_equal(ds.read().astype(np.float32), cols[key])
        # row blocks cover every row once, in order
        ds = [c, L.get_dataset("Density")]
        seen = 0
        for first, (xyz, rho) in I.stream_blocks(ds, 1000):
            assert first == seen
            assert np.array_equal(xyz, cols["pos"][first:first + len(rho)])
            assert np.array_equal(rho.astype(np.float32), cols["rho"][first:first + len(rho)])
            seen += len(rho)
        assert seen == n
        with pytest.raises(IndexError):
            c.read(n - 1, 2)
        with pytest.raises(OSError):
            L.get_dataset("NoSuchDataset")
