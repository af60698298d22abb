"""CPU oracle for illcrawl's projective SPH smoothing path (numpy bindings).

TEST INFRASTRUCTURE ONLY: this package is the checker.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to check or time the reference algorithm; the product
(``illcrawl_amd``) never imports it.

The arithmetic lives in ``oracle.c`` (each function cites the illcrawl source
it restates); this module only marshals numpy arrays through ctypes.

Pinning (which check fixes which function) is listed in ``PINNING`` below and
exercised by ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ILC_ORACLE_LIB selects another build of oracle.c (tests/sanitized_run.py)
_LIB_PATH = os.environ.get("ILC_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

K_QUARTIC_REF = 0
K_QUARTIC = 1
K_CUBIC = 2

# function -> the checks that pin it (tests/test_oracle.py)
PINNING = {
    "orc_camera_setup": "look_at=+z closed form (b0=x, b1=y); orthonormality and "
                        "roll rotation for random cameras",
    "orc_projection_matrix": "inverse of an orthonormal basis = its transpose (FP32 tol)",
    "orc_project_particles": "depth mask known answers; z := 0; transpose closed form",
    "orc_pixel_points": "corner convention min + i*ps*b (F8) closed form",
    "orc_grid_params": "hand-computed bbox/+-0.1/cbrt/ceil example",
    "orc_grid_build": "keys vs numpy float32 floor; order vs numpy stable argsort; "
                      "per-cell max vs numpy; cell ranges vs bincount",
    "orc_quartic_projection": "W(0,h)=7/(pi h^2); W(r>=h)=0; integral over the plane = 2",
    "orc_gather_ref": "uniform-h scenes where the F3 radius drops nothing equal the "
                      "brute-force exact sum; hand-built scene where F3 drops a known "
                      "particle; F4 alias scene",
    "orc_gather_exact": "single-particle closed forms; integral = 2 (quartic_ref), "
                        "1 (quartic, cubic)",
    "orc_cubic_projected_unit": "integral over the plane = 1; value at 0 = 3/(2 pi) "
                                "(closed form of the M4 column)",
    "orc_count_evals_lattice": "equals the per-point brute-force count",
    "orc_tabulated_1d": "hand-computed half-texel and zero-border answers (F7)",
    "orc_tabulated_2d": "hand-computed bilinear half-texel answers",
    "orc_temperature": "closed form (gamma-1) u mu 121.1475502",
    "orc_xray_bolometric": "closed form on table-node-free inputs",
    "orc_quartic3d": "W(0) = 1.04445/h^3, W(r >= h) = 0, 4 pi int r^2 W dr = 1.04445 * 32 pi / 105",
    "orc_smooth3d_exact": "single-particle closed form; lattice volume integral = w",
    "orc_smooth3d_ref": "uniform-h scenes where the search box drops nothing equal "
                        "orc_smooth3d_exact up to the 1.04445f constant; a hand-built scene "
                        "where the search box drops a known particle",
    "orc_backward_exact": "FP64 central differences of orc_gather_exact in w, h, x, y "
                          "(quartic_ref, quartic, cubic)",
    "orc_xray_flux": "constant Gaunt table: band integral of exp(-E/kT) in closed form "
                     "plus exactly the bias of the reference's 2197/4101 weight, "
                     "h (2197/4101 - 2197/4104) f(E + 12h/13) per step, on one- and "
                     "two-step bands; identical step sequences and integrals scaled by "
                     "lambda when the integrand is scaled by lambda (relative tolerance); "
                     "a table linear in its indices recovers the sampling coordinates "
                     "ln(gamma^2), ln(u) exactly",
    "orc_xray_spectral_flux": "closed form of the bremsstrahlung prefactor chain at one "
                              "energy with a constant table; the reference's van Hoof "
                              "table (gaunt_factor.dat): its entries at texel centres and "
                              "the mean of the four entries around each grid node",
    "orc_gather_clip": "a depth range holding the whole support equals orc_gather_exact; "
                       "complementary ranges add up to the whole; a cut through a "
                       "particle's centre gives half its column; a thin slab gives its "
                       "thickness times the 3-D quartic (1 - q^2)^2 K (15/16) / h^3",
    "orc_project_particles_depth": "equals orc_project_particles in x', y' and z' = the "
                                   "depth it masks on",
    "orc_crawl_block": "hand-built rows: periodic images moved in by one period (each "
                       "axis, both signs), border tolerance, inclusive volume faces, "
                       "input order kept; numpy brute force on random rows",
}


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile (gcc, no contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u64 = ctypes.c_uint64


def _declare(L):
    L.orc_camera_setup.argtypes = [_f64p, _f64p, ctypes.c_double, ctypes.c_double, _u64, _u64,
                                   _f64p, _f64p, _f64p, ctypes.POINTER(ctypes.c_double)]
    L.orc_camera_setup.restype = None
    L.orc_projection_matrix.argtypes = [_f32p, _f32p, _f32p, _f32p]
    L.orc_projection_matrix.restype = None
    L.orc_project_particles.argtypes = [_f32p, _u64, _f32p, _f32p, ctypes.c_float, _f32p]
    L.orc_project_particles.restype = None
    L.orc_project_particles_depth.argtypes = [_f32p, _u64, _f32p, _f32p, _f32p]
    L.orc_project_particles_depth.restype = None
    L.orc_gather_clip.argtypes = [_f32p, _f32p, _u64, _f32p, _u64, ctypes.c_double,
                                  ctypes.c_int, _f64p]
    L.orc_gather_clip.restype = None
    L.orc_pixel_points.argtypes = [_u64, _u64, ctypes.c_float, _f32p, _f32p, _f32p, _f32p,
                                   _f32p, _f32p]
    L.orc_pixel_points.restype = None
    L.orc_grid_params.argtypes = [_f32p, _u64, _u64, _i32p, _f32p, _f32p]
    L.orc_grid_params.restype = ctypes.c_int
    L.orc_cell_keys.argtypes = [_f32p, _u64, _i32p, _f32p, _f32p, _u64p]
    L.orc_cell_keys.restype = None
    L.orc_grid_build.argtypes = [_f32p, _f32p, _u64, _i32p, _f32p, _f32p, _u32p, _i32p,
                                 _f32p, _f32p, _f32p, ctypes.POINTER(ctypes.c_float)]
    L.orc_grid_build.restype = ctypes.c_int
    L.orc_quartic_projection.argtypes = [ctypes.c_float, ctypes.c_float]
    L.orc_quartic_projection.restype = ctypes.c_float
    L.orc_gather_ref.argtypes = [_i32p, _f32p, _f32p, _i32p, _f32p, _f32p, _f32p,
                                 ctypes.c_float, _f32p, _u64, ctypes.c_int, _f64p,
                                 ctypes.c_void_p]
    L.orc_gather_ref.restype = None
    L.orc_quartic3d.argtypes = [ctypes.c_float, ctypes.c_float]
    L.orc_quartic3d.restype = ctypes.c_float
    L.orc_smooth3d_ref.argtypes = [_i32p, _f32p, _f32p, _i32p, _f32p, _f32p, _f32p,
                                   ctypes.c_float, _f32p, _u64, ctypes.c_int, ctypes.c_int,
                                   _f64p]
    L.orc_smooth3d_ref.restype = None
    L.orc_smooth3d_exact.argtypes = [_f32p, _f32p, _u64, _f32p, _u64, _f64p]
    L.orc_smooth3d_exact.restype = None
    L.orc_gather_exact.argtypes = [_f32p, _f32p, _u64, _f32p, _u64, ctypes.c_int, _f64p,
                                   ctypes.c_void_p]
    L.orc_gather_exact.restype = None
    L.orc_cubic_projected_unit.argtypes = [ctypes.c_double]
    L.orc_cubic_projected_unit.restype = ctypes.c_double
    L.orc_count_evals_lattice.argtypes = [_f32p, _f32p, _u64, _f32p, _u64, _u64,
                                          ctypes.c_double]
    L.orc_count_evals_lattice.restype = ctypes.c_uint64
    L.orc_temperature.argtypes = [ctypes.c_float, ctypes.c_float]
    L.orc_temperature.restype = ctypes.c_float
    L.orc_tabulated_1d.argtypes = [_f32p, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float]
    L.orc_tabulated_1d.restype = ctypes.c_float
    L.orc_tabulated_2d.argtypes = [_f32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float]
    L.orc_tabulated_2d.restype = ctypes.c_float
    L.orc_xray_bolometric.argtypes = [_f32p, _f32p, _f32p, _u64, _f32p, ctypes.c_int32,
                                      ctypes.c_float, ctypes.c_float, _f32p]
    L.orc_xray_bolometric.restype = None
    L.orc_backward_exact.argtypes = [_f32p, _f32p, _u64, _f32p, _u64, _f64p, ctypes.c_int, _f64p,
                                     _f64p, _f64p, _f64p]
    L.orc_backward_exact.restype = None
    L.orc_crawl_block.argtypes = [ctypes.c_double * 3, ctypes.c_double * 3, ctypes.c_double * 3,
                                  ctypes.c_double, _f32p, _u64, ctypes.c_void_p, ctypes.c_int32,
                                  _f32p, ctypes.c_void_p, _u64p]
    L.orc_crawl_block.restype = ctypes.c_uint64
    L.orc_xray_flux.argtypes = [_f32p, _f32p, _f32p, _u64, ctypes.c_float, ctypes.c_float, _f32p,
                                ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                ctypes.c_int32, _f32p, _i32p]
    L.orc_xray_flux.restype = None
    L.orc_xray_spectral_flux.argtypes = [_f32p, _f32p, _f32p, _u64, ctypes.c_float,
                                         ctypes.c_float, _f32p, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_float, ctypes.c_float, _f32p]
    L.orc_xray_spectral_flux.restype = None
    L.orc_num_threads.argtypes = []
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_set_num_threads.argtypes = [ctypes.c_int]
    L.orc_set_num_threads.restype = None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- camera

class Camera:
    """Host camera state as the reference computes it (double precision)."""

    def __init__(self, position, look_at, roll, width, nx, ny):
        self.position = _c(position, np.float64).copy()
        self.look_at = _c(look_at, np.float64).copy()
        self.roll = float(roll)
        self.width = float(width)
        self.nx, self.ny = int(nx), int(ny)
        b0 = np.zeros(3)
        b1 = np.zeros(3)
        mp = np.zeros(3)
        ps = ctypes.c_double()
        lib().orc_camera_setup(self.position, self.look_at, self.roll, self.width,
                               self.nx, self.ny, b0, b1, mp, ctypes.byref(ps))
        self.b0, self.b1, self.min_position, self.pixel_size = b0, b1, mp, ps.value

    def projection_matrix(self):
        m = np.zeros(9, np.float32)
        lib().orc_projection_matrix(self.b0.astype(np.float32), self.b1.astype(np.float32),
                                    self.look_at.astype(np.float32), m)
        return m

    def pixel_points(self):
        """Projected pixel corner points, float32 (P, 4), index iy*nx+ix."""
        out = np.zeros((self.nx * self.ny, 4), np.float32)
        lib().orc_pixel_points(self.nx, self.ny, np.float32(self.pixel_size),
                               self.min_position.astype(np.float32),
                               self.b0.astype(np.float32), self.b1.astype(np.float32),
                               self.projection_matrix(), self.position.astype(np.float32),
                               out)
        return out

    def project_particles(self, p4, depth):
        p4 = _c(p4, np.float32)
        out = np.zeros_like(p4)
        lib().orc_project_particles(p4, p4.shape[0], self.projection_matrix(),
                                    self.position.astype(np.float32), np.float32(depth), out)
        return out

    def project_particles_depth(self, p4):
        """(x', y', z', w): the projection keeping the camera depth, w unmasked."""
        p4 = _c(p4, np.float32)
        out = np.zeros_like(p4)
        lib().orc_project_particles_depth(p4, p4.shape[0], self.projection_matrix(),
                                          self.position.astype(np.float32), out)
        return out


# ---------------------------------------------------------------- grid

class Grid:
    """Reference particle grid (restated): parameters, order, cells, max h."""

    def __init__(self, p4, h, target=1000, params=None):
        p4 = _c(p4, np.float32)
        h = _c(h, np.float32)
        n = p4.shape[0]
        if params is None:
            nc = np.zeros(3, np.int32)
            minc = np.zeros(3, np.float32)
            cs = np.zeros(3, np.float32)
            rc = lib().orc_grid_params(p4, n, int(target), nc, minc, cs)
            if rc != 0:
                raise ValueError("orc_grid_params failed (%d)" % rc)
        else:
            nc, minc, cs = (_c(params[0], np.int32), _c(params[1], np.float32),
                            _c(params[2], np.float32))
        self.nc, self.minc, self.cs = nc, minc, cs
        ncells = int(nc[0]) * int(nc[1]) * int(nc[2])
        self.perm = np.zeros(n, np.uint32)
        self.cells = np.zeros((ncells, 2), np.int32)
        self.sorted_p4 = np.zeros_like(p4)
        self.sorted_h = np.zeros_like(h)
        self.maxh_cell = np.zeros(ncells, np.float32)
        mh = ctypes.c_float()
        rc = lib().orc_grid_build(p4, h, n, nc, minc, cs, self.perm, self.cells,
                                  self.sorted_p4, self.sorted_h, self.maxh_cell,
                                  ctypes.byref(mh))
        if rc != 0:
            raise ValueError("orc_grid_build failed (%d)" % rc)
        self.max_h = np.float32(mh.value)

    @property
    def ncells(self):
        return int(self.nc[0]) * int(self.nc[1]) * int(self.nc[2])

    def keys(self, p4):
        p4 = _c(p4, np.float32)
        out = np.zeros(p4.shape[0], np.uint64)
        lib().orc_cell_keys(p4, p4.shape[0], self.nc, self.minc, self.cs, out)
        return out

    def gather_ref(self, pts4, acc_mode=1, with_pairs=False):
        """Reference gather (F3/F4 semantics).  acc_mode 0: FP32 accumulator
        (the reference); 1: FP64 accumulation of the same FP32 terms."""
        pts4 = _c(pts4, np.float32)
        out = np.zeros(pts4.shape[0], np.float64)
        pairs = np.zeros(pts4.shape[0], np.uint64) if with_pairs else None
        lib().orc_gather_ref(self.nc, self.minc, self.cs, self.cells, self.sorted_p4,
                             self.sorted_h, self.maxh_cell, np.float32(self.max_h), pts4,
                             pts4.shape[0], int(acc_mode), out,
                             pairs.ctypes.data if pairs is not None else None)
        return (out, pairs) if with_pairs else out


    def smooth3d_ref(self, pts4, acc_mode=1, form=0):
        """Volumetric smoothing with the reference's search (dm::grid).
        form 0: the reference's FP32 kernel expression; 1: the same kernel
        in FP64 from the stable form (1 - r^2/h^2)^2."""
        pts4 = _c(pts4, np.float32)
        out = np.zeros(pts4.shape[0], np.float64)
        lib().orc_smooth3d_ref(self.nc, self.minc, self.cs, self.cells, self.sorted_p4,
                               self.sorted_h, self.maxh_cell, np.float32(self.max_h), pts4,
                               pts4.shape[0], int(acc_mode), int(form), out)
        return out


def smooth3d_exact(p4, h, pts4):
    p4 = _c(p4, np.float32)
    h = _c(h, np.float32)
    pts4 = _c(pts4, np.float32)
    out = np.zeros(pts4.shape[0], np.float64)
    lib().orc_smooth3d_exact(p4, h, p4.shape[0], pts4, pts4.shape[0], out)
    return out


def quartic3d(r, h):
    return lib().orc_quartic3d(np.float32(r), np.float32(h))


# ---------------------------------------------------------------- exact

def gather_exact(p4, h, pts4, kind=K_QUARTIC_REF, with_counts=False):
    p4 = _c(p4, np.float32)
    h = _c(h, np.float32)
    pts4 = _c(pts4, np.float32)
    out = np.zeros(pts4.shape[0], np.float64)
    cnt = np.zeros(pts4.shape[0], np.uint64) if with_counts else None
    lib().orc_gather_exact(p4, h, p4.shape[0], pts4, pts4.shape[0], int(kind), out,
                           cnt.ctypes.data if cnt is not None else None)
    return (out, cnt) if with_counts else out


def gather_clip(p4z, h, pts4, depth, kind=K_QUARTIC_REF):
    """Line-of-sight clipped projection over [0, depth] (FP64 quadrature);
    p4z from Camera.project_particles_depth."""
    p4z = _c(p4z, np.float32)
    h = _c(h, np.float32)
    pts4 = _c(pts4, np.float32)
    out = np.zeros(pts4.shape[0], np.float64)
    lib().orc_gather_clip(p4z, h, p4z.shape[0], pts4, pts4.shape[0], float(depth), int(kind),
                          out)
    return out


def count_evals_lattice(p4, h, pts4, nx, ny, support_factor=1.0):
    p4 = _c(p4, np.float32)
    h = _c(h, np.float32)
    pts4 = _c(pts4, np.float32)
    return int(lib().orc_count_evals_lattice(p4, h, p4.shape[0], pts4, int(nx), int(ny),
                                             float(support_factor)))


def quartic_projection(r, h):
    return lib().orc_quartic_projection(np.float32(r), np.float32(h))


def cubic_projected_unit(b):
    return lib().orc_cubic_projected_unit(float(b))


# ---------------------------------------------------------------- quantities

def temperature(u, xe):
    return lib().orc_temperature(np.float32(u), np.float32(xe))


def tabulated_1d(y, x_min, dx, x):
    y = _c(y, np.float32)
    return lib().orc_tabulated_1d(y, y.shape[0], np.float32(x_min), np.float32(dx),
                                  np.float32(x))


def tabulated_2d(data, x_min, y_min, dx, dy, x, y):
    """data has shape (ny, nx) (row iy holds data[ix + nx*iy])."""
    data = _c(data, np.float32)
    ny, nx = data.shape
    return lib().orc_tabulated_2d(data.reshape(-1), nx, ny, np.float32(x_min),
                                  np.float32(y_min), np.float32(dx), np.float32(dy),
                                  np.float32(x), np.float32(y))


def xray_bolometric(rho, temp, xe, gtab, gx0, gdx):
    rho = _c(rho, np.float32)
    temp = _c(temp, np.float32)
    xe = _c(xe, np.float32)
    gtab = _c(gtab, np.float32)
    out = np.zeros_like(rho)
    lib().orc_xray_bolometric(rho, temp, xe, rho.shape[0], gtab, gtab.shape[0],
                              np.float32(gx0), np.float32(gdx), out)
    return out


def backward_exact(p4, h, pts4, G, kind=K_QUARTIC_REF):
    """FP64 gradients (dL/dw, dL/dh, dL/dx, dL/dy) of sum_t G[t] out[t]."""
    p4 = _c(p4, np.float32)
    h = _c(h, np.float32)
    pts4 = _c(pts4, np.float32)
    G = _c(G, np.float64)
    n = p4.shape[0]
    out = [np.zeros(n, np.float64) for _ in range(4)]
    lib().orc_backward_exact(p4, h, n, pts4, pts4.shape[0], G, int(kind), *out)
    return tuple(out)


def crawl_block(center, extent, period, border, xyz, cols=()):
    """Kept rows of one crawler block: (particles4, [columns], input indices)."""
    xyz = _c(xyz, np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    cols = [_c(c, np.float32).reshape(-1) for c in cols]
    out4 = np.zeros((n, 4), np.float32)
    outs = [np.zeros(n, np.float32) for _ in cols]
    idx = np.zeros(n, np.uint64)
    d3 = ctypes.c_double * 3
    cin = (ctypes.c_void_p * max(len(cols), 1))(*[c.ctypes.data for c in cols])
    cout = (ctypes.c_void_p * max(len(cols), 1))(*[c.ctypes.data for c in outs])
    m = lib().orc_crawl_block(d3(*center), d3(*extent), d3(*period), ctypes.c_double(border),
                              xyz, n, ctypes.cast(cin, ctypes.c_void_p), len(cols), out4,
                              ctypes.cast(cout, ctypes.c_void_p), idx)
    return out4[:m], [o[:m] for o in outs], idx[:m]


def xray_flux(rho, u, xe, z, dl, table, e_min, e_max, num_samples):
    """xray_flux: band-integrated bremsstrahlung flux per particle (RKF in
    energy); table is the 2-D Gaunt table as (rows = ln u, cols = ln gamma^2).
    Returns (flux, RKF steps)."""
    rho, u, xe = (_c(a, np.float32).reshape(-1) for a in (rho, u, xe))
    tab = _c(table, np.float32)
    ny, nx = tab.shape
    out = np.zeros(rho.shape[0], np.float32)
    steps = np.zeros(rho.shape[0], np.int32)
    lib().orc_xray_flux(rho, u, xe, rho.shape[0], float(z), float(dl), tab.reshape(-1), nx, ny,
                        float(e_min), float(e_max), int(num_samples), out, steps)
    return out, steps


def xray_spectral_flux(rho, u, xe, z, dl, table, E, dE):
    rho, u, xe = (_c(a, np.float32).reshape(-1) for a in (rho, u, xe))
    tab = _c(table, np.float32)
    ny, nx = tab.shape
    out = np.zeros(rho.shape[0], np.float32)
    lib().orc_xray_spectral_flux(rho, u, xe, rho.shape[0], float(z), float(dl), tab.reshape(-1),
                                 nx, ny, float(E), float(dE), out)
    return out


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))
