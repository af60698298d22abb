"""Host-side mirror of illcrawl's projective smoothing backend interface.

``ProjectiveSmoothingGrid`` has the method set of
``reconstruction_backends::dm::projective_smoothing_grid``
(reconstruction_backend.hpp:35-58, projective_smoothing_backend.hpp:35-42) and
``ProjectiveSmoothing`` the one of the camera wrapper
``reconstruction_backends::dm::projective_smoothing``
(projective_smoothing_reconstruction.hpp:33-89).  Buffers are torch CUDA
tensors (device memory and streams only); all work runs in the HIP library
through the C ABI.  Errors raise ``IllcrawlError`` (a RuntimeError, like the
reference's qcl_error / std::runtime_error).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N

_MODES = {"reference": N.MODE_REFERENCE, "exact": N.MODE_EXACT}
_KERNELS = {"quartic_ref": N.KERNEL_QUARTIC_REF, "quartic": N.KERNEL_QUARTIC,
            "cubic": N.KERNEL_CUBIC}


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev_f32(x, device, shape_last=None):
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
    t = x.to(device=device, dtype=torch.float32).contiguous()
    if shape_last is not None and (t.dim() != 2 or t.shape[1] != shape_last):
        raise ValueError(f"expected an (n, {shape_last}) array, got {tuple(t.shape)}")
    return t


class ProjectiveSmoothingGrid:
    """dm::projective_smoothing_grid on the HIP path (target 1000 particles
    per cell, projective_smoothing_backend_grid.cpp:27-31)."""

    def __init__(self, device=0, target_particles_per_cell=1000, mode="reference",
                 kernel="quartic_ref", tile=None):
        cfg = N.PsgConfig()
        cfg.device = int(device)
        cfg.target_particles_per_cell = int(target_particles_per_cell)
        cfg.mode = _MODES[mode]
        cfg.kernel = _KERNELS[kernel]
        if tile is not None:
            cfg.tile_w, cfg.tile_h = int(tile[0]), int(tile[1])
        self._h = ctypes.c_void_p()
        N.call("ilc_psg_create", ctypes.byref(cfg), ctypes.byref(self._h))
        self.device = torch.device("cuda", int(device))
        self.mode, self.kernel = mode, kernel
        self._keep = []       # borrowed buffers kept alive while in use
        self._npoints = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                N.lib().ilc_psg_destroy(h)
            except Exception:
                pass
            self._h = None

    # reconstruction_backend interface ---------------------------------
    def get_backend_name(self):
        return N.lib().ilc_psg_backend_name(self._h).decode()

    def init_backend(self, blocksize):
        N.call("ilc_psg_init_backend", self._h, int(blocksize))

    def setup_projected_particles(self, particles4, smoothing_lengths, stream=None):
        p = _dev_f32(particles4, self.device, 4)
        h = _dev_f32(smoothing_lengths, self.device).reshape(-1)
        if h.shape[0] != p.shape[0]:
            raise ValueError("one smoothing length per particle is required")
        N.call("ilc_psg_setup_projected_particles", self._h, _ptr(p), p.shape[0], _ptr(h),
               _stream(stream))

    def setup_particles(self, particles4, additional_datasets, stream=None):
        """reconstruction_backend::setup_particles: the one additional dataset
        of this backend is the smoothing length."""
        if len(additional_datasets) != 1:
            raise ValueError("the grid backend takes exactly one additional dataset (h)")
        self.setup_projected_particles(particles4, additional_datasets[0], stream)

    def setup_evaluation_points(self, points4, layout_nx=0, stream=None):
        pts = _dev_f32(points4, self.device, 4)
        self._keep = [pts]
        self._npoints = pts.shape[0]
        N.call("ilc_psg_setup_evaluation_points", self._h, _ptr(pts), pts.shape[0],
               int(layout_nx), _stream(stream))

    def run(self, stream=None):
        N.call("ilc_psg_run", self._h, _stream(stream))

    def retrieve_results(self, scale=1.0, stream=None):
        out = torch.empty(self._npoints, dtype=torch.float32, device=self.device)
        N.call("ilc_psg_retrieve_results", self._h, ctypes.c_float(scale), _ptr(out),
               _stream(stream))
        return out

    def retrieve_results_f64(self, stream=None):
        out = torch.empty(self._npoints, dtype=torch.float64, device=self.device)
        N.call("ilc_psg_retrieve_results_f64", self._h, _ptr(out), _stream(stream))
        return out

    def backward(self, dl_dvalue, scale=1.0, stream=None):
        """Adjoint of run + retrieve_results(scale) (exact mode): returns
        (dL/dw, dL/dh, dL/dx, dL/dy) per particle in input order, x and y in
        the projected frame."""
        g = _dev_f32(dl_dvalue, self.device).reshape(-1)
        if g.shape[0] != self._npoints:
            raise ValueError(f"expected {self._npoints} point gradients, got {g.shape[0]}")
        n = self.grid_info()["num_particles"]
        out = torch.empty((4, n), dtype=torch.float32, device=self.device)
        N.call("ilc_psg_backward", self._h, _ptr(g), ctypes.c_float(scale), _ptr(out[0]),
               _ptr(out[1]), _ptr(out[2]), _ptr(out[3]), _stream(stream))
        return out[0], out[1], out[2], out[3]

    # inspection ------------------------------------------------------------
    def grid_info(self):
        gi = N.GridInfo()
        N.call("ilc_psg_grid_info", self._h, ctypes.byref(gi))
        return dict(num_cells=list(gi.num_cells), min_corner=list(gi.min_corner),
                    cell_size=list(gi.cell_size), max_h=gi.max_h,
                    num_particles=gi.num_particles)

    def export_grid(self, stream=None):
        gi = self.grid_info()
        ncells = int(np.prod(gi["num_cells"]))
        cells = torch.empty((ncells, 2), dtype=torch.int32, device=self.device)
        maxh = torch.empty(ncells, dtype=torch.float32, device=self.device)
        perm = torch.empty(gi["num_particles"], dtype=torch.int32, device=self.device)
        N.call("ilc_psg_export_grid", self._h, _ptr(cells), _ptr(maxh), _ptr(perm),
               _stream(stream))
        return cells, maxh, perm

    def count_evals(self, per_point=False, stream=None):
        total = ctypes.c_uint64()
        pp = torch.empty(self._npoints, dtype=torch.int64, device=self.device) \
            if per_point else None
        N.call("ilc_psg_count_evals", self._h, _ptr(pp), ctypes.byref(total), _stream(stream))
        return (total.value, pp) if per_point else total.value


# ------------------------------------------------------------ camera wrapper

class GridSmoothing:
    """reconstruction_backends::dm::grid (dm_reconstruction_backend_grid.cpp:57-134):
    volumetric smoothing of 3-D particles at 3-D points with the 3-D quartic
    (smoothing_kernel.cl:27-37), e.g. for volumetric slices.  Results
    accumulate over run() calls until the next setup_evaluation_points.  The
    reference's default is 64 particles per cell
    (dm_reconstruction_backend_grid.hpp:39)."""

    def __init__(self, device=0, target_particles_per_cell=64, mode="reference",
                 kernel="quartic_ref"):
        cfg = N.PsgConfig()
        cfg.device = int(device)
        cfg.target_particles_per_cell = int(target_particles_per_cell)
        cfg.mode = _MODES[mode]
        cfg.kernel = _KERNELS[kernel]
        self._h = ctypes.c_void_p()
        N.call("ilc_grid3d_create", ctypes.byref(cfg), ctypes.byref(self._h))
        self.device = torch.device("cuda", int(device))
        self.mode, self.kernel = mode, kernel
        self._keep = []
        self._npoints = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                N.lib().ilc_grid3d_destroy(h)
            except Exception:
                pass
            self._h = None

    def get_backend_name(self):
        return "dm_smoothing/grid3d (hip gfx950)"

    def init_backend(self, blocksize):
        N.call("ilc_grid3d_init_backend", self._h, int(blocksize))

    def setup_particles(self, particles4, additional_datasets, stream=None):
        """particles4 (n, 4): x, y, z, w; additional_datasets: [h]."""
        if len(additional_datasets) != 1:
            raise ValueError("the grid backend takes exactly one additional dataset (h)")
        p = _dev_f32(particles4, self.device, 4)
        h = _dev_f32(additional_datasets[0], self.device).reshape(-1)
        if h.shape[0] != p.shape[0]:
            raise ValueError("one smoothing length per particle is required")
        N.call("ilc_grid3d_setup_particles", self._h, _ptr(p), p.shape[0], _ptr(h),
               _stream(stream))

    def setup_evaluation_points(self, points4, layout_nx=0, stream=None):
        pts = _dev_f32(points4, self.device, 4)
        self._keep = [pts]
        self._npoints = pts.shape[0]
        N.call("ilc_grid3d_setup_evaluation_points", self._h, _ptr(pts), pts.shape[0],
               int(layout_nx), _stream(stream))

    def run(self, stream=None):
        N.call("ilc_grid3d_run", self._h, _stream(stream))

    def retrieve_results(self, scale=1.0, stream=None):
        out = torch.empty(self._npoints, dtype=torch.float32, device=self.device)
        N.call("ilc_grid3d_retrieve_results", self._h, ctypes.c_float(scale), _ptr(out),
               _stream(stream))
        return out

    def retrieve_results_f64(self, stream=None):
        out = torch.empty(self._npoints, dtype=torch.float64, device=self.device)
        N.call("ilc_grid3d_retrieve_results_f64", self._h, _ptr(out), _stream(stream))
        return out

    def grid_info(self):
        gi = N.GridInfo()
        N.call("ilc_grid3d_grid_info", self._h, ctypes.byref(gi))
        return dict(num_cells=list(gi.num_cells), min_corner=list(gi.min_corner),
                    cell_size=list(gi.cell_size), max_h=gi.max_h,
                    num_particles=gi.num_particles)

    def count_evals(self, per_point=False, stream=None):
        """Exact count of (point, particle) pairs with w != 0 and the point
        inside the particle's support, for the current grid."""
        total = ctypes.c_uint64()
        pp = torch.empty(self._npoints, dtype=torch.int64, device=self.device) \
            if per_point else None
        N.call("ilc_grid3d_count_evals", self._h, _ptr(pp), ctypes.byref(total),
               _stream(stream))
        return (total.value, pp) if per_point else total.value


def slice_pixel_points(cam, num_additional_samples=0, seed=0):
    """Evaluation points of volumetric_slice::create_slice
    (volumetric_slice.cpp:51-76): camera::get_pixel_coordinate(x, y) =
    min_position + x ps b0 + y ps b1 in double, stored as float32 (w = 0),
    row-major; then num_additional_samples more copies jittered by up to half
    a pixel (numpy's seeded generator; the reference draws from
    std::random_device).  Returns (samples * nx * ny, 4) float32."""
    b0, b1, mn, ps = camera_basis(cam)
    nx, ny = int(cam["nx"]), int(cam.get("ny", cam["nx"]))
    xs, ys = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64))
    xs, ys = [xs.reshape(-1)], [ys.reshape(-1)]
    rng = np.random.default_rng(seed)
    for _ in range(int(num_additional_samples)):
        j = rng.uniform(-0.5, 0.5, (ny * nx, 2))
        xs.append(xs[0] + j[:, 0])
        ys.append(ys[0] + j[:, 1])
    x, y = np.concatenate(xs), np.concatenate(ys)
    pts = (mn[None, :] + (x * ps)[:, None] * b0[None, :]) + (y * ps)[:, None] * b1[None, :]
    out = np.zeros((x.size, 4), np.float32)
    out[:, :3] = pts
    return out


def tomography_points(cam, initial_z_step, num_steps):
    """Evaluation points of the slices of
    volumetric_tomography::create_tomographic_cube (volumetric_tomography.cpp:
    61-102): slice k has the camera moved to position + (initial_z_step + k)
    ps look_at.  Returns (num_steps * nx * ny, 4) float32, slice-major."""
    ps = float(cam["width"]) / int(cam["nx"])
    pos = np.asarray(cam["position"], np.float64)
    la = np.asarray(cam["look_at"], np.float64)
    out = []
    for k in range(int(num_steps)):
        c = dict(cam)
        c["position"] = tuple(pos + (float(initial_z_step + k) * ps) * la)
        out.append(slice_pixel_points(c))
    return np.concatenate(out, 0) if out else np.zeros((0, 4), np.float32)


def _smooth_blocks(points, nx, blocks, blocksize, device, grid_kwargs):
    """The crawler's loop: one grid per block, values accumulated."""
    blocks = blocks() if callable(blocks) else blocks
    if blocksize is None:
        blocks = list(blocks)
        blocksize = max([int(p.shape[0]) for p, _ in blocks] + [1])
    g = GridSmoothing(device=device, **grid_kwargs)
    g.init_backend(int(blocksize))
    g.setup_evaluation_points(points, layout_nx=nx)
    for p4, h in blocks:
        if int(p4.shape[0]) == 0:
            continue
        g.setup_particles(p4, [h])
        g.run()
    return g.retrieve_results_f64().cpu().numpy()


def volumetric_slice(cam, blocks, dV=1.0, num_additional_samples=0, seed=0, blocksize=None,
                     device=0, **grid_kwargs):
    """volumetric_slice::create_slice over particle blocks: ``blocks`` yields
    (particles4, h) per block (world frame, the crawler's output; or a
    callable returning such an iterable); every block builds its own grid
    and the values accumulate.  dV: the quantity's effective dV of a pixel
    volume; blocksize: the largest block (default: found by listing the
    blocks).  Returns the (ny, nx) float32 slice, the mean over the samples
    of dV times the reconstruction."""
    nx, ny = int(cam["nx"]), int(cam.get("ny", cam["nx"]))
    samples = int(num_additional_samples) + 1
    pts = slice_pixel_points(cam, num_additional_samples, seed)
    vals = _smooth_blocks(pts, nx, blocks, blocksize, device, grid_kwargs).astype(np.float32)
    out = np.zeros(nx * ny, np.float32)
    for i in range(samples):
        # float += double, as output[idx] += dV * contribution / samples
        out = (out.astype(np.float64) + float(dV) * vals[i * nx * ny:(i + 1) * nx * ny]
               .astype(np.float64) / samples).astype(np.float32)
    return out.reshape(ny, nx)


def volumetric_tomography(cam, blocks, z_range=None, initial_z_step=0, num_steps=None, dV=1.0,
                          blocksize=None, device=0, max_points_per_pass=1 << 27, **grid_kwargs):
    """volumetric_tomography::create_tomographic_cube: num_steps slices
    (default int(z_range / pixel size), at least 1), the camera moved one
    pixel along look_at per slice.  The points of many slices go to one pass
    over the blocks (at most max_points_per_pass points), so ``blocks`` must
    be re-iterable (a sequence or a callable) when the cube needs several.
    Returns (num_steps, ny, nx) float32."""
    nx, ny = int(cam["nx"]), int(cam.get("ny", cam["nx"]))
    ps = float(cam["width"]) / nx
    if num_steps is None:
        num_steps = max(1, int(float(z_range) / ps))
    npix = nx * ny
    per = max(1, int(max_points_per_pass) // npix)
    out = np.zeros((int(num_steps), ny, nx), np.float32)
    for z0 in range(0, int(num_steps), per):
        nz = min(per, int(num_steps) - z0)
        pts = tomography_points(cam, initial_z_step + z0, nz)
        vals = _smooth_blocks(pts, nx, blocks, blocksize, device, grid_kwargs).astype(np.float32)
        out[z0:z0 + nz] = (float(dV) * vals.astype(np.float64)).astype(np.float32) \
            .reshape(nz, ny, nx)
    return out


def slice_points(center, normal, width, nx, ny=None, up=(0.0, 0.0, 1.0)):
    """Pixel centres of a square slice of the volume (volumetric_slice.cpp):
    nx x ny points, row-major, on the plane through ``center`` normal to
    ``normal``, ``width`` wide.  Returns (nx*ny, 4) float32 (w = 0)."""
    ny = nx if ny is None else ny
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    u = np.cross(np.asarray(up, np.float64), n)
    if np.linalg.norm(u) < 1e-12:
        u = np.cross(np.array([0.0, 1.0, 0.0]), n)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    ps = float(width) / nx
    ix = (np.arange(nx) + 0.5 - 0.5 * nx) * ps
    iy = (np.arange(ny) + 0.5 - 0.5 * ny) * ps
    X, Y = np.meshgrid(ix, iy)
    pts = (np.asarray(center, np.float64)[None, :] + X.reshape(-1, 1) * u[None, :]
           + Y.reshape(-1, 1) * v[None, :])
    return np.concatenate([pts, np.zeros((pts.shape[0], 1))], 1).astype(np.float32)


def camera_basis(cam):
    c = N.camera_struct(cam)
    b0, b1, mp = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    ps = ctypes.c_double()
    N.call("ilc_camera_basis", ctypes.byref(c), ctypes.byref(b0), ctypes.byref(b1),
           ctypes.byref(mp), ctypes.byref(ps))
    return np.array(b0[:]), np.array(b1[:]), np.array(mp[:]), ps.value


def camera_projection(cam):
    """The FP32 matrix and centre of project_particles: x' = minv (x - center)."""
    c = N.camera_struct(cam)
    m, ctr = (ctypes.c_float * 9)(), (ctypes.c_float * 3)()
    N.call("ilc_camera_projection", ctypes.byref(c), ctypes.byref(m), ctypes.byref(ctr))
    return (np.array(m[:], dtype=np.float32).reshape(3, 3),
            np.array(ctr[:], dtype=np.float32))


def camera_pixel_points(cam, device=0, stream=None):
    """camera::generate_pixel_coordinates + project_evaluation_points."""
    c = N.camera_struct(cam)
    out = torch.empty((c.nx * c.ny, 4), dtype=torch.float32, device=torch.device("cuda", device))
    N.call("ilc_camera_pixel_points", ctypes.byref(c), _ptr(out), _stream(stream))
    return out


def camera_generate_pixel_coordinates(cam, device=0, stream=None):
    """camera::generate_pixel_coordinates: world-frame pixel corners (nx*ny, 4)."""
    c = N.camera_struct(cam)
    out = torch.empty((c.nx * c.ny, 4), dtype=torch.float32, device=torch.device("cuda", device))
    N.call("ilc_camera_generate_pixel_coordinates", ctypes.byref(c), _ptr(out), _stream(stream))
    return out


def project_evaluation_points(cam, points4, device=0, stream=None, out=None):
    """projective_smoothing::project_evaluation_points (z and w kept)."""
    c = N.camera_struct(cam)
    p = _dev_f32(points4, torch.device("cuda", device), 4)
    if out is None:
        out = torch.empty_like(p)
    N.call("ilc_project_evaluation_points", ctypes.byref(c), _ptr(p), p.shape[0], _ptr(out),
           _stream(stream))
    return out


def vector_scale(x, alpha, stream=None):
    """vector_scale: x *= alpha in FP32, in place (x: contiguous float32 CUDA tensor)."""
    if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
        raise ValueError("vector_scale needs a contiguous float32 CUDA tensor")
    N.call("ilc_vector_scale", _ptr(x), x.numel(), ctypes.c_float(alpha), _stream(stream))
    return x


def project_particles(cam, particles4, device=0, stream=None, out=None):
    """projective_smoothing::project_particles (depth mask, z := 0)."""
    c = N.camera_struct(cam)
    p = _dev_f32(particles4, torch.device("cuda", device), 4)
    if out is None:
        out = torch.empty_like(p)
    N.call("ilc_project_particles", ctypes.byref(c), _ptr(p), p.shape[0], _ptr(out),
           _stream(stream))
    return out


def crawl_block(center, extent, periodic_wraparound, border, xyz, columns=(), device=0,
                stream=None, with_index=False, static_scratch=False):
    """Block body of the data crawler (reconstructing_data_crawler.cpp:180-213):
    periodic correction towards the render volume, cutout, compaction in input
    order.  Returns (particles4 (m, 4) with w = 0, [columns (m,)], [index (m,)]).
    static_scratch: use ilc_crawl_block_static, which allocates nothing (blocks
    of at most 2^29 rows; calls overlapping on one GPU race)."""
    dev = torch.device("cuda", device)
    x = _dev_f32(xyz, dev, 3)
    n = x.shape[0]
    cols = [_dev_f32(c, dev).reshape(-1) for c in columns]
    if any(c.shape[0] != n for c in cols):
        raise ValueError("every column needs one value per row")
    out4 = torch.empty((n, 4), dtype=torch.float32, device=dev)
    outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in cols]
    idx = torch.empty(n, dtype=torch.int64, device=dev) if with_index else None
    vol = N.VolumeC((ctypes.c_double * 3)(*center), (ctypes.c_double * 3)(*extent),
                    (ctypes.c_double * 3)(*periodic_wraparound), float(border))
    k = max(len(cols), 1)
    cin = (ctypes.c_void_p * k)(*[c.data_ptr() for c in cols])
    cout = (ctypes.c_void_p * k)(*[c.data_ptr() for c in outs])
    m = ctypes.c_uint64()
    fn = "ilc_crawl_block_static" if static_scratch else "ilc_crawl_block"
    N.call(fn, ctypes.byref(vol), _ptr(x), n, ctypes.cast(cin, ctypes.c_void_p),
           len(cols), _ptr(out4), ctypes.cast(cout, ctypes.c_void_p), _ptr(idx), ctypes.byref(m),
           _stream(stream))
    m = m.value
    res = (out4[:m], [o[:m] for o in outs])
    return res + (idx[:m],) if with_index else res


def quantity_scaled(inp, scale, n=None, device=0, stream=None):
    """quantity_transformation with unprocessed_quantity (inp given) or
    constant_quantity (inp None): w = float(scale * inp)."""
    dev = torch.device("cuda", device)
    t = _dev_f32(inp, dev).reshape(-1) if inp is not None else None
    n = t.shape[0] if t is not None else int(n)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    N.call("ilc_quantity_scaled", _ptr(t), ctypes.c_double(scale), n, _ptr(out), _stream(stream))
    return out


def _gaunt2d(table):
    t = np.ascontiguousarray(table, dtype=np.float32)
    if t.ndim != 2:
        raise ValueError("the Gaunt table is 2-D: rows of constant ln(u), columns of ln(gamma^2)")
    return t, t.shape[1], t.shape[0]


def quantity_xray_flux(rho, u, xe, table, z, lum_dist, e_min, e_max, num_samples, device=0,
                       stream=None, with_steps=False):
    """xray_flux (quantities.cl:354-386): band-integrated bremsstrahlung per
    particle from density, internal energy and electron abundance."""
    dev = torch.device("cuda", device)
    r = _dev_f32(rho, dev).reshape(-1)
    uu = _dev_f32(u, dev).reshape(-1)
    x = _dev_f32(xe, dev).reshape(-1)
    t, nx, ny = _gaunt2d(table)
    out = torch.empty_like(r)
    steps = torch.empty(r.shape[0], dtype=torch.int32, device=dev) if with_steps else None
    N.call("ilc_quantity_xray_flux", _ptr(r), _ptr(uu), _ptr(x), r.shape[0],
           t.ctypes.data_as(ctypes.c_void_p), nx, ny, ctypes.c_float(z),
           ctypes.c_float(lum_dist), ctypes.c_float(e_min), ctypes.c_float(e_max),
           int(num_samples), _ptr(out), _ptr(steps), _stream(stream))
    return (out, steps) if with_steps else out


def quantity_xray_spectral_flux(rho, u, xe, table, z, lum_dist, e, de, device=0, stream=None):
    """xray_spectral_flux (quantities.cl:400-430): one energy bin."""
    dev = torch.device("cuda", device)
    r = _dev_f32(rho, dev).reshape(-1)
    uu = _dev_f32(u, dev).reshape(-1)
    x = _dev_f32(xe, dev).reshape(-1)
    t, nx, ny = _gaunt2d(table)
    out = torch.empty_like(r)
    N.call("ilc_quantity_xray_spectral_flux", _ptr(r), _ptr(uu), _ptr(x), r.shape[0],
           t.ctypes.data_as(ctypes.c_void_p), nx, ny, ctypes.c_float(z),
           ctypes.c_float(lum_dist), ctypes.c_float(e), ctypes.c_float(de), _ptr(out),
           _stream(stream))
    return out


def load_gaunt_table(path):
    """Reads a thermally averaged Gaunt table in the layout of van Hoof's
    distribution (gaunt_factor.dat): '#' comment lines, then one row of
    whitespace-separated values per ln(u) node.  Returns float32 (rows, cols)."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            rows.append([float(v) for v in line.split()])
    if not rows or any(len(r) != len(rows[0]) for r in rows):
        raise ValueError(f"{path}: not a rectangular table")
    return np.asarray(rows, np.float32)


def quantity_xray_bolometric(rho, temp, xe, table, x0, dx, device=0, stream=None):
    dev = torch.device("cuda", device)
    r = _dev_f32(rho, dev).reshape(-1)
    t = _dev_f32(temp, dev).reshape(-1)
    x = _dev_f32(xe, dev).reshape(-1)
    tab = np.ascontiguousarray(table, dtype=np.float32)
    out = torch.empty_like(r)
    N.call("ilc_quantity_xray_bolometric", _ptr(r), _ptr(t), _ptr(x), r.shape[0],
           tab.ctypes.data_as(ctypes.c_void_p), tab.shape[0], ctypes.c_float(x0),
           ctypes.c_float(dx), _ptr(out), _stream(stream))
    return out


class ProjectiveSmoothing:
    """dm::projective_smoothing: projects particles and pixels into the
    camera frame, forwards to the grid backend, and applies the quantity's
    effective dA at retrieval (projective_smoothing_reconstruction.cpp:49-142)."""

    def __init__(self, camera, max_integration_depth, effective_dA, backend):
        self.camera = dict(camera)
        self.depth = float(max_integration_depth)
        self.effective_dA = effective_dA   # callable(pixel_area, depth) -> scale, or a float
        self.backend = backend

    def get_backend_name(self):
        return "projective_smoothing via " + self.backend.get_backend_name()

    def init_backend(self, blocksize):
        self.backend.init_backend(blocksize)

    def set_camera(self, cam):
        self.camera = dict(cam)

    def set_integration_depth(self, depth):
        self.depth = float(depth)

    def setup_particles(self, particles4, additional_datasets, stream=None):
        cam = dict(self.camera, depth=self.depth)
        proj = project_particles(cam, particles4, self.backend.device.index, stream)
        self.backend.setup_projected_particles(proj, additional_datasets[0], stream)

    def setup_evaluation_points(self, stream=None):
        pts = camera_pixel_points(self.camera, self.backend.device.index, stream)
        self.backend.setup_evaluation_points(pts, layout_nx=int(self.camera["nx"]),
                                             stream=stream)

    def run(self, stream=None):
        self.backend.run(stream)

    def retrieve_results(self, stream=None):
        ps = float(self.camera["width"]) / float(self.camera["nx"])
        scale = self.effective_dA(ps * ps, self.depth) if callable(self.effective_dA) \
            else float(self.effective_dA)
        img = self.backend.retrieve_results(float(np.float32(scale)), stream)
        return img.reshape(int(self.camera.get("ny", self.camera["nx"])),
                           int(self.camera["nx"]))


# ------------------------------------------------------------ batched views

class BatchRenderer:
    """Many projections of one resident snapshot (ilc_batch_* ABI): one
    persistent gather over every (view, tile), optionally replayed from a
    hipGraph; ``render(plain=True)`` runs the one-view-per-launch path."""

    def __init__(self, device=0, mode="reference", kernel="quartic_ref",
                 target_particles_per_cell=1000, tile=None):
        cfg = N.PsgConfig()
        cfg.device = int(device)
        cfg.target_particles_per_cell = int(target_particles_per_cell)
        cfg.mode = _MODES[mode]
        cfg.kernel = _KERNELS[kernel]
        # (w, h) at creation; (w, h, workgroup) through set_tile
        if tile is not None and len(tile) == 2:
            cfg.tile_w, cfg.tile_h = int(tile[0]), int(tile[1])
        self._h = ctypes.c_void_p()
        N.call("ilc_batch_create", ctypes.byref(cfg), ctypes.byref(self._h))
        if tile is not None and len(tile) > 2:
            self.set_tile(*tile)
        self.device = torch.device("cuda", int(device))
        self._keep = []
        self.views = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                N.lib().ilc_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_particles(self, pos, mass, hsml, rho=None, temp=None, xe=None):
        d = self.device
        pos = _dev_f32(pos, d, 3)
        arrs = [pos] + [(_dev_f32(a, d).reshape(-1) if a is not None else None)
                        for a in (mass, hsml, rho, temp, xe)]
        self._keep = arrs
        self._n = pos.shape[0]
        N.call("ilc_batch_set_particles", self._h, *[_ptr(a) for a in arrs], pos.shape[0])

    def set_gaunt_table(self, table, x0, dx):
        tab = np.ascontiguousarray(table, dtype=np.float32)
        self._gaunt = tab
        N.call("ilc_batch_set_gaunt_table", self._h, tab.ctypes.data_as(ctypes.c_void_p),
               tab.shape[0], ctypes.c_float(x0), ctypes.c_float(dx))

    def set_views(self, views, stream=None):
        arr = (N.ViewC * len(views))()
        for k, v in enumerate(views):
            arr[k].cam = N.camera_struct(v)
            arr[k].quantity = int(v.get("quantity", 0))
            arr[k].scale = float(v.get("scale", 1.0))
        self.views = [dict(v) for v in views]
        N.call("ilc_batch_set_views", self._h, arr, len(views), _stream(stream))

    def render(self, plain=False, graph=False, stream=None):
        flags = (N.RENDER_PLAIN if plain else 0) | (N.RENDER_GRAPH if graph else 0)
        N.call("ilc_batch_render", self._h, flags, _stream(stream))

    def image(self, v, stream=None):
        cam = self.views[v]
        out = torch.empty((int(cam.get("ny", cam["nx"])), int(cam["nx"])), dtype=torch.float32,
                          device=self.device)
        N.call("ilc_batch_image", self._h, int(v), _ptr(out), _stream(stream))
        return out

    def image_f64(self, v, stream=None):
        cam = self.views[v]
        out = torch.empty((int(cam.get("ny", cam["nx"])), int(cam["nx"])), dtype=torch.float64,
                          device=self.device)
        N.call("ilc_batch_image_f64", self._h, int(v), _ptr(out), _stream(stream))
        return out

    def count_evals(self, stream=None):
        arr = (ctypes.c_uint64 * len(self.views))()
        N.call("ilc_batch_count_evals", self._h, arr, _stream(stream))
        return [int(x) for x in arr]

    def backward(self, dl_dimages, stream=None):
        """Adjoint of the last render (exact mode): dl_dimages[v] = dL/d(image
        of view v).  Returns (dL/dw, dL/dh, dL/dpos (n, 3)) per snapshot
        particle, summed over views."""
        if len(dl_dimages) != len(self.views):
            raise ValueError("one gradient image per view")
        dev = self.device
        gs = [_dev_f32(g, dev).reshape(-1) for g in dl_dimages]
        for g, v in zip(gs, self.views):
            if g.shape[0] != int(v["nx"]) * int(v.get("ny", v["nx"])):
                raise ValueError("a gradient image has the wrong size")
        ptrs = (ctypes.c_void_p * len(gs))(*[g.data_ptr() for g in gs])
        gw = torch.empty(self._n, dtype=torch.float32, device=dev)
        gh = torch.empty_like(gw)
        gp = torch.empty((self._n, 3), dtype=torch.float32, device=dev)
        N.call("ilc_batch_backward", self._h, ctypes.cast(ptrs, ctypes.c_void_p), _ptr(gw),
               _ptr(gh), _ptr(gp), _stream(stream))
        return gw, gh, gp

    def view_info(self, view):
        """Grid of one view after a render (num_particles: particles it holds)."""
        gi = N.GridInfo()
        N.call("ilc_batch_view_info", self._h, int(view), ctypes.byref(gi))
        return dict(num_cells=list(gi.num_cells), min_corner=list(gi.min_corner),
                    cell_size=list(gi.cell_size), max_h=gi.max_h,
                    num_particles=gi.num_particles)

    def set_tile(self, w, h, workgroup=256):
        """Gather tile of w x h points, one tile per workgroup of `workgroup`
        threads (256, or 64 for one-wave 16x16 tiles)."""
        N.call("ilc_batch_set_tile_wg", self._h, int(w), int(h), int(workgroup))

    def tile(self):
        """(tile_w, tile_h, workgroup) of the gather."""
        w, h, g = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        N.call("ilc_batch_get_tile", self._h, ctypes.byref(w), ctypes.byref(h), ctypes.byref(g))
        return w.value, h.value, g.value

    def set_depth_clip(self, clip=True):
        """Line-of-sight kernel clip of every view's depth range instead of the
        reference's centre mask (exact mode, quartic kernels, persistent
        renders)."""
        N.call("ilc_batch_set_depth_clip", self._h, 1 if clip else 0)

    def last_gather_ms(self):
        ms = ctypes.c_float()
        N.call("ilc_batch_last_gather_ms", self._h, ctypes.byref(ms))
        return ms.value

    def autotune(self, stream=None):
        w, h, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        N.call("ilc_batch_autotune", self._h, ctypes.byref(w), ctypes.byref(h), ctypes.byref(ms),
               _stream(stream))
        return (w.value, h.value), ms.value
