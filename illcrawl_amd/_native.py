"""ctypes binding of the C ABI in include/illcrawl_psg.h (libilcpsg.so).

The library must have been built in-tree (``python -m illcrawl_amd.build``);
there is no fallback: if it is missing or cannot be loaded, every entry point
raises ``NativeLibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ILC_LIB_PATH selects another build of the same library (e.g. the host-side
# sanitizer build of tests/sanitized_run.py)
LIB_PATH = os.environ.get("ILC_LIB_PATH") or os.path.join(HERE, "lib", "libilcpsg.so")

ILC_OK, ILC_EINVAL, ILC_ENOMEM, ILC_EHIP, ILC_ENOTREADY, ILC_ERANGE = range(6)
MODE_REFERENCE, MODE_EXACT = 0, 1
KERNEL_QUARTIC_REF, KERNEL_QUARTIC, KERNEL_CUBIC = 0, 1, 2

_STATUS_NAMES = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "ENOTREADY", 5: "ERANGE"}


class NativeLibraryMissing(RuntimeError):
    pass


class IllcrawlError(RuntimeError):
    """A failed ABI call: ``code`` is the ilc_status, the message the
    library's thread-local error text."""

    def __init__(self, code, msg):
        super().__init__(f"[{_STATUS_NAMES.get(code, code)}] {msg}")
        self.code = code


class PsgConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("target_particles_per_cell", ctypes.c_uint64),
                ("mode", ctypes.c_int), ("kernel", ctypes.c_int), ("tile_w", ctypes.c_int),
                ("tile_h", ctypes.c_int)]


class GridInfo(ctypes.Structure):
    _fields_ = [("num_cells", ctypes.c_int32 * 3), ("min_corner", ctypes.c_float * 3),
                ("cell_size", ctypes.c_float * 3), ("max_h", ctypes.c_float),
                ("num_particles", ctypes.c_uint64)]


class CameraC(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 3), ("look_at", ctypes.c_double * 3),
                ("roll", ctypes.c_double), ("width", ctypes.c_double), ("nx", ctypes.c_uint32),
                ("ny", ctypes.c_uint32), ("depth", ctypes.c_double)]


class VolumeC(ctypes.Structure):
    _fields_ = [("center", ctypes.c_double * 3), ("extent", ctypes.c_double * 3),
                ("periodic_wraparound", ctypes.c_double * 3), ("border", ctypes.c_double)]


class ViewC(ctypes.Structure):
    _fields_ = [("cam", CameraC), ("quantity", ctypes.c_int32), ("scale", ctypes.c_float)]


QUANTITY_MASS, QUANTITY_XRAY = 0, 1
RENDER_PLAIN, RENDER_GRAPH = 1, 2

# name -> (restype, argtypes); every symbol declared in include/illcrawl_psg.h
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_ST = ctypes.c_int
SIGNATURES = {
    "ilc_last_error": (ctypes.c_char_p, []),
    "ilc_abi_version": (ctypes.c_int, []),
    "ilc_psg_create": (_ST, [ctypes.POINTER(PsgConfig), ctypes.POINTER(_P)]),
    "ilc_psg_destroy": (_ST, [_P]),
    "ilc_psg_backend_name": (ctypes.c_char_p, [_P]),
    "ilc_psg_init_backend": (_ST, [_P, _U64]),
    "ilc_psg_setup_projected_particles": (_ST, [_P, _P, _U64, _P, _P]),
    "ilc_psg_setup_evaluation_points": (_ST, [_P, _P, _U64, ctypes.c_uint32, _P]),
    "ilc_psg_run": (_ST, [_P, _P]),
    "ilc_psg_retrieve_results": (_ST, [_P, ctypes.c_float, _P, _P]),
    "ilc_psg_retrieve_results_f64": (_ST, [_P, _P, _P]),
    "ilc_psg_backward": (_ST, [_P, _P, ctypes.c_float, _P, _P, _P, _P, _P]),
    "ilc_psg_grid_info": (_ST, [_P, ctypes.POINTER(GridInfo)]),
    "ilc_psg_export_grid": (_ST, [_P, _P, _P, _P, _P]),
    "ilc_psg_count_evals": (_ST, [_P, _P, ctypes.POINTER(_U64), _P]),
    "ilc_camera_basis": (_ST, [ctypes.POINTER(CameraC), ctypes.POINTER(ctypes.c_double * 3),
                               ctypes.POINTER(ctypes.c_double * 3),
                               ctypes.POINTER(ctypes.c_double * 3),
                               ctypes.POINTER(ctypes.c_double)]),
    "ilc_camera_projection": (_ST, [ctypes.POINTER(CameraC), ctypes.POINTER(ctypes.c_float * 9),
                                    ctypes.POINTER(ctypes.c_float * 3)]),
    "ilc_project_particles": (_ST, [ctypes.POINTER(CameraC), _P, _U64, _P, _P]),
    "ilc_camera_pixel_points": (_ST, [ctypes.POINTER(CameraC), _P, _P]),
    "ilc_camera_generate_pixel_coordinates": (_ST, [ctypes.POINTER(CameraC), _P, _P]),
    "ilc_project_evaluation_points": (_ST, [ctypes.POINTER(CameraC), _P, _U64, _P, _P]),
    "ilc_vector_scale": (_ST, [_P, _U64, ctypes.c_float, _P]),
    "ilc_grid3d_create": (_ST, [ctypes.POINTER(PsgConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "ilc_grid3d_destroy": (_ST, [_P]),
    "ilc_grid3d_init_backend": (_ST, [_P, _U64]),
    "ilc_grid3d_setup_particles": (_ST, [_P, _P, _U64, _P, _P]),
    "ilc_grid3d_setup_evaluation_points": (_ST, [_P, _P, _U64, ctypes.c_uint32, _P]),
    "ilc_grid3d_run": (_ST, [_P, _P]),
    "ilc_grid3d_retrieve_results": (_ST, [_P, ctypes.c_float, _P, _P]),
    "ilc_grid3d_retrieve_results_f64": (_ST, [_P, _P, _P]),
    "ilc_grid3d_grid_info": (_ST, [_P, ctypes.POINTER(GridInfo)]),
    "ilc_grid3d_count_evals": (_ST, [_P, _P, ctypes.POINTER(_U64), _P]),
    "ilc_particles_set_w": (_ST, [_P, _P, _U64, _P]),
    "ilc_crawl_block": (_ST, [ctypes.POINTER(VolumeC), _P, _U64, _P, ctypes.c_int32, _P, _P,
                              _P, ctypes.POINTER(_U64), _P]),
    "ilc_crawl_block_static": (_ST, [ctypes.POINTER(VolumeC), _P, _U64, _P, ctypes.c_int32, _P,
                                     _P, _P, ctypes.POINTER(_U64), _P]),
    "ilc_quantity_xray_flux": (_ST, [_P, _P, _P, _U64, _P, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_int32, _P, _P, _P]),
    "ilc_quantity_xray_spectral_flux": (_ST, [_P, _P, _P, _U64, _P, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                              ctypes.c_float, ctypes.c_float, _P, _P]),
    "ilc_quantity_scaled": (_ST, [_P, ctypes.c_double, _U64, _P, _P]),
    "ilc_quantity_stellar_mass": (_ST, [_P, _P, _U64, _P, _P]),
    "ilc_quantity_xray_bolometric": (_ST, [_P, _P, _P, _U64, _P, ctypes.c_int32, ctypes.c_float,
                                           ctypes.c_float, _P, _P]),
    "ilc_batch_create": (_ST, [ctypes.POINTER(PsgConfig), ctypes.POINTER(_P)]),
    "ilc_batch_destroy": (_ST, [_P]),
    "ilc_batch_set_particles": (_ST, [_P, _P, _P, _P, _P, _P, _P, _U64]),
    "ilc_batch_set_gaunt_table": (_ST, [_P, _P, ctypes.c_int32, ctypes.c_float, ctypes.c_float]),
    "ilc_batch_set_views": (_ST, [_P, ctypes.POINTER(ViewC), ctypes.c_int32, _P]),
    "ilc_batch_render": (_ST, [_P, ctypes.c_int32, _P]),
    "ilc_batch_image": (_ST, [_P, ctypes.c_int32, _P, _P]),
    "ilc_batch_image_f64": (_ST, [_P, ctypes.c_int32, _P, _P]),
    "ilc_batch_count_evals": (_ST, [_P, ctypes.POINTER(_U64), _P]),
    "ilc_batch_view_info": (_ST, [_P, ctypes.c_int32, ctypes.POINTER(GridInfo)]),
    "ilc_batch_backward": (_ST, [_P, _P, _P, _P, _P, _P]),
    "ilc_batch_set_tile": (_ST, [_P, ctypes.c_int32, ctypes.c_int32]),
    "ilc_batch_set_tile_wg": (_ST, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "ilc_batch_get_tile": (_ST, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32)]),
    "ilc_batch_set_depth_clip": (_ST, [_P, ctypes.c_int32]),
    "ilc_batch_last_gather_ms": (_ST, [_P, ctypes.POINTER(ctypes.c_float)]),
    "ilc_batch_autotune": (_ST, [_P, ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_int32),
                                 ctypes.POINTER(ctypes.c_double), _P]),
}

_lib = None


def lib():
    """Load libilcpsg.so (once).  Raises NativeLibraryMissing if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m illcrawl_amd.build` "
                "(there is no non-HIP fallback)")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeLibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != ILC_OK:
        msg = lib().ilc_last_error()
        raise IllcrawlError(code, msg.decode() if msg else "")


def call(name, *args):
    check(getattr(lib(), name)(*args))


def camera_struct(cam: dict) -> CameraC:
    c = CameraC()
    for i in range(3):
        c.position[i] = float(cam["position"][i])
        c.look_at[i] = float(cam["look_at"][i])
    c.roll = float(cam.get("roll", 0.0))
    c.width = float(cam["width"])
    c.nx = int(cam["nx"])
    c.ny = int(cam.get("ny", cam["nx"]))
    c.depth = float(cam.get("depth", 0.0))
    return c
