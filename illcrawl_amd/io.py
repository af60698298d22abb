"""Snapshot input and image output of the data crawler, for Python callers.

Mirrors include/illcrawl_io.hpp:

- ``IllustrisDataLoader`` reads a particle group (PartType<id>) of an
  Illustris/TNG HDF5 snapshot (hdf5_io.hpp:39-107).  It binds the HDF5 C
  library the image provides (``$ILC_IO_PREFIX/lib/libhdf5.so``, default
  /opt/conda) through ctypes.  Rows come back as float64, as the reference
  reads them before its cast to the device scalar.
- ``write_fits`` / ``read_fits`` are a numpy implementation of the FITS
  primary image the reference writes through cfitsio (fits.hpp:210-300):
  BITPIX -32 or -64, NAXIS1 the fastest index, big-endian data in 2880-byte
  records.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PREFIX = os.environ.get("ILC_IO_PREFIX", "/opt/conda")
_H5 = None


class HDF5Missing(RuntimeError):
    pass


def _h5():
    global _H5
    if _H5 is None:
        path = os.path.join(_PREFIX, "lib", "libhdf5.so")
        if not os.path.exists(path):
            raise HDF5Missing(f"{path} not found (set ILC_IO_PREFIX)")
        L = ctypes.CDLL(path)
        hid, herr = ctypes.c_int64, ctypes.c_int
        hsz = ctypes.POINTER(ctypes.c_ulonglong)
        sig = {
            "H5open": (herr, []),
            "H5Fopen": (hid, [ctypes.c_char_p, ctypes.c_uint, hid]),
            "H5Fclose": (herr, [hid]),
            "H5Gopen2": (hid, [hid, ctypes.c_char_p, hid]),
            "H5Gclose": (herr, [hid]),
            "H5Dopen2": (hid, [hid, ctypes.c_char_p, hid]),
            "H5Dclose": (herr, [hid]),
            "H5Dget_space": (hid, [hid]),
            "H5Dread": (herr, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [hid, hsz, hsz]),
            "H5Sselect_hyperslab": (herr, [hid, ctypes.c_int, hsz, hsz, hsz, hsz]),
            "H5Screate_simple": (hid, [ctypes.c_int, hsz, hsz]),
            "H5Sclose": (herr, [hid]),
            "H5Lexists": (ctypes.c_int, [hid, ctypes.c_char_p, hid]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        if L.H5open() < 0:
            raise HDF5Missing("H5open failed")
        L.native_double = ctypes.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
        _H5 = L
    return _H5


def _check(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


class Dataset:
    """One dataset of the group: ``shape`` and ``read(first, count)``."""

    def __init__(self, group_id, name):
        L = _h5()
        self.name = name
        self._id = _check(L.H5Dopen2(group_id, name.encode(), 0), f"open {name}")
        sp = _check(L.H5Dget_space(self._id), "get_space")
        try:
            nd = L.H5Sget_simple_extent_ndims(sp)
            if nd not in (1, 2):
                raise ValueError(f"{name}: rank {nd} (1 or 2 supported)")
            dims = (ctypes.c_ulonglong * nd)()
            L.H5Sget_simple_extent_dims(sp, dims, None)
            self.shape = tuple(int(d) for d in dims)
        finally:
            L.H5Sclose(sp)

    def __len__(self):
        return self.shape[0]

    def read(self, first=0, count=None):
        """Rows [first, first + count) as float64, shape (count,) or
        (count, columns)."""
        L = _h5()
        n = self.shape[0]
        count = n - first if count is None else int(count)
        if first < 0 or count < 0 or first + count > n:
            raise IndexError(f"{self.name}: rows [{first}, {first + count}) of {n}")
        nd = len(self.shape)
        cols = self.shape[1] if nd == 2 else 1
        out = np.empty((count, cols) if nd == 2 else (count,), dtype=np.float64)
        if count == 0:
            return out
        fs = _check(L.H5Dget_space(self._id), "get_space")
        off = (ctypes.c_ulonglong * nd)(first, *([0] * (nd - 1)))
        cnt = (ctypes.c_ulonglong * nd)(count, *([cols] * (nd - 1)))
        ms = -1
        try:
            _check(L.H5Sselect_hyperslab(fs, 0, off, None, cnt, None), "select_hyperslab")
            ms = _check(L.H5Screate_simple(nd, cnt, None), "create_simple")
            _check(L.H5Dread(self._id, L.native_double, ms, fs, 0,
                             out.ctypes.data_as(ctypes.c_void_p)), f"read {self.name}")
        finally:
            if ms >= 0:
                L.H5Sclose(ms)
            L.H5Sclose(fs)
        return out

    def close(self):
        if getattr(self, "_id", -1) >= 0:
            _h5().H5Dclose(self._id)
            self._id = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IllustrisDataLoader:
    """illustris_data_loader (hdf5_io.hpp:39-107)."""

    COORDINATES = "Coordinates"
    DENSITY = "Density"
    INTERNAL_ENERGY = "InternalEnergy"
    SMOOTHING_LENGTH = "SmoothingLength"
    VOLUME = "Volume"
    MASSES = "Masses"
    ELECTRON_ABUNDANCE = "ElectronAbundance"
    DM_SMOOTHING_LENGTH = "SubfindHsml"

    def __init__(self, filename, group=0):
        L = _h5()
        self.filename = filename
        self._file = _check(L.H5Fopen(os.fsencode(filename), 0, 0), f"open {filename}")
        self._group = -1
        self.select_group(group)

    def select_group(self, group):
        L = _h5()
        name = f"PartType{group}" if isinstance(group, int) else str(group)
        if self._group >= 0:
            L.H5Gclose(self._group)
            self._group = -1
        self._group = _check(L.H5Gopen2(self._file, name.encode(), 0), f"open group {name}")
        self.current_group_name = name

    def get_dataset(self, name):
        return Dataset(self._group, name)

    def has_dataset(self, name):
        return _h5().H5Lexists(self._group, name.encode(), 0) > 0

    def close(self):
        L = _h5()
        if self._group >= 0:
            L.H5Gclose(self._group)
            self._group = -1
        if self._file >= 0:
            L.H5Fclose(self._file)
            self._file = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_blocks(datasets, blocksize):
    """(first_row, [rows of each dataset as float64]) for consecutive row
    blocks of datasets with equal row counts."""
    n = len(datasets[0])
    if any(len(d) != n for d in datasets):
        raise ValueError("datasets have different row counts")
    for first in range(0, n, int(blocksize)):
        cnt = min(int(blocksize), n - first)
        yield first, [d.read(first, cnt) for d in datasets]


# ------------------------------------------------------------------- FITS

_BLOCK = 2880


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = "T" if value else "F"
    else:
        v = str(value)
    s = f"{key:<8}= {v:>20}"
    if comment:
        s += f" / {comment}"
    return s[:80].ljust(80)


def write_fits(path, data):
    """Primary image of ``data`` (float32 or float64; shape (ny, nx) for an
    image of nx columns, so NAXIS1 = nx as the reference's multi_array)."""
    a = np.asarray(data)
    if a.dtype == np.float64:
        bitpix, be = -64, ">f8"
    else:
        a = a.astype(np.float32, copy=False)
        bitpix, be = -32, ">f4"
    cards = [_card("SIMPLE", True, "file does conform to FITS standard"),
             _card("BITPIX", bitpix, "number of bits per data pixel"),
             _card("NAXIS", a.ndim, "number of data axes")]
    for k, n in enumerate(reversed(a.shape)):
        cards.append(_card(f"NAXIS{k + 1}", n, f"length of data axis {k + 1}"))
    cards.append(_card("EXTEND", True, "FITS dataset may contain extensions"))
    cards.append("END".ljust(80))
    head = "".join(cards).encode("ascii")
    head += b" " * (-len(head) % _BLOCK)
    body = np.ascontiguousarray(a).astype(be).tobytes()
    body += b"\0" * (-len(body) % _BLOCK)
    with open(path, "wb") as f:
        f.write(head + body)


def read_fits(path):
    """The primary image as a numpy array of shape (..., NAXIS2, NAXIS1)."""
    with open(path, "rb") as f:
        raw = f.read()
    hdr, pos = {}, 0
    while True:
        if pos + 80 > len(raw):
            raise ValueError(f"{path}: no END card")
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            hdr[key] = card[10:].split("/")[0].strip()
    pos += -pos % _BLOCK
    bitpix, naxis = int(hdr["BITPIX"]), int(hdr["NAXIS"])
    dims = [int(hdr[f"NAXIS{k + 1}"]) for k in range(naxis)]
    dt = {-32: ">f4", -64: ">f8", 8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8"}[bitpix]
    count = int(np.prod(dims)) if dims else 0
    a = np.frombuffer(raw, dtype=dt, count=count, offset=pos)
    return a.reshape(tuple(reversed(dims))).astype(a.dtype.newbyteorder("="))


def read_fits_header(path):
    """The primary header's keys as {name: value}: standard keys upper-case,
    long or dotted keys (cfitsio's HIERARCH convention) as written; string
    values unquoted, numbers as float.  COMMENT cards are returned in order
    under the key "COMMENT"."""
    with open(path, "rb") as f:
        raw = f.read()
    out, comments, pos = {}, [], 0
    while pos + 80 <= len(raw):
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        if card[:8].strip() == "END":
            break
        if card.startswith("COMMENT"):
            comments.append(card[8:].strip())
            continue
        if card.startswith("HIERARCH "):
            key, _, rest = card[9:].partition("=")
            key = key.strip()
        elif card[8:10] == "= ":
            key, rest = card[:8].strip(), card[10:]
        else:
            continue
        rest = rest.strip()
        if rest.startswith("'"):
            end = rest.find("'", 1)
            while end != -1 and rest[end + 1:end + 2] == "'":  # doubled quote
                end = rest.find("'", end + 2)
            out[key] = rest[1:end].replace("''", "'").rstrip()
        else:
            v = rest.split("/")[0].strip()
            try:
                out[key] = float(v.replace("D", "E"))
            except ValueError:
                out[key] = v
    out["COMMENT"] = comments
    return out

