"""The reference's HDF5 material files (cstool output) through
libnbl_mat.so (include/nbl_mat.h): open, read properties and raw tables,
prepare them for the device, and write synthetic materials in the same
layout.

libnbl_mat.so links HDF5 and is separate from libnbl_hip.so; it is loaded
on first use (NBL_MAT_LIB overrides its path)."""
from __future__ import annotations

import ctypes as C
import os

from . import _lib as L
from . import synth

_HERE = os.path.dirname(os.path.abspath(__file__))
_P, _I32, _U64 = C.c_void_p, C.c_int32, C.c_uint64
_mat = None


class MatError(RuntimeError):
    pass


def lib():
    global _mat
    if _mat is None:
        path = os.environ.get("NBL_MAT_LIB", os.path.join(_HERE, "_build", "libnbl_mat.so"))
        if not os.path.exists(path):
            raise ImportError(f"libnbl_mat.so not built ({path}); it needs the HDF5 headers at build time")
        m = C.CDLL(path)
        sig = {
            "nbl_mat_open": [C.c_char_p, C.POINTER(_P)],
            "nbl_mat_close": [_P],
            "nbl_mat_exists": [_P, C.c_char_p, C.POINTER(_I32)],
            "nbl_mat_get_string": [_P, C.c_char_p, C.c_char_p, _U64, C.POINTER(_U64)],
            "nbl_mat_get_quantity": [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(_I32)],
            "nbl_mat_parse_unit": [C.c_char_p, C.POINTER(C.c_double), C.POINTER(_I32)],
            "nbl_mat_kieft_raw": [_P, C.POINTER(L.KieftRaw)],
            "nbl_mat_penn_raw": [_P, C.POINTER(L.PennRaw)],
            "nbl_mat_write": [C.c_char_p, C.c_char_p, C.POINTER(L.KieftRaw), C.POINTER(L.PennRaw)],
        }
        for name, args in sig.items():
            f = getattr(m, name)
            f.argtypes = args
            f.restype = C.c_int
        m.nbl_mat_last_error.restype = C.c_char_p
        m.nbl_mat_last_error.argtypes = []
        _mat = m
    return _mat


def _call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != L.NBL_OK:
        raise MatError(f"{name} failed ({rc}): {lib().nbl_mat_last_error().decode(errors='replace')}")


def parse_unit(text):
    """(value in eV/nm/ns/K/e, powers of (energy, length, time, temperature, charge))."""
    v = C.c_double(); d = (_I32 * 5)()
    _call("nbl_mat_parse_unit", text.encode(), C.byref(v), d)
    return v.value, tuple(d)


class MatFile:
    """hdf5_file: one open material file."""

    def __init__(self, path):
        self.path = str(path)
        h = _P()
        _call("nbl_mat_open", self.path.encode(), C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().nbl_mat_close(self.h)
            self.h = _P()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exists(self, path) -> bool:
        out = _I32()
        _call("nbl_mat_exists", self.h, path.encode(), C.byref(out))
        return bool(out.value)

    def get_string(self, name) -> str:
        n = _U64()
        _call("nbl_mat_get_string", self.h, name.encode(), None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        _call("nbl_mat_get_string", self.h, name.encode(), buf, n.value + 1, C.byref(n))
        return buf.value.decode()

    def get_quantity(self, name):
        v = C.c_double(); d = (_I32 * 5)()
        _call("nbl_mat_get_quantity", self.h, name.encode(), C.byref(v), d)
        return v.value, tuple(d)

    def kieft_raw(self) -> L.KieftRaw:
        r = L.KieftRaw()
        _call("nbl_mat_kieft_raw", self.h, C.byref(r))
        return r

    def penn_raw(self) -> L.PennRaw:
        r = L.PennRaw()
        _call("nbl_mat_penn_raw", self.h, C.byref(r))
        return r


def load_material(path, inelastic="kieft"):
    """material::create over a file: the prepared tables (PreparedMaterial)
    with the Kieft-Bosch or full Penn inelastic model."""
    from .api import PreparedMaterial
    if inelastic not in ("kieft", "penn"):
        raise ValueError(f"unknown inelastic model {inelastic!r}")
    f = MatFile(path)
    kr = f.kieft_raw()
    pr = f.penn_raw() if inelastic == "penn" else None
    name = f.get_string("name") if f.exists("/") else os.path.basename(str(path))
    pm = PreparedMaterial(kieft_raw=kr, penn_raw=pr, keep=f, name=name)
    return pm


def write_material(path, spec: synth.MaterialSpec, name=None):
    """A synthetic material in the reference's HDF5 layout: the Kieft groups,
    and the full Penn groups (computed for the spec's Penn twin if the spec
    itself is Kieft)."""
    raw = synth.raw_material(spec)
    praw = raw if spec.inelastic == "penn" else synth.raw_material(synth.penn(spec))
    k = raw.raw_struct(L.KieftRaw)
    p = praw.penn_struct(L.PennRaw)
    _call("nbl_mat_write", str(path).encode(), (name or spec.name).encode(), C.byref(k), C.byref(p))
