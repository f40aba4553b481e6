"""The reference's file formats (io/load_tri_file.cpp, io/load_pri_file.cpp)
through libnbl_hip.so, plus writers for them and for the .nblmat material
files that the nbl_gpu command reads.

  .tri     text, one triangle per line: material_in material_out Ax Ay Az Bx By Bz Cx Cy Cz
  .pri     36-byte records: pos (3 float32), dir (3 float32), energy (float32), pixel x, y (2 int32)
  output   detected electrons, the same 36-byte records (gpu_main.cu:199-206)
  .nblmat  raw material tables (layout in nebula_amd/csrc/nbl_gpu.cpp)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import call, ptr

PRI_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("kin_energy", "<f4"), ("pixel", "<i4", 2)])
assert PRI_DTYPE.itemsize == 36


def load_tri_file(path) -> np.ndarray:
    """load_tri_file: triangles as an nbl_triangle array (empty if the file cannot be opened)."""
    n = C.c_uint32()
    b = str(path).encode()
    call("nbl_tri_file_load", b, None, 0, C.byref(n))
    out = np.zeros(n.value, L.TRIANGLE_DTYPE)
    if n.value:
        call("nbl_tri_file_load", b, ptr(out), n.value, C.byref(n))
    return out


def write_tri_file(path, mat_in, mat_out, a, b, c, comment=None):
    """Write vertices (N,3 each) as a .tri text file (repr precision)."""
    a, b, c = (np.asarray(x, np.float64).reshape(-1, 3) for x in (a, b, c))
    mi = np.broadcast_to(np.asarray(mat_in), (len(a),)); mo = np.broadcast_to(np.asarray(mat_out), (len(a),))
    with open(path, "w") as f:
        if comment:
            f.write(f"# {comment}\n")
        for k in range(len(a)):
            f.write(f"{int(mi[k])} {int(mo[k])} " + " ".join(repr(float(v)) for v in (*a[k], *b[k], *c[k])) + "\n")


def load_pri_file(path, aabb_min, aabb_max, max_energy):
    """load_pri_file: (particles, pixels [N,2] int32, warnings dict)."""
    n = C.c_uint64()
    b = str(path).encode()
    lo, hi = L.Vec3(*map(float, aabb_min)), L.Vec3(*map(float, aabb_max))
    call("nbl_pri_file_load", b, lo, hi, float(max_energy), None, None, 0, C.byref(n), None)
    parts = np.zeros(n.value, L.PARTICLE_DTYPE)
    pix = np.zeros((n.value, 2), np.int32)
    w = L.PriWarnings()
    call("nbl_pri_file_load", b, lo, hi, float(max_energy), ptr(parts), ptr(pix), n.value, C.byref(n), C.byref(w))
    return parts, pix, {k: getattr(w, k) for k, _ in L.PriWarnings._fields_}


def write_pri_file(path, particles, pixels):
    rec = np.zeros(len(particles), PRI_DTYPE)
    rec["pos"], rec["dir"], rec["kin_energy"] = particles["pos"], particles["dir"], particles["kin_energy"]
    rec["pixel"] = pixels
    rec.tofile(path)


def read_records(path_or_bytes) -> np.ndarray:
    """Detected-electron records (the nbl_gpu / nebula_gpu output format)."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        return np.frombuffer(path_or_bytes, PRI_DTYPE)
    return np.fromfile(path_or_bytes, PRI_DTYPE)


def sort_primaries(particles, pixels):
    """sort_pri_file: in place, Morton order of the quantised positions."""
    assert particles.flags["C_CONTIGUOUS"] and pixels.flags["C_CONTIGUOUS"]
    call("nbl_pri_sort_morton", ptr(particles), ptr(pixels), len(particles))


def prescan_shuffle(particles, pixels, prescan_size):
    """prescan_shuffle: in place."""
    assert particles.flags["C_CONTIGUOUS"] and pixels.flags["C_CONTIGUOUS"]
    call("nbl_pri_prescan_shuffle", ptr(particles), ptr(pixels), len(particles), int(prescan_size))
