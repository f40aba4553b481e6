"""ctypes bindings of the CPU oracle (oracle/_build/liborc.so) and of the
reference-code harness (oracle/_ref/libnbl_ref.so).  Test infrastructure.

The harness is compiled from the reference's own sources, so it is loaded
only where those sources are (REF_SOURCES): on a machine without them the
GPU tests check against the committed fixtures and the oracle alone, and a
stray copy of the library is never loaded."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "_build", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libnbl_ref.so")
REF_SOURCES = "/root/reference/source/geometry/octree.h"

_P, _U32, _I32, _U64, _F, _D = C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_double


class RunStats(C.Structure):
    _fields_ = [("primaries", _U64), ("events", _U64), ("detected", _U64), ("detected_se", _U64),
                ("detected_bse", _U64), ("lookups_1d", _U64), ("lookups_2d", _U64), ("rounddowns", _U64),
                ("node_steps", _U64), ("tri_tests", _U64), ("boundary_events", _U64),
                ("secondaries", _U64), ("seconds", _D), ("threads", _U32)]


def load_oracle(path=ORC_PATH):
    """The oracle library at path (liborc.so, or a test's mutated build of
    its source) with the argument types of nbl_oracle.h set."""
    if not os.path.exists(path):
        raise ImportError(f"oracle not built: {path} (run make -C oracle)")
    lib = C.CDLL(path)
    sig = {
        "orc_geometry_create": (_P, [_P, _U32]),
        "orc_geometry_destroy": (None, [_P]),
        "orc_geometry_info": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "orc_geometry_export": (None, [_P, _P, _P]),
        "orc_in_domain_batch": (None, [_P, _P, _U32, _P]),
        "orc_propagate_batch": (None, [_P, _P, _P, _P, _P, _P, _U32, _P, _P, _P, _P]),
        "orc_trilist_propagate_batch": (None, [_P, _U32, _P, _P, _P, _P, _P, _U32, _P, _P]),
        "orc_intersect_ray_batch": (None, [_P, _P, _P, _U32, _P]),
        "orc_kieft_prepare": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
        "orc_materials_create": (_P, [_P, _I32]),
        "orc_materials_destroy": (None, [_P]),
        "orc_materials_max_energy": (_F, [_P]),
        "orc_table1d_get_batch": (None, [_P, _I32, _I32, _P, _U32, _P]),
        "orc_table2d_get_batch": (None, [_P, _I32, _I32, _I32, _P, _P, _U32, _P]),
        "orc_penn_prepare": (C.c_int, [_P, _P, _P, _P, _P, _P]),
        "orc_table3d_get_batch": (None, [_P, _I32, _I32, _I32, _P, _P, _P, _U32, _P]),
        "orc_penn_plasmon_x": (None, [_F, _U32, _U64, _P]),
        "orc_physics_batch": (C.c_int, [_P, _P, _I32, _P, _P, _P, _U32, _P, _P, _P, _P, _P, _P, _P]),
        "orc_rng_seed": (None, [_U64, _U64, _U32, _P]),
        "orc_rng_uniform_batch": (None, [_P, _U32, _U32, _P]),
        "orc_slots_create": (_P, [_U32, _U64]),
        "orc_slots_destroy": (None, [_P]),
        "orc_slots_read": (None, [_P, _P]),
        "orc_slots_write": (None, [_P, _P]),
        "orc_slots_push": (_U32, [_P, _P, _P, _U32]),
        "orc_slots_iteration": (None, [_P, _P, _P, _F, _F]),
        "orc_sort_status": (None, [_P, _U32, _P]),
        "orc_cpu_mt_run": (C.c_int, [_P, _P, _P, _P, _U64, _F, _U32, _I32, _I32, _D, _P, _P, _U64, _P,
                                     C.POINTER(RunStats)]),
        "orc_cpu_mt_run_deposit": (C.c_int, [_P, _P, _P, _P, _U64, _F, _U32, _I32, _P, _P, _I32, _I32, _I32,
                                             _P, _P, _P, _P, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _load_ref():
    if not (os.path.exists(REF_PATH) and os.path.exists(REF_SOURCES)):
        return None
    lib = C.CDLL(REF_PATH)
    sig = {
        "ref_geometry_create": (_P, [_P, _U32]),
        "ref_geometry_destroy": (None, [_P]),
        "ref_geometry_has_base": (C.c_int, [_P]),
        "ref_geometry_info": (None, [_P, _P, _P, _P, _P, _P]),
        "ref_geometry_export": (None, [_P, _P, _P]),
        "ref_in_domain_batch": (None, [_P, _P, _U32, _P]),
        "ref_propagate_batch": (C.c_int, [_P, _P, _P, _P, _P, _P, _U32, _P, _P]),
        "ref_trilist_propagate_batch": (None, [_P, _U32, _P, _P, _P, _P, _P, _U32, _P, _P]),
        "ref_intersect_ray_batch": (None, [_P, _P, _P, _U32, _P]),
        "ref_table1d_get_batch": (None, [_I32, _F, _F, _P, _P, _U32, _P]),
        "ref_table2d_get_batch": (None, [_I32, _I32, _F, _F, _F, _F, _P, _I32, _P, _P, _U32, _P]),
        "ref_table3d_get_batch": (None, [_I32, _I32, _I32, _F, _F, _F, _F, _F, _F, _P, _I32, _P, _P, _P, _U32,
                                         _P]),
        "ref_tri_file_load": (C.c_int64, [C.c_char_p, _P]),
        "ref_pri_file_load": (C.c_int64, [C.c_char_p, _P, _P, _F, C.c_int, C.c_uint64, _P, _P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name, None)
        if f is None:
            return None                      # an older harness build: rebuild with make -C oracle
        f.restype = res
        f.argtypes = args
    return lib


orc = load_oracle()
ref = _load_ref()


def p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class OrcGeometry:
    def __init__(self, tris):
        self.tris = np.ascontiguousarray(tris)
        self.h = orc.orc_geometry_create(p(self.tris), len(self.tris))

    def __del__(self):
        if getattr(self, "h", None):
            orc.orc_geometry_destroy(self.h)

    def info(self):
        c = np.zeros(3, np.float32); h = np.zeros(3, np.float32)
        me = C.c_float(); ls = C.c_uint64(); nu = C.c_uint32(); nn = C.c_uint32(); nl = C.c_uint32(); md = C.c_uint32()
        orc.orc_geometry_info(self.h, p(c), p(h), C.byref(me), C.byref(ls), C.byref(nu), C.byref(nn), C.byref(nl), C.byref(md))
        return dict(center=c, halfsize=h, max_extent=me.value, lin_size=ls.value, n_unique=nu.value,
                    n_nodes=nn.value, n_leaves=nl.value, max_depth=md.value)

    def export(self):
        i = self.info()
        d = np.zeros(i["lin_size"], np.int32); o = np.zeros(i["n_unique"], np.uint32)
        orc.orc_geometry_export(self.h, p(d), p(o))
        return d, o

    def propagate(self, pos, dir, dist, ignore_tri, ignore_mat, counters=False):
        n = len(pos)
        od = np.zeros(n, np.float32); ot = np.zeros(n, np.uint32)
        ns = np.zeros(n, np.uint32) if counters else None
        tt = np.zeros(n, np.uint32) if counters else None
        orc.orc_propagate_batch(self.h, p(pos), p(dir), p(dist), p(ignore_tri), p(ignore_mat), n, p(od), p(ot), p(ns), p(tt))
        return (od, ot, ns, tt) if counters else (od, ot)

    def in_domain(self, pos):
        out = np.zeros(len(pos), np.uint8)
        orc.orc_in_domain_batch(self.h, p(pos), len(pos), p(out))
        return out


class RefGeometry:
    def __init__(self, tris):
        self.tris = np.ascontiguousarray(tris)
        self.h = ref.ref_geometry_create(p(self.tris), len(self.tris))

    def __del__(self):
        if getattr(self, "h", None):
            ref.ref_geometry_destroy(self.h)

    def info(self):
        c = np.zeros(3, np.float32); h = np.zeros(3, np.float32)
        me = C.c_float(); ls = C.c_uint64(); nu = C.c_uint32()
        ref.ref_geometry_info(self.h, p(c), p(h), C.byref(me), C.byref(ls), C.byref(nu))
        return dict(center=c, halfsize=h, max_extent=me.value, lin_size=ls.value, n_unique=nu.value)

    def export(self):
        i = self.info()
        d = np.zeros(i["lin_size"], np.int32); o = np.zeros(i["n_unique"], np.uint32)
        ref.ref_geometry_export(self.h, p(d), p(o))
        return d, o

    def propagate(self, pos, dir, dist, ignore_tri, ignore_mat):
        n = len(pos)
        od = np.zeros(n, np.float32); ot = np.zeros(n, np.uint32)
        rc = ref.ref_propagate_batch(self.h, p(pos), p(dir), p(dist), p(ignore_tri), p(ignore_mat), n, p(od), p(ot))
        assert rc == 0, "reference harness could not map ignore_tri"
        return od, ot

    def in_domain(self, pos):
        out = np.zeros(len(pos), np.uint8)
        ref.ref_in_domain_batch(self.h, p(pos), len(pos), p(out))
        return out


class OrcMaterials:
    """Oracle materials from prepared nbl_material_desc structs (kept alive by the caller)."""

    def __init__(self, descs):
        from nebula_amd._lib import MaterialDesc
        self.arr = (MaterialDesc * len(descs))(*descs)
        self.h = orc.orc_materials_create(C.addressof(self.arr), len(descs))

    def __del__(self):
        if getattr(self, "h", None):
            orc.orc_materials_destroy(self.h)
