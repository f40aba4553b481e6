"""ctypes bindings of libnbl_hip.so (the C ABI in include/nbl_hip.h).

The shared library is built in-tree by ``nebula_amd.build.build_library()``
(``make -C nebula_amd/csrc``) into ``nebula_amd/_build/libnbl_hip.so``.
Importing this module fails loudly if it is missing: there is no CPU
fallback for the transport path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# NBL_HIP_LIB selects another build of the library (e.g. an instrumented one)
LIB_PATH = os.environ.get("NBL_HIP_LIB") or os.path.join(HERE, "_build", "libnbl_hip.so")

NBL_OK = 0
NBL_ERR_INVALID = -1
NBL_ERR_ALLOC = -2
NBL_ERR_HIP = -3
NBL_ERR_RUNTIME = -4
NBL_ERR_NODEVICE = -5
NO_TRIANGLE = 0xFFFFFFFF

MAT_NOP, MAT_TERMINATOR, MAT_DETECTOR = -128, -127, -126
MAT_DETECTOR_LT50, MAT_DETECTOR_GE50, MAT_VACUUM, MAT_MIRROR = -125, -124, -123, -122

TAB_ELASTIC_IMFP, TAB_INELASTIC_IMFP = 0, 1
TAB_ELASTIC_ICDF, TAB_INELASTIC_W0_ICDF, TAB_IONISATION = 2, 3, 4
TAB_PENN_Q_ICDF, TAB_PENN_IONISATION = 5, 6
EVENT_SAMPLE_PATH, EVENT_ELASTIC, EVENT_INELASTIC, EVENT_INTERSECT = 0, 1, 2, 3
OUT_ALIVE, OUT_DETECTED, OUT_TERMINATED = 0, 1, 2
NUM_UNIFORMS = 8
HIST_BINS = 1024

ST_PENDING, ST_INELASTIC, ST_ELASTIC, ST_INTERSECT = 0x0, 0x4, 0x1, 0x2
ST_NO_EVENT, ST_DETECTED, ST_NEW_SECONDARY, ST_TERMINATED = 0x6, 0xA, 0xE, 0x3

# numpy dtypes with the C layouts of nbl_triangle / nbl_particle
TRIANGLE_DTYPE = np.dtype([("material_in", "<i4"), ("material_out", "<i4"),
                           ("r0", "<f4", 3), ("e1", "<f4", 3), ("e2", "<f4", 3)])
PARTICLE_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("kin_energy", "<f4")])
assert TRIANGLE_DTYPE.itemsize == 44 and PARTICLE_DTYPE.itemsize == 28


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Table1D(C.Structure):
    _fields_ = [("n", C.c_int32), ("x_min", C.c_float), ("x_max", C.c_float),
                ("data", C.POINTER(C.c_float))]


class Table2D(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("x_min", C.c_float),
                ("x_max", C.c_float), ("y_min", C.c_float), ("y_max", C.c_float),
                ("data", C.POINTER(C.c_float))]


class Table3D(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("depth", C.c_int32),
                ("x_min", C.c_float), ("x_max", C.c_float), ("y_min", C.c_float), ("y_max", C.c_float),
                ("z_min", C.c_float), ("z_max", C.c_float), ("data", C.POINTER(C.c_float))]


INELASTIC_KIEFT, INELASTIC_PENN = 0, 1


class MaterialDesc(C.Structure):
    _fields_ = [("barrier", C.c_float), ("fermi", C.c_float), ("band_gap", C.c_float),
                ("phonon_loss", C.c_float), ("recoil_const", C.c_float),
                ("elastic_log_imfp", Table1D), ("elastic_costheta_icdf", Table2D),
                ("inelastic_log_imfp", Table1D), ("inelastic_log_w0_icdf", Table2D),
                ("ionisation", Table2D), ("inelastic_model", C.c_int32),
                ("penn_q_icdf", Table3D), ("penn_ionisation", Table3D)]


class PennRaw(C.Structure):
    _fields_ = [("fermi", C.c_float), ("band_gap", C.c_float),
                ("n_k", C.c_int32), ("k_min", C.c_double), ("k_max", C.c_double),
                ("n_p", C.c_int32), ("p_min", C.c_double), ("p_max", C.c_double),
                ("imfp", C.POINTER(C.c_float)), ("omega_icdf", C.POINTER(C.c_float)),
                ("q_nk", C.c_int32), ("q_nw", C.c_int32), ("q_np", C.c_int32),
                ("q_k_min", C.c_double), ("q_k_max", C.c_double), ("q_w_min", C.c_double),
                ("q_w_max", C.c_double), ("q_p_min", C.c_double), ("q_p_max", C.c_double),
                ("q_icdf", C.POINTER(C.c_float)),
                ("i_nk", C.c_int32), ("i_nw", C.c_int32), ("i_np", C.c_int32),
                ("i_k_min", C.c_double), ("i_k_max", C.c_double), ("i_w_min", C.c_double),
                ("i_w_max", C.c_double), ("i_p_min", C.c_double), ("i_p_max", C.c_double),
                ("binding_icdf", C.POINTER(C.c_float))]


class KieftRaw(C.Structure):
    _fields_ = [("barrier", C.c_float), ("fermi", C.c_float), ("band_gap", C.c_float),
                ("phonon_loss", C.c_float), ("effective_a_grams", C.c_double),
                ("n_k", C.c_int32), ("k_min", C.c_double), ("k_max", C.c_double),
                ("n_p", C.c_int32), ("p_min", C.c_double), ("p_max", C.c_double),
                ("elastic_imfp", C.POINTER(C.c_float)), ("elastic_costheta", C.POINTER(C.c_float)),
                ("inelastic_imfp", C.POINTER(C.c_float)), ("inelastic_w0", C.POINTER(C.c_float)),
                ("binding_icdf", C.POINTER(C.c_float)), ("n_outer_shells", C.c_int32),
                ("outer_shells", C.POINTER(C.c_double))]


class GeometryInfo(C.Structure):
    _fields_ = [("aabb_min", Vec3), ("aabb_max", Vec3), ("max_extent", C.c_float),
                ("n_triangles_in", C.c_uint32), ("n_triangles_unique", C.c_uint32),
                ("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32), ("n_leaf_refs", C.c_uint32),
                ("max_depth", C.c_uint32), ("linearized_size", C.c_uint64),
                ("device_bytes", C.c_uint64)]


class DriverOpts(C.Structure):
    _fields_ = [("threads_per_block", C.c_uint32), ("reserved", C.c_uint32 * 7)]


class SlotView(C.Structure):
    _fields_ = [("status", C.c_void_p), ("particle_idx", C.c_void_p), ("particles", C.c_void_p),
                ("tags", C.c_void_p), ("material", C.c_void_p), ("last_tri", C.c_void_p),
                ("rng", C.c_void_p)]


class Pixel(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("energy", C.c_float),
                ("sigma", C.c_float), ("half_x", C.c_float), ("half_y", C.c_float),
                ("profile", C.c_uint32), ("n_primaries", C.c_uint32), ("reserved", C.c_uint32)]


PIXEL_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("energy", "<f4"),
                        ("sigma", "<f4"), ("half_x", "<f4"), ("half_y", "<f4"),
                        ("profile", "<u4"), ("n_primaries", "<u4"), ("reserved", "<u4")])
assert PIXEL_DTYPE.itemsize == C.sizeof(Pixel)


class PixelBatch(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("n_pixels", C.c_uint32), ("dir", Vec3),
                ("seed", C.c_uint64)]


class Tallies(C.Structure):
    _fields_ = [("primaries", C.c_uint64), ("detected_se", C.c_uint64), ("detected_bse", C.c_uint64),
                ("events", C.c_uint64), ("lookups_1d", C.c_uint64), ("lookups_2d", C.c_uint64),
                ("rounddowns", C.c_uint64), ("node_steps", C.c_uint64), ("tri_tests", C.c_uint64),
                ("boundary_events", C.c_uint64), ("secondaries", C.c_uint64),
                ("pending_retries", C.c_uint64), ("stopped", C.c_uint64), ("kernel_ms", C.c_double),
                ("pixel_se", C.c_void_p), ("pixel_bse", C.c_void_p), ("hist", C.c_void_p),
                ("deposit", C.c_void_p), ("detected", C.c_void_p), ("deposits", C.c_void_p)]


class DetectionBuffer(C.Structure):
    _fields_ = [("particles", C.c_void_p), ("tags", C.c_void_p), ("capacity", C.c_uint64), ("count", C.c_uint64)]


class DepositBuffer(C.Structure):
    _fields_ = [("records", C.c_void_p), ("tags", C.c_void_p), ("capacity", C.c_uint64), ("count", C.c_uint64)]


DEPOSIT_RECORD_DTYPE = np.dtype([("pos", "<f4", 3), ("energy", "<f4"), ("deposit", "<f4")])


class DepositGrid(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("lo", Vec3), ("hi", Vec3),
                ("energy", C.c_void_p), ("total", C.c_double), ("count", C.c_uint64)]


class PersistentOpts(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("stack_depth", C.c_uint32), ("refill_threshold", C.c_uint32),
                ("overflow_depth", C.c_uint32), ("inelastic_batch", C.c_uint32),
                ("inelastic_wait", C.c_uint32), ("allow_drops", C.c_uint32), ("count_work", C.c_uint32)]


class Int2(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32)]


class PriWarnings(C.Structure):
    _fields_ = [("outside_geometry", C.c_uint64), ("low_energy", C.c_uint64), ("high_energy", C.c_uint64),
                ("bad_direction", C.c_uint64)]


INT2_DTYPE = np.dtype([("x", np.int32), ("y", np.int32)])


class SamplerOpts(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("batch", C.c_uint32), ("lr", C.c_float),
                ("e_min", C.c_float), ("e_max", C.c_float), ("seed", C.c_uint64),
                ("kinds", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class SamplerReport(C.Structure):
    _fields_ = [("loss_elastic", C.c_double), ("loss_inelastic", C.c_double), ("train_ms", C.c_double),
                ("steps", C.c_uint64), ("samples", C.c_uint64), ("e_min", C.c_double), ("e_max", C.c_double)]


SAMPLER_ELASTIC, SAMPLER_INELASTIC = 0, 1
SAMPLER_PARAMS = 9988


DETECT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32, C.c_void_p)
POOL_FILL_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p)

_P = C.c_void_p
_U32, _I32, _U64, _F = C.c_uint32, C.c_int32, C.c_uint64, C.c_float

# name -> argtypes (every function returns int)
SIGNATURES = {
    "nbl_version": [C.POINTER(C.c_int)] * 3,
    "nbl_device_info": [C.c_char_p, C.c_size_t, C.POINTER(C.c_int)],
    "nbl_bench_copy": [_U64, _I32, C.POINTER(C.c_double)],
    "nbl_geometry_create": [_P, _U32, C.POINTER(_P)],
    "nbl_geometry_destroy": [_P],
    "nbl_geometry_get_info": [_P, C.POINTER(GeometryInfo)],
    "nbl_geometry_export_linearized": [_P, _P, _P],
    "nbl_octree_linearize": [_P, _U32, _P, C.POINTER(_U64), _P, C.POINTER(_U32), _P, _P],
    "nbl_geometry_in_domain_batch": [_P, _P, _U32, _P],
    "nbl_geometry_propagate_batch": [_P, _P, _P, _P, _P, _P, _U32, _P, _P],
    "nbl_geometry_propagate_batch_device": [_P, _P, _P, _P, _P, _P, _U32, _P, _P],
    "nbl_kieft_prepare": [C.POINTER(KieftRaw), _P, _P, _P, _P, _P, C.POINTER(MaterialDesc)],
    "nbl_materials_create": [C.POINTER(MaterialDesc), _I32, C.POINTER(_P)],
    "nbl_penn_prepare": [C.POINTER(PennRaw), _P, _P, _P, _P, C.POINTER(MaterialDesc)],
    "nbl_table3d_get_batch": [_P, _I32, _I32, _I32, _P, _P, _P, _U32, _P],
    "nbl_materials_destroy": [_P],
    "nbl_materials_get_max_energy": [_P, C.POINTER(_F)],
    "nbl_table1d_get_batch": [_P, _I32, _I32, _P, _U32, _P],
    "nbl_table2d_get_batch": [_P, _I32, _I32, _I32, _P, _P, _U32, _P],
    "nbl_physics_batch": [_P, _P, _I32, _P, _P, _P, _U32, _P, _P, _P, _P, _P, _P, _P],
    "nbl_driver_create": [_U32, _P, _P, _F, _F, _U64, C.POINTER(DriverOpts), C.POINTER(_P)],
    "nbl_driver_destroy": [_P],
    "nbl_driver_push": [_P, _P, _P, _U32, C.POINTER(_U32)],
    "nbl_driver_do_iteration": [_P],
    "nbl_driver_get_running_count": [_P, C.POINTER(_U32)],
    "nbl_driver_get_detected_count": [_P, C.POINTER(_U32)],
    "nbl_driver_flush_detected": [_P, DETECT_FN, _P],
    "nbl_driver_allocate_input_buffers": [_P, _U32],
    "nbl_driver_push_to_buffer": [_P, POOL_FILL_FN, _P],
    "nbl_driver_push_to_simulation": [_P],
    "nbl_driver_buffer_detected": [_P],
    "nbl_driver_flush_buffered": [_P, DETECT_FN, _P, C.POINTER(_U32)],
    "nbl_driver_read_slots": [_P, C.POINTER(SlotView)],
    "nbl_driver_write_slots": [_P, C.POINTER(SlotView)],
    "nbl_sort_status_batch": [_P, _U32, _P],
    "nbl_rng_seed": [_U64, _U64, _U32, _P],
    "nbl_rng_uniform_batch": [_P, _U32, _U32, _P],
    "nbl_driver_run_persistent": [_P, C.POINTER(PixelBatch), C.POINTER(PersistentOpts), C.POINTER(Tallies)],
    "nbl_driver_run_persistent_list": [_P, _P, _P, _U64, C.POINTER(PersistentOpts), C.POINTER(Tallies)],
    "nbl_driver_run_persistent_list_noalloc": [_P, _P, _P, _U64, _P, _P, _U64, C.POINTER(PersistentOpts),
                                               C.POINTER(Tallies)],
    "nbl_tri_file_load": [C.c_char_p, _P, _U32, C.POINTER(_U32)],
    "nbl_pri_file_load": [C.c_char_p, Vec3, Vec3, _F, _P, _P, _U64, C.POINTER(_U64), C.POINTER(PriWarnings)],
    "nbl_pri_sort_morton": [_P, _P, _U64],
    "nbl_pri_prescan_shuffle": [_P, _P, _U64, _U64],
    "nbl_sampler_create": [_P, C.POINTER(SamplerOpts), C.POINTER(_P)],
    "nbl_sampler_destroy": [_P],
    "nbl_sampler_get_report": [_P, C.POINTER(SamplerReport)],
    "nbl_sampler_get_params": [_P, _I32, _P, _U32, _P],
    "nbl_sampler_eval_batch": [_P, _I32, _P, _P, _P, _U32, _P],
    "nbl_sampler_selftest": [C.POINTER(C.c_double)],
    "nbl_driver_set_sampler": [_P, _P],
}


class NblError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libnbl_hip.so not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C nebula_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    lib.nbl_last_error.restype = C.c_char_p
    lib.nbl_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = C.c_int
        f.argtypes = args
    return lib


lib = _load()


def check(func_name, code):
    if code != NBL_OK:
        raise NblError(func_name, code, lib.nbl_last_error().decode(errors="replace"))


def call(name, *args):
    check(name, getattr(lib, name)(*args))


def ptr(a):
    """Raw pointer of a contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


def missing_symbols(header_path):
    """Names declared in include/nbl_hip.h that the library does not export."""
    import re
    text = open(header_path).read()
    names = set(re.findall(r"\b(nbl_[a-z0-9_]+)\s*\(", text))
    names -= {"nbl_detect_fn", "nbl_pool_fill_fn"}
    return sorted(n for n in names if not hasattr(lib, n))
