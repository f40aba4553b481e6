"""HDF5 material files (libnbl_mat.so, include/nbl_mat.h) without a GPU:
the unit parser against the reference's unit system by known answers; a
written file read back bit for bit and prepared identically to the same
tables given directly; the file's layout as an independent HDF5 reader
(h5dump) sees it; the reference loader's error rules (missing groups,
non-logarithmic axes, default band gap)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import synth

matfile = pytest.importorskip("nebula_amd.matfile")
MAT_SO = os.path.join(os.path.dirname(L.LIB_PATH), "libnbl_mat.so")
pytestmark = pytest.mark.skipif(not os.path.exists(MAT_SO), reason="libnbl_mat.so not built (no HDF5 headers)")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
HDF5_SO = "/opt/conda/lib/libhdf5.so"


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(MAT_SO)
    names = set(re.findall(r"\b(nbl_mat_[a-z0-9_]+)\s*\(", open(os.path.join(ROOT, "include", "nbl_mat.h")).read()))
    assert names and all(hasattr(lib, n) for n in names), sorted(n for n in names if not hasattr(lib, n))


@pytest.mark.parametrize("text, value, dims", [
    ("eV", 1.0, (1, 0, 0, 0, 0)),
    ("nm^-1", 1.0, (0, -1, 0, 0, 0)),
    ("m^-1", 1e-9, (0, -1, 0, 0, 0)),
    ("/nm", 1.0, (0, -1, 0, 0, 0)),
    ("g/cm^3", 6.2415e15 / 1e21, (1, -5, 2, 0, 0)),     # mass density: dimension.h:92
    ("g", 6.2415e15, (1, -2, 2, 0, 0)),
    ("eV*s", 1e9, (1, 0, 1, 0, 0)),
    ("C", 6.2415e18, (0, 0, 0, 0, 1)),
    ("K", 1.0, (0, 0, 0, 1, 0)),
    ("radian", 1.0, (0, 0, 0, 0, 0)),
    ("", 1.0, (0, 0, 0, 0, 0)),
])
def test_unit_parser_known_answers(text, value, dims):
    v, d = matfile.parse_unit(text)
    assert d == dims and v == pytest.approx(value, rel=1e-15)


@pytest.mark.parametrize("text", ["keV", "1/nm", "nm^x", "J"])
def test_unit_parser_rejects_unknown(text):
    with pytest.raises(matfile.MatError):
        matfile.parse_unit(text)


@pytest.fixture(scope="module")
def si_file(tmp_path_factory):
    p = tmp_path_factory.mktemp("mat") / "si.mat"
    matfile.write_material(p, synth.SI)
    return p


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_file_round_trip_is_bit_exact(si_file):
    raw = synth.raw_material(synth.SI)
    praw = synth.raw_material(synth.SI_PENN)
    with matfile.MatFile(si_file) as f:
        assert f.get_string("name") == "Si"
        assert f.exists("/kieft/elastic") and f.exists("kieft/inelastic") and f.exists("/full_penn")
        assert not f.exists("/nothing")
        assert f.get_quantity("fermi") == (float(np.float32(7.8)), (1, 0, 0, 0, 0))
        k = f.kieft_raw()
        nk, np_ = synth.N_K, synth.N_P
        assert (k.n_k, k.k_min, k.k_max, k.n_p, k.p_min, k.p_max) == (nk, synth.K_MIN, synth.K_MAX, np_, synth.P_MIN,
                                                                      synth.P_MAX)
        get = lambda ptr_, n: np.ctypeslib.as_array(ptr_, (n,)).copy()
        for name, ref in (("elastic_imfp", raw.el_imfp), ("elastic_costheta", raw.el_cos),
                          ("inelastic_imfp", raw.in_imfp), ("inelastic_w0", raw.in_w0), ("binding_icdf", raw.binding)):
            assert np.array_equal(_bits(get(getattr(k, name), ref.size)), _bits(ref.ravel())), name
        assert get(k.outer_shells, k.n_outer_shells).tolist() == raw.outer.tolist()
        for fld in ("barrier", "fermi", "band_gap", "phonon_loss"):
            assert getattr(k, fld) == np.float32(getattr(synth.SI, fld) if fld != "barrier" else synth.SI.barrier)
        assert k.effective_a_grams == pytest.approx(synth.SI.atomic_mass / synth.AVOGADRO, rel=1e-15)
        p = f.penn_raw()
        n = synth.PENN_N
        assert (p.q_nk, p.q_nw, p.q_np, p.i_nk, p.i_nw, p.i_np) == (n,) * 6
        assert (p.q_w_min, p.q_w_max, p.i_w_min, p.i_w_max) == (0.0, 1.0, synth.PENN_W_MIN, 1.0)
        for name, ref in (("imfp", praw.in_imfp), ("omega_icdf", praw.in_w0), ("q_icdf", praw.penn_q),
                          ("binding_icdf", praw.penn_binding)):
            assert np.array_equal(_bits(get(getattr(p, name), ref.size)), _bits(ref.ravel())), name


@pytest.mark.parametrize("model", ["kieft", "penn"])
def test_loaded_material_prepares_like_the_direct_tables(si_file, model):
    pm = matfile.load_material(si_file, model)
    ref = na.prepared(synth.SI if model == "kieft" else synth.SI_PENN)
    keys = ["el_imfp", "el_icdf", "in_imfp", "in_w0"] + (["penn_q", "penn_ion"] if model == "penn" else ["ion"])
    for k in keys:
        assert np.array_equal(_bits(getattr(pm, k)), _bits(getattr(ref, k))), k
    for f in ("barrier", "fermi", "band_gap", "phonon_loss", "recoil_const", "inelastic_model"):
        assert getattr(pm.desc, f) == getattr(ref.desc, f), f
    assert pm.inelastic == model and pm.name == "Si"


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_layout_as_h5dump_sees_it(si_file):
    """An independent reader: attribute values and units, table shapes and
    units, dimension scales attached to the tables."""
    def dump(*args):
        return subprocess.run([H5DUMP, *args, str(si_file)], capture_output=True, text=True, timeout=60,
                              check=True).stdout
    a = dump("-a", "/fermi")
    assert re.search(r"7\.8\d*,\s*\"eV\"", a), a
    assert '"g"' in dump("-a", "/effective_A")
    h = dump("-H", "-A")
    for path, shape in (("imfp", "( 1024 )"), ("costheta_icdf", "( 1024, 1024 )"), ("w0_icdf", "( 1024, 1024 )"),
                        ("binding_icdf", "( 1024, 1024 )"), ("q_icdf", "( 64, 64, 64 )")):
        assert re.search(rf'DATASET "{path}" {{.*?SIMPLE {{ {re.escape(shape)}', h, re.S), path
    assert h.count('ATTRIBUTE "DIMENSION_LIST"') == 5 + 4
    assert h.count('"DIMENSION_SCALE"') >= 16
    units = dump("-a", "/kieft/elastic/imfp/units")
    assert '"nm^-1"' in units
    part = dump("-d", "/kieft/inelastic/imfp", "-s", "500", "-c", "4", "-m", "%.9g")
    data = part[part.index("DATA {"):part.index("}", part.index("DATA {"))]
    vals = [float(x) for x in re.findall(r"\(\d+\): ([-+0-9.eE]+)", data)]
    assert np.array_equal(np.float32(vals), synth.raw_material(synth.SI).in_imfp[500:504])


def _h5():
    """The HDF5 C library through ctypes, for editing a written file."""
    if not os.path.exists(HDF5_SO):
        pytest.skip("libhdf5 not found")
    h = C.CDLL(HDF5_SO)
    h.H5open()
    for name in ("H5Fopen", "H5Dopen2"):
        getattr(h, name).restype = C.c_int64
    h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
    h.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    h.H5Ldelete.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
    h.H5Adelete.argtypes = [C.c_int64, C.c_char_p]
    h.H5Dwrite.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    h.H5Dclose.argtypes = [C.c_int64]
    h.H5Fclose.argtypes = [C.c_int64]
    return h


def _edit(path, fn):
    h = _h5()
    fid = h.H5Fopen(str(path).encode(), 1, 0)      # H5F_ACC_RDWR, H5P_DEFAULT
    assert fid >= 0
    try:
        fn(h, fid)
    finally:
        assert h.H5Fclose(fid) >= 0


def test_loader_error_rules(si_file, tmp_path):
    with pytest.raises(matfile.MatError, match="Unable to open"):
        matfile.MatFile(tmp_path / "missing.mat")
    # no /full_penn group: the Kieft part still loads, the Penn part fails
    p1 = tmp_path / "kieft_only.mat"
    shutil.copy(si_file, p1)
    _edit(p1, lambda h, fid: h.H5Ldelete(fid, b"/full_penn", 0))
    with matfile.MatFile(p1) as f:
        assert not f.exists("/full_penn")
        f.kieft_raw()
        with pytest.raises(matfile.MatError, match="Full Penn model not found"):
            f.penn_raw()
    # band_gap is optional and defaults to -1 eV (inelastic.h:221)
    p2 = tmp_path / "no_gap.mat"
    shutil.copy(si_file, p2)
    _edit(p2, lambda h, fid: h.H5Adelete(fid, b"band_gap"))
    with matfile.MatFile(p2) as f:
        assert f.kieft_raw().band_gap == -1.0 and f.penn_raw().band_gap == -1.0
    # the logarithmic-axis check is one-sided (hdf5_file.cpp:340-346: it fails
    # only where geomspace(front, back) exceeds the stored value by more than
    # EPSILON): an axis below the geometric one is rejected, a linear one
    # (above it everywhere) passes
    def set_axis(values):
        def edit(h, fid):
            d = h.H5Dopen2(fid, b"/kieft/elastic/imfp_energy", 0)
            assert d >= 0
            dbl = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
            v = np.ascontiguousarray(values, np.float64)
            assert h.H5Dwrite(d, dbl, 0, 0, 0, v.ctypes.data) >= 0
            h.H5Dclose(d)
        return edit
    t = np.linspace(0, 1, synth.N_K)
    p3 = tmp_path / "below.mat"
    shutil.copy(si_file, p3)
    _edit(p3, set_axis(synth.K_MIN * (synth.K_MAX / synth.K_MIN) ** (t ** 2)))
    with matfile.MatFile(p3) as f:
        with pytest.raises(matfile.MatError, match="Logarithmic dimension scale expected"):
            f.kieft_raw()
    p4 = tmp_path / "linear.mat"
    shutil.copy(si_file, p4)
    _edit(p4, set_axis(np.linspace(synth.K_MIN, synth.K_MAX, synth.N_K)))
    with matfile.MatFile(p4) as f:
        k = f.kieft_raw()
        assert (k.k_min, k.k_max) == (synth.K_MIN, synth.K_MAX)
