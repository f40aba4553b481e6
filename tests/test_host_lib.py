"""Host-side checks of libnbl_hip.so that need no GPU: the library loads and
exports every entry point of include/nbl_hip.h, the host octree builder
matches the reference (golden fixtures) and the oracle, and the table
preparation matches the oracle's restatement."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import scenes, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    missing = L.missing_symbols(os.path.join(ROOT, "include", "nbl_hip.h"))
    assert missing == [], f"declared but not exported: {missing}"


def test_version_and_error_paths():
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    L.call("nbl_version", C.byref(a), C.byref(b), C.byref(c))
    assert (a.value, b.value, c.value) == (1, 0, 2)
    h = C.c_void_p()
    # invalid arguments are rejected before any device work
    assert L.lib.nbl_geometry_create(None, 0, C.byref(h)) == L.NBL_ERR_INVALID
    assert b"no triangles" in L.lib.nbl_last_error()
    assert L.lib.nbl_materials_create(None, 0, C.byref(h)) == L.NBL_ERR_INVALID
    assert L.lib.nbl_geometry_destroy(None) == L.NBL_OK


@pytest.mark.parametrize("name", ["cube", "soup", "lines"])
def test_host_builder_matches_reference_fixture(name):
    f = np.load(os.path.join(GOLDEN, f"geometry_{name}.npz"), allow_pickle=False)
    d, o, c, h = na.octree_linearize(f["triangles"])
    assert np.array_equal(d, f["octree_data"]) and np.array_equal(o, f["triangle_order"])
    assert np.array_equal(c, f["center"]) and np.array_equal(h, f["halfsize"])


def test_host_builder_matches_oracle_on_rough_stripes():
    tris = scenes.rough_surface_stripes(n=160)
    d, o, c, h = na.octree_linearize(tris)
    od, oo = ol.OrcGeometry(tris).export()
    assert np.array_equal(d, od) and np.array_equal(o, oo)


def test_builder_rejects_bad_input():
    t = scenes.cube_slab().copy()
    t["material_in"][0] = 40000
    with pytest.raises(L.NblError):
        na.octree_linearize(t)
    t = scenes.cube_slab().copy()
    t["r0"][1, 0] = np.nan
    with pytest.raises(L.NblError):
        na.octree_linearize(t)


@pytest.mark.parametrize("spec", [synth.SI, synth.PMMA, synth.W_LIKE], ids=lambda s: s.name)
def test_kieft_prepare_matches_oracle(spec):
    pm = na.prepared(spec)
    raw = pm.raw_struct
    nk, np_ = synth.N_K, synth.N_P
    outs = [np.zeros(nk, np.float32), np.zeros((nk, np_), np.float32), np.zeros(nk, np.float32),
            np.zeros((nk, np_), np.float32), np.zeros((nk, np_), np.float32)]
    desc = L.MaterialDesc()
    assert ol.orc.orc_kieft_prepare(C.addressof(raw), *[ol.p(a) for a in outs], C.addressof(desc)) == 0
    mine = [pm.el_imfp, pm.el_icdf, pm.in_imfp, pm.in_w0, pm.ion]
    for a, b in zip(mine, outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert pm.desc.recoil_const == desc.recoil_const
    for k in ("elastic_costheta_icdf", "ionisation"):
        x, y = getattr(pm.desc, k), getattr(desc, k)
        assert (x.x_min, x.x_max, x.y_min, x.y_max) == (y.x_min, y.x_max, y.y_min, y.y_max)


def test_kieft_prepare_semantics():
    """Invariants of the prepared tables (elastic.h:150-157, inelastic.h:248-323):
    cos(theta) in [-1,1]; log w0 <= log(K - E_F); binding values are -1, an outer
    shell, or an inner-shell edge >= 50 eV; no inner shell below 100 eV."""
    spec = synth.SI
    pm = na.prepared(spec)
    k = synth.energy_grid()
    assert pm.el_icdf.min() >= -1 and pm.el_icdf.max() <= 1
    above = k > spec.fermi
    lim = np.log(k[above] - spec.fermi)[:, None]
    assert np.all(pm.in_w0[above] <= (lim + 1e-5).astype(np.float32))
    allowed = {-1.0, *[float(np.float32(s)) for s in spec.outer_shells], *[float(np.float32(b)) for b, _ in spec.inner_shells]}
    assert set(np.unique(pm.ion).tolist()) <= allowed
    assert np.all(pm.ion[k <= 100.0] < 50)
    # screened-Rutherford closed form reproduced by the raw table at u = P
    raw = synth.raw_material(spec)
    eta = synth.screening(spec, k)[:, None]
    u = synth.prob_grid()[None, :]
    assert np.allclose(raw.el_cos, 1 - 2 * eta * u / (1 + eta - u), atol=2e-6)
