"""The reference's file formats through libnbl_hip.so (host code, no GPU):
.tri text geometry and .pri primary records (io/load_tri_file.cpp,
io/load_pri_file.cpp), the optional Morton sort and the prescan shuffle,
each compared with the reference's own loader code built from its sources
(oracle/_ref), and against what the format fixes (closed forms: column
layout, record layout, warning rules)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import files

needs_ref = pytest.mark.skipif(ol.ref is None, reason="reference loaders not built (oracle/_ref)")


def _ref_tri(path):
    n = ol.ref.ref_tri_file_load(str(path).encode(), None)
    if n < 0:
        return None
    out = np.zeros(n, L.TRIANGLE_DTYPE)
    ol.ref.ref_tri_file_load(str(path).encode(), ol.p(out))
    return out


def _ref_pri(path, lo, hi, emax, sort=0, prescan=0):
    lo = np.asarray(lo, np.float32); hi = np.asarray(hi, np.float32)
    n = ol.ref.ref_pri_file_load(str(path).encode(), ol.p(lo), ol.p(hi), emax, 0, 0, None, None)
    parts = np.zeros(n, L.PARTICLE_DTYPE); pix = np.zeros((n, 2), np.int32)
    ol.ref.ref_pri_file_load(str(path).encode(), ol.p(lo), ol.p(hi), emax, sort, prescan, ol.p(parts), ol.p(pix))
    return parts, pix


def _random_tri_text(tmp_path, n=500, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-100, 100, (n, 3)); b = a + rng.normal(0, 3, (n, 3)); c = a + rng.normal(0, 3, (n, 3))
    mi = rng.integers(-126, 4, n); mo = rng.integers(-126, 4, n)
    path = tmp_path / "g.tri"
    with open(path, "w") as f:
        f.write("# comment line\n\n   # indented comment\n")
        for k in range(n):
            cols = [str(mi[k]), str(mo[k])] + [repr(float(v)) for v in (*a[k], *b[k], *c[k])]
            sep = "\t" if k % 3 == 0 else "  "
            f.write(sep.join(cols) + ("\r\n" if k % 7 == 0 else "\n"))
            if k % 50 == 0:
                f.write("\n")
    return path, mi, mo, a, b, c


def test_tri_file_layout(tmp_path):
    path, mi, mo, a, b, c = _random_tri_text(tmp_path)
    t = files.load_tri_file(path)
    assert len(t) == len(a)
    assert np.array_equal(t["material_in"], mi) and np.array_equal(t["material_out"], mo)
    a32, b32, c32 = (x.astype(np.float32) for x in (a, b, c))
    assert np.array_equal(t["r0"], a32)
    assert np.array_equal(t["e1"], b32 - a32) and np.array_equal(t["e2"], c32 - a32)


@needs_ref
def test_tri_file_matches_reference_loader(tmp_path):
    path = _random_tri_text(tmp_path, n=2000, seed=1)[0]
    ours, ref = files.load_tri_file(path), _ref_tri(path)
    assert ours.tobytes() == ref.tobytes()
    # unreadable file: both give no triangles
    assert len(files.load_tri_file(tmp_path / "missing.tri")) == 0 and len(_ref_tri(tmp_path / "missing.tri")) == 0


@pytest.mark.parametrize("line", ["0 1 2 3 4 5 6 7 8 9", "0 1 2 3 4 5 6 7 8 9 10 11", "   "])
def test_tri_file_rejects_bad_lines(tmp_path, line):
    path = tmp_path / "bad.tri"
    path.write_text("0 -123 0 0 0 1 0 0 0 1 0\n" + line + "\n")
    with pytest.raises(na.NblError) as e:
        files.load_tri_file(path)
    assert "line 2" in str(e.value)
    if ol.ref is not None:
        assert _ref_tri(path) is None          # the reference throws too


def _random_pri(tmp_path, n=3000, seed=2, dup=False):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, L.PARTICLE_DTYPE)
    p["pos"] = rng.uniform(-60, 60, (n, 3))
    if dup:                                     # many primaries per pixel: equal sort keys
        p["pos"] = np.repeat(p["pos"][: n // 100], 100, axis=0)[:n]
    d = rng.normal(0, 1, (n, 3)); p["dir"] = d
    p["kin_energy"] = rng.uniform(0, 1200, n)
    p["kin_energy"][::97] = 0.0                 # low energy
    p["dir"][::89] = 0.0                        # null direction
    p["pos"][::83, 2] = np.nan                  # NaN position
    pix = rng.integers(-5, 1000, (n, 2)).astype(np.int32)
    path = tmp_path / "p.pri"
    files.write_pri_file(path, p, pix)
    return path, p, pix


def test_pri_file_records_and_warnings(tmp_path):
    path, p, pix = _random_pri(tmp_path)
    assert os.path.getsize(path) == 36 * len(p)
    lo, hi, emax = (-50.0, -50.0, -50.0), (50.0, 50.0, 50.0), 1000.0
    q, qpix, w = files.load_pri_file(path, lo, hi, emax)
    assert q.tobytes() == p.tobytes() and np.array_equal(qpix, pix)
    inside = np.all((p["pos"] > lo) & (p["pos"] < hi), axis=1)
    eps = np.float32(1.1920928955078125e-06)
    assert w["outside_geometry"] == int((~inside).sum())
    assert w["low_energy"] == int((~(p["kin_energy"] > eps)).sum())
    assert w["high_energy"] == int((p["kin_energy"] > emax).sum())
    assert w["bad_direction"] == int((~np.any(np.abs(p["dir"]) > eps, axis=1)).sum())
    # a truncated file is an error
    with open(path, "ab") as f:
        f.write(b"\0" * 7)
    with pytest.raises(na.NblError):
        files.load_pri_file(path, lo, hi, emax)


@needs_ref
@pytest.mark.parametrize("dup", [False, True])
def test_pri_sort_and_prescan_shuffle_match_reference(tmp_path, dup):
    path, p, pix = _random_pri(tmp_path, n=5000, dup=dup)
    p["pos"][np.isnan(p["pos"])] = 0.0          # the sort's quantisation of NaN is undefined behaviour
    files.write_pri_file(path, p, pix)
    lo, hi, emax = (-100.0,) * 3, (100.0,) * 3, 1e4
    for sort in (0, 1):
        q, qpix, _ = files.load_pri_file(path, lo, hi, emax)
        if sort:
            files.sort_primaries(q, qpix)
        files.prescan_shuffle(q, qpix, 1000)
        r, rpix = _ref_pri(path, lo, hi, emax, sort=sort, prescan=1000)
        assert q.tobytes() == r.tobytes() and np.array_equal(qpix, rpix), sort


def test_morton_sort_orders_by_quantised_position(tmp_path):
    """Closed form: after the sort the Morton keys of the 16-bit quantised
    positions (bounding box of the primaries) ascend."""
    path, p, pix = _random_pri(tmp_path, n=4000, seed=9)
    p["pos"][np.isnan(p["pos"])] = 0.0
    q = np.ascontiguousarray(p.copy()); qpix = np.ascontiguousarray(pix.copy())
    files.sort_primaries(q, qpix)
    lo = q["pos"].min(0); size = q["pos"].max(0) - lo
    qq = ((q["pos"] - lo) / size * np.float32(65535)).astype(np.uint16).astype(np.uint64)
    key = np.zeros(len(q), np.uint64)
    for i in range(16):
        for ax in range(3):
            key |= ((qq[:, ax] >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i + ax)
    assert np.all(np.diff(key.astype(np.float64)) >= 0)
    # a permutation of the input, pixels moved with their primaries
    assert sorted(map(tuple, qpix.tolist())) == sorted(map(tuple, pix.tolist()))
