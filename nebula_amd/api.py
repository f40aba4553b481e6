"""Python mirror of the reference's host-side API for the transport path:
``octree<true>`` (Geometry), ``gpu_material_manager`` (Materials) and
``gpu_driver`` (Driver), all calling libnbl_hip.so through its C ABI.

Method names and argument meaning follow the reference
(geometry/octree.h:42-97, core/gpu_material_manager.h:37-74,
drivers/gpu/gpu_driver.h:58-182); errors are raised as NblError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import call, ptr
from . import synth


def as_vec3(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 3
    return a


class Geometry:
    """octree<true>: built on the host with the reference's rules, resident on the device."""

    def __init__(self, triangles: np.ndarray):
        self.triangles = np.ascontiguousarray(triangles, dtype=L.TRIANGLE_DTYPE)
        h = C.c_void_p()
        call("nbl_geometry_create", ptr(self.triangles), len(self.triangles), C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            L.lib.nbl_geometry_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def info(self) -> dict:
        i = L.GeometryInfo()
        call("nbl_geometry_get_info", self.h, C.byref(i))
        return {f: (np.array([getattr(i, f).x, getattr(i, f).y, getattr(i, f).z], np.float32)
                    if f.startswith("aabb") else getattr(i, f)) for f, _ in L.GeometryInfo._fields_}

    def export_linearized(self):
        inf = self.info()
        d = np.zeros(inf["linearized_size"], np.int32)
        o = np.zeros(inf["n_triangles_unique"], np.uint32)
        call("nbl_geometry_export_linearized", self.h, ptr(d), ptr(o))
        return d, o

    def in_domain(self, pos):
        pos = as_vec3(pos)
        out = np.zeros(len(pos), np.uint8)
        call("nbl_geometry_in_domain_batch", self.h, ptr(pos), len(pos), ptr(out))
        return out

    def propagate(self, pos, dir, dist, ignore_tri=None, ignore_mat=None):
        pos, dir = as_vec3(pos), as_vec3(dir)
        n = len(pos)
        dist = np.ascontiguousarray(dist, np.float32)
        it = np.full(n, L.NO_TRIANGLE, np.uint32) if ignore_tri is None else np.ascontiguousarray(ignore_tri, np.uint32)
        im = np.full(n, L.MAT_VACUUM, np.int32) if ignore_mat is None else np.ascontiguousarray(ignore_mat, np.int32)
        od = np.zeros(n, np.float32); ot = np.zeros(n, np.uint32)
        call("nbl_geometry_propagate_batch", self.h, ptr(pos), ptr(dir), ptr(dist), ptr(it), ptr(im), n, ptr(od), ptr(ot))
        return od, ot


def octree_linearize(triangles):
    """linearized_octree(octree_root(triangles)) on the host (no device needed)."""
    t = np.ascontiguousarray(triangles, dtype=L.TRIANGLE_DTYPE)
    size = C.c_uint64(); nu = C.c_uint32()
    call("nbl_octree_linearize", ptr(t), len(t), None, C.byref(size), None, C.byref(nu), None, None)
    d = np.zeros(size.value, np.int32); o = np.zeros(nu.value, np.uint32)
    c = np.zeros(3, np.float32); h = np.zeros(3, np.float32)
    call("nbl_octree_linearize", ptr(t), len(t), ptr(d), C.byref(size), ptr(o), C.byref(nu), ptr(c), ptr(h))
    return d, o, c, h


class PreparedMaterial:
    """A material's tables after the reference's create() post-processing
    (nbl_kieft_prepare, then nbl_penn_prepare for a full Penn material),
    held in numpy arrays; .desc points at them.  The raw tables come from a
    synthetic MaterialSpec or, as KieftRaw / PennRaw structs, from a material
    file (`keep` holds whatever owns their arrays)."""

    def __init__(self, spec: synth.MaterialSpec | None = None, kieft_raw=None, penn_raw=None, keep=None,
                 name=None):
        if spec is not None:
            if spec.inelastic not in ("kieft", "penn"):
                raise ValueError(f"unknown inelastic model {spec.inelastic!r}")
            keep = synth.raw_material(spec)
            kieft_raw = keep.raw_struct(L.KieftRaw)
            penn_raw = keep.penn_struct(L.PennRaw) if spec.inelastic == "penn" else None
        if kieft_raw is None:
            raise ValueError("PreparedMaterial needs a spec or raw tables")
        self.spec = spec
        self.name = name if name is not None else (spec.name if spec is not None else "")
        self._raw = keep
        nk, np_ = kieft_raw.n_k, kieft_raw.n_p
        self.el_imfp = np.zeros(nk, np.float32)
        self.el_icdf = np.zeros((nk, np_), np.float32)
        self.in_imfp = np.zeros(nk, np.float32)
        self.in_w0 = np.zeros((nk, np_), np.float32)
        self.ion = np.zeros((nk, np_), np.float32)
        self.raw_struct = kieft_raw
        self.desc = L.MaterialDesc()
        call("nbl_kieft_prepare", C.byref(self.raw_struct), ptr(self.el_imfp), ptr(self.el_icdf),
             ptr(self.in_imfp), ptr(self.in_w0), ptr(self.ion), C.byref(self.desc))
        if penn_raw is not None:
            r = penn_raw
            self.in_imfp = np.zeros(r.n_k, np.float32)
            self.in_w0 = np.zeros((r.n_k, r.n_p), np.float32)
            self.penn_q = np.zeros((r.q_nk, r.q_nw, r.q_np), np.float32)
            self.penn_ion = np.zeros((r.i_nk, r.i_nw, r.i_np), np.float32)
            self.penn_struct = penn_raw
            call("nbl_penn_prepare", C.byref(self.penn_struct), ptr(self.in_imfp), ptr(self.in_w0),
                 ptr(self.penn_q), ptr(self.penn_ion), C.byref(self.desc))

    @classmethod
    def from_desc(cls, desc, keep=None, name=""):
        """A material whose tables are already prepared: desc points at
        arrays that keep owns (they must outlive every Materials made from it)."""
        self = cls.__new__(cls)
        self.spec, self.name, self._raw, self.desc = None, name, keep, desc
        return self

    @property
    def inelastic(self) -> str:
        return "penn" if self.desc.inelastic_model == L.INELASTIC_PENN else "kieft"


_PREP: dict = {}


def prepared(spec: synth.MaterialSpec) -> PreparedMaterial:
    if spec.name not in _PREP or _PREP[spec.name].spec != spec:
        _PREP[spec.name] = PreparedMaterial(spec)
    return _PREP[spec.name]


class Materials:
    """gpu_material_manager: the listed materials uploaded to the device."""

    def __init__(self, specs):
        self.prepared = [s if isinstance(s, PreparedMaterial) else prepared(s) for s in specs]
        self.descs = (L.MaterialDesc * len(self.prepared))(*[p.desc for p in self.prepared])
        h = C.c_void_p()
        call("nbl_materials_create", self.descs, len(self.prepared), C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            L.lib.nbl_materials_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def max_energy(self) -> float:
        f = C.c_float()
        call("nbl_materials_get_max_energy", self.h, C.byref(f))
        return f.value

    def table1d(self, material, table, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(len(x), np.float32)
        call("nbl_table1d_get_batch", self.h, material, table, ptr(x), len(x), ptr(out))
        return out

    def table2d(self, material, table, x, y, rounddown=False):
        x = np.ascontiguousarray(x, np.float32); y = np.ascontiguousarray(y, np.float32)
        out = np.zeros(len(x), np.float32)
        call("nbl_table2d_get_batch", self.h, material, table, int(rounddown), ptr(x), ptr(y), len(x), ptr(out))
        return out

    def table3d(self, material, table, x, y, z, rounddown=False):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        out = np.zeros(len(x), np.float32)
        call("nbl_table3d_get_batch", self.h, material, table, int(rounddown), ptr(x), ptr(y), ptr(z), len(x),
             ptr(out))
        return out

    def physics(self, event, particles, material, uniforms, last_tri=None, geometry=None):
        """One event per particle with injected uniforms; returns a dict of outputs."""
        n = len(particles)
        io = np.ascontiguousarray(particles, L.PARTICLE_DTYPE).copy()
        mat = np.ascontiguousarray(material, np.int32)
        u = np.ascontiguousarray(uniforms, np.float32).reshape(n, L.NUM_UNIFORMS)
        lt = None if last_tri is None else np.ascontiguousarray(last_tri, np.uint32)
        om = np.zeros(n, np.int32); st = np.zeros(n, np.uint8)
        sec = np.zeros(n, L.PARTICLE_DTYPE); has = np.zeros(n, np.uint8)
        sc = np.zeros(n, np.float32); ty = np.zeros(n, np.uint8)
        call("nbl_physics_batch", self.h, geometry.h if geometry is not None else None, event, ptr(mat), ptr(lt),
             ptr(u), n, ptr(io), ptr(om), ptr(st), ptr(sec), ptr(has), ptr(sc), ptr(ty))
        return dict(particles=io, material=om, status=st, secondary=sec, has_secondary=has, scalar=sc, type=ty)


def ks_distance(a, b) -> float:
    """Two-sample Kolmogorov-Smirnov distance sup |F_a - F_b|."""
    a = np.sort(np.asarray(a, np.float64)); b = np.sort(np.asarray(b, np.float64))
    grid = np.concatenate([a, b])
    fa = np.searchsorted(a, grid, side="right") / len(a)
    fb = np.searchsorted(b, grid, side="right") / len(b)
    return float(np.max(np.abs(fa - fb)))


class Sampler:
    """Learned ICDF sampler of a material set: BF16 MLPs (elastic cos(theta),
    inelastic log omega') evaluated on the matrix cores, trained at
    construction on samples drawn from the material tables.  kinds selects
    which tables the networks replace ("both", "elastic", "inelastic"); the
    others stay tables in the transport."""

    KINDS = {L.SAMPLER_ELASTIC: L.TAB_ELASTIC_ICDF, L.SAMPLER_INELASTIC: L.TAB_INELASTIC_W0_ICDF}
    KIND_MASKS = {"both": 3, "elastic": 1 << L.SAMPLER_ELASTIC, "inelastic": 1 << L.SAMPLER_INELASTIC}

    def __init__(self, materials: Materials, samples=100_000_000, batch=4096, lr=4e-3, e_min=0.0, e_max=0.0,
                 seed=1, kinds="both"):
        self.materials = materials              # the tables must outlive the sampler
        self.kinds = kinds
        o = L.SamplerOpts(samples=int(samples), batch=int(batch), lr=float(lr), e_min=float(e_min),
                          e_max=float(e_max), seed=int(seed), kinds=self.KIND_MASKS[kinds])
        h = C.c_void_p()
        call("nbl_sampler_create", materials.h, C.byref(o), C.byref(h))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            L.lib.nbl_sampler_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def report(self) -> dict:
        r = L.SamplerReport()
        call("nbl_sampler_get_report", self.h, C.byref(r))
        return {k: getattr(r, k) for k, _ in L.SamplerReport._fields_}

    def params(self, kind):
        """(FP32 parameter block, [feature shift, feature scale, output shift, output scale])."""
        p = np.zeros(L.SAMPLER_PARAMS, np.float32)
        c = np.zeros(4, np.float32)
        call("nbl_sampler_get_params", self.h, kind, ptr(p), L.SAMPLER_PARAMS, ptr(c))
        return p, c

    def eval(self, kind, material, energy, u):
        """The network's value (cos theta or log omega') through the matrix-core path."""
        energy = np.ascontiguousarray(energy, np.float32)
        n = len(energy)
        mat = np.ascontiguousarray(np.broadcast_to(np.asarray(material, np.int32), (n,)))
        u = np.ascontiguousarray(u, np.float32)
        out = np.zeros(n, np.float32)
        call("nbl_sampler_eval_batch", self.h, kind, ptr(mat), ptr(energy), ptr(u), n, ptr(out))
        return out

    def table(self, kind, material, energy, u):
        """The same quantity from the material's table (the reference the sampler is checked against)."""
        return self.materials.table2d(material, self.KINDS[kind], np.log(np.asarray(energy, np.float32)), u)

    def ks_per_bin(self, kind, material, edges, n=20000, seed=0):
        """KS distance between learned and table samples in each energy bin
        [edges[i], edges[i+1]) (energies log-uniform within a bin, the same
        uniforms for both).  KS is invariant under the monotone map between
        log omega' and omega', so either may be compared.  Energies below the
        material's barrier are left out: an electron there has left the
        material or stopped (gpu_driver.inl:420-428), so the transport never
        draws from those cells; a bin entirely below it gives NaN."""
        rng = np.random.default_rng(seed)
        floor = float(self.materials.prepared[material].desc.barrier)
        out = []
        for lo, hi in zip(edges[:-1], edges[1:]):
            lo = max(lo, floor)
            if lo >= hi:
                out.append(float("nan"))
                continue
            e = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.float32)
            u = ((rng.integers(0, 1 << 24, n) + 1) * 2.0 ** -24).astype(np.float32)
            a, b = self.table(kind, material, e, u), self.eval(kind, material, e, u)
            ok = np.isfinite(a)          # table cells below the Fermi level hold no finite value
            out.append(ks_distance(a[ok], b[ok]) if ok.sum() >= 100 else float("nan"))
        return out


def device_info():
    name = C.create_string_buffer(128)
    cu = C.c_int()
    call("nbl_device_info", name, 128, C.byref(cu))
    return name.value.decode(), cu.value
