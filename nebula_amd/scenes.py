"""Synthetic geometries of the benchmark workloads, as nbl_triangle arrays.

Conventions (core/triangle.h:81-83): material_in is on the side the normal
e1 x e2 points away from, material_out on the side it points towards.
Vacuum is -123, a detector -126 (core/cpu_material_manager.h:82-91).  Every
scene has a detector plane above the sample so that escaping electrons are
detected rather than terminated at the domain edge.
"""
from __future__ import annotations

import numpy as np

from ._lib import TRIANGLE_DTYPE, MAT_VACUUM, MAT_DETECTOR


class TriangleSoup:
    def __init__(self):
        self.r0, self.r1, self.r2, self.mi, self.mo = [], [], [], [], []

    def add(self, r0, r1, r2, mat_in, mat_out, outward):
        """Triangles (N,3) x3; flips winding so the normal points along outward (N,3)."""
        r0, r1, r2 = (np.atleast_2d(np.asarray(a, dtype=np.float64)) for a in (r0, r1, r2))
        out = np.broadcast_to(np.asarray(outward, dtype=np.float64), r0.shape)
        n = np.cross(r1 - r0, r2 - r0)
        flip = np.einsum("ij,ij->i", n, out) < 0
        r1f = np.where(flip[:, None], r2, r1)
        r2f = np.where(flip[:, None], r1, r2)
        m = len(r0)
        self.r0.append(r0); self.r1.append(r1f); self.r2.append(r2f)
        self.mi.append(np.broadcast_to(np.asarray(mat_in, np.int32), (m,)))
        self.mo.append(np.broadcast_to(np.asarray(mat_out, np.int32), (m,)))

    def add_quads(self, a, b, c, d, mat_in, mat_out, outward):
        """Quads a-b-c-d (N,3 each) as triangles (a,b,c) and (a,c,d)."""
        self.add(a, b, c, mat_in, mat_out, outward)
        self.add(a, c, d, mat_in, mat_out, outward)

    def add_rect(self, origin, u, v, nu, nv, mat_in, mat_out, outward):
        """Rectangle origin + s*u + t*v, s,t in [0,1], split into nu x nv quads."""
        o, u, v = (np.asarray(x, dtype=np.float64) for x in (origin, u, v))
        s = np.linspace(0.0, 1.0, nu + 1)
        t = np.linspace(0.0, 1.0, nv + 1)
        S, T = np.meshgrid(s[:-1], t[:-1], indexing="ij")
        S1, T1 = np.meshgrid(s[1:], t[1:], indexing="ij")
        pt = lambda a, b: o + a.reshape(-1, 1) * u + b.reshape(-1, 1) * v
        self.add_quads(pt(S, T), pt(S1, T), pt(S1, T1), pt(S, T1), mat_in, mat_out, outward)

    def build(self) -> np.ndarray:
        r0 = np.concatenate(self.r0).astype(np.float32)
        r1 = np.concatenate(self.r1).astype(np.float32)
        r2 = np.concatenate(self.r2).astype(np.float32)
        tris = np.zeros(len(r0), dtype=TRIANGLE_DTYPE)
        tris["material_in"] = np.concatenate(self.mi)
        tris["material_out"] = np.concatenate(self.mo)
        tris["r0"] = r0
        # the reference stores edges r1 - r0, r2 - r0 in float (triangle.inl:1-4)
        tris["e1"] = r1 - r0
        tris["e2"] = r2 - r0
        return tris


def add_box(s: TriangleSoup, lo, hi, mat_in, mat_out, n=(1, 1, 1), skip=()):
    """Axis-aligned box with outward normals; faces named -x,+x,-y,+y,-z,+z."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float); ext = hi - lo
    nx, ny, nz = n
    ex, ey, ez = np.array([ext[0], 0, 0]), np.array([0, ext[1], 0]), np.array([0, 0, ext[2]])
    faces = {
        "-x": (lo, ey, ez, ny, nz, (-1, 0, 0)), "+x": (lo + ex, ey, ez, ny, nz, (1, 0, 0)),
        "-y": (lo, ex, ez, nx, nz, (0, -1, 0)), "+y": (lo + ey, ex, ez, nx, nz, (0, 1, 0)),
        "-z": (lo, ex, ey, nx, ny, (0, 0, -1)), "+z": (lo + ez, ex, ey, nx, ny, (0, 0, 1)),
    }
    for name, (o, u, v, nu, nv, out) in faces.items():
        if name not in skip:
            s.add_rect(o, u, v, nu, nv, mat_in, mat_out, out)


def add_detector(s: TriangleSoup, z, half):
    s.add_rect((-half, -half, z), (2 * half, 0, 0), (0, 2 * half, 0), 1, 1, MAT_DETECTOR, MAT_DETECTOR, (0, 0, 1))


def cube_slab(material=0):
    """Flat slab (material index `material`, Si in the workloads): 1 um cube
    x,y in [-500,500], z in [-1000,0] (12 triangles), plus a detector plane
    at z = +10 nm."""
    s = TriangleSoup()
    add_box(s, (-500, -500, -1000), (500, 500, 0), material, MAT_VACUUM)
    add_detector(s, 10.0, 1500.0)
    return s.build()


LINE_CENTERS = (-128.0, -64.0, 0.0, 64.0, 128.0)
LINE_HALF_WIDTH, LINE_HEIGHT = 16.0, 50.0
SUBSTRATE_HALF, SUBSTRATE_DEPTH = 256.0, 200.0


def line_trench(spacing=14.0, line_materials=(1, 1, 1, 1, 1)):
    """Si substrate (material 0) with 5 lines (32 nm wide, 64 nm pitch, 50 nm
    tall) along y.  Faces are tessellated with about `spacing` nm edges; the
    default gives about 1e4 triangles.  line_materials sets each line's material."""
    s = TriangleSoup()
    H, D = SUBSTRATE_HALF, SUBSTRATE_DEPTH
    seg = lambda length: max(1, int(round(length / spacing)))
    # substrate sides and bottom
    add_box(s, (-H, -H, -D), (H, H, 0), 0, MAT_VACUUM,
            n=(seg(2 * H), seg(2 * H), max(1, seg(D) // 4)), skip=("+z",))
    # substrate top split into exposed strips and PMMA/Si interfaces
    xs = [-H]
    for c in LINE_CENTERS:
        xs += [c - LINE_HALF_WIDTH, c + LINE_HALF_WIDTH]
    xs.append(H)
    for k in range(len(xs) - 1):
        x0, x1 = xs[k], xs[k + 1]
        under = k % 2 == 1
        out = line_materials[(k - 1) // 2] if under else MAT_VACUUM
        s.add_rect((x0, -H, 0), (x1 - x0, 0, 0), (0, 2 * H, 0), seg(x1 - x0), seg(2 * H), 0, out, (0, 0, 1))
    # lines: top, two side walls, two end caps
    for c, m in zip(LINE_CENTERS, line_materials):
        x0, x1 = c - LINE_HALF_WIDTH, c + LINE_HALF_WIDTH
        s.add_rect((x0, -H, LINE_HEIGHT), (x1 - x0, 0, 0), (0, 2 * H, 0), seg(x1 - x0), seg(2 * H), m, MAT_VACUUM, (0, 0, 1))
        s.add_rect((x0, -H, 0), (0, 2 * H, 0), (0, 0, LINE_HEIGHT), seg(2 * H), seg(LINE_HEIGHT), m, MAT_VACUUM, (-1, 0, 0))
        s.add_rect((x1, -H, 0), (0, 2 * H, 0), (0, 0, LINE_HEIGHT), seg(2 * H), seg(LINE_HEIGHT), m, MAT_VACUUM, (1, 0, 0))
        for y, o in ((-H, -1), (H, 1)):
            s.add_rect((x0, y, 0), (x1 - x0, 0, 0), (0, 0, LINE_HEIGHT), seg(x1 - x0), seg(LINE_HEIGHT), m, MAT_VACUUM, (0, o, 0))
    add_detector(s, LINE_HEIGHT + 10.0, H + 100.0)
    return s.build()


def gaussian_height_field(n=512, extent=1000.0, sigma=2.0, corr=10.0, seed=1):
    """Gaussian random height field: white noise filtered to the correlation
    C(r) = sigma^2 exp(-r^2/corr^2) on an n x n grid spanning `extent` nm."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n))
    dx = extent / (n - 1)
    k = np.fft.fftfreq(n, d=dx) * 2 * np.pi
    kx, ky = np.meshgrid(k, k, indexing="ij")
    # spectral density of exp(-r^2/l^2): proportional to exp(-k^2 l^2 / 4)
    filt = np.exp(-(kx ** 2 + ky ** 2) * corr ** 2 / 8.0)
    h = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    h -= h.mean()
    h *= sigma / h.std()
    return h


def height_field_scene(h, extent=1000.0, depth=100.0, column_material=None, detector_gap=15.0):
    """Triangulated height field z = h[i, j] over x,y in [-extent/2, extent/2]
    (2 (n-1)^2 surface triangles), closed by side walls and a bottom at
    z = -depth.  column_material(i) gives the material of grid column i (cells
    i..i+1 along x); vertical interfaces are added where it changes."""
    n = h.shape[0]
    half = extent / 2.0
    g = np.linspace(-half, half, n)
    mat_of = (lambda i: 0) if column_material is None else column_material
    col_mat = np.array([mat_of(i) for i in range(n - 1)], dtype=np.int32)
    s = TriangleSoup()
    X, Y = np.meshgrid(g, g, indexing="ij")
    P = np.stack([X, Y, h], axis=-1)
    a, b, c, d = P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:]
    mats = np.repeat(col_mat[:, None], n - 1, axis=1).reshape(-1)
    up = (0, 0, 1)
    s.add(a.reshape(-1, 3), b.reshape(-1, 3), c.reshape(-1, 3), mats, MAT_VACUUM, up)
    s.add(a.reshape(-1, 3), c.reshape(-1, 3), d.reshape(-1, 3), mats, MAT_VACUUM, up)
    zb = -depth
    # walls at y = -half and y = +half (segments along x)
    for j, o in ((0, -1), (n - 1, 1)):
        top0, top1 = P[:-1, j], P[1:, j]
        bot0, bot1 = top0.copy(), top1.copy(); bot0[:, 2] = zb; bot1[:, 2] = zb
        s.add_quads(bot0, bot1, top1, top0, col_mat, MAT_VACUUM, (0, o, 0))
    # walls at x = -half and x = +half (segments along y)
    for i, o, m in ((0, -1, col_mat[0]), (n - 1, 1, col_mat[-1])):
        top0, top1 = P[i, :-1], P[i, 1:]
        bot0, bot1 = top0.copy(), top1.copy(); bot0[:, 2] = zb; bot1[:, 2] = zb
        s.add_quads(bot0, bot1, top1, top0, m, MAT_VACUUM, (o, 0, 0))
    # interfaces between columns of different material (normal +x: in = left)
    for i in range(1, n - 1):
        if col_mat[i - 1] != col_mat[i]:
            top0, top1 = P[i, :-1], P[i, 1:]
            bot0, bot1 = top0.copy(), top1.copy(); bot0[:, 2] = zb; bot1[:, 2] = zb
            s.add_quads(bot0, bot1, top1, top0, col_mat[i - 1], col_mat[i], (1, 0, 0))
    # bottom, one strip per material run
    cuts = [0] + [i for i in range(1, n - 1) if col_mat[i] != col_mat[i - 1]] + [n - 1]
    for i0, i1 in zip(cuts[:-1], cuts[1:]):
        m = col_mat[i0]
        s.add_rect((g[i0], -half, zb), (g[i1] - g[i0], 0, 0), (0, extent, 0), 1, 1, m, MAT_VACUUM, (0, 0, -1))
    add_detector(s, float(h.max()) + detector_gap, half + 100.0)
    return s.build()


def rough_surface(n=512, seed=1):
    """512 x 512 Si height field over 1 um^2, sigma 2 nm, correlation 10 nm."""
    return height_field_scene(gaussian_height_field(n=n, seed=seed))


def rough_surface_stripes(n=512, seed=1, n_materials=8):
    """The rough surface split into n_materials stripes along x."""
    per = (n - 1) / n_materials
    return height_field_scene(gaussian_height_field(n=n, seed=seed),
                              column_material=lambda i: min(int(i // per), n_materials - 1))
