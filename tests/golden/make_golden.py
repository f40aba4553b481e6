"""Writes tests/golden/*.npz: input/output vectors of the reference's OWN
geometry and table code (oracle/_ref/libnbl_ref.so, built by
`make -C oracle` from the sources under /root/reference).

Run in the container that has /root/reference:
    make -C oracle && python tests/golden/make_golden.py

Each fixture holds triangles, the reference's linearized octree
(octree_data, triangle order), rays with the reference's propagate results,
in_domain results, and table queries with the reference's table values.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as ol  # noqa: E402
from nebula_amd import scenes  # noqa: E402
from nebula_amd._lib import TRIANGLE_DTYPE, NO_TRIANGLE  # noqa: E402


def random_soup(n, seed):
    """Random triangles in a 100 nm box, with exact and permuted duplicates,
    two materials and the special detector/mirror materials."""
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(-50, 50, (n, 3))
    r1 = r0 + rng.normal(0, 12, (n, 3))
    r2 = r0 + rng.normal(0, 12, (n, 3))
    t = np.zeros(n, TRIANGLE_DTYPE)
    t["material_in"] = rng.choice([0, 1, -123], n)
    t["material_out"] = rng.choice([0, 1, -123, -126, -122], n)
    t["r0"] = r0.astype(np.float32)
    t["e1"] = (r1.astype(np.float32) - r0.astype(np.float32))
    t["e2"] = (r2.astype(np.float32) - r0.astype(np.float32))
    t[n // 2] = t[3]                        # exact duplicate
    a = t[5].copy()                         # same triangle, vertices rotated
    v0 = a["r0"]; v1 = v0 + a["e1"]; v2 = v0 + a["e2"]
    b = a.copy(); b["r0"] = v1; b["e1"] = v2 - v1; b["e2"] = v0 - v1
    t[n // 2 + 1] = b
    return t


def rays(tris_box_lo, tris_box_hi, n, seed, n_tri, mats):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(tris_box_lo), np.asarray(tris_box_hi)
    pos = rng.uniform(lo - 5, hi + 5, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d[: n // 10] = np.round(d[: n // 10])   # axis-aligned and diagonal directions, zeros included
    d[np.linalg.norm(d, axis=1) == 0] = (0, 0, -1)
    d = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    dist = rng.exponential(0.3 * np.linalg.norm(hi - lo), n).astype(np.float32)
    it = np.where(rng.random(n) < 0.3, rng.integers(0, n_tri, n), NO_TRIANGLE).astype(np.uint32)
    im = rng.choice(mats, n).astype(np.int32)
    return pos, d, dist, it, im


def geometry_fixture(name, tris, n_rays, seed, mats):
    g = ol.RefGeometry(tris)
    od, order = g.export()
    info = g.info()
    lo = info["center"] - info["halfsize"]; hi = info["center"] + info["halfsize"]
    pos, d, dist, it, im = rays(lo, hi, n_rays, seed, len(tris), mats)
    # rays that start on triangles and leave along their normals (edge cases)
    k = min(len(tris), n_rays // 10)
    sel = np.arange(k)
    c = tris["r0"][sel] + (tris["e1"][sel] + tris["e2"][sel]) / np.float32(3)
    nrm = np.cross(tris["e1"][sel], tris["e2"][sel]).astype(np.float64)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1), 1e-30)[:, None]
    pos[:k] = c.astype(np.float32)
    d[:k] = np.where(np.arange(k)[:, None] % 2 == 0, nrm, -nrm).astype(np.float32)
    it[:k] = sel.astype(np.uint32)
    out_d, out_t = g.propagate(pos, d, dist, it, im)
    ind = g.in_domain(pos)
    np.savez_compressed(os.path.join(HERE, f"geometry_{name}.npz"), triangles=tris, octree_data=od,
                        triangle_order=order, center=info["center"], halfsize=info["halfsize"],
                        max_extent=np.float32(info["max_extent"]), pos=pos, dir=d, dist=dist,
                        ignore_tri=it, ignore_mat=im, out_dist=out_d, out_tri=out_t, in_domain=ind)
    print(name, len(tris), "tris,", len(od), "octree ints,", (out_t != NO_TRIANGLE).mean(), "hit fraction")


def table_fixture():
    rng = np.random.default_rng(11)
    n1, w, h = 37, 29, 41
    d1 = rng.normal(0, 1, n1).astype(np.float32)
    d2 = rng.normal(0, 1, (w, h)).astype(np.float32)
    x1 = np.concatenate([rng.uniform(-3, 5, 4000), np.linspace(0.1, 2.3, n1)]).astype(np.float32)
    x2 = rng.uniform(-1, 4, 4000).astype(np.float32)
    y2 = rng.uniform(-0.2, 1.2, 4000).astype(np.float32)
    o1 = np.zeros(len(x1), np.float32); o2 = np.zeros(len(x2), np.float32); o3 = np.zeros(len(x2), np.float32)
    ol.ref.ref_table1d_get_batch(n1, np.float32(0.1), np.float32(2.3), ol.p(d1), ol.p(x1), len(x1), ol.p(o1))
    ol.ref.ref_table2d_get_batch(w, h, np.float32(0.0), np.float32(3.0), np.float32(0.0), np.float32(1.0),
                                 ol.p(np.ascontiguousarray(d2)), 0, ol.p(x2), ol.p(y2), len(x2), ol.p(o2))
    ol.ref.ref_table2d_get_batch(w, h, np.float32(0.0), np.float32(3.0), np.float32(0.0), np.float32(1.0),
                                 ol.p(np.ascontiguousarray(d2)), 1, ol.p(x2), ol.p(y2), len(x2), ol.p(o3))
    # triangle intersections: random and edge-grazing rays
    tris = random_soup(3000, 12)
    pos = rng.uniform(-60, 60, (3000, 3)).astype(np.float32)
    aim = tris["r0"] + rng.uniform(0, 1, (3000, 1)).astype(np.float32) * tris["e1"] \
        + rng.uniform(-0.05, 0.05, (3000, 1)).astype(np.float32) * tris["e2"]
    dv = (aim - pos); dv = (dv / np.linalg.norm(dv, axis=1)[:, None]).astype(np.float32)
    ot = np.zeros(3000, np.float32)
    ol.ref.ref_intersect_ray_batch(ol.p(tris), ol.p(pos), ol.p(dv), 3000, ol.p(ot))
    np.savez_compressed(os.path.join(HERE, "tables.npz"), d1=d1, d2=d2, x1=x1, x2=x2, y2=y2, o1=o1, o2=o2, o3=o3,
                        t1=np.array([0.1, 2.3], np.float32), t2=np.array([0.0, 3.0, 0.0, 1.0], np.float32),
                        tri=tris, tri_pos=pos, tri_dir=dv, tri_t=ot)
    print("tables + intersect_ray fixture written")


T3_SCALE = (-1.0, 4.0, 0.0, 1.0, -0.5, 0.5)      # x_min, x_max, y_min, y_max, z_min, z_max


def table3d_fixture():
    """table_3D get / get_rounddown on a random 13 x 17 x 11 table, queries
    inside, on and outside every face (clamped extrapolation)."""
    rng = np.random.default_rng(21)
    w, h, dp = 13, 17, 11
    d3 = rng.normal(0, 1, (w, h, dp)).astype(np.float32)
    n = 6000
    x = rng.uniform(-1.5, 4.5, n).astype(np.float32)
    y = rng.uniform(-0.2, 1.2, n).astype(np.float32)
    z = rng.uniform(-0.7, 0.7, n).astype(np.float32)
    x[:200] = np.linspace(-1.0, 4.0, 200, dtype=np.float32)      # on grid lines and the ends
    y[200:400] = np.float32(1.0); z[400:600] = np.float32(-0.5)
    sc = [np.float32(v) for v in T3_SCALE]
    out = {}
    for rd, key in ((0, "o_get"), (1, "o_rounddown")):
        o = np.zeros(n, np.float32)
        ol.ref.ref_table3d_get_batch(w, h, dp, *sc, ol.p(d3), rd, ol.p(x), ol.p(y), ol.p(z), n, ol.p(o))
        out[key] = o
    np.savez_compressed(os.path.join(HERE, "tables3d.npz"), d3=d3, x=x, y=y, z=z,
                        scale=np.array(T3_SCALE, np.float32), **out)
    print("table_3D fixture written")


def trilist_fixture():
    """trilist::propagate (geometry/trilist.inl:52-85), the brute-force
    nearest hit, on a random soup with duplicates: random rays, rays with an
    ignored triangle or material, and rays started on a triangle along its
    normal (the ignore rule decides them)."""
    tris = random_soup(400, 31)
    lo = np.min(np.concatenate([tris["r0"], tris["r0"] + tris["e1"], tris["r0"] + tris["e2"]]), axis=0)
    hi = np.max(np.concatenate([tris["r0"], tris["r0"] + tris["e1"], tris["r0"] + tris["e2"]]), axis=0)
    pos, d, dist, it, im = rays(lo, hi, 8000, 32, len(tris), [0, 1, -123])
    k = len(tris)
    c = tris["r0"][:k] + (tris["e1"][:k] + tris["e2"][:k]) / np.float32(3)
    nrm = np.cross(tris["e1"][:k], tris["e2"][:k]).astype(np.float64)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1), 1e-30)[:, None]
    pos[:k] = c.astype(np.float32)
    d[:k] = np.where(np.arange(k)[:, None] % 2 == 0, nrm, -nrm).astype(np.float32)
    it[:k] = np.arange(k, dtype=np.uint32)
    n = len(pos)
    out_d = np.zeros(n, np.float32); out_t = np.zeros(n, np.uint32)
    ol.ref.ref_trilist_propagate_batch(ol.p(tris), len(tris), ol.p(pos), ol.p(d), ol.p(dist), ol.p(it), ol.p(im), n,
                                       ol.p(out_d), ol.p(out_t))
    np.savez_compressed(os.path.join(HERE, "trilist.npz"), triangles=tris, pos=pos, dir=d, dist=dist, ignore_tri=it,
                        ignore_mat=im, out_dist=out_d, out_tri=out_t)
    print("trilist fixture written:", (out_t != NO_TRIANGLE).mean(), "hit fraction")


FIXTURES = {
    "geometry": lambda: (geometry_fixture("cube", scenes.cube_slab(), 4000, 1, [0, -123]),
                         geometry_fixture("soup", random_soup(600, 2), 6000, 2, [0, 1, -123]),
                         geometry_fixture("lines", scenes.line_trench(spacing=40.0), 6000, 3, [0, 1, -123])),
    "tables": table_fixture,
    "tables3d": table3d_fixture,
    "trilist": trilist_fixture,
}


def main():
    """python tests/golden/make_golden.py [fixture ...] (default: all)."""
    if ol.ref is None:
        sys.exit("oracle/_ref/libnbl_ref.so is not built (needs /root/reference): make -C oracle")
    for name in sys.argv[1:] or list(FIXTURES):
        FIXTURES[name]()


if __name__ == "__main__":
    main()
