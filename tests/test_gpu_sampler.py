"""GPU tests of the learned ICDF sampler (BF16 MLPs on the matrix cores):
the matrix-core operand layout (a known-answer 16x16x16 product), the
network path against a NumPy restatement of the same BF16 forward pass,
the sampler's KS distance to its own tables per energy bin, and persistent
transport with the learned sampler against the table path (yields).

The sampler is an approximation of the tables by construction, so its
checks are distances with stated bounds, not equality; the bounds are the
measured accuracy of this network size and training budget with margin."""
import ctypes as C

import numpy as np
import pytest

import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import scenes, synth
from nebula_amd.driver import Driver, raster_scan

pytestmark = pytest.mark.gpu

E_LO, E_HI = 1.0, 600.0


def test_matrix_core_layout_known_answer():
    print("[gpu] sampler selftest", flush=True)
    err = C.c_double(-1)
    L.call("nbl_sampler_selftest", C.byref(err))
    assert err.value == 0.0


@pytest.fixture(scope="module")
def trained():
    print("[gpu] sampler training (8 stripe materials, 1e8 samples per network)", flush=True)
    m = na.Materials(synth.stripe_materials(8))
    s = na.Sampler(m, samples=100_000_000, e_min=E_LO, e_max=E_HI, seed=3)
    r = s.report()
    print(f"  report {r}", flush=True)
    return m, s


def _bf16(x):
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32)


def _numpy_forward(p, c, kind, mat, energy, u):
    """The network of mlp_device.h in NumPy: BF16 weights (the output layer's
    too) and activations, FP32 sums (parameter layout of NBL_SAMPLER_PARAMS in include/nbl_hip.h)."""
    n = len(energy)
    f = np.zeros((n, 16), np.float32)
    f[:, 0] = (np.log(energy).astype(np.float32) - c[0]) * c[1]
    f[:, 1] = 2 * u - 1
    f[:, 2] = np.log(u) * 0.125
    f[:, 3] = np.log(np.maximum(1 - u, np.float32(2.0 ** -24))) * 0.125
    f[:, 4:12] = p[9476 + 8 * mat: 9476 + 8 * mat + 8]
    w1, b1 = p[0:1024].reshape(64, 16), p[1024:1088]
    w2, b2 = p[1088:5184].reshape(64, 64), p[5184:5248]
    w3, b3 = p[5248:9344].reshape(64, 64), p[9344:9408]
    w4, b4 = p[9408:9472], p[9472]
    h = np.maximum(_bf16(f) @ _bf16(w1).T + b1, 0)
    h = np.maximum(_bf16(h) @ _bf16(w2).T + b2, 0)
    h = np.maximum(_bf16(h) @ _bf16(w3).T + b3, 0)
    v = (_bf16(h) @ _bf16(w4) + b4) * c[3] + c[2]
    return 1.000001 - np.exp(v) if kind == L.SAMPLER_ELASTIC else v


@pytest.mark.parametrize("kind", [L.SAMPLER_ELASTIC, L.SAMPLER_INELASTIC])
def test_network_matches_numpy_bf16_forward(trained, kind):
    print(f"[gpu] sampler eval vs numpy kind={kind}", flush=True)
    m, s = trained
    rng = np.random.default_rng(kind)
    n = 5000                                   # not a multiple of 64: a ragged last wavefront
    e = np.exp(rng.uniform(np.log(2.0), np.log(E_HI), n)).astype(np.float32)
    u = ((rng.integers(0, 1 << 24, n) + 1) * 2.0 ** -24).astype(np.float32)
    p, c = s.params(kind)
    for mat in (0, 5):
        got = s.eval(kind, mat, e, u)
        ref = _numpy_forward(p, c, kind, mat, e, u)
        assert np.all(np.isfinite(got))
        scale = float(np.std(ref))
        err = np.abs(got - ref)
        print(f"  mat {mat}: max |gpu - numpy| {err.max():.3g} (spread {scale:.3g})")
        # same BF16 operands; only the FP32 summation order differs
        assert np.quantile(err, 0.999) < 1e-3 * max(scale, 1.0) and err.max() < 2e-2 * max(scale, 1.0)


@pytest.mark.parametrize("kind,bound", [(L.SAMPLER_ELASTIC, 0.03), (L.SAMPLER_INELASTIC, 0.08)])
def test_ks_per_energy_bin(trained, kind, bound):
    """The learned distribution vs the table's, per energy bin (from each
    material's barrier up) and material.  Measured with this network and
    budget (1e8 samples in batches of 4096, three training seeds): <= 0.0056
    elastic, <= 0.016 inelastic."""
    print(f"[gpu] sampler KS kind={kind}", flush=True)
    m, s = trained
    edges = np.geomspace(5.0, 550.0, 9)
    worst = 0.0
    for mat in range(8):
        ks = s.ks_per_bin(kind, mat, edges, n=20000, seed=mat)
        print(f"  mat {mat}: " + " ".join(f"{k:.4f}" for k in ks))
        worst = max(worst, np.nanmax(ks))
    assert worst < bound


def test_persistent_learned_vs_table_yields(trained):
    """Whole cascades with the learned sampler vs the tables (same kernel,
    same geometry): SE and BSE yields within 3 sigma, the variance from the
    spread of 400 pixels' yields."""
    print("[gpu] persistent learned vs table", flush=True)
    m, s = trained
    g = na.Geometry(scenes.cube_slab(material=3))
    beam = raster_scan(np.zeros(20), np.zeros(20), 5.0, 500.0, 500, 0.0, seed=4)
    out = {}
    for name, smp in (("table", None), ("learned", s)):
        with Driver(g, m, capacity=1 << 16, seed=21) as drv:
            drv.set_sampler(smp)
            out[name] = drv.run_persistent(beam, per_pixel=True)
        assert out[name]["primaries"] == 200_000 and out[name]["stopped"] == 0
    for key in ("pixel_se", "pixel_bse"):
        a, b = out["table"][key] / 500.0, out["learned"][key] / 500.0
        se = np.hypot(a.std(ddof=1), b.std(ddof=1)) / np.sqrt(len(a))
        z = (b.mean() - a.mean()) / se
        print(f"  {key}: table {a.mean():.4f} learned {b.mean():.4f} z {z:+.2f}")
        assert abs(z) < 3, (key, z)


def test_elastic_only_sampler(trained):
    """kinds="elastic": only the elastic network is trained and used; the
    inelastic draws stay on the tables (the untrained network is refused),
    and the transport's yields match the table path's within 3 sigma."""
    print("[gpu] elastic-only sampler", flush=True)
    m, _ = trained
    s = na.Sampler(m, samples=30_000_000, e_min=E_LO, e_max=E_HI, seed=5, kinds="elastic")
    r = s.report()
    assert r["loss_inelastic"] == 0 and r["loss_elastic"] > 0
    with pytest.raises(na.NblError):
        s.eval(L.SAMPLER_INELASTIC, 0, np.full(4, 100.0, np.float32), np.full(4, 0.5, np.float32))
    assert np.all(np.isfinite(s.eval(L.SAMPLER_ELASTIC, 0, np.full(4, 100.0, np.float32),
                                     np.full(4, 0.5, np.float32))))
    g = na.Geometry(scenes.cube_slab(material=3))
    beam = raster_scan(np.zeros(20), np.zeros(20), 5.0, 500.0, 1000, 0.0, seed=6)
    out = {}
    for name, smp in (("table", None), ("learned", s)):
        with Driver(g, m, capacity=1 << 16, seed=23) as drv:
            drv.set_sampler(smp)
            out[name] = drv.run_persistent(beam, per_pixel=True)
        assert out[name]["primaries"] == 400_000 and out[name]["stopped"] == 0
    for key in ("pixel_se", "pixel_bse"):
        a, b = out["table"][key] / 1000.0, out["learned"][key] / 1000.0
        z = (b.mean() - a.mean()) / (np.hypot(a.std(ddof=1), b.std(ddof=1)) / np.sqrt(len(a)))
        print(f"  {key}: table {a.mean():.4f} learned (elastic) {b.mean():.4f} z {z:+.2f}")
        assert abs(z) < 3, (key, z)


def test_set_sampler_rejects_other_material_set(trained):
    m, s = trained
    g = na.Geometry(scenes.cube_slab())
    m2 = na.Materials([synth.SI])
    with Driver(g, m2, capacity=1024, seed=1) as drv:
        with pytest.raises(na.NblError):
            drv.set_sampler(s)
