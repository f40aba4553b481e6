"""Synthetic cross-section tables in the raw form a cross-section tool writes.

There is no network and no cstool output here, so every workload uses
materials generated from closed-form models, fixed once in ``MANIFEST``:

* elastic: screened Rutherford,
  dsigma/dOmega = (1.44 eV nm * Z / (2E))^2 / (1 - cos(theta) + 2 eta)^2,
  eta = 3.4 Z^(2/3) / E[eV]; total sigma = pi (1.44 Z / 2E)^2 / (eta (1 + eta));
  the ICDF has the closed form cos(theta) = 1 - 2 eta u / (1 + eta - u).
* inelastic: single-pole Drude energy-loss function
  ELF(w) = wp^2 g w / ((wp^2 - w^2)^2 + (g w)^2), optical approximation
  dlambda^-1/dw = ELF(w) ln(q+/q-) / (2 pi E a0) with ln(q+/q-) =
  ln((1 + sqrt(1 - w/E))^2 / (w/E)), w in (w_lo, E - E_F).
* binding ICDF: inner shells ionised with probabilities rising as
  f (1 - B/w') above each edge B; the remainder is "no inner shell" (-1).
* full Penn (materials with inelastic="penn"): IMFP and omega ICDF as above
  from the Drude ELF plus a core edge f A (B/w)^2.5 above each inner shell B
  (weight f, A = 0.02); a q ICDF with density 1/q between the kinematic limits
  q+- = sqrt(K) +- sqrt(K - omega) (q in units of hbar/sqrt(2m), eV^1/2,
  stored in nm^-1 as a cross-section tool writes it); a 3D binding ICDF over
  (K, omega/K, P) where every shell with B < omega is open (inner shells with
  weight f, outer shells with weight 0.3, both times (1 - B/omega)).

Grids: 1024 log-spaced kinetic energies from 1 eV to 10 keV and 1024
probabilities from 0 to 1, float32; the full Penn 3D tables are 64^3 over
1 eV..10 keV, omega/K in [0, 1] (q) or log-spaced in [1e-4, 1] (binding),
and P in [0, 1].
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

HARTREE_EV = 27.211386245988
BOHR_NM = 0.0529177210903
AVOGADRO = 6.02214076e23

N_K = 1024
N_P = 1024
K_MIN, K_MAX = 1.0, 1.0e4
P_MIN, P_MAX = 0.0, 1.0
PENN_N = 64                         # points per axis of the full Penn 3D tables
PENN_W_MIN = 1.0e-4                 # lowest omega/K of the binding table
HBAR_SQRT2M = 0.19519               # hbar / sqrt(2 m_e), nm eV^1/2
OUTER_SHELL_WEIGHT = 0.3
PENN_EDGE_AMP = 0.02                # core-edge strength added to the full Penn ELF


@dataclass
class MaterialSpec:
    name: str
    z: float                  # screening charge of the elastic model
    density: float            # atoms per nm^3
    plasmon: float            # Drude pole, eV
    damping: float            # Drude width, eV
    fermi: float              # eV
    work_function: float      # eV; barrier = fermi + work_function
    band_gap: float           # eV; < 0 for a metal
    phonon_loss: float        # eV
    atomic_mass: float        # g/mol of the effective atom
    inner_shells: list = field(default_factory=list)   # [(binding eV, weight)]
    outer_shells: list = field(default_factory=list)   # eV
    inelastic: str = "kieft"  # "kieft" (Kieft-Bosch) or "penn" (full Penn)

    @property
    def barrier(self) -> float:
        return self.fermi + self.work_function


SI = MaterialSpec("Si", z=14, density=49.94, plasmon=16.7, damping=4.0, fermi=7.8,
                  work_function=4.8, band_gap=1.12, phonon_loss=0.063, atomic_mass=28.0855,
                  inner_shells=[(99.2, 0.6), (149.7, 0.2), (1839.0, 0.5)], outer_shells=[7.0])
PMMA = MaterialSpec("PMMA", z=3.6, density=107.4, plasmon=19.0, damping=6.0, fermi=0.0,
                    work_function=4.0, band_gap=5.6, phonon_loss=0.1, atomic_mass=6.67,
                    inner_shells=[(284.2, 0.5), (543.1, 0.3)], outer_shells=[8.0])
# third material of the batched scan: a metal-like layer
W_LIKE = MaterialSpec("W-like", z=74, density=63.2, plasmon=22.0, damping=8.0, fermi=9.0,
                      work_function=4.5, band_gap=-1.0, phonon_loss=0.03, atomic_mass=183.84,
                      inner_shells=[(245.0, 0.3), (1809.0, 0.3)], outer_shells=[3.0, 8.0])


def penn(spec: MaterialSpec) -> MaterialSpec:
    """The same material with the full Penn inelastic model."""
    return replace(spec, name=spec.name + " (Penn)", inelastic="penn")


def stripe_materials(n=8):
    """n materials with Drude pole 10..25 eV, Fermi 0..10 eV, Z 6..79."""
    out = []
    for k in range(n):
        f = k / max(n - 1, 1)
        fermi = 10.0 * f
        out.append(MaterialSpec(
            f"stripe{k}", z=round(6 + 73 * f), density=50.0, plasmon=10.0 + 15.0 * f,
            damping=3.0 + 3.0 * f, fermi=fermi, work_function=4.5,
            band_gap=-1.0 if fermi > 5 else 2.0, phonon_loss=0.05,
            atomic_mass=2.0 * round(6 + 73 * f),
            inner_shells=[(120.0 + 200.0 * f, 0.4)], outer_shells=[6.0]))
    return out


SI_PENN = penn(SI)
PMMA_PENN = penn(PMMA)
W_LIKE_PENN = penn(W_LIKE)

MANIFEST = {
    "grid": {"n_k": N_K, "k_min_eV": K_MIN, "k_max_eV": K_MAX, "n_p": N_P, "p": [P_MIN, P_MAX]},
    "elastic": "screened Rutherford, eta = 3.4 Z^(2/3)/E[eV]",
    "inelastic": "single-pole Drude ELF, optical approximation, 1/(2 pi E a0)",
    "penn_3d": {"n": PENN_N, "q": "density 1/q on [q-, q+]", "binding": "shells with B < omega open"},
    "materials": {m.name: m.__dict__ for m in (SI, PMMA, W_LIKE)},
}


def energy_grid():
    ks = np.log(K_MAX / K_MIN) / (N_K - 1)
    return K_MIN * np.exp(np.arange(N_K) * ks)


def prob_grid():
    return P_MIN + np.arange(N_P) * ((P_MAX - P_MIN) / (N_P - 1))


def screening(spec: MaterialSpec, e):
    return 3.4 * spec.z ** (2.0 / 3.0) / e


def elastic_tables(spec: MaterialSpec):
    k = energy_grid()
    p = prob_grid()
    eta = screening(spec, k)
    sigma = np.pi * (1.44 * spec.z / (2.0 * k)) ** 2 / (eta * (1.0 + eta))   # nm^2
    imfp = (spec.density * sigma).astype(np.float32)                         # nm^-1
    e = eta[:, None]
    u = p[None, :]
    cos_t = 1.0 - 2.0 * e * u / (1.0 + e - u)
    return imfp, cos_t.astype(np.float32)


def drude_elf(spec: MaterialSpec, w):
    wp, g = spec.plasmon, spec.damping
    return wp * wp * g * w / ((wp * wp - w * w) ** 2 + (g * w) ** 2)


def loss_function(spec: MaterialSpec, w):
    """Drude ELF; full Penn materials add an edge f A (B/w)^2.5 above each
    inner shell B (A = PENN_EDGE_AMP), so that core losses occur."""
    elf = drude_elf(spec, w)
    if spec.inelastic == "penn":
        for b, f in spec.inner_shells:
            elf = elf + np.where(w > b, f * PENN_EDGE_AMP * (b / np.maximum(w, b)) ** 2.5, 0.0)
    return elf


def inelastic_tables(spec: MaterialSpec, n_w=2048, w_lo=0.05):
    k = energy_grid()
    p = prob_grid()
    imfp = np.full(N_K, 1e-10, dtype=np.float64)
    icdf = np.empty((N_K, N_P), dtype=np.float64)
    for i, e in enumerate(k):
        w_hi = e - spec.fermi
        if w_hi <= w_lo * 1.01:
            icdf[i, :] = w_lo
            continue
        w = np.geomspace(w_lo, w_hi, n_w)
        x = np.minimum(w / e, 1.0)
        lq = np.log((1.0 + np.sqrt(1.0 - x)) ** 2 / x)
        dens = loss_function(spec, w) * lq / (2.0 * np.pi * (e / HARTREE_EV) * BOHR_NM) / HARTREE_EV  # nm^-1 eV^-1
        seg = 0.5 * (dens[1:] + dens[:-1]) * np.diff(w)
        cdf = np.concatenate([[0.0], np.cumsum(seg)])
        total = cdf[-1]
        if not total > 0:
            icdf[i, :] = w_lo
            continue
        imfp[i] = max(total, 1e-10)
        icdf[i, :] = np.interp(p * total, cdf, w)
    return imfp.astype(np.float32), icdf.astype(np.float32)


def binding_table(spec: MaterialSpec):
    """Raw binding ICDF over (omega', P): -1 = no inner shell."""
    k = energy_grid()
    p = prob_grid()
    out = np.full((N_K, N_P), -1.0, dtype=np.float64)
    shells = sorted(spec.inner_shells)
    for i, w in enumerate(k):
        open_ = [(b, f * (1.0 - b / w)) for b, f in shells if w > b]
        if not open_:
            continue
        q = np.array([x for _, x in open_])
        total = q.sum()
        if total > 1.0:
            q, total = q / total, 1.0
        # P in [0, 1-total) -> no inner shell; then one interval per open shell
        bounds = np.concatenate([[1.0 - total], 1.0 - total + np.cumsum(q)])
        labels = np.array([-1.0] + [b for b, _ in open_])
        idx = np.minimum(np.searchsorted(bounds, p, side="right"), len(open_))
        out[i] = labels[idx]
    return out.astype(np.float32)


def penn_q_table(spec: MaterialSpec, n=PENN_N):
    """Raw q ICDF [K][omega/K][P] in nm^-1: q = q_lo (q+/q_lo)^P, q_lo = max(q-, 1e-3)."""
    k = np.geomspace(K_MIN, K_MAX, n)[:, None, None]
    w = np.linspace(0.0, 1.0, n)[None, :, None]
    p = np.linspace(P_MIN, P_MAX, n)[None, None, :]
    om = w * k
    rest = np.sqrt(np.maximum(k - om, 0.0))
    q_hi = np.sqrt(k) + rest
    q_lo = np.maximum(np.sqrt(k) - rest, 1e-3)
    q = q_lo * (q_hi / q_lo) ** p
    return (q / HBAR_SQRT2M).astype(np.float32)


def penn_binding_table(spec: MaterialSpec, n=PENN_N):
    """Raw binding ICDF [K][omega/K][P] in eV (-1 = no shell)."""
    k = np.geomspace(K_MIN, K_MAX, n)
    w = np.geomspace(PENN_W_MIN, 1.0, n)
    p = np.linspace(P_MIN, P_MAX, n)
    shells = sorted(list(spec.inner_shells) + [(b, OUTER_SHELL_WEIGHT) for b in spec.outer_shells])
    out = np.full((n, n, n), -1.0)
    for i in range(n):
        for j in range(n):
            om = w[j] * k[i]
            open_ = [(b, f * (1.0 - b / om)) for b, f in shells if om > b]
            if not open_:
                continue
            q = np.array([x for _, x in open_])
            total = q.sum()
            if total > 1.0:
                q, total = q / total, 1.0
            bounds = np.concatenate([[1.0 - total], 1.0 - total + np.cumsum(q)])
            labels = np.array([-1.0] + [b for b, _ in open_])
            out[i, j] = labels[np.minimum(np.searchsorted(bounds, p, side="right"), len(open_))]
    return out.astype(np.float32)


class RawMaterial:
    """Raw tables of one material, kept alive for the C struct that points at them."""

    def __init__(self, spec: MaterialSpec):
        self.spec = spec
        self.el_imfp, self.el_cos = elastic_tables(spec)
        self.in_imfp, self.in_w0 = inelastic_tables(spec)
        self.binding = binding_table(spec)
        self.outer = np.ascontiguousarray(spec.outer_shells, dtype=np.float64)
        if spec.inelastic == "penn":
            self.penn_q = penn_q_table(spec)
            self.penn_binding = penn_binding_table(spec)

    def raw_struct(self, raw_cls):
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        s = self.spec
        return raw_cls(
            barrier=s.barrier, fermi=s.fermi, band_gap=s.band_gap, phonon_loss=s.phonon_loss,
            effective_a_grams=s.atomic_mass / AVOGADRO, n_k=N_K, k_min=K_MIN, k_max=K_MAX,
            n_p=N_P, p_min=P_MIN, p_max=P_MAX, elastic_imfp=fp(self.el_imfp),
            elastic_costheta=fp(self.el_cos), inelastic_imfp=fp(self.in_imfp),
            inelastic_w0=fp(self.in_w0), binding_icdf=fp(self.binding),
            n_outer_shells=len(self.outer),
            outer_shells=self.outer.ctypes.data_as(C.POINTER(C.c_double)))


    def penn_struct(self, raw_cls):
        """The full Penn raw tables (omega ICDF and IMFP are the Drude ones)."""
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        s, n = self.spec, PENN_N
        return raw_cls(
            fermi=s.fermi, band_gap=s.band_gap, n_k=N_K, k_min=K_MIN, k_max=K_MAX, n_p=N_P, p_min=P_MIN,
            p_max=P_MAX, imfp=fp(self.in_imfp), omega_icdf=fp(self.in_w0),
            q_nk=n, q_nw=n, q_np=n, q_k_min=K_MIN, q_k_max=K_MAX, q_w_min=0.0, q_w_max=1.0,
            q_p_min=P_MIN, q_p_max=P_MAX, q_icdf=fp(self.penn_q),
            i_nk=n, i_nw=n, i_np=n, i_k_min=K_MIN, i_k_max=K_MAX, i_w_min=PENN_W_MIN, i_w_max=1.0,
            i_p_min=P_MIN, i_p_max=P_MAX, binding_icdf=fp(self.penn_binding))


_CACHE: dict = {}


def raw_material(spec: MaterialSpec) -> RawMaterial:
    key = tuple(sorted((k, str(v)) for k, v in spec.__dict__.items()))
    if key not in _CACHE:
        _CACHE[key] = RawMaterial(spec)
    return _CACHE[key]


def write_material_file(spec: MaterialSpec, path):
    """The material's raw tables as an .nblmat file (read by the nbl_gpu
    command; layout documented in nebula_amd/csrc/nbl_gpu.cpp)."""
    raw = raw_material(spec)
    s = spec
    with open(path, "wb") as f:
        f.write(b"NBLMAT1\0")
        np.array([N_K, N_P, len(raw.outer), 0], "<u4").tofile(f)
        np.array([K_MIN, K_MAX, P_MIN, P_MAX, s.atomic_mass / AVOGADRO], "<f8").tofile(f)
        np.array([s.barrier, s.fermi, s.band_gap, s.phonon_loss], "<f4").tofile(f)
        for a in (raw.el_imfp, raw.el_cos, raw.in_imfp, raw.in_w0, raw.binding):
            np.ascontiguousarray(a, "<f4").tofile(f)
        np.ascontiguousarray(raw.outer, "<f8").tofile(f)
