"""Closed-form pins of the oracle's transport physics (CPU only).

The oracle's boundary crossing, Kieft elastic and inelastic events and free
path sampling (oracle/nbl_oracle.cpp, restating physics/boundary_intersect.h,
physics/kieft/elastic.h, physics/kieft/inelastic.h, core/scatter_list.h)
must pass every check of tests/physics_pins.py, which fixes their outputs by
the physics rather than by the code's formulas.  To show the checks have
teeth, each of a list of plausible mistakes (a dropped factor, a flipped
sign or comparison, swapped draws or branches) is planted into a copy of the
oracle's source, compiled on the CPU, and must fail at least one check.
"""
import concurrent.futures as cf
import os
import subprocess

import pytest

import oracle_lib as ol
import physics_pins as pp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "nbl_oracle.cpp")

# (name, text in nbl_oracle.cpp, replacement, occurrence, checks that must catch it)
MUTANTS = [
    ("T_factor", "const float T = 4 * s / ((1 + s) * (1 + s));", "const float T = 2 * s / ((1 + s) * (1 + s));", 0,
     ["boundary"]),
    ("refraction_sign", "p.dir = (normalised_dir - n * cos_theta) + s * n * cos_theta;",
     "p.dir = (normalised_dir - n * cos_theta) - s * n * cos_theta;", 0, ["boundary"]),
    ("refraction_no_s", "p.dir = (normalised_dir - n * cos_theta) + s * n * cos_theta;",
     "p.dir = normalised_dir;", 0, ["boundary"]),
    ("barrier_sign", "if (mat_in >= 0) dU -= mats[mat_in].barrier;", "if (mat_in >= 0) dU += mats[mat_in].barrier;",
     0, ["boundary"]),
    ("energy_not_shifted", "p.kin_energy += dU;\n            mat = mat_out;", "mat = mat_out;", 0, ["boundary"]),
    ("mirror_half", "case NBL_MAT_MIRROR: p.dir = normalised_dir - 2 * n * cos_theta;",
     "case NBL_MAT_MIRROR: p.dir = normalised_dir - n * cos_theta;", 0, ["boundary"]),
    ("side_swapped", "if (cos_theta > 0) { mat_in = tri.material_in;", "if (cos_theta < 0) { mat_in = tri.material_in;",
     0, ["boundary"]),
    ("lt50_inclusive", "p.kin_energy < 50 ? NBL_OUT_DETECTED", "p.kin_energy <= 50 ? NBL_OUT_DETECTED", 0,
     ["boundary"]),
    ("nop_as_detector", "case NBL_MAT_NOP: return NBL_OUT_ALIVE;", "case NBL_MAT_NOP: return NBL_OUT_DETECTED;", 0,
     ["boundary"]),
    ("phonon_inclusive", "if (p.kin_energy < 200)", "if (p.kin_energy <= 200)", 0, ["elastic"]),
    ("phonon_max", "p.kin_energy -= minr(50e-3f, m.phonon_loss);", "p.kin_energy -= maxr(50e-3f, m.phonon_loss);", 0,
     ["elastic"]),
    ("recoil_sign", "p.kin_energy -= m.recoil_const * (1 - cos_theta) * p.kin_energy;",
     "p.kin_energy -= m.recoil_const * (1 + cos_theta) * p.kin_energy;", 0, ["elastic"]),
    ("no_clamp", "cos_theta = clampr(m.el_icdf.get(x, y), -1, 1);", "cos_theta = m.el_icdf.get(x, y);", 0,
     ["elastic"]),
    ("azimuth_swapped", "return unit_u * cos_phi + unit_v * sin_phi;", "return unit_u * sin_phi + unit_v * cos_phi;", 0,
     ["elastic", "inelastic_metal"]),
    ("tie_to_inelastic", "if (path < dist) { type = 1; dist = path; }", "if (path <= dist) { type = 1; dist = path; }",
     0, ["sample_path"]),
    ("path_not_inverted", "const float path = rng.exponential(1 / imfp);", "const float path = rng.exponential(imfp);",
     0, ["sample_path"]),
    ("ashley_u_flipped", "omega = -w0 / std::expm1(w0 * (f_min * (1 - U) + f_max * U));",
     "omega = -w0 / std::expm1(w0 * (f_min * U + f_max * (1 - U)));", 0, ["inelastic_metal"]),
    ("fallback_u_flipped", "omega = omega0 * omega_max / (omega0 * (1 - U) + omega_max * U);",
     "omega = omega0 * omega_max / (omega0 * U + omega_max * (1 - U));", 0, ["inelastic_metal"]),
    ("omega_min_root", "omega_min = 0.5f * (p.kin_energy + omega0 - std::sqrt(",
     "omega_min = 0.5f * (p.kin_energy - omega0 - std::sqrt(", 0, ["inelastic_metal"]),
    ("gap_binding_zero", "binding = m.band_gap;", "binding = 0;", 0, ["inelastic_band_gap"]),
    ("phonon_branch_keeps_energy", "p.kin_energy -= omega0;   /* optical phonon loss", "p.kin_energy -= 0;   /*", 0,
     ["inelastic_band_gap"]),
    ("se_energy_no_binding", "sec.kin_energy = m.fermi + omega - binding;", "sec.kin_energy = m.fermi + omega;", 0,
     ["inelastic_shell"]),
    ("cone_single_binding", "const float _K = p.kin_energy - m.fermi + 2 * binding;",
     "const float _K = p.kin_energy - m.fermi + binding;", 0, ["inelastic_shell"]),
    ("momentum_term_over_K", "secondary_dir += std::sqrt(binding / dK) * rng.uniform_vector();",
     "secondary_dir += std::sqrt(binding / _K) * rng.uniform_vector();", 0, ["inelastic_shell"]),
    ("primary_not_recoiled", "p.dir -= cos_theta * secondary_dir;", "p.dir -= secondary_dir;", 0,
     ["inelastic_metal", "inelastic_shell"]),
    ("uniform_vector_order", "const float c = 2 * this->unit() - 1; const float p = phi();",
     "const float p = phi(); const float c = 2 * this->unit() - 1;", 0, ["inelastic_shell"]),
    ("binary_cone_sqrt_dropped", "const float cos_theta = std::sqrt(dK / _K);", "const float cos_theta = dK / _K;", 0,
     ["inelastic_metal"]),
]


def _mutate(text, old, new, occurrence):
    parts = text.split(old)
    assert len(parts) > occurrence + 1, f"mutation site not found: {old!r}"
    return old.join(parts[:occurrence + 1]) + new + old.join(parts[occurrence + 1:])


@pytest.fixture(scope="module")
def pins():
    pm = pp.PinMaterials()
    tris = pp.pin_triangles()
    return pm, tris


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    text = open(SRC).read()

    def build(m):
        name, old, new, occ, _ = m
        src = d / f"{name}.cpp"
        src.write_text(_mutate(text, old, new, occ))
        so = d / f"{name}.so"
        r = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                            "-I", os.path.join(ROOT, "oracle"), "-o", str(so), str(src), "-lpthread"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        return name, str(so)

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.parametrize("check", list(pp.CHECKS))
def test_oracle_physics_meets_closed_forms(pins, check):
    pm, tris = pins
    run = pp.OracleRunner(ol.orc, pm, tris)
    pp.CHECKS[check](run, pm, tris)


@pytest.mark.parametrize("mutant", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_physics_mistakes_are_caught(pins, mutant_libs, mutant):
    pm, tris = pins
    name, _, _, _, checks = mutant
    run = pp.OracleRunner(ol.load_oracle(mutant_libs[name]), pm, tris)
    caught = []
    for c in checks:
        try:
            pp.CHECKS[c](run, pm, tris)
        except AssertionError as e:
            caught.append(f"{c}: {str(e).splitlines()[0]}")
    assert caught, f"planted mistake {name} passed the checks {checks}"
