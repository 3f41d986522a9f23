"""The oracle's pins catch plausible slips (CPU only).

Each mutation below plants one plausible mistake in a scratch copy of
oracle/oracle.cpp, builds it with g++, and runs the oracle's independent
pins (closed forms, direct restatements, library routines; never the golden
vectors the oracle wrote itself) against the broken build. Every mutant must
fail at least one pin; the unmutated oracle passes them all in the normal
suite. The mutants are test artefacts in a temporary directory and are never
loaded by anything but these subprocesses.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SRC = ROOT / "oracle" / "oracle.cpp"
PINS = ["tests/test_oracle_closed_forms.py", "tests/test_oracle_pins.py", "tests/test_oracle.py",
        "tests/test_oracle_surface.py"]

# name -> (exact text in oracle.cpp, replacement, occurrence to replace (0-based))
MUTATIONS = {
    "coriolis_sign_u": ("+ fc * (vavg - vref) + relax", "- fc * (vavg - vref) + relax", 0),
    "coriolis_sign_v": ("- fc * (uavg - uref) + relax", "+ fc * (uavg - uref) + relax", 0),
    "rk3_stage2_fraction": ("{dt / 3, dt / 2, dt}", "{dt / 3, dt / 2.2, dt}", 0),
    "rk3_stage3_substeps": ("std::max(1, ns / 2), ns}", "std::max(1, ns / 2), ns - 1}", 0),
    "pgf_map_factor_inverted": ("-kCp * thbu * mu * grad", "-kCp * thbu / mu * grad", 0),
    "dh_map_factor_dropped": ("const double Dh = mc * (", "const double Dh = 1.0 * (", 0),
    "divergence_damping_sign": ("s.pip[c] + p.acoustic_smdiv * (s.pip[c] - piprev[c])",
                                "s.pip[c] - p.acoustic_smdiv * (s.pip[c] - piprev[c])", 0),
    "offcentring_swapped": ("c1 = (1.0 + p.acoustic_beta) * 0.5", "c1 = (1.0 - p.acoustic_beta) * 0.5",
                            0),
    "evaporation_constant": ("+ 5.4e5)", "+ 5.4e4)", 0),
    "fall_speed_density_ratio": ("std::sqrt(rho_sfc / rhol[k]) *\n                                    fallfac(c)",
                                 "std::sqrt(rhol[k] / rho_sfc) *\n                                    fallfac(c)", 0),
    "ice_fall_factor": ("T < kTfrz ? 0.35 : 1.0", "T < kTfrz ? 0.5 : 1.0", 0),
    "psi_m_unstable_power": ("std::pow(1 - 16 * zeta, 0.25);", "std::pow(1 - 16 * zeta, 0.5);", 0),
    "psi_h_stable_slope": ("if (zeta >= 0) return -5 * std::fmin(zeta, 2.0);\n    const double x = "
                           "std::pow(1 - 16 * zeta, 0.25);\n    return 2 * std::log((1 + x * x)",
                           "if (zeta >= 0) return -4 * std::fmin(zeta, 2.0);\n    const double x = "
                           "std::pow(1 - 16 * zeta, 0.25);\n    return 2 * std::log((1 + x * x)", 0),
    "diffusion_upper_conductance": ("      B[k] = m + a + c;", "      B[k] = m + a + 0.5 * c;", 0),
    "declination_day_offset": ("(284 + doy) / 365.0", "(280 + doy) / 365.0", 0),
    "hour_angle_rate": ("0.2618 * (hsol - 12)", "0.25 * (hsol - 12)", 0),
    "solar_time_longitude": ("utc + lon / 15", "utc + lon / 12", 0),
    "lw_surface_emissivity": ("Fup[0] = 0.97 * kSigma", "Fup[0] = 0.95 * kSigma", 0),
    "sw_beam_path": ("std::exp(-(0.008 * du[k] + 0.15 * dl[k]) / mu)",
                     "std::exp(-(0.008 * du[k] + 0.15 * dl[k]) / std::sqrt(mu))", 0),
    "log_law_10m": ("std::log(10 / z0)", "std::log(9 / z0)", 0),
    "heat_roughness": ("const double z0h = z0 * 0.1;", "const double z0h = z0 * 0.2;", 0),
    "soil_layer_depth": ("{0.1, 0.3, 0.6, 1.0}", "{0.1, 0.3, 0.5, 1.0}", 0),
    "louis_neutral_kappa": ("double a2 = kKappa / std::log(z1 / z0);",
                            "double a2 = 0.41 / std::log(z1 / z0);", 0),
}


def _mutate(text: str, old: str, new: str, occ: int) -> str:
    pos = -1
    for _ in range(occ + 1):
        pos = text.find(old, pos + 1)
        assert pos >= 0, f"mutation anchor not found: {old!r}"
    return text[:pos] + new + text[pos + len(old):]


def _build_and_pin(name: str, workdir: Path) -> tuple[str, int, str]:
    old, new, occ = MUTATIONS[name]
    src = workdir / f"{name}.cpp"
    src.write_text(_mutate(ORACLE_SRC.read_text(), old, new, occ))
    so = workdir / f"lib{name}.so"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-shared",
                        f"-I{ROOT / 'oracle'}", str(src), str(ROOT / "oracle" / "case_builder.cpp"),
                        "-o", str(so)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return name, -100, r.stderr[-2000:]
    env = dict(os.environ, WFE_ORACLE_LIB=str(so), OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        *PINS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return name, r.returncode, (r.stdout + r.stderr)[-3000:]


@pytest.fixture(scope="module")
def mutant_results(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        res = list(ex.map(lambda n: _build_and_pin(n, d), MUTATIONS))
    return {n: (rc, out) for n, rc, out in res}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_pins_catch_mutation(name, mutant_results):
    rc, out = mutant_results[name]
    assert rc != -100, f"mutant did not build:\n{out}"
    # pytest exit code 1: tests ran and at least one failed
    assert rc == 1, f"mutation {name} passed every pin (rc={rc}):\n{out}"
