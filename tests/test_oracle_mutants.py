"""The oracle's pins catch planted mistakes (no GPU).

Each mutant is a copy of oracle/fe_oracle.cpp with one plausible restatement mistake (a
flipped sign, the wrong density index, a buffer read from the other ping-pong half),
compiled with the oracle's own flags into a temporary directory. The pins are then run
against it in a child process (FE_ORACLE_LIB points the oracle binding at the mutant), and
the named pin must fail. The mistakes are planted only in these CPU copies.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "fe_oracle.cpp")
PINS = "tests/test_oracle_pins.py"
SEQ = "tests/test_oracle_sequence.py::test_convergence_band_against_reference"

# name -> ([(text, replacement), ...], pytest target, the test that must fail)
MUTANTS = {
    # viscosity.comp:54: v_j - v_i
    "viscosity_sign": ([("sub(s.v3(s.vel_b, j), s.v3(s.vel_b, i))", "sub(s.v3(s.vel_b, i), s.v3(s.vel_b, j))")],
                       PINS, "test_viscosity_quadratic_field_closed_form"),
    # viscosity.comp:55: m / rho_j
    "viscosity_own_density": ([("mass / s.dens[j]", "mass / s.dens[i]")],
                              PINS, "test_viscosity_uses_neighbour_density"),
    # density_source_term.comp:57: v_i - v_j
    "density_source_sign": ([("V3<R> vd = sub(vi, s.v3(s.vel_a, j));", "V3<R> vd = sub(s.v3(s.vel_a, j), vi);")],
                            PINS, "test_source_terms_linear_field_closed_form"),
    # divergence_source_term.comp:53: v_i - v_j
    "divergence_source_sign": ([("V3<R> vd = sub(vi, s.v3(vel, j));", "V3<R> vd = sub(s.v3(vel, j), vi);")],
                               PINS, "test_source_terms_linear_field_closed_form"),
    # pressure_integration.comp:29: v + a dt
    "integration_accel_sign": ([("nv[k] = vel[k] + ap[k] * dt", "nv[k] = vel[k] - ap[k] * dt")],
                               PINS, "test_pressure_integration_nonzero_acceleration_closed_form"),
    # pressure_integration.comp:30: x_b + v' dt
    "integration_old_velocity": ([("np[k] = pos[k] + nv[k] * dt", "np[k] = pos[k] + vel[k] * dt")],
                                 PINS, "test_pressure_integration_nonzero_acceleration_closed_form"),
    # divergence_integration.comp: v + a dt
    "divergence_integration_sign": ([("V3<R> nv = add(v, scale_r(a, dt));", "V3<R> nv = sub(v, scale_r(a, dt));")],
                                    PINS, "test_divergence_integration_nonzero_acceleration_closed_form"),
    # viscosity.rs:49 binds velocity_b (Q2: the divergence correction is not fed back)
    "viscosity_reads_velocity_a": ([("sub(s.v3(s.vel_b, j), s.v3(s.vel_b, i))", "sub(s.v3(s.vel_a, j), s.v3(s.vel_a, i))"),
                                    ("V3<R> v = s.v3(s.vel_b, i);", "V3<R> v = s.v3(s.vel_a, i);")],
                                   SEQ + "[k2]", "test_convergence_band_against_reference[k2]"),
    # divergence_source_term.rs:46-47 binds velocity_a in the previous order (Q1)
    "divergence_source_reads_velocity_b": (
        [("const std::vector<R>& vel = s.o.divergence_order == FE_DIVERGENCE_CONSISTENT ? s.vel_b : s.vel_a;",
          "const std::vector<R>& vel = s.vel_b;")],
        SEQ + "[k1]", "test_convergence_band_against_reference[k1]"),
    # density_source_term.comp:73: the density solve starts from p = 0 (Q3)
    "density_solve_keeps_pressure": ([("s.pres[i] = R(0);", "(void)0;")],
                                     SEQ + "[k1]", "test_convergence_band_against_reference[k1]"),
}


def _compile(name, edits, out_dir):
    text = open(SRC).read()
    for old, new in edits:
        assert text.count(old) == 1, (name, old)
        text = text.replace(old, new)
    src = os.path.join(out_dir, f"{name}.cpp")
    with open(src, "w") as f:
        f.write(text)
    lib = os.path.join(out_dir, f"{name}.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", lib, src], check=True)
    return lib


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("oracle_mutants"))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = {name: ex.submit(_compile, name, spec[0], out) for name, spec in MUTANTS.items()}
        return {name: f.result() for name, f in futs.items()}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_mutant(mutant_libs, name):
    _, target, must_fail = MUTANTS[name]
    env = dict(os.environ)
    env["FE_ORACLE_LIB"] = mutant_libs[name]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu", target],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    failed = [ln for ln in r.stdout.splitlines() if ln.startswith("FAILED")]
    assert r.returncode == 1, (name, r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert any(must_fail in ln for ln in failed), (name, failed)
