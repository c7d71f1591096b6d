"""The oracle's pins catch plausible mistakes: each mutant below is a copy of the
oracle with one such mistake planted (on the CPU, in a temporary build), and at
least one pin of tests/test_oracle_pins.py must fail on it."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

import test_oracle_pins as pins
from oracle import Oracle

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle"

# (name, original text, mutated text) in euler_oracle.hpp
MUTANTS = [
    ("ke_prim_no_half", "(q.ux2 * q.ux2)) / 2;", "(q.ux2 * q.ux2));"),
    ("velocity_divides", "u.mx0 * vol_spe", "u.mx0 / u.d"),
    ("davis_speed_plus", "kmin(unL, unR) - kmax(cL, cR)", "kmin(unL, unR) + kmax(cL, cR)"),
    ("ustar_sign", "(qR.p - qL.p + rcL * unL - rcR * unR)", "(qL.p - qR.p + rcL * unL - rcR * unR)"),
    ("pstar_drop_term", "+ rcR * (ustar - unR));", ");"),
    ("upwind_side_swapped", "ustar > 0 ? qL : qR", "ustar > 0 ? qR : qL"),
    ("energy_denominator", "/ (S - ustar));", "/ (S - un));"),
    ("normal_flux_no_pressure", "f.mx1 = (d_o * un_o * un_o) + ptot_o;", "f.mx1 = (d_o * un_o * un_o);"),
    ("godunov_y_area", "flux.d += ds1 * (fR.d - fL.d);", "flux.d += ds0 * (fR.d - fL.d);"),
    ("godunov_z_stride", "load(prim, c + s2));", "load(prim, c + s1));"),
    ("time_step_no_abs", "cs + std::abs(q.ux1)", "cs + q.ux1"),
    ("time_step_dx", "(cx2 * invdx2)", "(cx2 * dx2)"),
]

# (name, original text, mutated text) in oracle_capi.cpp: the step loop of simulate
LOOP_MUTANTS = [
    ("simulate_no_cfl", "cfl * dt, n0, n1, n2);", "dt, n0, n1, n2);"),
    ("simulate_dt_recorded_with_cfl", "dts[it] = dt;", "dts[it] = cfl * dt;"),
    ("simulate_no_conversion", "oracle::cons_to_prim(U, P, eos, n);", ";"),
]

CHECKS = [
    lambda o: pins.test_prim_to_cons_closed_form(o),
    lambda o: pins.test_cons_to_prim_closed_form(o),
    lambda o: pins.test_velocity_is_momentum_times_reciprocal_density(o),
    lambda o: pins.test_time_step_uniform_closed_form(o),
    lambda o: pins.test_time_step_anisotropic_closed_form(o),
    lambda o: [pins.test_hllc_consistency(o, d) for d in range(3)],
    lambda o: [pins.test_hllc_moving_contact_resolved(o, d) for d in range(3)],
    lambda o: [pins.test_hllc_matches_toro_form(o, d) for d in range(3)],
    lambda o: pins.test_godunov_telescoped_conservation(o),
    lambda o: pins.test_godunov_axis_permutation(o, (1, 0, 2)),
    lambda o: pins.test_riemann_problems_converge_to_exact(o, "sod"),
    lambda o: pins.test_simulate_is_the_reference_loop(o),
]


def build_mutant(tmp: Path, name: str, old: str, new: str, fname: str = "euler_oracle.hpp") -> Path:
    text = (SRC / fname).read_text()
    assert text.count(old) >= 1, f"mutation anchor for {name} not found"
    d = tmp / name
    d.mkdir()
    for f in ("euler_oracle.hpp", "oracle_capi.cpp"):
        if f != fname:
            shutil.copy(SRC / f, d / f)
    (d / fname).write_text(text.replace(old, new, 1))
    lib = d / "liboracle.so"
    subprocess.run(["g++", "-std=c++20", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                    "-o", str(lib), str(d / "oracle_capi.cpp")], check=True)
    return lib


@pytest.mark.parametrize("name,old,new,fname", [m + ("euler_oracle.hpp",) for m in MUTANTS]
                         + [m + ("oracle_capi.cpp",) for m in LOOP_MUTANTS],
                         ids=[m[0] for m in MUTANTS + LOOP_MUTANTS])
def test_mutant_is_caught(tmp_path, name, old, new, fname):
    mutant = Oracle(lib_path=build_mutant(tmp_path, name, old, new, fname))
    caught = []
    for i, check in enumerate(CHECKS):
        try:
            check(mutant)
        except (AssertionError, ValueError, FloatingPointError):
            caught.append(i)
    assert caught, f"no pin catches mutant {name}"
