"""CMake callers: a project written like the reference's (add_subdirectory, link the
`euler_operators` target, include "godunov.hpp" and friends, the reference's type and
function names) configures and builds against this repository, drivers included.
CPU only: compile and link, nothing runs on a GPU."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

CONSUMER_CMAKE = """cmake_minimum_required(VERSION 3.25)
project(consumer CXX)
add_subdirectory({root} euler)
add_executable(consumer consumer.cpp)
target_link_libraries(consumer PRIVATE euler_operators)
"""

# the reference driver's loop (simulations/euler_simulation.cpp:50-67) with only the
# Kokkos view and execution-space spellings changed
CONSUMER_CPP = """#include "cons_to_prim.hpp"
#include "euler_arrays.hpp"
#include "godunov.hpp"
#include "init_implode.hpp"
#include "prim_to_cons.hpp"
#include "time_step.hpp"

using view = eb::mdspan<double, eb::dextents<int, 3>, eb::layout_left>;

int main()
{
    eb::DefaultExecutionSpace const exec_space;
    PerfectGas<double> const eos(1.4);
    UniformMesh3d<double> const mesh(0.1, 0.1, 0.1);
    auto const prims_alloc = create_prim_arrays_1d<double>(exec_space, 1000);
    EulerPrimArrays<view> const prim_arrays = to_mdspan<view>(prims_alloc, 10, 10, 10);
    auto const cons_alloc = create_cons_arrays_1d<double>(exec_space, 1000);
    EulerConsArrays<view> const cons_arrays = to_mdspan<view>(cons_alloc, 10, 10, 10);
    init_implode(exec_space, prim_arrays, mesh);
    prim_to_cons(exec_space, as_const(prim_arrays), cons_arrays, eos);
    for (int it = 0; it < 3; ++it) {
        double const dt = time_step(exec_space, as_const(prim_arrays), eos, mesh);
        godunov(exec_space, as_const(prim_arrays), cons_arrays, eos, mesh, hllc {}, 0.49 * dt);
        cons_to_prim(exec_space, as_const(cons_arrays), prim_arrays, eos);
    }
    exec_space.fence();
    return 0;
}
"""


def test_reference_style_cmake_project_builds(tmp_path):
    cmake = shutil.which("cmake")
    if cmake is None:
        pytest.skip("cmake not available")
    src = tmp_path / "consumer"
    src.mkdir()
    (src / "CMakeLists.txt").write_text(CONSUMER_CMAKE.format(root=ROOT.as_posix()))
    (src / "consumer.cpp").write_text(CONSUMER_CPP)
    build = tmp_path / "build"
    r = subprocess.run([cmake, "-S", str(src), "-B", str(build)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    r = subprocess.run([cmake, "--build", str(build), "-j", "4"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    for exe in (build / "consumer", build / "euler" / "euler_simulation", build / "euler" / "euler_benchmarks"):
        assert exe.exists(), exe
    ldd = subprocess.run(["ldd", str(build / "consumer")], capture_output=True, text=True)
    line = [l for l in ldd.stdout.splitlines() if "libeuler_hip.so" in l]
    assert line and "not found" not in line[0], ldd.stdout
