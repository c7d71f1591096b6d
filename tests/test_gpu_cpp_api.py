"""C++ API (include/euler/*.hpp) over the HIP library: the compiled check
program, and the step-loop driver's .npy output against the CPU oracle's
reference-order loop (euler_simulation.cpp:50-80), bitwise, for the reference
loop, the fused loop and the graph-replayed loop."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "benchmark-euler-kernels_amd" / "bin"


def _run(args, cwd=None, timeout=120):
    exe = BIN / args[0]
    if not exe.exists():
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    print("running", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([str(exe)] + [str(a) for a in args[1:]], cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    return r


def test_cpp_api_checks():
    r = _run(["test_cpp_api"])
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout
    assert "FAIL" not in r.stdout


@pytest.mark.parametrize("real", ["f64", "f32"])
def test_driver_matches_oracle(orc, tmp_path, real):
    nx, nt, freq, cfl, gamma = 16, 20, 10, 0.49, 1.4
    dtype = np.float64 if real == "f64" else np.float32
    outs = {}
    for loop in ("reference", "fused", "graph"):
        d = tmp_path / loop
        d.mkdir()
        r = _run(["euler_simulation", "--nx", nx, "--nt", nt, "--output-freq", freq, "--cfl", cfl, "--real", real,
                  "--loop", loop, "--out", d])
        assert r.returncode == 0
        outs[loop] = [np.load(d / f"test_{it:010d}.npy") for it in (10, 20)]
    dx = 1.0 / nx
    na = nx + 2
    prim = orc.init_implode(na, na, na, dx, dx, dx, dtype=dtype)
    cons = orc.prim_to_cons(prim, gamma)
    for s, it in enumerate((10, 20)):
        orc.simulate(prim, cons, gamma, dx, dx, dx, cfl, 10)
        ref = prim[1].transpose(2, 1, 0)  # (n2, n1, n0) -> [i, j, k]
        for loop, arrs in outs.items():
            got = arrs[s]
            assert got.dtype == dtype and got.shape == (na, na, na)
            assert np.array_equal(got, ref), (loop, it)


def test_benchmarks_write_google_benchmark_json(tmp_path):
    """euler_benchmarks (the counterpart of the reference's benchmarks/ binary)
    writes Google-Benchmark-shaped JSON that plot.py's loader reads: names
    <Family>/<size>, per-iteration cpu_time in ns and the iteration-invariant
    counters cells_per_second / bytes_per_second = per-iteration cells or bytes
    divided by the CPU time, with the reference's sizes and byte formulas."""
    import json

    out = tmp_path / "bm.json"
    r = _run(["euler_benchmarks", "--benchmark_filter=^(Godunov|TimeStep|ConsToPrim|PrimToCons)/(8|16|32)$|"
              "^EulerSimulation/16$", "--benchmark_min_time=0.02s", f"--benchmark_out={out}",
              "--benchmark_out_format=json"])
    assert r.returncode == 0
    data = json.loads(out.read_text())
    assert data["context"]["real_type"] == "double" and data["context"]["index_type"] == "int"
    rows = {b["name"]: b for b in data["benchmarks"]}
    expected = {f"{f}/{n}" for f in ("Godunov", "TimeStep", "ConsToPrim", "PrimToCons") for n in (8, 16, 32)}
    expected.add("EulerSimulation/16")
    assert set(rows) == expected
    s = 8
    for name, b in rows.items():
        family, size = name.split("/")[0], int(name.split("/")[-1])  # plot.py's parse
        assert b["time_unit"] == "ns" and b["iterations"] >= 1 and b["cpu_time"] > 0 and b["real_time"] > 0
        cpu_s = b["cpu_time"] * 1e-9
        if family == "Godunov":
            cells, nbytes = (size + 2) ** 3, 15 * s * (size + 2) ** 3
        elif family == "TimeStep":
            cells, nbytes = size ** 3, 5 * s * size ** 3
        elif family in ("ConsToPrim", "PrimToCons"):
            cells, nbytes = size ** 3, 10 * s * size ** 3
        else:
            cells, nbytes = (size + 2) ** 3, None
        assert b["cells_per_second"] == pytest.approx(cells / cpu_s, rel=1e-9)
        if nbytes is None:
            assert "bytes_per_second" not in b
        else:
            assert b["bytes_per_second"] == pytest.approx(nbytes / cpu_s, rel=1e-9)
        # the host spins while it waits, so CPU time tracks real time
        assert 0.5 < b["cpu_time"] / b["real_time"] < 1.5, name
