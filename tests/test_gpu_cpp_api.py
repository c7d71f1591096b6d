"""C++ API (include/euler/*.hpp) over the HIP library: the compiled check
programs (host and device sides of the API, index types past the 32-bit
kernels' limit), and the step-loop driver's .npy files against the CPU oracle's
reference-order loop (euler_simulation.cpp:50-80) byte for byte, header
included, for the reference loop, the fused loop and the graph-replayed loop,
at the reference driver's own settings (nx 128, nt 200, a file every 10 steps)."""
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


def test_device_api_checks():
    """hllc functor, get<Dir>, load / store and subspan views inside kernels."""
    r = _run(["test_device_api"])
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout
    assert "FAIL" not in r.stdout


def test_plug_points():
    """include/euler/plugins.hpp: the generic kernels with the reference's solver,
    EOS and mesh equal the library's kernels bitwise; a caller's Rusanov solver,
    stiffened-gas EOS and mesh type run on the GPU and equal a host evaluation of
    the reference's loops through the same functors, bitwise (FP64 and FP32)."""
    r = _run(["test_plugins"])
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout
    assert "FAIL" not in r.stdout


def test_index_types_past_32_bit_kernels():
    """int (1290^3 = 2,146,689,000 cells) and unsigned (1300^3 = 2,197,000,000
    cells, past 2^31) through the C++ operators, which route them to the 64-bit
    kernels: cells past 2^31 checked against the reference's formulas evaluated
    on the host through the same public point functions. FP32, about 88 GB of HBM
    per case, one case at a time."""
    r = _run(["test_device_api", "--large"], timeout=300)
    assert r.returncode == 0
    assert "ALL PASS" in r.stdout
    assert "FAIL" not in r.stdout


def npy_bytes(field: np.ndarray) -> bytes:
    """The reference writer's file (simulations/save_npy.cpp:34-75) for one field
    of an oracle state, restated here: header dict with every extent followed by
    ', ', padded with spaces so that 10 + dict + newline is a multiple of 64, then
    the layout_left payload. `field` is (n2, n1, n0), so its C order is the
    Fortran order of [i, j, k]."""
    n2, n1, n0 = field.shape
    descr = "<f8" if field.dtype == np.float64 else "<f4"
    shape = "".join(f"{e}, " for e in (n0, n1, n2))
    d = "{'descr': '" + descr + "', 'fortran_order': True, 'shape': (" + shape + "), }"
    total = 10 + len(d) + 1
    d += " " * ((total + 63) // 64 * 64 - total) + "\n"
    head = b"\x93NUMPY\x01\x00" + len(d).to_bytes(2, "little") + d.encode("latin1")
    return head + np.ascontiguousarray(field).tobytes()


@pytest.mark.parametrize("real,nx,nt,index", [("f64", 128, 200, "int"), ("f32", 16, 20, "unsigned"),
                                              ("f64", 16, 20, "long")])
def test_driver_matches_oracle(orc, tmp_path, real, nx, nt, index):
    """The reference driver's settings (nx 128, nt 200, CFL 0.49, a file every 10
    steps, FP64, int) and smaller FP32 / unsigned and int64 runs: every file equals
    the oracle's byte for byte."""
    freq, cfl, gamma = 10, 0.49, 1.4
    dtype = np.float64 if real == "f64" else np.float32
    dx = 1.0 / nx
    na = nx + 2
    prim = orc.init_implode(na, na, na, dx, dx, dx, dtype=dtype)
    cons = orc.prim_to_cons(prim, gamma)
    expected = {}
    for it in range(freq, nt + 1, freq):
        orc.simulate(prim, cons, gamma, dx, dx, dx, cfl, freq)
        expected[it] = npy_bytes(prim[1])
    for loop in ("reference", "fused", "graph"):
        d = tmp_path / loop
        d.mkdir()
        r = _run(["euler_simulation", "--nx", nx, "--nt", nt, "--output-freq", freq, "--cfl", cfl, "--real", real,
                  "--index", index, "--loop", loop, "--out", d])
        assert r.returncode == 0
        for it, want in expected.items():
            path = d / f"test_{it:010d}.npy"
            got = path.read_bytes()
            assert got[:128] == want[:128], (loop, it, got[:128])
            assert got == want, (loop, it)
            path.unlink()
    # the expected bytes are a valid .npy of the pressure indexed [i, j, k]
    import io

    back = np.load(io.BytesIO(expected[nt]))
    assert back.shape == (na, na, na) and np.array_equal(back, prim[1].transpose(2, 1, 0))


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
