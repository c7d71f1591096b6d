"""The C-ABI library loads without a GPU and exports every entry point that
include/euler_capi.h declares (no compute calls)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "euler_capi.h"
LIB = ROOT / "benchmark-euler-kernels_amd" / "libeuler_hip.so"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for op in ("cons_to_prim", "prim_to_cons", "time_step", "time_step_sync", "godunov", "godunov_dev",
               "step_fused", "ghost_refresh", "init_implode", "init_diagonal", "init_random"):
        for suf in ("f64", "f32"):
            assert f"eb_{op}_{suf}" in names


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "build the library first (make -C benchmark-euler-kernels_amd)"
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_without_gpu():
    lib = ctypes.CDLL(str(LIB))
    lib.eb_status_string.restype = ctypes.c_char_p
    assert lib.eb_status_string(0) == b"ok"
    assert lib.eb_status_string(2) == b"index overflow"


def test_python_package_binds_all_symbols():
    """The host package's ctypes table covers the header (import needs no GPU)."""
    from __graft_entry__ import load_package

    eb = load_package()
    from benchmark_euler_kernels_amd import _lib

    bound = set(_lib.SIGNATURES) | {"eb_status_string", "eb_last_error"}
    assert set(declared_symbols()) <= bound, sorted(set(declared_symbols()) - bound)
    assert eb.godunov is not None


def test_no_undefined_kernel_stubs():
    """Every kernel launched from host code has its host stub in the library (a
    device-only construct in a kernel body can make the host pass drop it)."""
    import shutil
    import subprocess

    nm = shutil.which("nm")
    if nm is None:
        import pytest

        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", str(LIB)], capture_output=True, text=True, check=True).stdout
    undefined = [l for l in out.splitlines() if " U " in l and "__device_stub__" in l]
    assert not undefined, undefined[:5]


# Headers with device code, for callers that build with hipcc.
DEVICE_HEADERS = {"plugins.hpp"}


def test_cpp_headers_compile_standalone(tmp_path):
    """Each public C++ header compiles on its own: with the host compiler, or with
    hipcc for gfx950 for the headers that hold device code."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    inc = ROOT / "include"
    hipcc = Path("/opt/rocm/bin/hipcc")
    for h in sorted((inc / "euler").glob("*.hpp")):
        if h.name in DEVICE_HEADERS:
            src = tmp_path / f"inc_{h.stem}.hip"
            src.write_text(f"#include <euler/{h.name}>\nint main() {{ return 0; }}\n")
            cmd = [str(hipcc), "-std=c++20", "-fsyntax-only", "--offload-arch=gfx950", "-Wall", "-Wextra",
                   "-Werror", "-Wno-unused-command-line-argument", f"-I{inc}", str(src)]
        else:
            src = tmp_path / f"inc_{h.stem}.cpp"
            src.write_text(f"#include <euler/{h.name}>\nint main() {{ return 0; }}\n")
            cmd = [gxx, "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                   "-D__HIP_PLATFORM_AMD__", f"-I{inc}", "-I/opt/rocm/include", str(src)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, (h.name, r.stderr[-2000:])


def test_argument_checks_without_gpu():
    """Host-side argument validation answers before any device work: bad
    configurations, a negative fill count, a null fill target, an unknown real
    width for the tile-height query, a fused step without dt_out."""
    lib = ctypes.CDLL(str(LIB))
    lib.eb_last_error.restype = ctypes.c_char_p
    EB_OK = 0
    assert lib.eb_set_godunov_config(1, 5, 0) != EB_OK
    assert lib.eb_set_godunov_config(1, 4, -1) != EB_OK
    assert lib.eb_set_godunov_config(5, 0, 0) != EB_OK
    v, by, kc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.eb_get_godunov_config(ctypes.byref(v), ctypes.byref(by), ctypes.byref(kc)) == EB_OK
    assert (v.value, by.value, kc.value) == (4, 0, 0)
    h = ctypes.c_int()
    assert lib.eb_get_godunov_tile_height(8, 0, ctypes.byref(h)) == EB_OK and h.value == 16
    assert lib.eb_get_godunov_tile_height(2, 0, ctypes.byref(h)) != EB_OK
    assert lib.eb_set_godunov_config(1, 0, 0) == EB_OK
    assert lib.eb_get_godunov_tile_height(8, 0, ctypes.byref(h)) == EB_OK and h.value == 4
    assert lib.eb_get_godunov_tile_height(8, 1, ctypes.byref(h)) == EB_OK and h.value == 8
    assert lib.eb_get_godunov_tile_height(4, 0, ctypes.byref(h)) == EB_OK and h.value == 8
    assert lib.eb_set_godunov_config(1, 16, 8) == EB_OK
    assert lib.eb_get_godunov_tile_height(8, 0, ctypes.byref(h)) == EB_OK and h.value == 16
    assert lib.eb_set_godunov_config(4, 0, 0) == EB_OK
    lib.eb_fill_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double]
    assert lib.eb_fill_f64(None, None, -1, 0.0) != EB_OK
    assert b"negative" in lib.eb_last_error()
    assert lib.eb_fill_f64(None, None, 10, 0.0) != EB_OK
    assert b"null" in lib.eb_last_error()
    assert lib.eb_fill_f64(None, None, 0, 0.0) == EB_OK  # nothing to do
    # dt_out may be null only for the box step, and then without time_acc (the
    # checks come before the context or any array is touched on the device)
    fake = (ctypes.c_double * 8)()
    ptrs = (ctypes.c_void_p * 5)(*[ctypes.addressof(fake)] * 5)
    dt_in = ctypes.c_double(1.0)
    region = (ctypes.c_int64 * 6)(1, 2, 1, 2, 1, 2)
    acc = ctypes.c_double(0.0)
    common = [ctypes.byref(fake), None, ptrs, ptrs, (ctypes.c_void_p * 5)(*[ctypes.addressof(fake) + 8] * 5),
              ctypes.c_int64(4), ctypes.c_int64(4), ctypes.c_int64(4), 64, ctypes.c_double(1.4),
              ctypes.c_double(0.5), ctypes.c_double(0.5), ctypes.c_double(0.5), ctypes.c_double(0.5),
              ctypes.byref(dt_in), None, None]
    assert lib.eb_step_fused_f64(*common, None) != EB_OK
    assert b"null" in lib.eb_last_error()
    assert lib.eb_step_fused_region_f64(*common, ctypes.byref(acc), region) != EB_OK
    assert b"time_acc" in lib.eb_last_error()


REFERENCE_SPELLED_CALLER = r"""
#include <euler/euler.hpp>

using view = eb::mdspan<double, eb::dextents<int, 3>, eb::layout_left>;

int main()
{
    eb::DefaultExecutionSpace const exec_space;
    PerfectGas<double> const eos(1.4);
    UniformMesh3d<double> const mesh(0.1, 0.1, 0.1);
    auto const prim_alloc = create_prim_arrays_1d<double>(exec_space, 1000);
    EulerPrimArrays<view> const prim = to_mdspan<view>(prim_alloc, 10, 10, 10);
    auto const cons_alloc = create_cons_arrays_1d<double>(exec_space, 1000);
    EulerConsArrays<view> const cons = to_mdspan<view>(cons_alloc, 10, 10, 10);
    init_implode(exec_space, prim, mesh);
    prim_to_cons(exec_space, as_const(prim), cons, eos);
    double const dt = time_step(exec_space, as_const(prim), eos, mesh);
    godunov(exec_space, as_const(prim), cons, eos, mesh, hllc(), 0.5 * dt);
    cons_to_prim(exec_space, as_const(cons), prim, eos);
    EulerPrim<double> const q {1.0, 1.0, 0.0, 0.0, 0.0};
    EulerCons<double> const u = to_cons(q, eos.internal_energy(q.d, q.p));
    EulerPrim<double> const q2 = to_prim(u, eos.pressure(u.d, internal_energy(u)));
    EulerFlux<double> const f = hllc()(dir_t<0>(), eos, q, q2);
    double const g = get(dir_t<1>(), q) + kinetic_energy(q) + f.d;
    std::size_t const b = size_bytes(prim) + size(cons);
    auto const dh = data_handle(prim);
    auto const sub = subspan(prim, 1, eb::full_extent, eb::full_extent);
    EulerPrim<double> const q3 = load(prim, 1, 1, 1);
    store(q3, prim, 2, 2, 2);
    save_npy("x.npy", exec_space, prim.d);
    int const n = int_cast<int>(std::size_t(10));
    (void)g; (void)b; (void)dh; (void)sub; (void)n; (void)as<view>(prim);
    return 0;
}
"""


def test_reference_spellings_compile(tmp_path):
    """A caller written with the reference's global names (EulerPrimArrays,
    PerfectGas, hllc, dir_t, godunov, ... unqualified, as its drivers spell them;
    only the Kokkos view and execution-space types are spelled eb::) compiles
    against the headers; with EB_NO_GLOBAL_NAMES the same source does not."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "caller.cpp"
    src.write_text(REFERENCE_SPELLED_CALLER)
    cmd = [gxx, "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROOT / 'include'}", "-I/opt/rocm/include", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run(cmd + ["-DEB_NO_GLOBAL_NAMES"], capture_output=True, text=True)
    assert r.returncode != 0 and "PerfectGas" in r.stderr


def test_config_reads_are_consistent_under_concurrent_sets():
    """eb_set_godunov_config replaces (variant, by, kchunk) as one value: readers in
    other threads see one of the triples that were set, never a mix of two."""
    import threading

    lib = ctypes.CDLL(str(LIB))
    triples = [(4, 8, 3), (1, 16, 77), (3, 4, 0), (0, 2, 1000)]
    stop = threading.Event()
    bad = []

    def writer(t):
        while not stop.is_set():
            assert lib.eb_set_godunov_config(*t) == 0

    def reader():
        v, by, kc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        for _ in range(20000):
            lib.eb_get_godunov_config(ctypes.byref(v), ctypes.byref(by), ctypes.byref(kc))
            got = (v.value, by.value, kc.value)
            if got not in triples and got != (4, 0, 0):
                bad.append(got)

    ws = [threading.Thread(target=writer, args=(t,)) for t in triples]
    rs = [threading.Thread(target=reader) for _ in range(2)]
    for t in ws + rs:
        t.start()
    for t in rs:
        t.join()
    stop.set()
    for t in ws:
        t.join()
    assert lib.eb_set_godunov_config(4, 0, 0) == 0
    assert not bad, bad[:5]
