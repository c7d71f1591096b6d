"""Operator-by-operator parity of the HIP kernels against the CPU oracle.

Tolerances (stated bar): FP64 results within 1e-13 relative, FP32 within 1e-5
relative, dt within 1 ulp. Both sides evaluate the reference's expression trees
with contraction off and correctly rounded division and square root, so the
results are expected to be bit-identical, and the tests also assert that.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import random_prim
from gpu_util import max_rel, to_cons, to_host, to_prim

pytestmark = pytest.mark.gpu

RTOL = {np.float64: 1e-13, np.float32: 1e-5}
TDTYPE = {np.float64: torch.float64, np.float32: torch.float32}
SHAPES = [(7, 5, 3), (66, 10, 7), (70, 19, 13), (3, 3, 3), (131, 4, 5)]


def assert_parity(got, ref, dtype, what):
    err = max_rel(got, ref)
    assert err <= RTOL[dtype], f"{what}: max rel err {err:.3e}"
    mism = int(np.sum(got != ref))
    assert mism == 0, f"{what}: {mism} cells differ in the last bits (max rel {err:.3e})"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("index_bits", [32, 64])
@pytest.mark.parametrize("shape", SHAPES[:3])
def test_conversions(eb, ex, orc, dtype, index_bits, shape):
    print(f"conversions dtype={dtype.__name__} bits={index_bits} shape={shape}", flush=True)
    gamma = 1.4
    prim = random_prim(shape, seed=11, dtype=dtype)
    cons_ref = orc.prim_to_cons(prim, gamma)
    P = to_prim(eb, prim, index_bits)
    C = to_cons(eb, np.zeros_like(prim), index_bits)
    eb.prim_to_cons(ex, P, C, eb.PerfectGas(gamma))
    assert_parity(to_host(C), cons_ref, dtype, "prim_to_cons")

    prim_ref = orc.cons_to_prim(cons_ref, gamma)
    P2 = to_prim(eb, np.zeros_like(prim), index_bits)
    eb.cons_to_prim(ex, C, P2, eb.PerfectGas(gamma))
    assert_parity(to_host(P2), prim_ref, dtype, "cons_to_prim")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("index_bits", [32, 64])
@pytest.mark.parametrize("shape", [(7, 5, 3), (66, 10, 7), (97, 33, 41)])
def test_time_step(eb, ex, orc, dtype, index_bits, shape):
    print(f"time_step dtype={dtype.__name__} bits={index_bits} shape={shape}", flush=True)
    gamma = 1.4
    dx = (1.0 / 7, 1.0 / 5, 1.0 / 3)
    prim = random_prim(shape, seed=12, dtype=dtype)
    ref = orc.time_step(prim, gamma, *dx)
    P = to_prim(eb, prim, index_bits)
    got = eb.time_step(ex, P, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx))
    assert got == ref
    # Device-resident form, launched twice to exercise the self-resetting scratch.
    out = eb.time_step_async(ex, P, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx))
    out2 = eb.time_step_async(ex, P, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx))
    ex.fence()
    assert float(out.item()) == ref and float(out2.item()) == ref


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_time_step_large_and_misaligned(eb, ex, orc, dtype):
    """Enough cells for four 16-byte groups in flight per thread (6.5 M cells, an
    odd count, so a partial group is left), and the same fields at an offset of one
    element (not 16-byte aligned, so the one-cell-per-load kernel runs): both equal
    the oracle's dt."""
    gamma = 1.4
    dx = (1.0 / 254, 1.0 / 158, 1.0 / 159)
    shape = (256, 160, 161)
    prim = random_prim(shape, seed=14, dtype=dtype)
    ref = orc.time_step(prim, gamma, *dx)
    P = to_prim(eb, prim, 64)
    assert all(t.data_ptr() % 16 == 0 for t in P.tensors())
    got = eb.time_step(ex, P, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx))
    assert got == ref
    shifted = []
    for f in range(5):
        base = torch.empty(prim[f].size + 1, dtype=P.d.dtype, device="cuda")
        view = base[1:].view(prim[f].shape)
        view.copy_(torch.from_numpy(np.ascontiguousarray(prim[f])))
        shifted.append(view)
    Q = eb.EulerPrimArrays(*shifted, index_bits=64)
    assert all(t.data_ptr() % 16 != 0 for t in Q.tensors())
    assert eb.time_step(ex, Q, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx)) == ref


def test_time_step_nan_cells_skipped(eb, ex, orc):
    """Kokkos::Min keeps the incumbent when the candidate is NaN (time_step.hpp:37)."""
    prim = random_prim((9, 8, 7), seed=13)
    prim[1, 3, 4, 5] = -1.0  # negative pressure -> NaN sound speed -> NaN candidate
    ref = orc.time_step(prim, 1.4, 0.1, 0.1, 0.1)
    assert np.isfinite(ref)
    got = eb.time_step(ex, to_prim(eb, prim), eb.PerfectGas(1.4), eb.UniformMesh3d(0.1, 0.1, 0.1))
    assert got == ref


def test_time_step_uniform_closed_form(eb, ex):
    """rho = p = 1, u = 0, dx = 1: dt = 1/(3 sqrt(1.4))."""
    n = 16
    prim = np.zeros((5, n, n, n))
    prim[0] = 1.0
    prim[1] = 1.0
    got = eb.time_step(ex, to_prim(eb, prim), eb.PerfectGas(1.4), eb.UniformMesh3d(1.0, 1.0, 1.0))
    c = np.sqrt(1.4)
    assert got == 1.0 / ((c + c) + c)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("index_bits", [32, 64])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tile", [(0, 4, 64), (0, 2, 3), (1, 8, 64), (1, 4, 3), (1, 16, 1), (2, 4, 64), (2, 8, 3),
                                  (2, 4, 1), (3, 4, 64), (3, 8, 3), (3, 4, 1), (1, 0, 0), (3, 0, 5), (4, 8, 64),
                                  (4, 8, 3), (4, 16, 1), (4, 16, 5), (1, 4, 64), (4, 0, 0)])
def test_godunov_tiled(eb, ex, orc, dtype, index_bits, shape, tile):
    print(f"godunov dtype={dtype.__name__} bits={index_bits} shape={shape} tile={tile}", flush=True)
    gamma = 1.4
    n0, n1, n2 = shape
    dx = (1.0 / n0, 1.0 / n1, 1.0 / n2)
    prim = random_prim(shape, seed=21, dtype=dtype)
    cons = orc.prim_to_cons(prim, gamma)
    dt = 0.5 * orc.time_step(prim, gamma, *dx)
    ref = cons.copy()
    orc.godunov(prim, ref, gamma, *dx, dt)

    eb.set_godunov_config(*tile)
    try:
        P = to_prim(eb, prim, index_bits)
        C = to_cons(eb, cons, index_bits)
        eb.godunov(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), eb.hllc(), float(dt))
        got = to_host(C)
    finally:
        eb.set_godunov_config(4, 0, 0)
    assert_parity(got, ref, dtype, "godunov")
    # Ghost layer untouched.
    for sl in (np.s_[:, 0], np.s_[:, -1], np.s_[:, :, 0], np.s_[:, :, -1], np.s_[..., 0], np.s_[..., -1]):
        assert np.array_equal(got[sl], cons[sl])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(130, 70, 100), (190, 50, 77)])
def test_godunov_block_counts_off_the_xcd_multiple(eb, ex, orc, dtype, shape):
    """Default configuration where the launch has 8q + r blocks with r != 0 (250
    and 225 blocks of 64 x 16 tiles): the block-to-tile map over the eight XCDs
    must still cover every tile once. One godunov step, then two fused steps."""
    print(f"godunov + fused steps default config {dtype.__name__} shape={shape}", flush=True)
    gamma, cfl = 1.4, 0.5
    n0, n1, n2 = shape
    dx = (1.0 / (n0 - 2), 1.0 / (n1 - 2), 1.0 / (n2 - 2))
    prim = random_prim(shape, seed=25, dtype=dtype)
    cons = orc.prim_to_cons(prim, gamma)
    dt = 0.5 * orc.time_step(prim, gamma, *dx)
    ref = cons.copy()
    orc.godunov(prim, ref, gamma, *dx, dt)
    P, C = to_prim(eb, prim), to_cons(eb, cons)
    eb.godunov(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), eb.hllc(), float(dt))
    assert_parity(to_host(C), ref, dtype, "godunov")

    p_ref, c_ref = prim.copy(), cons.copy()
    orc.simulate(p_ref, c_ref, gamma, *dx, cfl, 2)
    P, C = to_prim(eb, prim, 64), to_cons(eb, cons, 64)
    sim = eb.simulation.Simulation(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), cfl)
    sim.run(2)
    sim.finish()
    assert_parity(to_host(C), c_ref, dtype, "fused steps (cons)")
    assert_parity(to_host(P), p_ref, dtype, "fused steps (prim)")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_godunov_reference_order(eb, ex, orc, dtype):
    gamma = 1.4
    shape = (70, 19, 13)
    dx = (1.0 / 68, 1.0 / 17, 1.0 / 11)
    prim = random_prim(shape, seed=22, dtype=dtype)
    cons = orc.prim_to_cons(prim, gamma)
    dt = 0.5 * orc.time_step(prim, gamma, *dx)
    ref = cons.copy()
    orc.godunov(prim, ref, gamma, *dx, dt)
    P, C = to_prim(eb, prim), to_cons(eb, cons)
    eb.godunov(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), eb.hllc(), float(dt), reference_order=True)
    assert_parity(to_host(C), ref, dtype, "godunov reference order")


def test_godunov_device_dt(eb, ex, orc):
    """dt read on the device as cfl * dt_dev[0] gives the same bits as a host dt."""
    gamma, cfl = 1.4, 0.49
    shape = (40, 12, 9)
    dx = (1.0 / 38, 1.0 / 10, 1.0 / 7)
    prim = random_prim(shape, seed=23)
    cons = orc.prim_to_cons(prim, gamma)
    ref = cons.copy()
    orc.godunov(prim, ref, gamma, *dx, cfl * orc.time_step(prim, gamma, *dx))
    P, C = to_prim(eb, prim), to_cons(eb, cons)
    eos, mesh = eb.PerfectGas(gamma), eb.UniformMesh3d(*dx)
    dt_dev = eb.time_step_async(ex, P, eos, mesh)
    eb.godunov(ex, P, C, eos, mesh, eb.hllc(), dt_dev, cfl=cfl)
    assert_parity(to_host(C), ref, np.float64, "godunov device dt")


@pytest.mark.parametrize("shape", [(2, 5, 5), (5, 1, 5), (0, 4, 4)])
def test_godunov_no_interior_is_noop(eb, ex, shape):
    prim = random_prim(shape, seed=24)
    cons = prim * 3.0
    C = to_cons(eb, cons)
    eb.godunov(ex, to_prim(eb, prim), C, eb.PerfectGas(1.4), eb.UniformMesh3d(1.0, 1.0, 1.0), eb.hllc(), 0.1)
    assert np.array_equal(to_host(C), cons)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("index_bits", [32, 64])
def test_init_implode_and_diagonal(eb, ex, orc, dtype, index_bits):
    n0, n1, n2 = 18, 12, 10
    dx = (1.0 / 16, 1.0 / 10, 1.0 / 8)
    P = to_prim(eb, np.zeros((5, n2, n1, n0), dtype=dtype), index_bits)
    eb.init_implode(ex, P, eb.UniformMesh3d(*dx))
    assert np.array_equal(to_host(P), orc.init_implode(n0, n1, n2, *dx, dtype=dtype))
    eb.init_diagonal(ex, P, eb.UniformMesh3d(*dx), offset=0.45)
    assert np.array_equal(to_host(P), orc.init_diagonal(n0, n1, n2, *dx, offset=0.45, dtype=dtype))


def test_error_paths(eb, ex):
    from benchmark_euler_kernels_amd._lib import EulerError  # loaded by the eb fixture

    prim = random_prim((6, 6, 6), seed=25)
    P, C = to_prim(eb, prim), to_cons(eb, prim)
    with pytest.raises(EulerError):
        eb.set_godunov_config(0, 8, 8)
    with pytest.raises(EulerError):
        eb.set_godunov_config(1, 5, 8)
    with pytest.raises(EulerError):
        eb.set_godunov_config(0, 4, -1)
    with pytest.raises(EulerError):
        eb.set_godunov_config(3, 16, 8)
    with pytest.raises(EulerError):
        eb.set_godunov_config(4, 4, 8)  # paired rows: 8 or 16 rows per tile
    # linear tiles need rows of at most 64*by cells
    wide = random_prim((300, 4, 4), seed=26)
    eb.set_godunov_config(3, 4, 8)
    try:
        with pytest.raises(EulerError):
            eb.godunov(ex, to_prim(eb, wide), to_cons(eb, wide), eb.PerfectGas(1.4), eb.UniformMesh3d(1, 1, 1),
                       eb.hllc(), 0.1)
    finally:
        eb.set_godunov_config(4, 0, 0)
    assert eb.get_godunov_config() == (4, 0, 0)
    with pytest.raises(TypeError):
        eb.godunov(ex, P, to_cons(eb, prim.astype(np.float32)), eb.PerfectGas(1.4),
                   eb.UniformMesh3d(1, 1, 1), eb.hllc(), 0.1)
    with pytest.raises(TypeError):
        eb.godunov(ex, P, C, eb.PerfectGas(1.4), eb.UniformMesh3d(1, 1, 1), object(), 0.1)
    with pytest.raises(ValueError):
        eb.godunov(ex, P, to_cons(eb, prim), eb.PerfectGas(1.4), eb.UniformMesh3d(1, 1, 1), eb.hllc(), 0.1,
                   cfl=0.5)


def test_index_overflow_rejected_before_launch(eb, ex):
    """32-bit indexing of more cells than int can address is refused (int_cast rule,
    benchmark_utils.hpp:10-17); the check precedes any launch, so the small buffers
    behind the pointers are never touched."""
    import ctypes

    from benchmark_euler_kernels_amd import _lib

    prim = random_prim((4, 4, 4), seed=26)
    P, C = to_prim(eb, prim), to_cons(eb, prim)
    pp, cp = eb.operators.prim_ptrs(P), eb.operators.cons_ptrs(C)
    rc = _lib.lib.eb_godunov_f64(ex.handle, ctypes.byref(pp), ctypes.byref(cp), 2000, 2000, 2000, 32,
                                 1.4, 1.0, 1.0, 1.0, 0.1)
    assert rc == _lib.EB_ERR_INDEX_OVERFLOW
    rc = _lib.lib.eb_godunov_f64(ex.handle, ctypes.byref(pp), ctypes.byref(cp), -1, 4, 4, 64,
                                 1.4, 1.0, 1.0, 1.0, 0.1)
    assert rc == _lib.EB_ERR_INVALID_ARGUMENT
    rc = _lib.lib.eb_godunov_f64(ex.handle, None, ctypes.byref(cp), 4, 4, 4, 64, 1.4, 1.0, 1.0, 1.0, 0.1)
    assert rc == _lib.EB_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_stream_probe_adds_fields(eb, ex, dtype):
    """The bandwidth probe behind bench.py's streaming ceiling: b.f += a.f for the
    five fields over the first n cells, the rest untouched; bad arguments refused."""
    import ctypes

    from benchmark_euler_kernels_amd import _lib

    suf = "f64" if dtype == np.float64 else "f32"
    a = random_prim((9, 7, 5), seed=31, dtype=dtype)
    b = random_prim((9, 7, 5), seed=32, dtype=dtype)
    A, B = to_prim(eb, a), to_cons(eb, b)
    pa, pb = eb.operators.prim_ptrs(A), eb.operators.cons_ptrs(B)
    n = 9 * 7 * 5 - 4
    fn = getattr(_lib.lib, f"eb_stream_probe_{suf}")
    assert fn(ex.handle, ctypes.byref(pa), ctypes.byref(pb), n) == _lib.EB_OK
    got = to_host(B)
    flat_a = a.reshape(5, -1)  # (5, n2, n1, n0) C order: memory order i + n0*(j + n1*k)
    flat_b = b.reshape(5, -1)
    want = flat_b.copy()
    want[:, :n] = flat_b[:, :n] + flat_a[:, :n]
    assert np.array_equal(got.reshape(5, -1), want)
    assert fn(ex.handle, ctypes.byref(pa), ctypes.byref(pb), 0) == _lib.EB_OK
    assert fn(ex.handle, ctypes.byref(pa), ctypes.byref(pb), -1) == _lib.EB_ERR_INVALID_ARGUMENT
    assert fn(ex.handle, None, ctypes.byref(pb), n) == _lib.EB_ERR_INVALID_ARGUMENT
    assert fn(ex.handle, ctypes.byref(pa), None, n) == _lib.EB_ERR_INVALID_ARGUMENT
