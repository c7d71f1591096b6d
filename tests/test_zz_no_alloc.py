"""eb_time_step_noalloc_f64 (the context-free, allocation-free time step) through
the C ABI against eb_time_step_f64 (context scratch): two calls in a row, the
second on a smaller grid, each with 32- and 64-bit indices; every dt bitwise equal
(the minimum is exact, so both reductions give the same bits), and argument
errors reported without a launch."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from conftest import random_prim
from gpu_util import to_prim

pytestmark = pytest.mark.gpu


def _dt_with_context(eb, ex, P, shape, dx, index_bits):
    from benchmark_euler_kernels_amd import _lib

    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    pp = eb.operators.prim_ptrs(P)
    rc = _lib.lib.eb_time_step_f64(ex.context, ex.handle, ctypes.byref(pp), *shape, index_bits, 1.4, *dx,
                                   out.data_ptr())
    assert rc == _lib.EB_OK
    return out


def _dt_noalloc(eb, ex, P, shape, dx, index_bits):
    from benchmark_euler_kernels_amd import _lib

    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    pp = eb.operators.prim_ptrs(P)
    rc = _lib.lib.eb_time_step_noalloc_f64(ex.handle, ctypes.byref(pp), *shape, index_bits, 1.4, *dx,
                                           out.data_ptr())
    assert rc == _lib.EB_OK, _lib.lib.eb_last_error()
    return out


@pytest.mark.parametrize("index_bits", [32, 64])
def test_noalloc_time_step_matches_context_time_step(eb, ex, index_bits):
    cases = [((258, 130, 97), 41), ((19, 7, 5), 42)]  # the second call on a smaller grid
    for shape, seed in cases:
        print(f"noalloc time step shape={shape} bits={index_bits}", flush=True)
        prim = random_prim(shape, seed=seed)
        P = to_prim(eb, prim, index_bits)
        dx = (1.0 / (shape[0] - 2), 1.0 / (shape[1] - 2), 1.0 / (shape[2] - 2))
        want = _dt_with_context(eb, ex, P, shape, dx, index_bits)
        got = _dt_noalloc(eb, ex, P, shape, dx, index_bits)
        ex.fence()
        assert got.item() == want.item() and want.item() > 0, (shape, got.item(), want.item())


def test_noalloc_time_step_misaligned_fields_and_errors(eb, ex):
    """Fields at an offset of one element (the one-cell-per-load kernel) give the
    context path's dt; a null output or oversized 32-bit extents are refused."""
    from benchmark_euler_kernels_amd import _lib

    shape = (33, 20, 11)
    prim = random_prim(shape, seed=43)
    views = []
    for f in range(5):
        base = torch.empty(prim[f].size + 1, dtype=torch.float64, device="cuda")
        v = base[1:].view(prim[f].shape)
        v.copy_(torch.from_numpy(np.ascontiguousarray(prim[f])))
        views.append(v)
    P = eb.EulerPrimArrays(*views, index_bits=64)
    dx = (0.1, 0.2, 0.3)
    want = _dt_with_context(eb, ex, P, shape, dx, 64)
    got = _dt_noalloc(eb, ex, P, shape, dx, 64)
    ex.fence()
    assert got.item() == want.item()
    pp = eb.operators.prim_ptrs(P)
    rc = _lib.lib.eb_time_step_noalloc_f64(ex.handle, ctypes.byref(pp), *shape, 64, 1.4, *dx, None)
    assert rc == _lib.EB_ERR_INVALID_ARGUMENT
    rc = _lib.lib.eb_time_step_noalloc_f64(ex.handle, ctypes.byref(pp), 2000, 2000, 2000, 32, 1.4, *dx,
                                           want.data_ptr())
    assert rc == _lib.EB_ERR_INDEX_OVERFLOW
