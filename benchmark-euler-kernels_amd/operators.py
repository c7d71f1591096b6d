"""Host-side mirror of the reference's euler_operators API over the HIP C ABI.

Names, argument order and meaning follow the reference headers:

    cons_to_prim(exec_space, cons_arrays, prim_arrays, eos)          cons_to_prim.hpp:7-18
    prim_to_cons(exec_space, prim_arrays, cons_arrays, eos)          prim_to_cons.hpp:7-18
    time_step(exec_space, prim_arrays, eos, mesh) -> float           time_step.hpp:8-16
    godunov(exec_space, prim_arrays, cons_arrays, eos, mesh, riemann_solver, dt)
                                                                     godunov.hpp:15-29
    init_implode(exec_space, prim_arrays, mesh)                      init_implode.hpp:7-14

with the data types EulerPrimArrays / EulerConsArrays (euler_arrays.hpp:77-85,
228-236), PerfectGas (perfect_gas.hpp:6-29), UniformMesh3d (uniform_mesh.hpp:5-52)
and the hllc tag (hllc.hpp:50). Arrays are torch tensors on the GPU holding the
reference's layout_left storage: a field of extents (n0, n1, n2) is a contiguous
tensor of shape (n2, n1, n0), so element (i, j, k) is t[k, j, i].

Every operator enqueues on the exec space's HIP stream and returns immediately,
except time_step, which returns the host value as the reference does.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Union

import torch

from . import _lib
from ._lib import ConsPtrs, PrimPtrs, call

# real_type / index_type switches (benchmarks/real_type.hpp:3-4,
# benchmarks/index_type.hpp:3-4): double and int by default.
REAL_TYPES = {torch.float64: "f64", torch.float32: "f32"}
INDEX_BITS_DEFAULT = 32


@dataclass(frozen=True)
class PerfectGas:
    """perfect_gas.hpp:6-29 (gamma only; the relations live in the kernels)."""

    gamma: float


@dataclass(frozen=True)
class UniformMesh3d:
    """uniform_mesh.hpp:5-52."""

    dx0: float
    dx1: float
    dx2: float

    def ds0(self) -> float:
        return self.dx1 * self.dx2

    def ds1(self) -> float:
        return self.dx2 * self.dx0

    def ds2(self) -> float:
        return self.dx0 * self.dx1

    def dv(self) -> float:
        return self.dx0 * self.dx1 * self.dx2


class hllc:  # noqa: N801 - the reference's type name
    """Tag selecting the HLLC Riemann solver (hllc.hpp:50-105), the only one."""


@dataclass
class EulerPrimArrays:
    """SoA primitive arrays {d, p, ux0, ux1, ux2} (euler_arrays.hpp:77-85).

    index_bits plays the role of the mdspan's IndexType: 32 (int/unsigned) or 64.
    """

    d: torch.Tensor
    p: torch.Tensor
    ux0: torch.Tensor
    ux1: torch.Tensor
    ux2: torch.Tensor
    index_bits: int = INDEX_BITS_DEFAULT

    def tensors(self):
        return (self.d, self.p, self.ux0, self.ux1, self.ux2)


@dataclass
class EulerConsArrays:
    """SoA conservative arrays {d, e, mx0, mx1, mx2} (euler_arrays.hpp:228-236)."""

    d: torch.Tensor
    e: torch.Tensor
    mx0: torch.Tensor
    mx1: torch.Tensor
    mx2: torch.Tensor
    index_bits: int = INDEX_BITS_DEFAULT

    def tensors(self):
        return (self.d, self.e, self.mx0, self.mx1, self.mx2)


Arrays = Union[EulerPrimArrays, EulerConsArrays]


class ExecSpace:
    """Stand-in for Kokkos::DefaultExecutionSpace: a HIP stream plus the device
    scratch of the CFL reduction (an eb_context)."""

    def __init__(self, device: Optional[torch.device] = None, stream: Optional[torch.cuda.Stream] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("ExecSpace needs a HIP device")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            call("eb_context_create", ctypes.byref(self._ctx))

    @property
    def handle(self) -> int:
        return self.stream.cuda_stream

    @property
    def context(self) -> ctypes.c_void_p:
        return self._ctx

    def fence(self) -> None:
        self.stream.synchronize()

    def close(self) -> None:
        if self._ctx:
            self.fence()
            _lib.lib.eb_context_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: nothing left to release into
            pass


# ----------------------------------------------------------------- helpers ---


def _suffix(t: torch.Tensor) -> str:
    try:
        return REAL_TYPES[t.dtype]
    except KeyError:
        raise TypeError(f"real type must be float64 or float32, got {t.dtype}") from None


def extents(arrays: Arrays):
    """(n0, n1, n2) of the layout_left arrays."""
    t = arrays.d
    if t.dim() != 3:
        raise ValueError("fields must be 3-D tensors of shape (n2, n1, n0)")
    return int(t.shape[2]), int(t.shape[1]), int(t.shape[0])


def _validate(*groups: Arrays) -> str:
    ref = groups[0].d
    suf = _suffix(ref)
    for g in groups:
        for t in g.tensors():
            if t.dtype != ref.dtype:
                raise TypeError("all fields must share one real type")
            if t.device.type != "cuda":
                raise ValueError("fields must live on the GPU")
            if not t.is_contiguous():
                raise ValueError("fields must be contiguous (layout_left storage)")
            if t.shape != ref.shape:
                raise ValueError("all fields must share the same extents")
        if g.index_bits not in (32, 64):
            raise ValueError("index_bits must be 32 or 64")
    return suf


def prim_ptrs(a: EulerPrimArrays) -> PrimPtrs:
    return PrimPtrs(*(t.data_ptr() for t in a.tensors()))


def cons_ptrs(a: EulerConsArrays) -> ConsPtrs:
    return ConsPtrs(*(t.data_ptr() for t in a.tensors()))


def _real(suf: str, x: float):
    return ctypes.c_double(x) if suf == "f64" else ctypes.c_float(x)


# ------------------------------------------------------------- allocation ---


def create_prim_arrays_1d(exec_space: ExecSpace, n: int, dtype=torch.float64) -> EulerPrimArrays:
    """Five zero-initialised 1-D fields (euler_arrays.hpp:216-226)."""
    mk = lambda: torch.zeros(n, dtype=dtype, device=exec_space.device)  # noqa: E731
    return EulerPrimArrays(mk(), mk(), mk(), mk(), mk())


def create_cons_arrays_1d(exec_space: ExecSpace, n: int, dtype=torch.float64) -> EulerConsArrays:
    """Five zero-initialised 1-D fields (euler_arrays.hpp:401-411)."""
    mk = lambda: torch.zeros(n, dtype=dtype, device=exec_space.device)  # noqa: E731
    return EulerConsArrays(mk(), mk(), mk(), mk(), mk())


def to_mdspan(arrays: Arrays, n0: int, n1: int, n2: int, index_bits: int = INDEX_BITS_DEFAULT) -> Arrays:
    """3-D layout_left views of 1-D fields (euler_arrays.hpp:130-142, 281-293)."""
    views = [t.view(n2, n1, n0) for t in arrays.tensors()]
    return type(arrays)(*views, index_bits=index_bits)


def as_const(arrays: Arrays) -> Arrays:
    """Read-only view marker (euler_arrays.hpp:98-104); tensors are shared."""
    return arrays


def size(arrays: Arrays) -> int:
    """Elements per field (euler_arrays.hpp:367-379)."""
    return arrays.d.numel()


def size_bytes(arrays: Arrays) -> int:
    """5 * sizeof(T) * size (euler_arrays.hpp:381-399)."""
    return 5 * arrays.d.element_size() * size(arrays)


# -------------------------------------------------------------- operators ---


def cons_to_prim(exec_space: ExecSpace, cons_arrays: EulerConsArrays, prim_arrays: EulerPrimArrays, eos: PerfectGas):
    """cons_to_prim.hpp:7-33: all cells, ghosts included."""
    suf = _validate(cons_arrays, prim_arrays)
    n0, n1, n2 = extents(cons_arrays)
    c, p = cons_ptrs(cons_arrays), prim_ptrs(prim_arrays)
    call(f"eb_cons_to_prim_{suf}", exec_space.handle, ctypes.byref(c), ctypes.byref(p),
         n0, n1, n2, prim_arrays.index_bits, _real(suf, eos.gamma))


def prim_to_cons(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, cons_arrays: EulerConsArrays, eos: PerfectGas):
    """prim_to_cons.hpp:7-33: all cells, ghosts included."""
    suf = _validate(prim_arrays, cons_arrays)
    n0, n1, n2 = extents(prim_arrays)
    p, c = prim_ptrs(prim_arrays), cons_ptrs(cons_arrays)
    call(f"eb_prim_to_cons_{suf}", exec_space.handle, ctypes.byref(p), ctypes.byref(c),
         n0, n1, n2, cons_arrays.index_bits, _real(suf, eos.gamma))


def time_step_async(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, eos: PerfectGas, mesh: UniformMesh3d,
                    out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """time_step.hpp:8-41 without the host round trip: the raw dt (no CFL factor)
    lands in the one-element device tensor `out` (allocated if None)."""
    suf = _validate(prim_arrays)
    n0, n1, n2 = extents(prim_arrays)
    if out is None:
        out = torch.empty(1, dtype=prim_arrays.d.dtype, device=prim_arrays.d.device)
    if out.dtype != prim_arrays.d.dtype or out.numel() < 1 or out.device.type != "cuda":
        raise ValueError("out must be a device tensor of the arrays' real type")
    p = prim_ptrs(prim_arrays)
    r = lambda x: _real(suf, x)  # noqa: E731
    call(f"eb_time_step_{suf}", exec_space.context, exec_space.handle, ctypes.byref(p), n0, n1, n2,
         prim_arrays.index_bits, r(eos.gamma), r(mesh.dx0), r(mesh.dx1), r(mesh.dx2), out.data_ptr())
    return out


def time_step(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, eos: PerfectGas, mesh: UniformMesh3d) -> float:
    """time_step.hpp:8-41: min over all cells of 1/sum_d (c+|u_d|)/dx_d, returned
    to the host (blocking, like the reference)."""
    suf = _validate(prim_arrays)
    n0, n1, n2 = extents(prim_arrays)
    p = prim_ptrs(prim_arrays)
    r = lambda x: _real(suf, x)  # noqa: E731
    host = ctypes.c_double() if suf == "f64" else ctypes.c_float()
    call(f"eb_time_step_sync_{suf}", exec_space.context, exec_space.handle, ctypes.byref(p), n0, n1, n2,
         prim_arrays.index_bits, r(eos.gamma), r(mesh.dx0), r(mesh.dx1), r(mesh.dx2), ctypes.byref(host))
    return float(host.value)


def godunov(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, cons_arrays: EulerConsArrays, eos: PerfectGas,
            mesh: UniformMesh3d, riemann_solver: hllc, dt: Union[float, torch.Tensor], cfl: float = 1.0,
            reference_order: bool = False):
    """godunov.hpp:15-92: cons -= dt/dv * sum_d ds_d (F_{d+1/2} - F_{d-1/2}) on
    interior cells, F the HLLC flux of the (const) primitives.

    dt is a host float, as in the reference, or a one-element device tensor, in
    which case cfl * dt[0] is used on the device with no host synchronisation.
    reference_order=True runs the six-solves-per-cell kernel (validation only).
    """
    if not isinstance(riemann_solver, hllc) and riemann_solver is not hllc:
        raise TypeError("the Riemann solver must be hllc")
    suf = _validate(prim_arrays, cons_arrays)
    if prim_arrays.d.numel() > 0 and prim_arrays.d.data_ptr() == cons_arrays.d.data_ptr():
        raise ValueError("prim and cons must not alias")
    n0, n1, n2 = extents(prim_arrays)
    p, c = prim_ptrs(prim_arrays), cons_ptrs(cons_arrays)
    r = lambda x: _real(suf, x)  # noqa: E731
    common = (exec_space.handle, ctypes.byref(p), ctypes.byref(c), n0, n1, n2, cons_arrays.index_bits,
              r(eos.gamma), r(mesh.dx0), r(mesh.dx1), r(mesh.dx2))
    if isinstance(dt, torch.Tensor):
        if reference_order:
            raise ValueError("the reference-order kernel takes a host dt")
        if dt.dtype != prim_arrays.d.dtype or dt.device.type != "cuda":
            raise ValueError("device dt must match the arrays' real type and live on the GPU")
        call(f"eb_godunov_dev_{suf}", *common, dt.data_ptr(), r(cfl))
    else:
        if cfl != 1.0:
            raise ValueError("scale a host dt before the call, as the reference's caller does")
        name = f"eb_godunov_reference_order_{suf}" if reference_order else f"eb_godunov_{suf}"
        call(name, *common, r(float(dt)))


def init_implode(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, mesh: UniformMesh3d):
    """init_implode.hpp:7-45: sphere r^2 < 0.25^2 about the centre, ghosts included."""
    suf = _validate(prim_arrays)
    n0, n1, n2 = extents(prim_arrays)
    p = prim_ptrs(prim_arrays)
    r = lambda x: _real(suf, x)  # noqa: E731
    call(f"eb_init_implode_{suf}", exec_space.handle, ctypes.byref(p), n0, n1, n2, prim_arrays.index_bits,
         r(mesh.dx0), r(mesh.dx1), r(mesh.dx2))


def init_diagonal(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, mesh: UniformMesh3d, offset: float = 0.45,
                  d_hi: float = 1.0, p_hi: float = 1.0, d_lo: float = 0.125, p_lo: float = 0.14):
    """Diagonal implosion: x0+x1+x2 > offset -> (d_hi, p_hi) else (d_lo, p_lo), u = 0,
    cell centres ((i-1)+0.5)*dx on the unit cube (no reference counterpart)."""
    suf = _validate(prim_arrays)
    n0, n1, n2 = extents(prim_arrays)
    p = prim_ptrs(prim_arrays)
    r = lambda x: _real(suf, x)  # noqa: E731
    call(f"eb_init_diagonal_{suf}", exec_space.handle, ctypes.byref(p), n0, n1, n2, prim_arrays.index_bits,
         r(mesh.dx0), r(mesh.dx1), r(mesh.dx2), r(offset), r(d_hi), r(p_hi), r(d_lo), r(p_lo))


def init_random(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, seed: int = 1234, d_range=(0.5, 2.0),
                p_range=(0.5, 2.0), u_sigma: float = 0.3):
    """Seeded synthetic state: rho ~ U[d_range), p ~ U[p_range), u_k ~ N(0, u_sigma),
    counter-based so it does not depend on the launch shape (no reference counterpart)."""
    suf = _validate(prim_arrays)
    p = prim_ptrs(prim_arrays)
    call(f"eb_init_random_{suf}", exec_space.handle, ctypes.byref(p), size(prim_arrays), prim_arrays.index_bits,
         ctypes.c_uint64(seed), d_range[0], d_range[1], p_range[0], p_range[1], u_sigma)


def conserved_totals(exec_space: ExecSpace, cons_arrays: EulerConsArrays,
                     out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Interior totals of (d, e, mx0, mx1, mx2) in FP64 with Neumaier-compensated,
    fixed-order sums (device tensor of 5 doubles; no reference counterpart)."""
    suf = _validate(cons_arrays)
    n0, n1, n2 = extents(cons_arrays)
    if out is None:
        out = torch.empty(5, dtype=torch.float64, device=cons_arrays.d.device)
    call(f"eb_conserved_totals_{suf}", exec_space.context, exec_space.handle, ctypes.byref(cons_ptrs(cons_arrays)),
         n0, n1, n2, cons_arrays.index_bits, out.data_ptr())
    return out


def boundary_flux(exec_space: ExecSpace, prim_arrays: EulerPrimArrays, eos: PerfectGas, mesh: UniformMesh3d,
                  dt_dev: torch.Tensor, cfl: float, acc: torch.Tensor) -> None:
    """Adds to acc (10 doubles: (sum, compensation) per conservative field) the
    amount by which a godunov step with dt = cfl * dt_dev[0] lowers the interior
    totals: dt/dv * sum_d ds_d (boundary flux out - boundary flux in)."""
    suf = _validate(prim_arrays)
    n0, n1, n2 = extents(prim_arrays)
    if acc.dtype != torch.float64 or acc.numel() < 10 or not acc.is_cuda:
        raise ValueError("acc must be a device tensor of 10 float64")
    r = lambda x: _real(suf, x)  # noqa: E731
    call(f"eb_boundary_flux_{suf}", exec_space.context, exec_space.handle, ctypes.byref(prim_ptrs(prim_arrays)),
         n0, n1, n2, prim_arrays.index_bits, r(eos.gamma), r(mesh.dx0), r(mesh.dx1), r(mesh.dx2),
         dt_dev.data_ptr(), r(cfl), acc.data_ptr())


def set_godunov_config(variant: int, by: int, kchunk: int) -> None:
    """Godunov kernel variant (0 face-parallel, 1 cell-parallel, 2 streamed,
    3 linear tiles, 4 cell-parallel with two rows per wave), tile height (0:
    chosen per kernel) and planes marched per block (0: sized per launch from the
    grid). Default (4, 0, 0). Results do not depend on it."""
    call("eb_set_godunov_config", variant, by, kchunk)


def get_godunov_config():
    v, by, kc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("eb_get_godunov_config", ctypes.byref(v), ctypes.byref(by), ctypes.byref(kc))
    return v.value, by.value, kc.value


def godunov_tile_height(dtype=torch.float64, fused: bool = False) -> int:
    """Tile height the current configuration launches for this real type, in the
    flux kernel or (fused=True) the fused step."""
    by = ctypes.c_int()
    call("eb_get_godunov_tile_height", 8 if dtype == torch.float64 else 4, int(fused), ctypes.byref(by))
    return by.value


__all__ = [
    "PerfectGas", "UniformMesh3d", "hllc", "EulerPrimArrays", "EulerConsArrays", "ExecSpace",
    "create_prim_arrays_1d", "create_cons_arrays_1d", "to_mdspan", "as_const", "size", "size_bytes",
    "cons_to_prim", "prim_to_cons", "time_step", "time_step_async", "godunov", "init_implode",
    "init_diagonal", "init_random", "set_godunov_config", "get_godunov_config", "godunov_tile_height", "extents",
    "conserved_totals", "boundary_flux",
]
