"""Python face of the CPU oracle (test infrastructure only: used by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the product).

States are numpy arrays of shape (5, n2, n1, n0): the five SoA fields in the
reference's order, each in layout_left storage (element (i, j, k) at [f, k, j, i]).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"

_PP = ctypes.c_void_p * 5


def _fields(a: np.ndarray):
    if a.ndim != 4 or a.shape[0] != 5 or not a.flags.c_contiguous:
        raise ValueError("state must be a C-contiguous (5, n2, n1, n0) array")
    return _PP(*(a[f].ctypes.data for f in range(5)))


class Oracle:
    def __init__(self, threads: int | None = None, lib_path: Path | None = None):
        path = Path(lib_path) if lib_path is not None else LIB_PATH
        if not path.exists():
            raise ImportError(f"{path} missing: run `make -C {LIB_PATH.parent}`")
        self.lib = ctypes.CDLL(os.fspath(path))
        i64, vp = ctypes.c_int64, ctypes.c_void_p
        for suf, real in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            L = self.lib
            getattr(L, f"oracle_hllc_{suf}").argtypes = [ctypes.c_int, real, vp, vp, vp]
            getattr(L, f"oracle_cons_to_prim_{suf}").argtypes = [_PP, _PP, real, i64]
            getattr(L, f"oracle_prim_to_cons_{suf}").argtypes = [_PP, _PP, real, i64]
            getattr(L, f"oracle_time_step_{suf}").argtypes = [_PP, real, real, real, real, i64]
            getattr(L, f"oracle_time_step_{suf}").restype = real
            getattr(L, f"oracle_godunov_{suf}").argtypes = [_PP, _PP, real, real, real, real, real, i64, i64, i64]
            getattr(L, f"oracle_init_implode_{suf}").argtypes = [_PP, real, real, real, i64, i64, i64]
            getattr(L, f"oracle_init_diagonal_{suf}").argtypes = [_PP] + [real] * 8 + [i64, i64, i64]
            getattr(L, f"oracle_simulate_{suf}").argtypes = [_PP, _PP] + [real] * 5 + [i64, i64, i64, i64, vp]
        self.lib.oracle_set_threads.argtypes = [ctypes.c_int]
        self.lib.oracle_max_threads.restype = ctypes.c_int
        if threads is not None:
            self.lib.oracle_set_threads(threads)

    @staticmethod
    def _suf(a: np.ndarray):
        if a.dtype == np.float64:
            return "f64", ctypes.c_double
        if a.dtype == np.float32:
            return "f32", ctypes.c_float
        raise TypeError(a.dtype)

    def threads(self) -> int:
        return self.lib.oracle_max_threads()

    def set_threads(self, n: int) -> None:
        self.lib.oracle_set_threads(n)

    def hllc(self, direction: int, gamma: float, qL, qR, dtype=np.float64) -> np.ndarray:
        """Flux (d, e, mx0, mx1, mx2) of hllc.hpp:53-104 across a face normal to `direction`."""
        qL = np.ascontiguousarray(qL, dtype=dtype)
        qR = np.ascontiguousarray(qR, dtype=dtype)
        out = np.zeros(5, dtype=dtype)
        suf, real = self._suf(out)
        getattr(self.lib, f"oracle_hllc_{suf}")(direction, real(gamma), qL.ctypes.data, qR.ctypes.data, out.ctypes.data)
        return out

    def cons_to_prim(self, cons: np.ndarray, gamma: float, out: np.ndarray | None = None) -> np.ndarray:
        prim = np.zeros_like(cons) if out is None else out
        suf, real = self._suf(cons)
        getattr(self.lib, f"oracle_cons_to_prim_{suf}")(_fields(cons), _fields(prim), real(gamma), cons[0].size)
        return prim

    def prim_to_cons(self, prim: np.ndarray, gamma: float, out: np.ndarray | None = None) -> np.ndarray:
        cons = np.zeros_like(prim) if out is None else out
        suf, real = self._suf(prim)
        getattr(self.lib, f"oracle_prim_to_cons_{suf}")(_fields(prim), _fields(cons), real(gamma), prim[0].size)
        return cons

    def time_step(self, prim: np.ndarray, gamma: float, dx0: float, dx1: float, dx2: float):
        suf, real = self._suf(prim)
        return getattr(self.lib, f"oracle_time_step_{suf}")(
            _fields(prim), real(gamma), real(dx0), real(dx1), real(dx2), prim[0].size
        )

    def godunov(self, prim: np.ndarray, cons: np.ndarray, gamma: float, dx0, dx1, dx2, dt) -> None:
        """In place on cons (godunov.hpp:15-92)."""
        suf, real = self._suf(prim)
        n2, n1, n0 = prim.shape[1:]
        getattr(self.lib, f"oracle_godunov_{suf}")(
            _fields(prim), _fields(cons), real(gamma), real(dx0), real(dx1), real(dx2), real(dt), n0, n1, n2
        )

    def init_implode(self, n0: int, n1: int, n2: int, dx0, dx1, dx2, dtype=np.float64) -> np.ndarray:
        prim = np.zeros((5, n2, n1, n0), dtype=dtype)
        suf, real = self._suf(prim)
        getattr(self.lib, f"oracle_init_implode_{suf}")(_fields(prim), real(dx0), real(dx1), real(dx2), n0, n1, n2)
        return prim

    def init_diagonal(self, n0, n1, n2, dx0, dx1, dx2, offset=0.45, d_hi=1.0, p_hi=1.0, d_lo=0.125, p_lo=0.14,
                      dtype=np.float64) -> np.ndarray:
        prim = np.zeros((5, n2, n1, n0), dtype=dtype)
        suf, real = self._suf(prim)
        args = [real(v) for v in (dx0, dx1, dx2, offset, d_hi, p_hi, d_lo, p_lo)]
        getattr(self.lib, f"oracle_init_diagonal_{suf}")(_fields(prim), *args, n0, n1, n2)
        return prim

    def simulate(self, prim: np.ndarray, cons: np.ndarray, gamma, dx0, dx1, dx2, cfl, nsteps) -> np.ndarray:
        """euler_simulation.cpp:56-68 loop, in place; returns the raw dt of each step."""
        suf, real = self._suf(prim)
        n2, n1, n0 = prim.shape[1:]
        dts = np.zeros(nsteps, dtype=prim.dtype)
        getattr(self.lib, f"oracle_simulate_{suf}")(
            _fields(prim), _fields(cons), real(gamma), real(dx0), real(dx1), real(dx2), real(cfl),
            nsteps, n0, n1, n2, dts.ctypes.data
        )
        return dts
