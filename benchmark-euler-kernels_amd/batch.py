"""Batched ensemble: many independent Godunov problems (own extents, gamma, cell
widths, CFL factor and time step each) advanced by one persistent launch per
step over a queue of tiles drawn from every instance (eb_batch_* in the C ABI).

Each instance follows the reference step loop (simulations/euler_simulation.cpp:56-67)
with its own raw dt kept on the device and its own FP64 simulated time; results
are bitwise those of running the instance alone through simulation.Simulation.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import operators as ops
from ._lib import BatchInstanceF32, BatchInstanceF64, call


class Ensemble:
    """prims[i] holds instance i's initial primitives (and, after finish(), its
    current ones); conss[i] must hold prim_to_cons(prims[i]). All instances share
    one real type and index width."""

    def __init__(self, exec_space: ops.ExecSpace, prims: Sequence[ops.EulerPrimArrays],
                 conss: Sequence[ops.EulerConsArrays], eoss: Sequence[ops.PerfectGas],
                 meshes: Sequence[ops.UniformMesh3d], cfls: Sequence[float]):
        n = len(prims)
        if n == 0 or not (n == len(conss) == len(eoss) == len(meshes) == len(cfls)):
            raise ValueError("one prim, cons, eos, mesh and cfl per instance")
        self.ex = exec_space
        sufs = {ops._validate(p, c) for p, c in zip(prims, conss)}
        if len(sufs) != 1:
            raise TypeError("all instances must share one real type")
        self.suf = sufs.pop()
        bits = {p.index_bits for p in prims} | {c.index_bits for c in conss}
        if len(bits) != 1:
            raise ValueError("all instances must share one index width")
        self.index_bits = bits.pop()
        self.prims = list(prims)
        self.conss = list(conss)
        self.second = [ops.EulerPrimArrays(*(torch.empty_like(t) for t in p.tensors()), index_bits=self.index_bits)
                       for p in prims]
        Inst = BatchInstanceF64 if self.suf == "f64" else BatchInstanceF32
        arr = (Inst * n)()
        for i, (p, c, e, m, cfl) in enumerate(zip(prims, conss, eoss, meshes, cfls)):
            n0, n1, n2 = ops.extents(p)
            if ops.extents(c) != (n0, n1, n2):
                raise ValueError(f"instance {i}: prim and cons extents differ")
            arr[i].prim_a = ops.prim_ptrs(p)
            arr[i].prim_b = ops.prim_ptrs(self.second[i])
            arr[i].cons = ops.cons_ptrs(c)
            arr[i].n0, arr[i].n1, arr[i].n2 = n0, n1, n2
            arr[i].gamma, arr[i].dx0, arr[i].dx1, arr[i].dx2, arr[i].cfl = e.gamma, m.dx0, m.dx1, m.dx2, cfl
        self._handle = ctypes.c_void_p()
        with torch.cuda.device(exec_space.device):
            torch.cuda.synchronize()  # the instances' arrays are ready before the batch reads them
            call(f"eb_batch_create_{self.suf}", ctypes.byref(self._handle), arr, n, self.index_bits)
        self.n = n
        self.steps_done = 0
        self._started = False
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.per_instance_graph: Optional[torch.cuda.CUDAGraph] = None

    # ------------------------------------------------------------- config ---
    def set_tiles(self, by: int, kchunk: int, layout: int = 0) -> None:
        """Tile layout 0: 64 x by columns (by 4, 8, 16); 1: runs of 64*by cells of a
        plane (by 4, 8, and 16 in FP32; every n0 <= 64*by); 2: 64 x by columns with two
        rows per lane (by 8, 16). kchunk planes per tile, 0 = whole columns."""
        call("eb_batch_set_tiles", self._handle, layout, by, kchunk)
        self.graph = self.per_instance_graph = None

    def tiles(self):
        lay, by, kc, nt, grid = (ctypes.c_int() for _ in range(5))
        call("eb_batch_get_tiles", self._handle, ctypes.byref(lay), ctypes.byref(by), ctypes.byref(kc),
             ctypes.byref(nt), ctypes.byref(grid))
        return {"layout": lay.value, "by": by.value, "kchunk": kc.value, "tiles": nt.value, "grid": grid.value}

    def autotune(self, reps: int = 3) -> float:
        """Times the candidate tile shapes on the live state (left unchanged) and
        keeps the fastest; returns its milliseconds per launch."""
        if not self._started:
            self.start()
        ms = ctypes.c_float()
        call("eb_batch_autotune", self._handle, self.ex.handle, self.steps_done % 2, reps, ctypes.byref(ms))
        self.graph = self.per_instance_graph = None
        return ms.value

    # -------------------------------------------------------------- steps ---
    def start(self):
        call("eb_batch_start", self._handle, self.ex.handle)
        self._started = True

    def _launch(self, parity: int, stream: int, per_instance: bool):
        name = "eb_batch_step_per_instance" if per_instance else "eb_batch_step"
        call(name, self._handle, stream, parity)

    def step(self, per_instance: bool = False):
        if not self._started:
            self.start()
        self._launch(self.steps_done % 2, self.ex.handle, per_instance)
        if self.steps_done == 0:
            call("eb_batch_refresh_ghosts", self._handle, self.ex.handle, 0)
        self.steps_done += 1

    def run(self, nsteps: int, use_graph: bool = False, per_instance: bool = False):
        """Advance every instance nsteps. With use_graph, pairs of steps are
        replayed from a captured graph (one launch per step, or one per instance
        and step with per_instance)."""
        while self.steps_done < 1 and nsteps > 0:
            self.step(per_instance)
            nsteps -= 1
        if use_graph and nsteps >= 2 and self.steps_done % 2 == 0:
            self.step(per_instance)  # captured pairs start at odd parity
            nsteps -= 1
        if use_graph and nsteps >= 2:
            g = self._pair_graph(per_instance)
            while nsteps >= 2:
                g.replay()
                self.steps_done += 2
                nsteps -= 2
        while nsteps > 0:
            self.step(per_instance)
            nsteps -= 1

    def _pair_graph(self, per_instance: bool) -> torch.cuda.CUDAGraph:
        g = self.per_instance_graph if per_instance else self.graph
        if g is None:
            s = torch.cuda.Stream(self.ex.device)
            s.wait_stream(self.ex.stream)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                self._launch(1, s.cuda_stream, per_instance)
                self._launch(0, s.cuda_stream, per_instance)
            self.ex.stream.wait_stream(s)
            if per_instance:
                self.per_instance_graph = g
            else:
                self.graph = g
        return g

    # -------------------------------------------------------------- state ---
    def current_prim(self, i: int) -> ops.EulerPrimArrays:
        return self.prims[i] if self.steps_done % 2 == 0 else self.second[i]

    def state(self):
        """(dt, time): dt is (n, 3) raw [ping, pong, ghost min], time is (n,) FP64."""
        dtype = torch.float64 if self.suf == "f64" else torch.float32
        dt = torch.empty(self.n, 3, dtype=dtype, device=self.ex.device)
        t = torch.empty(self.n, dtype=torch.float64, device=self.ex.device)
        call("eb_batch_copy_state", self._handle, self.ex.handle, dt.data_ptr(), t.data_ptr())
        return dt, t

    def next_dt(self) -> torch.Tensor:
        """Raw dt each instance's next step uses."""
        dt, _ = self.state()
        return dt[:, self.steps_done % 2].clone()

    def finish(self):
        """Leave every instance's current primitives in the caller's prim arrays."""
        if self.steps_done % 2 == 1:
            for dst, src in zip(self.prims, self.second):
                for a, b in zip(dst.tensors(), src.tensors()):
                    a.copy_(b)
        self.ex.fence()

    def close(self):
        if self._handle:
            self.ex.fence()
            call("eb_batch_destroy", self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_ensemble(ex: ops.ExecSpace, count: int, seed: int = 1234, lo: int = 32, hi: int = 160,
                    dtype=torch.float32, index_bits: int = 32, cfl: float = 0.5):
    """The measured ensemble: instance i (seed + i) has interior extents drawn
    from U{lo..hi} per axis, gamma ~ U[1.2, 1.67], a diagonal implosion with
    interface offset c ~ U[0.3, 0.6], inner rho ~ U[0.05, 0.3] and inner
    p ~ U[0.05, 0.3] (outer state rho = p = 1), u = 0, dx_d = 1/n_d."""
    import numpy as np

    prims: List[ops.EulerPrimArrays] = []
    conss: List[ops.EulerConsArrays] = []
    eoss, meshes, cfls, params = [], [], [], []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        n = [int(v) for v in rng.integers(lo, hi + 1, size=3)]
        gamma = float(rng.uniform(1.2, 1.67))
        c = float(rng.uniform(0.3, 0.6))
        d_in = float(rng.uniform(0.05, 0.3))
        p_in = float(rng.uniform(0.05, 0.3))
        na = [v + 2 for v in n]
        mesh = ops.UniformMesh3d(1.0 / n[0], 1.0 / n[1], 1.0 / n[2])
        size = na[0] * na[1] * na[2]
        prim = ops.to_mdspan(ops.create_prim_arrays_1d(ex, size, dtype), *na, index_bits=index_bits)
        cons = ops.to_mdspan(ops.create_cons_arrays_1d(ex, size, dtype), *na, index_bits=index_bits)
        eos = ops.PerfectGas(gamma)
        ops.init_diagonal(ex, prim, mesh, offset=c, d_hi=1.0, p_hi=1.0, d_lo=d_in, p_lo=p_in)
        ops.prim_to_cons(ex, prim, cons, eos)
        prims.append(prim)
        conss.append(cons)
        eoss.append(eos)
        meshes.append(mesh)
        cfls.append(cfl)
        params.append({"n": n, "gamma": gamma, "offset": c, "rho_in": d_in, "p_in": p_in})
    return prims, conss, eoss, meshes, cfls, params
