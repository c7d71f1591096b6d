"""Step loop of the reference driver (simulations/euler_simulation.cpp:50-80) on
the fused, host-sync-free path.

Reference order per step:  dt = time_step(prim); godunov(prim, cons, cfl*dt);
cons_to_prim(cons -> prim).  Here each step is ONE launch of the fused kernel
(godunov + cons_to_prim + next time_step) over double-buffered primitives, with
dt kept on the device. The ghost primitives are refreshed from the (frozen) ghost
conservatives by a small kernel, which also yields the ghost cells' share of the
dt minimum; because ghost conservatives never change (godunov.hpp:40-43) that share
is constant after the first refresh.

Exactly reproduced semantics:
  * the first dt comes from time_step on the initial primitives, ghosts included;
  * step 0 reads the initial ghost primitives, every later step reads ghost
    primitives round-tripped through cons_to_prim (euler_simulation.cpp:50-67);
  * simulated time t = sum of cfl*dt is accumulated in FP64 on the device.

With track_conservation, each step also adds its boundary outflow (the amount by
which it lowers the interior totals of the conservative fields) to an FP64
compensated accumulator on the device, so that totals(U_n) + outflow = totals(U_0)
can be checked to rounding (conservation() reports it).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import operators as ops
from ._lib import call


class Simulation:
    """Fused-step driver. prim / cons are the caller's arrays (prim is read as
    the initial state and, after run(), holds the current primitives)."""

    def __init__(self, exec_space: ops.ExecSpace, prim: ops.EulerPrimArrays, cons: ops.EulerConsArrays,
                 eos: ops.PerfectGas, mesh: ops.UniformMesh3d, cfl: float, track_conservation: bool = False):
        self.suf = ops._validate(prim, cons)
        self.ex = exec_space
        self.eos = eos
        self.mesh = mesh
        self.cfl = cfl
        self.shape = ops.extents(prim)
        n0, n1, n2 = self.shape
        if min(self.shape) < 3:
            raise ValueError("the fused step needs at least one interior cell")
        dtype = prim.d.dtype
        dev = prim.d.device
        self.buf = [prim, ops.EulerPrimArrays(*(torch.empty_like(t) for t in prim.tensors()),
                                              index_bits=prim.index_bits)]
        self.cons = cons
        # dt[0], dt[1]: ping-pong raw dt; dt[2]: ghost-cell minimum.
        self.dt = torch.zeros(3, dtype=dtype, device=dev)
        self.time = torch.zeros(1, dtype=torch.float64, device=dev)
        self.steps_done = 0
        self._ptrs = [ops.prim_ptrs(b) for b in self.buf]
        self._cons_ptrs = ops.cons_ptrs(cons)
        self._started = False
        self.graph: Optional[torch.cuda.CUDAGraph] = None  # one pair of steps
        self.chunk_graph: Optional[torch.cuda.CUDAGraph] = None  # CHUNK_PAIRS pairs
        self.track = track_conservation
        self.outflow = torch.zeros(10, dtype=torch.float64, device=dev)
        self.totals0: Optional[torch.Tensor] = None

    # ------------------------------------------------------------ launches ---
    def _r(self, x):
        return ops._real(self.suf, x)

    def _ghost_refresh(self, target: int, stream: int):
        n0, n1, n2 = self.shape
        m = self.mesh
        call(f"eb_ghost_refresh_{self.suf}", self.ex.context, stream, ctypes.byref(self._cons_ptrs),
             ctypes.byref(self._ptrs[target]), n0, n1, n2, self.buf[0].index_bits, self._r(self.eos.gamma),
             self._r(m.dx0), self._r(m.dx1), self._r(m.dx2), self.dt[2:].data_ptr())

    def _fused(self, src: int, stream: int):
        n0, n1, n2 = self.shape
        m = self.mesh
        dst = 1 - src
        call(f"eb_step_fused_{self.suf}", self.ex.context, stream, ctypes.byref(self._ptrs[src]),
             ctypes.byref(self._cons_ptrs), ctypes.byref(self._ptrs[dst]), n0, n1, n2, self.buf[0].index_bits,
             self._r(self.eos.gamma), self._r(m.dx0), self._r(m.dx1), self._r(m.dx2), self._r(self.cfl),
             self.dt[src:].data_ptr(), self.dt[2:].data_ptr(), self.dt[dst:].data_ptr(), self.time.data_ptr())
        if self.track:
            call(f"eb_boundary_flux_{self.suf}", self.ex.context, stream, ctypes.byref(self._ptrs[src]), n0, n1, n2,
                 self.buf[0].index_bits, self._r(self.eos.gamma), self._r(m.dx0), self._r(m.dx1), self._r(m.dx2),
                 self.dt[src:].data_ptr(), self._r(self.cfl), self.outflow.data_ptr())

    def start(self):
        """First dt from the initial primitives; ghost primitives of the second
        buffer from the ghost conservatives."""
        h = self.ex.handle
        if self.track:
            self.totals0 = ops.conserved_totals(self.ex, self.cons)
        ops.time_step_async(self.ex, self.buf[0], self.eos, self.mesh, out=self.dt[0:1])
        self._ghost_refresh(1, h)
        self._started = True

    def step(self):
        if not self._started:
            self.start()
        h = self.ex.handle
        src = self.steps_done % 2
        self._fused(src, h)
        if self.steps_done == 0:
            # Buffer 0 still holds the initial ghost primitives; from now on every
            # read must see the round-tripped ones.
            self._ghost_refresh(0, h)
        self.steps_done += 1

    # Steps per replay of the long graph: short graphs (two steps) measured ~4%
    # slower than eager launches at 128^3, so bulk steps replay a longer one.
    CHUNK_PAIRS = 16

    def run(self, nsteps: int, use_graph: bool = False):
        """Advance nsteps. With use_graph, runs of 2*CHUNK_PAIRS steps and pairs
        of steps are each captured once into a hipGraph and replayed (the first two
        steps always run eagerly)."""
        while self.steps_done < 2 and nsteps > 0:
            self.step()
            nsteps -= 1
        if use_graph and nsteps >= 2 and self.steps_done % 2 == 1:
            self.step()  # the captured pairs start from buffer 0
            nsteps -= 1
        if use_graph and nsteps >= 2 * self.CHUNK_PAIRS:
            g = self._chunk_graph()
            while nsteps >= 2 * self.CHUNK_PAIRS:
                g.replay()
                self.steps_done += 2 * self.CHUNK_PAIRS
                nsteps -= 2 * self.CHUNK_PAIRS
        if use_graph and nsteps >= 2:
            g = self._pair_graph()
            while nsteps >= 2:
                g.replay()
                self.steps_done += 2
                nsteps -= 2
        while nsteps > 0:
            self.step()
            nsteps -= 1

    def _capture(self, pairs: int) -> torch.cuda.CUDAGraph:
        assert self.steps_done % 2 == 0
        s = torch.cuda.Stream(self.ex.device)
        s.wait_stream(self.ex.stream)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(pairs):
                self._fused(0, s.cuda_stream)
                self._fused(1, s.cuda_stream)
        self.ex.stream.wait_stream(s)
        return g

    def _pair_graph(self) -> torch.cuda.CUDAGraph:
        if self.graph is None:
            self.graph = self._capture(1)
        return self.graph

    def _chunk_graph(self) -> torch.cuda.CUDAGraph:
        if self.chunk_graph is None:
            self.chunk_graph = self._capture(self.CHUNK_PAIRS)
        return self.chunk_graph

    def conservation(self) -> dict:
        """Interior totals now and at start, the accumulated boundary outflow and
        the residual totals + outflow - totals0 (FP64, host floats)."""
        if not self.track or self.totals0 is None:
            raise RuntimeError("construct with track_conservation=True and run first")
        now = ops.conserved_totals(self.ex, self.cons)
        out = self.outflow[0::2] + self.outflow[1::2]
        t0, t1, o = (x.cpu().tolist() for x in (self.totals0, now, out))
        fields = ("mass", "energy", "mom0", "mom1", "mom2")
        res = {f: {"total0": a, "total": b, "outflow": c, "residual": b + c - a} for f, a, b, c in zip(fields, t0, t1, o)}
        return res

    # -------------------------------------------------------------- state ---
    def current_prim(self) -> ops.EulerPrimArrays:
        return self.buf[self.steps_done % 2]

    def last_dt(self) -> torch.Tensor:
        """Raw dt that the next step will use (device tensor, one element)."""
        return self.dt[self.steps_done % 2:][:1]

    def finish(self):
        """Leave the current primitives in the caller's prim arrays."""
        if self.steps_done % 2 == 1:
            for dst, src in zip(self.buf[0].tensors(), self.buf[1].tensors()):
                dst.copy_(src)
        self.ex.fence()
