"""Adjoint of a Godunov rollout (reference step loop, euler_simulation.cpp:56-67)
with the time step frozen per step, exposed as a torch.autograd.Function.

Forward:  U_0 = p2c(P_0);  for n < N:  dt_n = time_step(P_n) (recorded, then held
constant for differentiation);  U_{n+1} = godunov(P_n, U_n, cfl*dt_n);
P_{n+1} = c2p(U_{n+1}).  The first step reads the given P_0, later steps read
c2p(U_n), exactly as the reference loop does.

Backward: conservative states are checkpointed every `checkpoint_every` steps;
each segment is recomputed from its checkpoint (with the recorded dt, so it is
bitwise the forward) and then reversed step by step through the hand-written HIP
vector-Jacobian products (eb_godunov_vjp*, eb_cons_to_prim_vjp). The recompute
of the next segment runs on a second stream while the current segment's adjoint
steps run, the two alternating between two sets of segment buffers. The forward
pass writes the last two segments' primitive states straight into those two sets,
so the backward starts with both sets holding states (no fill, no recompute for
them) and needs no checkpoint for them. Memory: one state per checkpoint plus two
primitive states per step of a segment.

Locality: the adjoint of P_n is zero outside the support box of dL/dP_N grown by
N - n cells (one cell per step, the stencil's reach). Each adjoint step only
processes that box, and each recomputed step only advances the cells the adjoint
step will read (the box grown by one more cell); both are exact, so a localised
loss costs less than a full-grid one.
"""
from __future__ import annotations

import ctypes
import weakref
from typing import List, Optional, Sequence

import torch

from . import operators as ops
from ._lib import call


def _clone_prim(a: ops.EulerPrimArrays) -> ops.EulerPrimArrays:
    return ops.EulerPrimArrays(*(t.clone() for t in a.tensors()), index_bits=a.index_bits)


def _clone_cons(a: ops.EulerConsArrays) -> ops.EulerConsArrays:
    return ops.EulerConsArrays(*(t.clone() for t in a.tensors()), index_bits=a.index_bits)


def _empty_cons_like(a: ops.Arrays) -> ops.EulerConsArrays:
    return ops.EulerConsArrays(*(torch.empty_like(t) for t in a.tensors()), index_bits=a.index_bits)


def _empty_prim_like(a: ops.Arrays) -> ops.EulerPrimArrays:
    return ops.EulerPrimArrays(*(torch.empty_like(t) for t in a.tensors()), index_bits=a.index_bits)


def _zeros_cons_like(a: ops.Arrays) -> ops.EulerConsArrays:
    return ops.EulerConsArrays(*(torch.zeros_like(t) for t in a.tensors()), index_bits=a.index_bits)


def _zeros_prim_like(a: ops.Arrays) -> ops.EulerPrimArrays:
    return ops.EulerPrimArrays(*(torch.zeros_like(t) for t in a.tensors()), index_bits=a.index_bits)


class Rollout:
    """N-step rollout with checkpoints and its adjoint (dt frozen per step).

    Steps run on the fused kernel (godunov + cons_to_prim + next time step, the
    launch sequence of simulation.Simulation), so the forward costs what the plain
    step loop costs plus one copy of the conservative state per checkpoint."""

    def __init__(self, exec_space: ops.ExecSpace, eos: ops.PerfectGas, mesh: ops.UniformMesh3d, cfl: float,
                 nsteps: int, checkpoint_every: int = 20):
        if nsteps < 1 or checkpoint_every < 1:
            raise ValueError("nsteps and checkpoint_every must be >= 1")
        self.ex = exec_space
        self.eos = eos
        self.mesh = mesh
        self.cfl = cfl
        self.nsteps = nsteps
        self.k = checkpoint_every
        self.dts: Optional[torch.Tensor] = None
        self.checkpoints: dict = {}
        self.prim0: Optional[ops.EulerPrimArrays] = None
        self.prim_final: Optional[ops.EulerPrimArrays] = None
        self._slots: Optional[torch.Tensor] = None  # [0]: discarded dt, [1]: ghost-cell dt minimum
        # buffer sets of the last segment (P_tail ...) and the one before (P_kept ...)
        self._sets: List[List[ops.EulerPrimArrays]] = [[], []]
        self.tail = 0
        self.kept = 0
        self._tail_ready = False
        self._geom_args = None
        self._ptrs: dict = {}
        self._cfl_arg = None

    # ---------------------------------------------------------------- launches
    def _geom(self):
        # extents, index width, gamma and cell widths as C arguments, built once per forward
        if self._geom_args is None:
            n0, n1, n2 = ops.extents(self.prim0)
            m = self.mesh
            r = lambda x: ops._real(self.suf, x)  # noqa: E731
            self._geom_args = (n0, n1, n2, self.prim0.index_bits, r(self.eos.gamma), r(m.dx0), r(m.dx1), r(m.dx2))
        return self._geom_args

    def _p(self, a: ops.Arrays):
        """The C pointer struct of one of the rollout's states, passed by reference;
        built once per state (the launches of a rollout reuse a few buffers, and
        rebuilding the struct per call costs more host time than the small adjoint
        steps take on the GPU)."""
        hit = self._ptrs.get(id(a))
        if hit is None or hit[0]() is not a:
            st = ops.prim_ptrs(a) if isinstance(a, ops.EulerPrimArrays) else ops.cons_ptrs(a)
            hit = (weakref.ref(a), st, ctypes.byref(st))
            self._ptrs[id(a)] = hit
        return hit[2]

    def _slot_ptr(self, i: int) -> int:
        # [0]: discarded dt, [1]: ghost-cell dt minimum
        return self._slots.data_ptr() + i * self._slots.element_size()

    def _dt_ptr(self, n: int) -> int:
        return self.dts.data_ptr() + n * self.dts.element_size()

    def _ghost_refresh(self, cons: ops.EulerConsArrays, dst: ops.EulerPrimArrays, with_dt: bool = True):
        # with_dt=False: the ghost primitives only (segment buffers; the ghost dt
        # minimum, slot 1, comes from the first refresh and never changes)
        call(f"eb_ghost_refresh_{self.suf}", self.ex.context, self.ex.handle, self._p(cons), self._p(dst),
             *self._geom(), self._slot_ptr(1) if with_dt else None)

    def _fused(self, src: ops.EulerPrimArrays, cons: ops.EulerConsArrays, dst: ops.EulerPrimArrays, dt_in: int,
               dt_out: Optional[int], region=None, ex: Optional[ops.ExecSpace] = None):
        ex = self.ex if ex is None else ex
        args = (ex.context, ex.handle, self._p(src), self._p(cons), self._p(dst), *self._geom(),
                self._cfl_arg, dt_in, self._slot_ptr(1), dt_out, None)
        if region is None:
            call(f"eb_step_fused_{self.suf}", *args)
        else:
            call(f"eb_step_fused_region_{self.suf}", *args, region)

    # ---------------------------------------------------------------- forward
    def forward(self, prim0: ops.EulerPrimArrays, frozen_dts: Optional[torch.Tensor] = None,
                keep: bool = True) -> ops.EulerPrimArrays:
        """Runs the rollout from P_0; returns P_N (new tensors). With frozen_dts
        (raw time steps of an earlier rollout, one per step) those time steps are
        used instead of the CFL ones, which is the map the adjoint differentiates.
        keep=False stores neither checkpoints nor states (no backward follows)."""
        self.suf = ops._validate(prim0)
        if min(ops.extents(prim0)) < 3:
            raise ValueError("a rollout needs at least one interior cell")
        self._geom_args = None
        self._ptrs = {}
        self._cfl_arg = ops._real(self.suf, self.cfl)
        dtype, dev = prim0.d.dtype, prim0.d.device
        self.prim0 = _clone_prim(prim0)
        self._slots = torch.zeros(2, dtype=dtype, device=dev)
        self._sets = [[], []]
        cons = _empty_cons_like(prim0)
        ops.prim_to_cons(self.ex, self.prim0, cons, self.eos)
        if frozen_dts is not None:
            if frozen_dts.numel() != self.nsteps or frozen_dts.dtype != dtype:
                raise ValueError("frozen_dts must hold one raw time step per step in the arrays' real type")
            dt = frozen_dts.to(dev).clone()
        else:
            dt = torch.empty(self.nsteps + 1, dtype=dtype, device=dev)
            ops.time_step_async(self.ex, self.prim0, self.eos, self.mesh, out=dt[0:1])
        self.dts = dt[:self.nsteps]
        dt_base, esz = dt.data_ptr(), dt.element_size()
        # The last two segments' states (P_n, kept <= n < nsteps) are written straight
        # into the backward's two buffer sets, so the backward recomputes neither (and
        # needs no checkpoint for them).
        self.tail = ((self.nsteps - 1) // self.k) * self.k if keep else self.nsteps
        self.kept = max(self.tail - self.k, 0) if keep else self.nsteps
        self._ensure_segment(cons, self.nsteps - self.tail, valid=False, bufs=self._sets[0])
        self._ensure_segment(cons, self.tail - self.kept, valid=False, bufs=self._sets[1])
        buf = [_clone_prim(prim0), _empty_prim_like(prim0)]
        self._ghost_refresh(cons, buf[1])
        self.checkpoints = {}
        cur = buf[0]
        for n in range(self.nsteps):
            if keep and n % self.k == 0 and n < self.kept:
                self.checkpoints[n] = _clone_cons(cons)
            p = n + 1  # the step writes P_p
            if self.tail <= p < self.nsteps:
                nxt = self._sets[0][p - self.tail]
            elif self.kept <= p < self.tail:
                nxt = self._sets[1][p - self.kept]
            else:
                nxt = buf[1] if cur is buf[0] else buf[0]
            out = self._slot_ptr(0) if frozen_dts is not None else dt_base + (n + 1) * esz
            self._fused(cur, cons, nxt, dt_base + n * esz, out)
            if n == 0:
                # later steps read ghost primitives round-tripped through cons_to_prim
                self._ghost_refresh(cons, buf[0])
            cur = nxt
        self.prim_final = cur
        self._tail_ready = keep
        return _clone_prim(self.prim_final)

    def _ensure_segment(self, cons: ops.EulerConsArrays, need: int, valid: bool, bufs: list):
        """Grows a set of segment buffers to `need`.
        valid=True fills new buffers with a finite state (rho = p = 1, u = 0): a
        box-restricted recompute leaves the cells outside its box as they were, and
        the adjoint steps read them with a zero adjoint weight, which needs finite
        values (fills write each field once, where a copy of P_0 would also read it).
        The forward pass writes every interior cell of the buffers it asks for, so
        it passes valid=False."""
        while len(bufs) < need:
            b = _empty_prim_like(cons)
            if valid:
                b.d.fill_(1.0)
                b.p.fill_(1.0)
                for u in (b.ux0, b.ux1, b.ux2):
                    u.zero_()
            self._ghost_refresh(cons, b, with_dt=False)  # ghost conservatives are frozen: set once
            bufs.append(b)

    # --------------------------------------------------------------- backward
    def _recompute(self, start: int, stop: int, bufs: list, needed=None,
                   ex: Optional[ops.ExecSpace] = None) -> List[ops.EulerPrimArrays]:
        """P_n for n in [start, stop), recomputed from the checkpoint at `start`
        with the recorded time steps (bitwise the forward's states). The backward
        visits each segment once, so the checkpoint itself becomes the working state
        (and its memory is released with the segment).

        needed(n) -> region (six int64, inclusive interior box) or None: P_n is
        only required on that box. The boxes must shrink by at least one cell per
        step (needed(n+1) grown by one cell lies inside needed(n)); P_n is then
        exact on needed(n) and holds some earlier valid state elsewhere.

        bufs: the buffer set to fill (k valid buffers); ex: the execution space
        (stream) to run on."""
        ex = self.ex if ex is None else ex
        cons = self.checkpoints.pop(start)
        if start == 0:
            prims = [self.prim0]
        else:
            ops.cons_to_prim(ex, cons, bufs[0], self.eos)
            prims = [bufs[0]]
        for n in range(start, stop - 1):
            nxt = bufs[n - start + 1]
            if needed is None:
                self._fused(prims[-1], cons, nxt, self._dt_ptr(n), self._slot_ptr(0), None, ex)
            else:
                # the recorded time steps are replayed: the box step computes none
                self._fused(prims[-1], cons, nxt, self._dt_ptr(n), None, needed(n + 1), ex)
            prims.append(nxt)
        return prims

    def backward(self, gprim_final: Sequence[torch.Tensor]) -> ops.EulerPrimArrays:
        """dL/dP_0 from dL/dP_N (five tensors in the primitive field order). Consumes
        the checkpoints: one backward per forward."""
        if self.prim_final is None:
            raise RuntimeError("backward before forward")
        if not self._tail_ready or len(self.checkpoints) != self.kept // self.k:
            raise RuntimeError("the checkpoints of this rollout were already consumed by a backward pass")
        self._tail_ready = False
        suf = self.suf
        r = lambda x: ops._real(suf, x)  # noqa: E731
        geom = self._geom()
        n0, n1, n2, bits = geom[:4]
        h = self.ex.handle

        gP = ops.EulerPrimArrays(*(g.contiguous() for g in gprim_final), index_bits=bits)
        # The adjoint spreads by at most one cell per step from the support of the
        # incoming gradient: each step only processes that box (outputs elsewhere stay
        # exactly zero, hence the zero-filled buffers).
        nz = torch.stack([g != 0 for g in gP.tensors()]).any(dim=0)
        idx = nz[1:-1, 1:-1, 1:-1].nonzero()
        box = None
        if idx.numel() > 0:
            lo = (idx.min(dim=0).values + 1).tolist()  # (k, j, i), array coordinates
            hi = (idx.max(dim=0).values + 1).tolist()
            box = [lo[2], hi[2], lo[1], hi[1], lo[0], hi[0]]
        mu = _zeros_cons_like(self.prim0)
        mu_next = _zeros_cons_like(self.prim0)
        call(f"eb_cons_to_prim_vjp_{suf}", h, self._p(self.prim_final), ctypes.byref(ops.prim_ptrs(gP)),
             self._p(mu), n0, n1, n2, bits, r(self.eos.gamma))
        gprim0 = _zeros_prim_like(self.prim0)
        step = 0

        def grown(g):
            if box is None:
                return (ctypes.c_int64 * 6)(1, 0, 1, 0, 1, 0)  # empty: ghost cells only
            return (ctypes.c_int64 * 6)(box[0] - g, box[1] + g, box[2] - g, box[3] + g, box[4] - g, box[5] + g)

        def region():
            # the adjoint of P_n is nonzero at most on the box grown by N - n cells
            return grown(step)

        def needed(n):
            # the adjoint step of P_n reads P_n on its output box grown by one cell
            return grown(self.nsteps - n + 1)
        # Segments last to first. The next segment's states are recomputed on a side
        # stream into the other of two buffer sets while this segment's adjoint
        # steps run on the main stream; events order each set's reuse.
        # Segment i uses buffer set i % 2: the last segment the forward's set 0, the
        # one before it set 1 (both written by the forward), the others recomputed.
        stored = {self.tail: 0}
        if self.kept < self.tail:
            stored[self.kept] = 1
        segs = list(reversed(sorted(self.checkpoints) + sorted(stored)))
        sets = self._sets
        side = _side_exec(self.ex) if self.checkpoints else None
        if self.checkpoints:
            # ghost conservatives are frozen, so any state gives the buffers' ghosts
            any_cons = self.checkpoints[max(self.checkpoints)]
            self._ensure_segment(any_cons, self.k, valid=True, bufs=sets[0])
            self._ensure_segment(any_cons, self.k, valid=True, bufs=sets[1])
        main = self.ex.stream
        if side is not None:
            # the side stream starts after everything issued so far on the main
            # stream (the forward's checkpoints, the buffers set up above)
            ready = torch.cuda.Event()
            ready.record(main)
            side.stream.wait_event(ready)
        released = [None, None]  # event after the last adjoint step that read set j
        keep = []  # states used on the side stream stay referenced until the end

        def recompute_on_side(i):
            start = segs[i]
            stop = min(start + self.k, self.nsteps)
            if released[i % 2] is not None:
                side.stream.wait_event(released[i % 2])
            keep.append(self.checkpoints[start])
            prims = self._recompute(start, stop, sets[i % 2], needed, side)
            done = torch.cuda.Event()
            done.record(side.stream)
            return prims, done

        pending = None
        for i, start in enumerate(segs):
            stop = min(start + self.k, self.nsteps)
            if start in stored:
                # written by the forward pass
                s = sets[stored[start]]
                head = self.prim0 if start == 0 else s[0]
                prims = [head] + s[1:stop - start]
            else:
                prims, done = pending
                main.wait_event(done)
            pending = recompute_on_side(i + 1) if i + 1 < len(segs) and segs[i + 1] not in stored else None
            for n in range(stop - 1, start - 1, -1):
                P = prims[n - start]
                dt = self._dt_ptr(n)
                step += 1
                if n == 0:
                    call(f"eb_godunov_vjp_first_{suf}", h, self._p(P), self._p(mu), self._p(gprim0), *geom, dt,
                         self._cfl_arg, region())
                else:
                    call(f"eb_godunov_vjp_{suf}", h, self._p(P), self._p(mu), self._p(mu_next), *geom, dt,
                         self._cfl_arg, region())
                    mu, mu_next = mu_next, mu
            released[i % 2] = torch.cuda.Event()
            released[i % 2].record(main)
        del keep
        # every side-stream launch has been waited for by the main stream
        self._sets = [[], []]
        return gprim0


_SIDE: dict = {}


def _side_exec(ex: ops.ExecSpace) -> ops.ExecSpace:
    """A second execution space (own stream at the default priority, own reduction
    scratch) on ex's device, kept for the process: the backward recomputes on it
    while the adjoint steps run on ex's stream."""
    key = (ex.device.index, ex.handle)
    side = _SIDE.get(key)
    if side is None:
        side = ops.ExecSpace(ex.device, torch.cuda.Stream(ex.device))
        _SIDE[key] = side
    return side


class _RolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rollout: Rollout, index_bits: int, frozen_dts, d, p, u0, u1, u2):
        P0 = ops.EulerPrimArrays(d.contiguous(), p.contiguous(), u0.contiguous(), u1.contiguous(), u2.contiguous(),
                                 index_bits=index_bits)
        # without a gradient to propagate, nothing is kept for a backward pass
        PN = rollout.forward(P0, frozen_dts, keep=any(ctx.needs_input_grad[3:]))
        ctx.rollout = rollout
        return PN.tensors()

    @staticmethod
    def backward(ctx, *grads):
        rollout: Rollout = ctx.rollout
        ref = rollout.prim0.d
        g = [torch.zeros_like(ref) if x is None else x for x in grads]
        g0 = rollout.backward(g)
        return (None, None, None) + tuple(g0.tensors())


def rollout(prim0: ops.EulerPrimArrays, eos: ops.PerfectGas, mesh: ops.UniformMesh3d, cfl: float, nsteps: int,
            checkpoint_every: int = 20, exec_space: Optional[ops.ExecSpace] = None,
            frozen_dts: Optional[torch.Tensor] = None, tape: Optional[list] = None):
    """Differentiable N-step rollout: returns the five final primitive fields as
    torch tensors whose gradient w.r.t. prim0's fields is the hand-written adjoint
    (dt frozen per step). prim0's tensors may require grad. frozen_dts replays the
    time steps of an earlier rollout; `tape`, if given, receives the Rollout
    (its .dts holds the raw time steps used)."""
    ex = exec_space if exec_space is not None else ops.ExecSpace(prim0.d.device)
    ro = Rollout(ex, eos, mesh, cfl, nsteps, checkpoint_every)
    if tape is not None:
        tape.append(ro)
    return _RolloutFunction.apply(ro, prim0.index_bits, frozen_dts, *prim0.tensors())
