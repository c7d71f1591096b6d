"""Hand-written adjoint (HIP VJP kernels) against torch autograd of a plain FP64
tensor implementation of the same scheme, and against central finite
differences, with the time step frozen per step."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import torch_reference as tr

pytestmark = pytest.mark.gpu

G = 1.4


def _faces(n, seed, umag=0.8):
    g = torch.Generator().manual_seed(seed)
    q = torch.empty(5, n, dtype=torch.float64)
    q[0] = 0.3 + 2.0 * torch.rand(n, generator=g)
    q[1] = 0.3 + 2.0 * torch.rand(n, generator=g)
    q[2:] = umag * torch.randn(3, n, generator=g)
    return q.cuda()


def _prim(eb, q):
    return eb.EulerPrimArrays(*[q[f].contiguous() for f in range(5)], index_bits=64)


def _call_hllc_vjp(eb, d, qL, qR, w):
    from benchmark_euler_kernels_amd._lib import call

    n = qL.shape[1]
    gL, gR = torch.empty_like(qL), torch.empty_like(qL)
    PL, PR = _prim(eb, qL), _prim(eb, qR)
    W = eb.EulerConsArrays(*[w[f].contiguous() for f in range(5)])
    GL, GR = _prim(eb, gL), _prim(eb, gR)
    ops = eb.operators
    call("eb_hllc_vjp_f64", torch.cuda.current_stream().cuda_stream, d, ctypes.byref(ops.prim_ptrs(PL)),
         ctypes.byref(ops.prim_ptrs(PR)), ctypes.byref(ops.cons_ptrs(W)), ctypes.byref(ops.prim_ptrs(GL)),
         ctypes.byref(ops.prim_ptrs(GR)), n, ctypes.c_double(G))
    torch.cuda.synchronize()
    return torch.stack(GL.tensors()), torch.stack(GR.tensors())


def _switch_margin(qL, qR, d):
    cL = torch.sqrt(G * qL[1] / qL[0])
    cR = torch.sqrt(G * qR[1] / qR[0])
    unL, unR = qL[2 + d], qR[2 + d]
    SL = torch.minimum(unL, unR) - torch.maximum(cL, cR)
    SR = torch.maximum(unL, unR) + torch.maximum(cL, cR)
    F = tr.hllc(list(qL), list(qR), d, G)  # noqa: F841 (forward for parity only)
    rcL, rcR = qL[0] * (SL - unL), qR[0] * (SR - unR)
    us = (qR[1] - qL[1] + rcL * unL - rcR * unR) / (rcL - rcR)
    return torch.stack([(cL - cR).abs(), (unL - unR).abs(), us.abs(), (SL * SR).abs()]).min(dim=0).values


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_vjp_matches_autograd(eb, ex, d):
    qL, qR = _faces(4000, 100 + d), _faces(4000, 200 + d)
    w = _faces(4000, 300 + d)
    gL, gR = _call_hllc_vjp(eb, d, qL, qR, w)
    aL, aR = qL.clone().requires_grad_(True), qR.clone().requires_grad_(True)
    F = tr.hllc(list(aL), list(aR), d, G)
    s = sum((f * w[i]).sum() for i, f in enumerate(F))
    rL, rR = torch.autograd.grad(s, (aL, aR))
    scale = max(rL.abs().max().item(), rR.abs().max().item())
    assert torch.allclose(gL, rL, rtol=1e-11, atol=1e-12 * scale)
    assert torch.allclose(gR, rR, rtol=1e-11, atol=1e-12 * scale)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_vjp_central_differences(eb, ex, d):
    """Directional derivatives of the HIP flux by central FD (eps = 1e-6) agree
    with the VJP away from branch switches; every branch is exercised."""
    from benchmark_euler_kernels_amd._lib import call

    n = 3000
    qL, qR = _faces(n, 400 + d, umag=1.5), _faces(n, 500 + d, umag=1.5)
    w = _faces(n, 600 + d)
    gL, gR = _call_hllc_vjp(eb, d, qL, qR, w)
    ops = eb.operators

    def flux(a, b):
        out = torch.empty_like(a)
        Fo = eb.EulerConsArrays(*[out[f] for f in range(5)])
        call("eb_hllc_flux_f64", torch.cuda.current_stream().cuda_stream, d, ctypes.byref(ops.prim_ptrs(_prim(eb, a))),
             ctypes.byref(ops.prim_ptrs(_prim(eb, b))), ctypes.byref(ops.cons_ptrs(Fo)), n, ctypes.c_double(G))
        torch.cuda.synchronize()
        return out

    g = torch.Generator().manual_seed(700 + d)
    vL = torch.randn(5, n, generator=g, dtype=torch.float64).cuda()
    vR = torch.randn(5, n, generator=g, dtype=torch.float64).cuda()
    eps = 1e-6
    fd = ((flux(qL + eps * vL, qR + eps * vR) - flux(qL - eps * vL, qR - eps * vR)) * w).sum(0) / (2 * eps)
    an = (gL * vL).sum(0) + (gR * vR).sum(0)
    keep = _switch_margin(qL, qR, d) > 1e-3
    assert keep.float().mean() > 0.9
    err = (fd - an).abs()[keep] / (an.abs()[keep] + 1e-3 * an.abs().max())
    assert err.max().item() < 1e-5
    sup = (qL[2 + d] - torch.sqrt(G * qL[1] / qL[0]) > 0) | (qR[2 + d] + torch.sqrt(G * qR[1] / qR[0]) < 0)
    assert sup.any() and (~sup).any()


def _smooth_state(shape, seed):
    n0, n1, n2 = shape
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    x, y, z = (x + 0.5) / n0, (y + 0.5) / n1, (z + 0.5) / n2
    a = rng.uniform(0.1, 0.3, 8)
    P = np.empty((5, n2, n1, n0))
    P[0] = 1.0 + a[0] * np.sin(2 * np.pi * (x + 2 * y)) + a[1] * np.cos(2 * np.pi * z)
    P[1] = 1.0 + a[2] * np.cos(2 * np.pi * (y - z)) + a[3] * np.sin(2 * np.pi * x)
    P[2] = a[4] * np.sin(2 * np.pi * (y + z))
    P[3] = a[5] * np.cos(2 * np.pi * (x - z))
    P[4] = a[6] * np.sin(2 * np.pi * (x + y + z))
    return P


def _octant_mask(shape, dx):
    n0, n1, n2 = shape
    c = [((np.arange(n) - 1) + 0.5) * h for n, h in zip((n0, n1, n2), dx)]
    m = (c[2][:, None, None] < 0.25) & (c[1][None, :, None] < 0.25) & (c[0][None, None, :] < 0.25)
    m &= np.ones((n2, n1, n0), bool)
    inner = np.zeros_like(m)
    inner[1:-1, 1:-1, 1:-1] = True
    return torch.from_numpy(m & inner).cuda()


def _hip_grad(eb, ex, P0, dx, nsteps, k, cfl=0.5, tape=None):
    ts = [torch.from_numpy(P0[f].copy()).cuda().requires_grad_(True) for f in range(5)]
    prim = eb.EulerPrimArrays(*ts, index_bits=32)
    out = eb.adjoint.rollout(prim, eb.PerfectGas(G), eb.UniformMesh3d(*dx), cfl, nsteps, k, exec_space=ex,
                             tape=tape)
    return ts, out


@pytest.mark.parametrize("k", [1, 4, 100])
def test_rollout_gradient_matches_torch_reference(eb, ex, k):
    shape = (18, 14, 12)
    dx = (1.0 / 16, 1.0 / 12, 1.0 / 10)
    nsteps, cfl = 6, 0.5
    P0 = _smooth_state(shape, 1)
    mask = _octant_mask(shape, dx)
    ts, out = _hip_grad(eb, ex, P0, dx, nsteps, k, cfl)
    loss = (out[0] * out[0] * mask).sum()
    loss.backward()
    ours = torch.stack([t.grad for t in ts])

    ref_in = [torch.from_numpy(P0[f].copy()).cuda().requires_grad_(True) for f in range(5)]
    PN = tr.rollout(ref_in, G, dx, cfl, nsteps)
    ref_loss = (PN[0] * PN[0] * mask).sum()
    ref_loss.backward()
    ref = torch.stack([t.grad for t in ref_in])
    assert abs(loss.item() - ref_loss.item()) <= 1e-13 * abs(ref_loss.item())
    err = (ours - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-11, err
    assert ref.abs().sum() > 0


def test_rollout_gradient_fp32(eb, ex):
    """The FP32 rollout and its adjoint (FP32 state, adjoint kernels and box
    recompute) against the FP64 tensor reference of the same scheme: the loss to
    FP32 rounding, the gradient to 1e-4 of its largest entry (normwise), and the
    gradient bitwise independent of the checkpoint interval."""
    shape = (18, 14, 12)
    dx = (1.0 / 16, 1.0 / 12, 1.0 / 10)
    nsteps, cfl = 6, 0.5
    P0 = _smooth_state(shape, 1)
    mask = _octant_mask(shape, dx)
    grads = []
    for k in (2, 100):
        ts = [torch.from_numpy(P0[f].copy()).to("cuda", torch.float32).requires_grad_(True) for f in range(5)]
        out = eb.adjoint.rollout(eb.EulerPrimArrays(*ts, index_bits=32), eb.PerfectGas(G), eb.UniformMesh3d(*dx),
                                 cfl, nsteps, k, exec_space=ex)
        loss = (out[0] * out[0] * mask).sum()
        loss.backward()
        grads.append(torch.stack([t.grad for t in ts]))
    assert torch.equal(grads[0], grads[1])
    ref_in = [torch.from_numpy(P0[f].copy()).cuda().requires_grad_(True) for f in range(5)]
    PN = tr.rollout(ref_in, G, dx, cfl, nsteps)
    ref_loss = (PN[0] * PN[0] * mask).sum()
    ref_loss.backward()
    ref = torch.stack([t.grad for t in ref_in])
    assert abs(loss.item() - ref_loss.item()) <= 1e-5 * abs(ref_loss.item())
    err = (grads[0].double() - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-4, err


def test_rollout_checkpoint_interval_is_bitwise_invariant(eb, ex):
    shape = (18, 14, 12)
    dx = (1.0 / 16, 1.0 / 12, 1.0 / 10)
    P0 = _smooth_state(shape, 2)
    mask = _octant_mask(shape, dx)
    grads = []
    for k in (1, 3, 7, 50):
        ts, out = _hip_grad(eb, ex, P0, dx, 7, k)
        (out[0] * out[0] * mask).sum().backward()
        grads.append(torch.stack([t.grad for t in ts]))
    for g in grads[1:]:
        assert torch.equal(g, grads[0])


def _monotone_state(shape, dx):
    """Smooth state whose normal velocity and sound speed increase along every
    axis and whose flow is subsonic with u* > 0, so that away from the frozen
    ghost layer no HLLC face sits near a branch switch."""
    n0, n1, n2 = shape
    z, y, x = np.meshgrid(*[(np.arange(n) - 0.5) * h for n, h in zip((n2, n1, n0), (dx[2], dx[1], dx[0]))],
                          indexing="ij")
    P = np.empty((5, n2, n1, n0))
    P[0] = 1.0 + 0.2 * np.sin(2 * np.pi * (x - y)) + 0.1 * np.cos(2 * np.pi * z)
    P[1] = P[0] * (1.0 + 0.3 * (x + y + z))
    P[2] = 0.4 + 0.3 * x + 0.05 * np.sin(2 * np.pi * (y + z))
    P[3] = 0.4 + 0.3 * y + 0.05 * np.cos(2 * np.pi * (x + z))
    P[4] = 0.4 + 0.3 * z + 0.05 * np.sin(2 * np.pi * (x + y))
    return P


def test_rollout_central_differences(eb, ex):
    """<dL/dP0, v> against (L(P0 + eps v) - L(P0 - eps v)) / (2 eps), eps = 1e-6,
    dt frozen at the unperturbed rollout's values, on 6 random directions over
    all five fields. The loss is sum rho^2 over a 6^3 box in the middle of the
    grid; the frozen ghost layer drives faces near it through min/max switches
    (|unL - unR| reaches 3e-7 there), so the test first checks, with the FP64
    tensor implementation, that every face inside the loss's domain of
    dependence (the box grown by one cell per remaining step) stays at least
    1e-3 away from each switch, i.e. FD never straddles a kink."""
    shape = (40, 36, 32)
    dx = (1.0 / 38, 1.0 / 34, 1.0 / 30)
    nsteps, cfl = 6, 0.5
    P0 = _monotone_state(shape, dx)
    centre = [s // 2 for s in (shape[2], shape[1], shape[0])]
    mask = torch.zeros(shape[2], shape[1], shape[0], dtype=torch.bool)
    mask[tuple(slice(c - 3, c + 3) for c in centre)] = True
    mask = mask.cuda()

    # premise: no switch within reach of the loss
    P = [torch.from_numpy(P0[f].copy()) for f in range(5)]
    U = tr.prim_to_cons(P, G)
    for n in range(nsteps):
        r = 3 + nsteps - n
        m = tr.switch_margins(P, G, [(c - r, c + r) for c in centre])
        assert min(m.values()) > 1e-3, (n, m)
        U = tr.godunov(P, U, G, dx, cfl * tr.time_step(P, G, dx))
        P = tr.cons_to_prim(U, G)

    tape = []
    ts, out = _hip_grad(eb, ex, P0, dx, nsteps, 4, cfl, tape=tape)
    (out[0] * out[0] * mask).sum().backward()
    g = torch.stack([t.grad for t in ts]).cpu().numpy()
    dts = tape[0].dts.clone()

    def loss(Pp):
        with torch.no_grad():
            prim = eb.EulerPrimArrays(*[torch.from_numpy(Pp[f].copy()).cuda() for f in range(5)])
            o = eb.adjoint.rollout(prim, eb.PerfectGas(G), eb.UniformMesh3d(*dx), cfl, nsteps, 4, exec_space=ex,
                                   frozen_dts=dts)
            return float((o[0] * o[0] * mask).sum().item())

    rng = np.random.default_rng(5678)
    scale = np.abs(P0).mean(axis=(1, 2, 3))[:, None, None, None]
    for _ in range(6):
        v = rng.normal(0.0, 1.0, P0.shape) * scale
        eps = 1e-6
        fd = (loss(P0 + eps * v) - loss(P0 - eps * v)) / (2 * eps)
        an = float((g * v).sum())
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_implosion_gradient_matches_torch_reference_32(eb, ex):
    """The measured adjoint configuration at 32^3: diagonal implosion with the
    density perturbed by N(0,0.01) * bump, 200 steps, checkpoints every 20; loss =
    sum rho^2 over the octant x, y, z < 0.25 at the final step; gradient w.r.t.
    all five initial fields against autograd of the FP64 tensor implementation."""
    n, nsteps, cfl = 32, 200, 0.5
    dx = (1.0 / n,) * 3
    P0 = eb.workloads.adjoint_state(ex, n, index_bits=32)
    c = (torch.arange(n + 2, dtype=torch.float64) - 0.5) / n
    inner = (c > 0) & (c < 0.25)
    mask = (inner[:, None, None] & inner[None, :, None] & inner[None, None, :]).cuda()

    ts = [t.clone().requires_grad_(True) for t in P0.tensors()]
    out = eb.adjoint.rollout(eb.EulerPrimArrays(*ts, index_bits=32), eb.PerfectGas(G), eb.UniformMesh3d(*dx), cfl,
                             nsteps, 20, exec_space=ex)
    loss = (out[0] * out[0] * mask).sum()
    loss.backward()
    ours = torch.stack([t.grad for t in ts])

    ref_in = [t.clone().requires_grad_(True) for t in P0.tensors()]
    PN = tr.rollout(ref_in, G, dx, cfl, nsteps)
    ref_loss = (PN[0] * PN[0] * mask).sum()
    ref_loss.backward()
    ref = torch.stack([t.grad for t in ref_in])
    assert abs(loss.item() - ref_loss.item()) <= 1e-12 * abs(ref_loss.item())
    err = (ours - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-10, err
    assert (ref[0].abs() > 0).sum() > 1000


@pytest.mark.parametrize("first", [False, True])
def test_vjp_region_is_exact_for_localised_adjoint(eb, ex, first):
    """An adjoint step restricted to the support of mu_in grown by one cell equals
    the full step bitwise (zero-filled outputs), for both step forms."""
    from benchmark_euler_kernels_amd._lib import call

    shape = (40, 30, 26)
    P0 = _smooth_state(shape, 9)
    prim = eb.EulerPrimArrays(*[torch.from_numpy(P0[f].copy()).cuda() for f in range(5)])
    g = torch.Generator().manual_seed(3)
    mu = torch.zeros(5, shape[2], shape[1], shape[0], dtype=torch.float64)
    mu[:, 8:13, 10:18, 5:21] = torch.randn(5, 5, 8, 16, generator=g, dtype=torch.float64)
    mu = mu.cuda()
    MU = eb.EulerConsArrays(*[mu[f] for f in range(5)])
    dt = torch.tensor([2e-3], dtype=torch.float64, device="cuda")
    ops = eb.operators
    name = "eb_godunov_vjp_first_f64" if first else "eb_godunov_vjp_f64"
    outs = []
    # support: i 5..20, j 10..17, k 8..12 (array coordinates); grown by one cell
    for region in (None, (ctypes.c_int64 * 6)(4, 21, 9, 18, 7, 13)):
        out = torch.zeros_like(mu)
        O = (eb.EulerPrimArrays if first else eb.EulerConsArrays)(*[out[f] for f in range(5)])
        ptrs = ops.prim_ptrs(O) if first else ops.cons_ptrs(O)
        call(name, torch.cuda.current_stream().cuda_stream, ctypes.byref(ops.prim_ptrs(prim)),
             ctypes.byref(ops.cons_ptrs(MU)), ctypes.byref(ptrs), *shape, 32, ctypes.c_double(G),
             ctypes.c_double(1 / 38), ctypes.c_double(1 / 28), ctypes.c_double(1 / 24), dt.data_ptr(),
             ctypes.c_double(0.5), region)
        torch.cuda.synchronize()
        outs.append(out)
    assert torch.equal(outs[0], outs[1])
    assert outs[0].abs().sum() > 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_step_fused_region_is_exact_on_its_box(eb, ex, dtype):
    """The box-restricted fused step (eb_step_fused_region_*, FP64 and FP32) writes
    exactly the box's cells, bitwise equal to the full step's, and leaves every other cell of
    the output primitives and of the conservatives as it was; its dt is the
    minimum over the box (not below the full step's). Without dt_out the box step
    computes no time step and writes the same cells, and the context's reduction
    scratch stays usable (the next full step's dt is unchanged)."""
    from benchmark_euler_kernels_amd._lib import call

    ops = eb.operators
    shape = (150, 21, 23)  # three 64-wide tiles in i, ragged in every axis
    dx = (1 / 148, 1 / 19, 1 / 21)
    P0 = _smooth_state(shape, 4)
    prim = eb.EulerPrimArrays(*[torch.from_numpy(P0[f].copy()).to("cuda", dtype) for f in range(5)])
    suf = "f64" if dtype == torch.float64 else "f32"
    real = ctypes.c_double if dtype == torch.float64 else ctypes.c_float
    cons0 = eb.EulerConsArrays(*[torch.empty_like(t) for t in prim.tensors()])
    ops.prim_to_cons(ex, prim, cons0, eb.PerfectGas(G))
    dt_in = torch.tensor([1e-3], dtype=dtype, device="cuda")
    big = torch.tensor([1e30], dtype=dtype, device="cuda")
    box = (60, 131, 3, 12, 5, 19)  # i_lo, i_hi, j_lo, j_hi, k_lo, k_hi
    sentinel = [torch.full_like(t, 7.0) for t in prim.tensors()]
    results = []
    for region, with_dt in ((None, True), ((ctypes.c_int64 * 6)(*box), True), ((ctypes.c_int64 * 6)(*box), False),
                            (None, True)):
        cons = eb.EulerConsArrays(*[t.clone() for t in cons0.tensors()])
        out = eb.EulerPrimArrays(*[t.clone() for t in sentinel])
        dt_out = torch.zeros(1, dtype=dtype, device="cuda")
        args = (ex.context, ex.handle, ctypes.byref(ops.prim_ptrs(prim)), ctypes.byref(ops.cons_ptrs(cons)),
                ctypes.byref(ops.prim_ptrs(out)), *shape, 32, real(G), *map(real, dx),
                real(0.5), dt_in.data_ptr(), big.data_ptr(), dt_out.data_ptr() if with_dt else None, None)
        if region is None:
            call(f"eb_step_fused_{suf}", *args)
        else:
            call(f"eb_step_fused_region_{suf}", *args, region)
        torch.cuda.synchronize()
        results.append((torch.stack(cons.tensors()), torch.stack(out.tensors()), dt_out.item()))
    (c_full, p_full, dt_full), (c_box, p_box, dt_box), (c_nodt, p_nodt, dt_nodt), (c_again, p_again, dt_again) = results
    assert torch.equal(c_nodt, c_box) and torch.equal(p_nodt, p_box)
    assert dt_nodt == 0.0  # not written
    assert torch.equal(c_again, c_full) and torch.equal(p_again, p_full) and dt_again == dt_full
    inside = torch.zeros(shape[2], shape[1], shape[0], dtype=torch.bool, device="cuda")
    inside[box[4]:box[5] + 1, box[2]:box[3] + 1, box[0]:box[1] + 1] = True
    assert torch.equal(p_box[:, inside], p_full[:, inside])
    assert torch.equal(c_box[:, inside], c_full[:, inside])
    assert torch.equal(p_box[:, ~inside], torch.stack(sentinel)[:, ~inside])
    assert torch.equal(c_box[:, ~inside], torch.stack(cons0.tensors())[:, ~inside])
    assert dt_box >= dt_full > 0


@pytest.mark.parametrize("nsteps,k", [(10, 4), (12, 5)])
def test_localised_rollout_gradient_is_bitwise_checkpoint_invariant(eb, ex, nsteps, k):
    """A loss over a small corner box of a 70 x 40 x 36 grid: the backward with
    box-restricted recomputes (checkpoints every k steps; the last segment shorter
    than k, so recompute buffers are added in the backward) equals the backward
    of the fully stored rollout bitwise."""
    shape = (72, 42, 38)
    dx = (1 / 70, 1 / 40, 1 / 36)
    P0 = _smooth_state(shape, 6)
    mask = torch.zeros(shape[2], shape[1], shape[0], dtype=torch.bool, device="cuda")
    mask[30:35, 2:6, 60:68] = True
    grads = []
    for kk in (nsteps, k):
        ts, out = _hip_grad(eb, ex, P0, dx, nsteps, kk)
        (out[0] * out[0] * mask).sum().backward()
        grads.append(torch.stack([t.grad for t in ts]))
    assert torch.equal(grads[0], grads[1])
    assert grads[0].abs().sum() > 0


def test_rollout_without_grad_keeps_nothing_and_matches(eb, ex):
    """Under no_grad the rollout stores no checkpoints or states, and its result
    is bitwise that of the differentiable rollout. With gradients, 9 steps and
    k = 4 keep P_8 and P_4..P_7 (the last two segments) and one checkpoint (step 0)."""
    shape = (18, 14, 12)
    dx = (1.0 / 16, 1.0 / 12, 1.0 / 10)
    P0 = _smooth_state(shape, 3)
    tape = []
    ts, out = _hip_grad(eb, ex, P0, dx, 9, 4, tape=tape)
    assert sorted(tape[0].checkpoints) == [0] and [len(b) for b in tape[0]._sets] == [1, 4]
    tape2 = []
    with torch.no_grad():
        prim = eb.EulerPrimArrays(*[torch.from_numpy(P0[f].copy()).cuda() for f in range(5)], index_bits=32)
        out2 = eb.adjoint.rollout(prim, eb.PerfectGas(G), eb.UniformMesh3d(*dx), 0.5, 9, 4, exec_space=ex,
                                  tape=tape2)
    assert not tape2[0].checkpoints and not any(tape2[0]._sets)
    for a, b in zip(out, out2):
        assert torch.equal(a.detach(), b)
    with pytest.raises(RuntimeError):
        tape2[0].backward([torch.zeros_like(t) for t in out2])


@pytest.mark.parametrize("box", [(5, 4, 1, 30, 1, 30),       # empty (lo > hi): no work
                                 (40, 60, 1, 5, 1, 5),       # outside the interior: no work
                                 (-10, 3, 18, 99, -5, 2)])   # clipped to the interior
def test_step_fused_region_edges(eb, ex, box):
    """Boxes that are empty or outside the interior launch nothing and leave every
    array as it was; a box reaching past the interior is clipped to it."""
    from benchmark_euler_kernels_amd._lib import call

    ops = eb.operators
    shape = (24, 20, 22)
    dx = (1 / 22, 1 / 18, 1 / 20)
    P0 = _smooth_state(shape, 8)
    prim = eb.EulerPrimArrays(*[torch.from_numpy(P0[f].copy()).cuda() for f in range(5)])
    cons0 = eb.EulerConsArrays(*[torch.empty_like(t) for t in prim.tensors()])
    ops.prim_to_cons(ex, prim, cons0, eb.PerfectGas(G))
    dt_in = torch.tensor([1e-3], dtype=torch.float64, device="cuda")
    big = torch.tensor([1e300], dtype=torch.float64, device="cuda")
    results = []
    for region in (None, (ctypes.c_int64 * 6)(*box)):
        cons = eb.EulerConsArrays(*[t.clone() for t in cons0.tensors()])
        out = eb.EulerPrimArrays(*[torch.full_like(t, 3.0) for t in prim.tensors()])
        dt_out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
        args = (ex.context, ex.handle, ctypes.byref(ops.prim_ptrs(prim)), ctypes.byref(ops.cons_ptrs(cons)),
                ctypes.byref(ops.prim_ptrs(out)), *shape, 64, ctypes.c_double(G), *map(ctypes.c_double, dx),
                ctypes.c_double(0.5), dt_in.data_ptr(), big.data_ptr(), dt_out.data_ptr(), None)
        if region is None:
            call("eb_step_fused_f64", *args)
        else:
            call("eb_step_fused_region_f64", *args, region)
        torch.cuda.synchronize()
        results.append((torch.stack(cons.tensors()), torch.stack(out.tensors()), dt_out.item()))
    (c_full, p_full, _), (c_box, p_box, dt_box) = results
    lo = [max(box[2 * d], 1) for d in range(3)]
    hi = [min(box[2 * d + 1], shape[d] - 2) for d in range(3)]
    inside = torch.zeros(shape[2], shape[1], shape[0], dtype=torch.bool, device="cuda")
    if all(lo[d] <= hi[d] for d in range(3)):
        inside[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
        assert torch.equal(p_box[:, inside], p_full[:, inside])
        assert torch.equal(c_box[:, inside], c_full[:, inside])
        assert dt_box > 0
    else:
        assert dt_box == -1.0  # nothing launched, dt_out untouched
    assert torch.equal(p_box[:, ~inside], torch.full_like(p_box[:, ~inside], 3.0))
    assert torch.equal(c_box[:, ~inside], torch.stack(cons0.tensors())[:, ~inside])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ghost_refresh_without_dt_writes_the_same_ghosts(eb, ex, dtype):
    """eb_ghost_refresh with a null ghost_min (the rollout's segment buffers) writes
    the same ghost primitives, bitwise, as with one, leaves the interior as it was,
    and leaves the ghost dt minimum of an earlier refresh untouched."""
    from benchmark_euler_kernels_amd._lib import call

    ops = eb.operators
    shape = (21, 17, 13)
    dx = (1 / 19, 1 / 15, 1 / 11)
    P0 = _smooth_state(shape, 9)
    prim = eb.EulerPrimArrays(*[torch.from_numpy(P0[f].copy()).to("cuda", dtype) for f in range(5)])
    cons = eb.EulerConsArrays(*[torch.empty_like(t) for t in prim.tensors()])
    ops.prim_to_cons(ex, prim, cons, eb.PerfectGas(G))
    suf = "f64" if dtype == torch.float64 else "f32"
    R = ctypes.c_double if dtype == torch.float64 else ctypes.c_float
    gmin = torch.full((1,), -1.0, dtype=dtype, device="cuda")
    outs = []
    for with_dt in (True, False):
        out = eb.EulerPrimArrays(*[torch.full_like(t, 7.0) for t in prim.tensors()])
        call(f"eb_ghost_refresh_{suf}", ex.context, ex.handle, ctypes.byref(ops.cons_ptrs(cons)),
             ctypes.byref(ops.prim_ptrs(out)), *shape, 32, R(G), *map(R, dx), gmin.data_ptr() if with_dt else None)
        torch.cuda.synchronize()
        outs.append(torch.stack(out.tensors()))
    assert torch.equal(outs[0], outs[1])
    ghost = torch.ones(shape[2], shape[1], shape[0], dtype=torch.bool, device="cuda")
    ghost[1:-1, 1:-1, 1:-1] = False
    assert torch.all(outs[1][:, ~ghost] == 7.0)
    # ghost primitives are cons_to_prim(prim_to_cons(P0)): P0 to rounding
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert torch.allclose(outs[1][:, ghost], torch.stack(prim.tensors())[:, ghost], rtol=tol, atol=tol)
    assert gmin.item() > 0
