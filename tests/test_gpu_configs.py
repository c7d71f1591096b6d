"""The measured configurations at their stated sizes, against the CPU oracle.

  cfg0  diagonal implosion 64^3 (x+y+z > 0.45), FP64, CFL 0.5, 100 steps: the
        fused step loop (eager and graph-replayed, 32- and 64-bit indices)
        bitwise equal to the oracle's reference-order loop, time steps included.
  cfg1  per-operator kernels at 512^3, FP64 and FP32, int64 indices, seeded random
        state (seed 1234): prim_to_cons, cons_to_prim and time_step on 512^3
        arrays (the reference benchmarks' n^3 arrays), godunov and two fused steps
        on the 512^3 interior of 514^3 arrays with the default configuration (128-
        plane chunks, 1024 blocks, XCD-ordered tiles): every cell bitwise equal.
  cfg2  diagonal implosion 512^3, 1000 steps, FP64 and FP32: the first two steps
        bitwise equal to the oracle; at t_end the FP64 conservation residual, the
        FP32 error against FP64, simulated times and axis-permutation symmetry
        within the stated bounds.
  cfg3  the 256-instance ensemble (extents U{32..160} per axis, FP32, autotuned
        tiles, graph-replayed): three steps, every instance bitwise equal to the
        oracle's loop on its own, its time steps and simulated time included.
  cfg4  forward + adjoint of the 200-step FP64 rollout at 128^3: the adjoint's
        <grad, v> on 64 density directions against central differences, within
        the rounding floor of the 200-step loss, and between the one-sided ones.

Size-independent bars: bitwise equality wherever the oracle runs; the
conservation, FP32-error and FD bounds are stated in each test.
"""
from __future__ import annotations

import statistics

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GAMMA = 1.4
TORCH = {np.float64: torch.float64, np.float32: torch.float32}


def _host(arrays, shape, dtype) -> np.ndarray:
    """(5, n2, n1, n0) host copy of a device state of extents (n0, n1, n2)."""
    n0, n1, n2 = shape
    out = np.empty((5, n2, n1, n0), dtype=dtype)
    for f, t in enumerate(arrays.tensors()):
        out[f] = t.view(n2, n1, n0).cpu().numpy()
    return out


def _assert_bitwise(arrays, ref: np.ndarray, what: str):
    n2, n1, n0 = ref.shape[1:]
    for f, t in enumerate(arrays.tensors()):
        got = t.view(n2, n1, n0).cpu().numpy()
        bad = got != ref[f]
        count = int(np.count_nonzero(bad))
        if count:
            k, j, i = (int(v[0]) for v in np.nonzero(bad))
            pytest.fail(f"{what}: field {f}, {count} cells differ, first (i, j, k) = ({i}, {j}, {k}): "
                        f"{got[k, j, i]!r} vs {ref[f, k, j, i]!r}")


def _arrays(eb, ex, shape, dtype, index_bits=64):
    n = shape[0] * shape[1] * shape[2]
    prim = eb.to_mdspan(eb.create_prim_arrays_1d(ex, n, TORCH[dtype]), *shape, index_bits=index_bits)
    cons = eb.to_mdspan(eb.create_cons_arrays_1d(ex, n, TORCH[dtype]), *shape, index_bits=index_bits)
    return prim, cons


# ------------------------------------------------------------------ cfg0 ---

@pytest.mark.parametrize("index_bits", [32, 64])
@pytest.mark.parametrize("use_graph", [False, True])
def test_cfg0_diagonal_64_fp64_100_steps(eb, ex, orc, index_bits, use_graph):
    n, nsteps, cfl = 64, 100, 0.5
    na = n + 2
    shape = (na, na, na)
    dx = 1.0 / n
    print(f"cfg0: diagonal {n}^3 FP64 CFL {cfl} {nsteps} steps bits={index_bits} graph={use_graph}", flush=True)
    prim0 = orc.init_diagonal(na, na, na, dx, dx, dx, offset=0.45)
    cons0 = orc.prim_to_cons(prim0, GAMMA)
    p_ref, c_ref = prim0.copy(), cons0.copy()
    dts = orc.simulate(p_ref, c_ref, GAMMA, dx, dx, dx, cfl, nsteps)

    P, C = _arrays(eb, ex, shape, np.float64, index_bits)
    eos, mesh = eb.PerfectGas(GAMMA), eb.UniformMesh3d(dx, dx, dx)
    eb.init_diagonal(ex, P, mesh, offset=0.45)
    eb.prim_to_cons(ex, P, C, eos)
    _assert_bitwise(P, prim0, "init_diagonal")
    _assert_bitwise(C, cons0, "prim_to_cons")
    sim = eb.simulation.Simulation(ex, P, C, eos, mesh, cfl)
    seen = []
    if use_graph:
        sim.run(nsteps, use_graph=True)
    else:
        sim.start()
        for _ in range(nsteps):
            seen.append(sim.last_dt().item())
            sim.step()
    sim.finish()
    _assert_bitwise(C, c_ref, "cons after 100 steps")
    _assert_bitwise(P, p_ref, "prim after 100 steps")
    if seen:
        assert np.array_equal(np.array(seen), dts)
    t_expected = sum(float(cfl * d) for d in dts)
    assert abs(sim.time.item() - t_expected) <= 1e-15 * t_expected * nsteps


# ------------------------------------------------------------------ cfg1 ---

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cfg1_pointwise_512(eb, ex, orc, dtype):
    """prim_to_cons, cons_to_prim and time_step on 512^3 arrays, int64."""
    n = 512
    shape = (n, n, n)
    dx = 1.0 / n
    print(f"cfg1 pointwise {n}^3 {dtype.__name__} int64 seed 1234", flush=True)
    eos, mesh = eb.PerfectGas(GAMMA), eb.UniformMesh3d(dx, dx, dx)
    P, C = _arrays(eb, ex, shape, dtype)
    eb.init_random(ex, P, seed=1234)
    prim = _host(P, shape, dtype)
    eb.prim_to_cons(ex, P, C, eos)
    cons_ref = orc.prim_to_cons(prim, GAMMA)
    _assert_bitwise(C, cons_ref, "prim_to_cons 512^3")
    dt_ref = orc.time_step(prim, GAMMA, dx, dx, dx)
    assert eb.time_step(ex, P, eos, mesh) == dt_ref
    del prim
    P2 = eb.to_mdspan(eb.create_prim_arrays_1d(ex, n**3, TORCH[dtype]), *shape, index_bits=64)
    eb.cons_to_prim(ex, C, P2, eos)
    _assert_bitwise(P2, orc.cons_to_prim(cons_ref, GAMMA), "cons_to_prim 512^3")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cfg1_godunov_512(eb, ex, orc, dtype):
    """One godunov step on the 512^3 interior of 514^3 arrays, int64, default
    configuration, dt = 0.5 * time_step: every cell bitwise (the chunk seams at
    k = 128/129, 256/257, 384/385 and the tile edges included)."""
    n = 512
    na = n + 2
    shape = (na, na, na)
    dx = 1.0 / n
    assert eb.get_godunov_config() == (4, 0, 0)
    print(f"cfg1 godunov {n}^3 {dtype.__name__} int64 seed 1234", flush=True)
    eos, mesh = eb.PerfectGas(GAMMA), eb.UniformMesh3d(dx, dx, dx)
    P, C = _arrays(eb, ex, shape, dtype)
    eb.init_random(ex, P, seed=1234)
    eb.prim_to_cons(ex, P, C, eos)
    prim = _host(P, shape, dtype)
    cons = orc.prim_to_cons(prim, GAMMA)
    _assert_bitwise(C, cons, "prim_to_cons 514^3")
    dt_ref = orc.time_step(prim, GAMMA, dx, dx, dx)
    dt = eb.time_step(ex, P, eos, mesh)
    assert dt == dt_ref
    step = 0.5 * dt
    eb.godunov(ex, P, C, eos, mesh, eb.hllc(), step)
    orc.godunov(prim, cons, GAMMA, dx, dx, dx, float(np.asarray(step, dtype=dtype)))
    _assert_bitwise(C, cons, "godunov 512^3")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cfg1_fused_steps_512(eb, ex, orc, dtype):
    """Two fused steps (godunov + cons_to_prim + next time step in one launch) on
    the same 512^3 state against two steps of the oracle's reference-order loop."""
    n, cfl, nsteps = 512, 0.5, 2
    na = n + 2
    shape = (na, na, na)
    dx = 1.0 / n
    print(f"cfg1 fused {n}^3 {dtype.__name__} int64 seed 1234, {nsteps} steps", flush=True)
    eos, mesh = eb.PerfectGas(GAMMA), eb.UniformMesh3d(dx, dx, dx)
    P, C = _arrays(eb, ex, shape, dtype)
    eb.init_random(ex, P, seed=1234)
    eb.prim_to_cons(ex, P, C, eos)
    prim = _host(P, shape, dtype)
    cons = orc.prim_to_cons(prim, GAMMA)
    dts = orc.simulate(prim, cons, GAMMA, dx, dx, dx, cfl, nsteps)
    sim = eb.simulation.Simulation(ex, P, C, eos, mesh, cfl)
    sim.start()
    seen = []
    for _ in range(nsteps):
        seen.append(sim.last_dt().item())
        sim.step()
    seen.append(sim.last_dt().item())
    sim.finish()
    assert np.array_equal(np.array(seen[:nsteps], dtype=dtype), dts)
    _assert_bitwise(C, cons, "fused steps 512^3 (cons)")
    _assert_bitwise(P, prim, "fused steps 512^3 (prim)")
    # the dt the fused launch reduced for the next step
    assert seen[-1] == orc.time_step(prim, GAMMA, dx, dx, dx)


# ------------------------------------------------------------------ cfg2 ---

def test_cfg2_implosion_512_1000_steps(eb, ex, orc):
    """Diagonal implosion 512^3, CFL 0.5, 1000 steps, FP64 and FP32 (its own time
    steps, graph-replayed). Bars:
      * first two steps of each real type bitwise equal to the oracle at 512^3;
      * FP64 conservation residual (totals(U_n) + boundary outflow - totals(U_0),
        compensated FP64 sums) at most 1e-13 of the totals;
      * FP32 conservation residual at most 1e-6 of the totals;
      * FP32 t_end within 1e-6 relative of FP64's;
      * FP32 rho, p and u_k against FP64 at t_end: L1 at most 1e-6, Linf at most
        1e-4 (measured ~3e-8 and ~2.5e-6);
      * the FP64 state is symmetric under exchanging the x and z axes (the initial
        state is) to 1e-10: the scheme's x, y, z accumulation order breaks the
        symmetry only at round-off."""
    n, nsteps, cfl = 512, 1000, 0.5
    na = n + 2
    shape = (na, na, na)
    dx = 1.0 / n
    eos, mesh = eb.PerfectGas(GAMMA), eb.UniformMesh3d(dx, dx, dx)

    def state(dtype):
        P, C = _arrays(eb, ex, shape, dtype)
        eb.init_diagonal(ex, P, mesh, offset=0.45)
        eb.prim_to_cons(ex, P, C, eos)
        return P, C

    for dtype in (np.float64, np.float32):
        print(f"cfg2 first steps vs oracle {dtype.__name__}", flush=True)
        P, C = state(dtype)
        prim = _host(P, shape, dtype)
        assert np.array_equal(prim, orc.init_diagonal(na, na, na, dx, dx, dx, offset=0.45, dtype=dtype))
        cons = orc.prim_to_cons(prim, GAMMA)
        orc.simulate(prim, cons, GAMMA, dx, dx, dx, cfl, 2)
        sim = eb.simulation.Simulation(ex, P, C, eos, mesh, cfl)
        sim.run(2)
        sim.finish()
        _assert_bitwise(C, cons, f"cfg2 {dtype.__name__} step 2 (cons)")
        _assert_bitwise(P, prim, f"cfg2 {dtype.__name__} step 2 (prim)")
        del P, C, prim, cons, sim
        torch.cuda.empty_cache()

    runs = {}
    for dtype in (np.float64, np.float32):
        print(f"cfg2 {nsteps} steps {dtype.__name__}", flush=True)
        P, C = state(dtype)
        sim = eb.simulation.Simulation(ex, P, C, eos, mesh, cfl, track_conservation=True)
        sim.run(nsteps, use_graph=True)
        sim.finish()
        runs[dtype] = (P, sim.time.item(), sim.conservation())
        del C, sim
    P64, t64, cons64 = runs[np.float64]
    P32, t32, cons32 = runs[np.float32]
    print(f"t_end FP64 {t64!r} FP32 {t32!r}", flush=True)
    for name, r in cons64.items():
        scale = max(abs(r["total0"]), abs(r["total"]), abs(r["outflow"]))
        print(f"FP64 conservation {name}: residual {r['residual']:.3e} of {scale:.3e}", flush=True)
        assert abs(r["residual"]) <= 1e-13 * scale, (name, r)
    for name, r in cons32.items():
        scale = max(abs(r["total0"]), abs(r["total"]), abs(r["outflow"]))
        print(f"FP32 conservation {name}: residual {r['residual']:.3e} of {scale:.3e}", flush=True)
        assert abs(r["residual"]) <= 1e-6 * scale, (name, r)
    assert abs(t32 - t64) <= 1e-6 * t64
    assert 0.16 < t64 < 0.18  # 1000 steps of CFL 0.5 on this implosion: t_end = 0.16810
    inner = (slice(1, -1),) * 3
    for name, a, b in zip(("rho", "p", "u0", "u1", "u2"), P32.tensors(), P64.tensors()):
        a3 = a.view(na, na, na)[inner].double()
        b3 = b.view(na, na, na)[inner]
        d = (a3 - b3).abs()
        l1, linf = d.mean().item(), d.max().item()
        print(f"FP32 vs FP64 {name}: L1 {l1:.3e} Linf {linf:.3e}", flush=True)
        assert l1 <= 1e-6 and linf <= 1e-4, name
    rho = P64.d.view(na, na, na)
    asym = (rho - rho.permute(2, 1, 0)).abs().max().item()
    print(f"FP64 x<->z asymmetry of rho at t_end: {asym:.3e}", flush=True)
    assert asym <= 1e-10
    del runs, P64, P32
    torch.cuda.empty_cache()


# ------------------------------------------------------------------ cfg3 ---

def test_cfg3_ensemble_256_instances(eb, ex, orc):
    """256 uneven FP32 implosions (instance i from seed 1234+i: extents U{32..160}
    per axis, gamma U[1.2,1.67], offset U[0.3,0.6], inner rho, p U[0.05,0.3]),
    CFL 0.5, autotuned tiles, one eager step then a graph-replayed pair: every
    instance bitwise equal to three steps of the oracle's loop, and its dt and
    FP64 simulated time those of the oracle."""
    count, nsteps, cfl = 256, 3, 0.5
    prims, conss, eoss, meshes, cfls, params = eb.batch.random_ensemble(ex, count, seed=1234, dtype=torch.float32,
                                                                        index_bits=32, cfl=cfl)
    print(f"cfg3: {count} instances, {sum(np.prod(p['n']) for p in params)} interior cells", flush=True)
    shapes = [tuple(v + 2 for v in p["n"]) for p in params]
    init = [_host(p, s, np.float32) for p, s in zip(prims, shapes)]
    ens = eb.batch.Ensemble(ex, prims, conss, eoss, meshes, cfls)
    ens.start()
    ens.autotune(reps=1)
    print(f"tiles {ens.tiles()}", flush=True)
    ens.run(1)
    ens.run(2, use_graph=True)
    assert ens.steps_done == nsteps
    _, times = ens.state()
    times = times.cpu().numpy()
    ens.finish()
    for i, (s, p0, par) in enumerate(zip(shapes, init, params)):
        n0, n1, n2 = par["n"]
        dx = (1.0 / n0, 1.0 / n1, 1.0 / n2)
        prim = p0.copy()
        cons = orc.prim_to_cons(prim, par["gamma"])
        dts = orc.simulate(prim, cons, par["gamma"], *dx, cfl, nsteps)
        _assert_bitwise(conss[i], cons, f"instance {i} {s} cons")
        _assert_bitwise(prims[i], prim, f"instance {i} {s} prim")
        t_ref = sum(float(np.float32(cfl) * d) for d in dts)
        assert abs(times[i] - t_ref) <= 1e-12 * t_ref, (i, times[i], t_ref)
    ens.close()


# ------------------------------------------------------------------ cfg4 ---

def test_cfg4_adjoint_128_central_differences(eb, ex):
    """The 200-step FP64 rollout of the 128^3 diagonal implosion with rho +=
    N(0, 0.01) * bump (seed 1234), loss L = sum rho^2 over the octant
    x, y, z < 0.25, checkpoints every 20 steps, dt frozen at the rollout's values.
    For each of 64 density directions v_j = N(0,1) * bump (seeds 5678+j), eps 1e-6:
    fd = (L(+) - L(-)) / 2eps and the one-sided fd+ = (L(+) - L(0)) / eps,
    fd- = (L(0) - L(-)) / eps. The FD error is limited by the rounding of L (about
    1e3, after 200 steps of FP64 arithmetic) divided by eps: measured |fd - an|
    5e-5 to 1.2e-4 in absolute terms on every direction, i.e. 1e-6 relative where
    |an| ~ 50 and 1e-4 where |an| ~ 0.7. Bars, for every direction:
      * |fd - <grad, v>| <= 3e-6 |<grad, v>| + 2e-4;
      * <grad, v> lies between fd- and fd+ (a kink inside [-eps, eps] would show
        as fd+ != fd-, and the adjoint takes the unperturbed forward's branch, so
        its derivative is one of the one-sided ones) within the same allowance;
    and the median relative error is at most 2e-6."""
    n, nsteps, k, ndirs, eps = 128, 200, 20, 64, 1e-6
    eos = eb.PerfectGas(GAMMA)
    mesh = eb.UniformMesh3d(1.0 / n, 1.0 / n, 1.0 / n)
    P0 = eb.workloads.adjoint_state(ex, n)
    c = (torch.arange(n + 2, dtype=torch.float64, device=P0.d.device) - 0.5) / n
    inner = (c > 0) & (c < 0.25)
    mask = inner[:, None, None] & inner[None, :, None] & inner[None, None, :]
    ts = [t.clone().requires_grad_(True) for t in P0.tensors()]
    tape = []
    out = eb.adjoint.rollout(eb.EulerPrimArrays(*ts, index_bits=P0.index_bits), eos, mesh, 0.5, nsteps, k,
                             exec_space=ex, tape=tape)
    (out[0] * out[0] * mask).sum().backward()
    grad = ts[0].grad.detach()
    dts = tape[0].dts.clone()

    def loss(rho):
        with torch.no_grad():
            P = eb.EulerPrimArrays(rho, *(t.clone() for t in P0.tensors()[1:]), index_bits=P0.index_bits)
            o = eb.adjoint.rollout(P, eos, mesh, 0.5, nsteps, nsteps, exec_space=ex, frozen_dts=dts)
            return float((o[0] * o[0] * mask).sum().item())

    l0 = loss(P0.d.clone())
    print(f"cfg4: L(0) = {l0!r}", flush=True)
    print("   j  <grad,v>              |fd-an|       |fd-an|/|an|  |fd+ - fd-|/|an|", flush=True)
    errs = []
    for j in range(ndirs):
        v = eb.workloads.density_direction(n, j, P0.d.device)
        lp, lm = loss(P0.d + eps * v), loss(P0.d - eps * v)
        an = float((grad * v).sum())
        fd, fp, fm = (lp - lm) / (2 * eps), (lp - l0) / eps, (l0 - lm) / eps
        print(f"{j:4d}  {an:+.15e}  {abs(fd - an):.3e}     {abs(fd - an) / abs(an):.3e}     "
              f"{abs(fp - fm) / abs(an):.3e}", flush=True)
        allow = 3e-6 * abs(an) + 2e-4
        assert abs(fd - an) <= allow, (j, an, fd)
        assert min(fp, fm) - allow <= an <= max(fp, fm) + allow, (j, an, fp, fm)
        errs.append(abs(fd - an) / abs(an))
    print(f"cfg4 FD: median rel err {statistics.median(errs):.2e}, max {max(errs):.2e}", flush=True)
    assert statistics.median(errs) <= 2e-6
