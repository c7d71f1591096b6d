"""Fused step loop (godunov + cons_to_prim + next dt in one launch, dt on the
device) against the oracle's reference-order loop (euler_simulation.cpp:56-68).
Expected bit-identical: same expression trees, same operands, exact minimum."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import random_prim
from gpu_util import to_cons, to_host, to_prim

pytestmark = pytest.mark.gpu


def _setup(eb, ex, orc, kind, shape, dtype, index_bits=32):
    n0, n1, n2 = shape
    dx = (1.0 / (n0 - 2), 1.0 / (n1 - 2), 1.0 / (n2 - 2))
    if kind == "implode":
        prim = orc.init_implode(n0, n1, n2, *dx, dtype=dtype)
    elif kind == "diagonal":
        prim = orc.init_diagonal(n0, n1, n2, *dx, offset=0.45, dtype=dtype)
    else:
        prim = random_prim(shape, seed=31, dtype=dtype)
    cons = orc.prim_to_cons(prim, 1.4)
    return prim, cons, dx


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind,shape,nsteps", [
    ("implode", (18, 18, 18), 12),
    ("diagonal", (34, 20, 12), 9),
    ("random", (70, 11, 9), 5),
])
def test_fused_loop_matches_reference_loop(eb, ex, orc, dtype, kind, shape, nsteps):
    print(f"fused loop {kind} {shape} {dtype.__name__} steps={nsteps}", flush=True)
    gamma, cfl = 1.4, 0.49
    prim, cons, dx = _setup(eb, ex, orc, kind, shape, dtype)
    p_ref, c_ref = prim.copy(), cons.copy()
    dts = orc.simulate(p_ref, c_ref, gamma, *dx, cfl, nsteps)

    P, C = to_prim(eb, prim), to_cons(eb, cons)
    sim = eb.simulation.Simulation(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), cfl)
    for _ in range(nsteps):
        sim.step()
    sim.finish()
    assert np.array_equal(to_host(C), c_ref)
    assert np.array_equal(to_host(P), p_ref)
    # dt used by each step (the first is read after start()).
    sim2 = eb.simulation.Simulation(ex, to_prim(eb, prim), to_cons(eb, cons), eb.PerfectGas(gamma),
                                    eb.UniformMesh3d(*dx), cfl)
    sim2.start()
    seen = []
    for _ in range(nsteps):
        seen.append(sim2.last_dt().item())
        sim2.step()
    assert np.array_equal(np.array(seen, dtype=dtype), dts)
    t_expected = sum(float(np.asarray(cfl, dtype=dtype) * d) for d in dts)
    assert abs(sim2.time.item() - t_expected) <= 1e-15 * t_expected * nsteps


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fused_loop_negative_cell_width(eb, ex, orc, dtype):
    """A negative cell width in x: the fused step's time step then divides every
    cell's candidate out instead of folding the candidates into their largest
    denominator (the fold needs every 1/dx_d >= 0), and equals the reference loop
    bitwise as the folded form does in the other tests of this file."""
    gamma, cfl = 1.4, 0.49
    prim, cons, dx = _setup(eb, ex, orc, "random", (40, 13, 11), dtype)
    dx = (-1.0, dx[1], dx[2])  # finite, positive time steps for a few steps
    p_ref, c_ref = prim.copy(), cons.copy()
    dts = orc.simulate(p_ref, c_ref, gamma, *dx, cfl, 3)
    assert np.all(np.isfinite(p_ref)) and np.all(dts > 0)
    P, C = to_prim(eb, prim), to_cons(eb, cons)
    sim = eb.simulation.Simulation(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), cfl)
    for _ in range(3):
        sim.step()
    sim.finish()
    assert np.array_equal(to_host(C), c_ref)
    assert np.array_equal(to_host(P), p_ref)


def test_fused_loop_int64_and_tiles(eb, ex, orc):
    gamma, cfl = 1.4, 0.49
    prim, cons, dx = _setup(eb, ex, orc, "random", (67, 21, 10), np.float64)
    p_ref, c_ref = prim.copy(), cons.copy()
    orc.simulate(p_ref, c_ref, gamma, *dx, cfl, 4)
    for tile in ((0, 2, 3), (0, 4, 64), (1, 4, 3), (1, 16, 5), (2, 4, 3), (2, 8, 64), (3, 4, 3), (3, 8, 64), (0, 0, 7),
                 (4, 8, 3), (4, 16, 64), (1, 0, 0)):
        eb.set_godunov_config(*tile)
        try:
            P, C = to_prim(eb, prim, 64), to_cons(eb, cons, 64)
            sim = eb.simulation.Simulation(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), cfl)
            sim.run(4)
            sim.finish()
        finally:
            eb.set_godunov_config(4, 0, 0)
        assert np.array_equal(to_host(C), c_ref), tile
        assert np.array_equal(to_host(P), p_ref), tile


def test_graph_replays_match_eager(eb, ex, orc):
    """Pairs of fused steps captured in a hipGraph: every replay, with eager work
    and synchronisation in between, equals the eager path at the same step."""
    gamma, cfl = 1.4, 0.49
    prim, cons, dx = _setup(eb, ex, orc, "implode", (34, 34, 34), np.float64)
    eos, mesh = eb.PerfectGas(gamma), eb.UniformMesh3d(*dx)
    Pe, Ce = to_prim(eb, prim), to_cons(eb, cons)
    eager = eb.simulation.Simulation(ex, Pe, Ce, eos, mesh, cfl)
    Pg, Cg = to_prim(eb, prim), to_cons(eb, cons)
    graphed = eb.simulation.Simulation(ex, Pg, Cg, eos, mesh, cfl)
    eager.run(2)
    graphed.run(2)
    for rep in range(5):
        graphed.run(2, use_graph=True)
        eager.run(2)
        torch.cuda.synchronize()
        # eager torch work between replays
        _ = (Ce.d * 2.0).sum().item()
        assert np.array_equal(to_host(Cg), to_host(Ce)), f"replay {rep}"
        assert np.array_equal(to_host(graphed.current_prim()), to_host(eager.current_prim())), f"replay {rep}"
        assert graphed.time.item() == eager.time.item()
    assert graphed.graph is not None


def test_long_graph_replays_match_eager(eb, ex, orc):
    """Runs long enough for the multi-pair graph (2*CHUNK_PAIRS steps per replay,
    then pairs, then an odd last step): every run, with eager work and
    synchronisation in between, equals the eager loop at the same step."""
    gamma, cfl = 1.4, 0.49
    prim, cons, dx = _setup(eb, ex, orc, "implode", (26, 22, 18), np.float64)
    eos, mesh = eb.PerfectGas(gamma), eb.UniformMesh3d(*dx)
    Pe, Ce = to_prim(eb, prim), to_cons(eb, cons)
    eager = eb.simulation.Simulation(ex, Pe, Ce, eos, mesh, cfl)
    Pg, Cg = to_prim(eb, prim), to_cons(eb, cons)
    graphed = eb.simulation.Simulation(ex, Pg, Cg, eos, mesh, cfl)
    chunk = 2 * eb.simulation.Simulation.CHUNK_PAIRS
    for rep, n in enumerate((chunk + 5, chunk, 2 * chunk + 3)):
        graphed.run(n, use_graph=True)
        eager.run(n)
        torch.cuda.synchronize()
        _ = (Ce.d * 2.0).sum().item()
        assert np.array_equal(to_host(Cg), to_host(Ce)), f"run {rep}"
        assert np.array_equal(to_host(graphed.current_prim()), to_host(eager.current_prim())), f"run {rep}"
        assert graphed.time.item() == eager.time.item()
    assert graphed.chunk_graph is not None and graphed.graph is not None


def test_graph_run_after_odd_step_count(eb, ex, orc):
    """A graph-replayed run that starts after an odd number of steps (the captured
    pair starts from buffer 0, so one eager step realigns it) equals the eager loop."""
    gamma, cfl = 1.4, 0.49
    prim, cons, dx = _setup(eb, ex, orc, "implode", (20, 18, 22), np.float64)
    eos, mesh = eb.PerfectGas(gamma), eb.UniformMesh3d(*dx)
    Pe, Ce = to_prim(eb, prim), to_cons(eb, cons)
    eager = eb.simulation.Simulation(ex, Pe, Ce, eos, mesh, cfl)
    Pg, Cg = to_prim(eb, prim), to_cons(eb, cons)
    graphed = eb.simulation.Simulation(ex, Pg, Cg, eos, mesh, cfl)
    eager.run(14)
    graphed.run(3)
    graphed.run(4, use_graph=True)
    graphed.run(1)
    graphed.run(6, use_graph=True)
    assert graphed.steps_done == 14
    assert np.array_equal(to_host(Cg), to_host(Ce))
    assert np.array_equal(to_host(graphed.current_prim()), to_host(eager.current_prim()))
