"""FP64 compensated conservation diagnostics: interior totals against exactly
rounded sums (math.fsum), the boundary outflow of one step against fluxes from
the oracle's HLLC on the boundary faces, and the telescoped balance
totals(U_n) + outflow = totals(U_0) over a tracked fused-step run (the ghost
layer is frozen, godunov.hpp:40-43)."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from conftest import random_prim
from gpu_util import to_cons, to_host, to_prim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(19, 7, 5), (3, 3, 3), (70, 33, 9)])
def test_conserved_totals_match_fsum(eb, ex, orc, dtype, shape):
    prim = random_prim(shape, seed=41, dtype=dtype)
    cons = orc.prim_to_cons(prim, 1.4)
    got = eb.conserved_totals(ex, to_cons(eb, cons, 64)).cpu().numpy()
    inner = cons[:, 1:-1, 1:-1, 1:-1].astype(np.float64)
    for f in range(5):
        ref = math.fsum(inner[f].ravel().tolist())
        scale = math.fsum(np.abs(inner[f]).ravel().tolist())
        assert abs(got[f] - ref) <= 2e-16 * max(scale, 1e-300), (f, got[f], ref)


def test_conserved_totals_empty_interior(eb, ex):
    c = eb.EulerConsArrays(*[torch.ones(2, 4, 4, dtype=torch.float64, device="cuda") for _ in range(5)])
    assert torch.equal(eb.conserved_totals(ex, c).cpu(), torch.zeros(5, dtype=torch.float64))


def _oracle_outflow(orc, prim, gamma, dx, dt):
    """dt/dv * sum_d ds_d (sum F(high boundary faces) - sum F(low boundary faces)),
    fluxes from the oracle's HLLC, face by face."""
    ds = (dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1])
    dv = dx[0] * dx[1] * dx[2]
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    total = np.zeros(5)
    # direction d <-> array axis 3 - d of (5, n2, n1, n0)
    for d in range(3):
        ax = 3 - d
        def face(lo, hi):
            sl_l = list(inner)
            sl_r = list(inner)
            sl_l[ax] = slice(lo, lo + 1)
            sl_r[ax] = slice(hi, hi + 1)
            qL = prim[tuple(sl_l)].reshape(5, -1)
            qR = prim[tuple(sl_r)].reshape(5, -1)
            return np.stack([orc.hllc(d, gamma, qL[:, m], qR[:, m]) for m in range(qL.shape[1])], axis=1)
        n = prim.shape[ax]
        lo = face(0, 1)
        hi = face(n - 2, n - 1)
        total += (dt / dv) * ds[d] * (hi.sum(axis=1) - lo.sum(axis=1))
    return total


def test_boundary_flux_matches_oracle(eb, ex, orc):
    shape, gamma = (21, 12, 9), 1.4
    dx = (1.0 / 19, 1.0 / 10, 1.0 / 7)
    prim = random_prim(shape, seed=42)
    cons = orc.prim_to_cons(prim, gamma)
    dt = 0.4 * orc.time_step(prim, gamma, *dx)
    acc = torch.zeros(10, dtype=torch.float64, device="cuda")
    dt_dev = torch.tensor([dt / 0.5], dtype=torch.float64, device="cuda")
    eb.boundary_flux(ex, to_prim(eb, prim), eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), dt_dev, 0.5, acc)
    got = (acc[0::2] + acc[1::2]).cpu().numpy()
    ref = _oracle_outflow(orc, prim, gamma, dx, 0.5 * (dt / 0.5))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max())
    # telescoped balance of one step: totals(U1) + B == totals(U0) up to the update's rounding
    C = to_cons(eb, cons)
    t0 = eb.conserved_totals(ex, C).cpu().numpy()
    eb.godunov(ex, to_prim(eb, prim), C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), eb.hllc(), dt_dev, cfl=0.5)
    t1 = eb.conserved_totals(ex, C).cpu().numpy()
    scale = np.abs(cons[:, 1:-1, 1:-1, 1:-1]).reshape(5, -1).sum(axis=1)
    assert np.all(np.abs(t1 + got - t0) <= 1e-14 * scale)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_tracked_run_balances_and_matches_untracked(eb, ex, orc, dtype, tol):
    shape, gamma, cfl, nsteps = (34, 30, 26), 1.4, 0.5, 20
    dx = tuple(1.0 / (n - 2) for n in shape)
    prim = orc.init_diagonal(*shape, *dx, offset=0.45, dtype=dtype)
    cons = orc.prim_to_cons(prim, gamma)
    runs = []
    for track, graph in ((False, False), (True, False), (True, True)):
        P, C = to_prim(eb, prim), to_cons(eb, cons)
        sim = eb.simulation.Simulation(ex, P, C, eb.PerfectGas(gamma), eb.UniformMesh3d(*dx), cfl,
                                       track_conservation=track)
        sim.run(nsteps, use_graph=graph)
        sim.finish()
        runs.append((to_host(C), to_host(P), sim.conservation() if track else None))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[1][0], runs[2][0])
    for res in (runs[1][2], runs[2][2]):
        scale = np.abs(cons[:, 1:-1, 1:-1, 1:-1].astype(np.float64)).reshape(5, -1).sum(axis=1)
        for f, (name, r) in enumerate(res.items()):
            # momenta start at zero: measure against the magnitudes that moved
            ref = max(scale[f], abs(r["total"]), abs(r["outflow"]))
            assert abs(r["residual"]) <= tol * ref, (name, r)
        assert abs(res["mass"]["outflow"]) > 0  # the diagonal interface meets the walls
    assert runs[1][2] == runs[2][2]  # graph replays accumulate the same bits
