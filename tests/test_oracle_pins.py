"""Pins of the CPU oracle (CPU only).

The reference ships no golden vectors and cannot be built here (it needs Kokkos,
which is absent), so each oracle function is pinned by checks that a plausible
mistake (dropped term, wrong sign, transposed operand, wrong operation order)
fails: hand-derived closed forms, brute force in exact rational arithmetic,
invariants of the HLLC flux, an independently written Toro-form HLLC, the
telescoped conservation balance, axis symmetry, and 1-D Riemann problems against
the exact solution (tests/riemann_exact.py).
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from conftest import random_prim
from riemann_exact import sample

G = 1.4


def state(*cells, dtype=np.float64):
    """(5, 1, 1, n) state from per-cell tuples."""
    a = np.array(cells, dtype=np.float64).T.reshape(5, 1, 1, len(cells))
    return np.ascontiguousarray(a.astype(dtype))


# ---------------------------------------------------------------- conversions


def test_prim_to_cons_closed_form(orc):
    # gamma = 1.5: p/(gamma-1) = 2p exactly; rho=2, p=0.5, u=(1,2,3): ke = 2*14/2 = 14.
    c = orc.prim_to_cons(state((2.0, 0.5, 1.0, 2.0, 3.0)), 1.5)
    assert c[:, 0, 0, 0].tolist() == [2.0, 15.0, 2.0, 4.0, 6.0]


def test_cons_to_prim_closed_form(orc):
    p = orc.cons_to_prim(state((2.0, 15.0, 2.0, 4.0, 6.0)), 1.5)
    assert p[:, 0, 0, 0].tolist() == [2.0, 0.5, 1.0, 2.0, 3.0]


def test_cons_to_prim_reference_benchmark_state(orc):
    """benchmark_cons_to_prim.cpp:29 state: rho=1, E=1/0.4, m=0 -> p = (gamma-1)*(1/0.4)."""
    p = orc.cons_to_prim(state((1.0, 1 / 0.4, 0.0, 0.0, 0.0)), G)
    assert p[1, 0, 0, 0] == (G - 1) * (1 / 0.4)
    assert abs(p[1, 0, 0, 0] - 1.0) < 4e-16


def test_velocity_is_momentum_times_reciprocal_density(orc):
    """euler_arrays.hpp:69-72: u = m * (1/rho), not m/rho. 5*(1/3) != 5/3 in FP64."""
    assert 5.0 * (1.0 / 3.0) != 5.0 / 3.0
    p = orc.cons_to_prim(state((3.0, 10.0, 5.0, -5.0, 0.0)), G)
    assert p[2, 0, 0, 0] == 5.0 * (1.0 / 3.0)
    assert p[3, 0, 0, 0] == -5.0 * (1.0 / 3.0)


def test_kinetic_energy_operation_order(orc):
    """ke(cons) = ((m0^2+m1^2)+m2^2)/rho/2: division by rho before the halving."""
    d, m = 3.0, (0.1, 0.2, 0.7)
    e = 1.0
    ke = ((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]) / d / 2
    p = orc.cons_to_prim(state((d, e) + m), G)
    assert p[1, 0, 0, 0] == (G - 1) * (e - ke)


def test_round_trip(orc):
    prim = random_prim((9, 7, 5), seed=1)
    back = orc.cons_to_prim(orc.prim_to_cons(prim, G), G)
    assert np.max(np.abs(back - prim) / np.abs(prim)) < 1e-14


# ------------------------------------------------------------------ time step


def test_time_step_uniform_closed_form(orc):
    prim = np.zeros((5, 4, 4, 4))
    prim[0] = prim[1] = 1.0
    c = np.sqrt(1.4)
    assert orc.time_step(prim, G, 1.0, 1.0, 1.0) == 1.0 / ((c + c) + c)
    assert abs(orc.time_step(prim, G, 1.0, 1.0, 1.0) - 0.28171808490950555) < 1e-16


def test_time_step_anisotropic_closed_form(orc):
    # rho = 1, p = 1/1.4 -> c = 1 up to rounding; u = (0.5, -1, 0),
    # dx = (0.5, 0.25, 0.125): invdt = 1.5/0.5 + 2/0.25 + 1/0.125 = 19 -> dt = 1/19.
    prim = np.zeros((5, 2, 2, 2))
    prim[0] = 1.0
    prim[1] = 1.0 / 1.4
    prim[2] = 0.5
    prim[3] = -1.0
    dt = orc.time_step(prim, G, 0.5, 0.25, 0.125)
    assert abs(dt - 1.0 / 19.0) < 1e-15
    # The minimum picks the fastest cell.
    prim[2, 1, 1, 1] = 3.0  # invdt = 4/0.5 + 8 + 8 = 24
    assert abs(orc.time_step(prim, G, 0.5, 0.25, 0.125) - 1.0 / 24.0) < 1e-15


def test_time_step_skips_nan_and_identity(orc):
    prim = random_prim((5, 4, 3), seed=2)
    ref = orc.time_step(prim, G, 0.1, 0.1, 0.1)
    prim[1, 1, 1, 1] = -1.0  # NaN candidate is never taken (Kokkos::Min join)
    assert orc.time_step(prim, G, 0.1, 0.1, 0.1) >= ref
    empty = np.zeros((5, 0, 3, 3))
    assert orc.time_step(empty, G, 0.1, 0.1, 0.1) == np.finfo(np.float64).max


# ----------------------------------------------------------------------- hllc


def euler_flux(q, d, g=G):
    """Physical Euler flux normal to direction d (textbook)."""
    rho, p, u = q[0], q[1], q[2:5]
    un = u[d]
    e = p / (g - 1) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    f = np.array([rho * un, (e + p) * un, rho * un * u[0], rho * un * u[1], rho * un * u[2]])
    f[2 + d] += p
    return f


def cons_vec(q, g=G):
    rho, p, u = q[0], q[1], q[2:5]
    e = p / (g - 1) + 0.5 * rho * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    return np.array([rho, e, rho * u[0], rho * u[1], rho * u[2]])


def toro_hllc(qL, qR, d, g=G):
    """HLLC in Toro's flux form F*_K = F_K + S_K (U*_K - U_K) with Davis speeds,
    S* from the standard contact-speed formula; written independently."""
    rl, pl, ul = qL[0], qL[1], qL[2 + d]
    rr, pr, ur = qR[0], qR[1], qR[2 + d]
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    sl = min(ul, ur) - max(cl, cr)
    sr = max(ul, ur) + max(cl, cr)
    ss = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))

    def ustar(q, s, un):
        rho, p = q[0], q[1]
        uk = cons_vec(q, g)
        fac = rho * (s - un) / (s - ss)
        vel = q[2:5].copy()
        vel[d] = ss
        e = fac * (uk[1] / rho + (ss - un) * (ss + p / (rho * (s - un))))
        return np.array([fac, e, fac * vel[0], fac * vel[1], fac * vel[2]])

    if sl > 0:
        return euler_flux(qL, d, g)
    if sr < 0:
        return euler_flux(qR, d, g)
    if ss > 0:
        return euler_flux(qL, d, g) + sl * (ustar(qL, sl, ul) - cons_vec(qL, g))
    return euler_flux(qR, d, g) + sr * (ustar(qR, sr, ur) - cons_vec(qR, g))


def _rand_q(rng, n, umag=0.5):
    q = np.empty((n, 5))
    q[:, 0] = rng.uniform(0.1, 3.0, n)
    q[:, 1] = rng.uniform(0.1, 3.0, n)
    q[:, 2:] = rng.normal(0.0, umag, (n, 3))
    return q


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_consistency(orc, d):
    """F(q, q) is the physical flux."""
    rng = np.random.default_rng(3 + d)
    for q in _rand_q(rng, 200, umag=1.5):
        f = orc.hllc(d, G, q, q)
        ref = euler_flux(q, d)
        assert np.allclose(f, ref, rtol=1e-13, atol=1e-13 * np.max(np.abs(ref)))


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_supersonic_upwinding(orc, d):
    rng = np.random.default_rng(7 + d)
    for qL, qR in zip(_rand_q(rng, 100), _rand_q(rng, 100)):
        for sign in (1.0, -1.0):
            a, b = qL.copy(), qR.copy()
            a[2 + d] = sign * (5.0 + abs(a[2 + d]))
            b[2 + d] = sign * (5.0 + abs(b[2 + d]))
            f = orc.hllc(d, G, a, b)
            ref = euler_flux(a if sign > 0 else b, d)
            assert np.allclose(f, ref, rtol=1e-14, atol=1e-14 * np.max(np.abs(ref)))


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_stationary_contact_exact(orc, d):
    """rho jump, u = 0, equal p: mass and energy flux 0, normal momentum flux p."""
    qL = np.array([1.0, 0.7, 0.0, 0.0, 0.0])
    qR = np.array([0.125, 0.7, 0.0, 0.0, 0.0])
    f = orc.hllc(d, G, qL, qR)
    ref = np.zeros(5)
    ref[2 + d] = 0.7
    assert f.tolist() == ref.tolist()


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_moving_contact_resolved(orc, d):
    """Equal normal velocity and pressure, different rho and transverse velocity:
    HLLC returns the upwind physical flux (an HLL flux would smear it)."""
    for un in (0.3, -0.3):
        qL = np.array([1.0, 0.5, 0.2, -0.4, 0.1])
        qR = np.array([0.2, 0.5, -0.7, 0.3, 0.9])
        qL[2 + d] = qR[2 + d] = un
        f = orc.hllc(d, G, qL, qR)
        ref = euler_flux(qL if un > 0 else qR, d)
        assert np.allclose(f, ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_mirror_symmetry(orc, d):
    rng = np.random.default_rng(11 + d)
    for qL, qR in zip(_rand_q(rng, 200), _rand_q(rng, 200)):
        f = orc.hllc(d, G, qL, qR)
        mL, mR = qR.copy(), qL.copy()
        mL[2 + d] *= -1
        mR[2 + d] *= -1
        g = orc.hllc(d, G, mL, mR)
        expect = -f
        expect[2 + d] = f[2 + d]
        assert np.allclose(g, expect, rtol=1e-12, atol=1e-13 * np.max(np.abs(f)))


@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_matches_toro_form(orc, d):
    """Same flux as the textbook HLLC form to round-off over random states,
    including strong jumps and every wave configuration."""
    rng = np.random.default_rng(17 + d)
    regions = set()
    for qL, qR in zip(_rand_q(rng, 1500, umag=1.2), _rand_q(rng, 1500, umag=1.2)):
        f = orc.hllc(d, G, qL, qR)
        t = toro_hllc(qL, qR, d)
        scale = max(np.max(np.abs(euler_flux(qL, d))), np.max(np.abs(euler_flux(qR, d))))
        assert np.allclose(f, t, rtol=1e-11, atol=1e-12 * scale), (qL, qR, f, t)
        cl, cr = np.sqrt(G * qL[1] / qL[0]), np.sqrt(G * qR[1] / qR[0])
        sl = min(qL[2 + d], qR[2 + d]) - max(cl, cr)
        sr = max(qL[2 + d], qR[2 + d]) + max(cl, cr)
        regions.add("L" if sl > 0 else "R" if sr < 0 else "star")
    assert regions == {"L", "R", "star"}


# -------------------------------------------------------------------- godunov


def test_godunov_uniform_state_is_identity(orc):
    prim = np.zeros((5, 6, 7, 8))
    prim[0], prim[1], prim[2], prim[3], prim[4] = 1.3, 0.8, 0.2, -0.1, 0.05
    cons = orc.prim_to_cons(prim, G)
    out = cons.copy()
    orc.godunov(prim, out, G, 0.1, 0.2, 0.3, 0.01)
    assert np.array_equal(out, cons)


def test_godunov_telescoped_conservation(orc):
    """sum_interior (U_new - U_old) dv = -dt * sum_d ds_d (sum of high boundary
    face fluxes - sum of low boundary face fluxes), anisotropic spacing."""
    n0, n1, n2 = 9, 7, 6
    dx = (0.11, 0.07, 0.13)
    prim = random_prim((n0, n1, n2), seed=5)
    cons = orc.prim_to_cons(prim, G)
    out = cons.copy()
    dt = 0.3 * orc.time_step(prim, G, *dx)
    orc.godunov(prim, out, G, *dx, dt)
    dv = dx[0] * dx[1] * dx[2]
    ds = (dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1])
    lhs = (out - cons)[:, 1:-1, 1:-1, 1:-1].sum(axis=(1, 2, 3)) * dv
    q = lambda i, j, k: prim[:, k, j, i]  # noqa: E731
    rhs = np.zeros(5)
    for k in range(1, n2 - 1):
        for j in range(1, n1 - 1):
            rhs += ds[0] * (orc.hllc(0, G, q(n0 - 2, j, k), q(n0 - 1, j, k)) - orc.hllc(0, G, q(0, j, k), q(1, j, k)))
    for k in range(1, n2 - 1):
        for i in range(1, n0 - 1):
            rhs += ds[1] * (orc.hllc(1, G, q(i, n1 - 2, k), q(i, n1 - 1, k)) - orc.hllc(1, G, q(i, 0, k), q(i, 1, k)))
    for j in range(1, n1 - 1):
        for i in range(1, n0 - 1):
            rhs += ds[2] * (orc.hllc(2, G, q(i, j, n2 - 2), q(i, j, n2 - 1)) - orc.hllc(2, G, q(i, j, 0), q(i, j, 1)))
    rhs *= -dt
    scale = np.abs((out - cons)).sum(axis=(1, 2, 3)) * dv
    assert np.all(np.abs(lhs - rhs) <= 1e-12 * scale + 1e-18)


def test_simulate_is_the_reference_loop(orc):
    """oracle.simulate against the pinned operators composed as in
    euler_simulation.cpp:56-68: the raw dt from the current primitives (returned
    without the CFL factor), godunov with cfl*dt, then cons_to_prim on all cells.
    A loop that took dt after the update, applied the CFL factor twice or not at
    all, converted before the update or touched the ghost conservatives differs."""
    n0, n1, n2 = 9, 7, 6
    dx = (0.11, 0.07, 0.13)
    cfl = 0.4
    nsteps = 4
    prim = random_prim((n0, n1, n2), seed=8)
    prim0 = prim.copy()
    cons = orc.prim_to_cons(prim, G)
    ghosts = np.ones((n2, n1, n0), dtype=bool)
    ghosts[1:-1, 1:-1, 1:-1] = False
    cons_ghosts0 = cons[:, ghosts].copy()
    p_ref, c_ref = prim.copy(), cons.copy()
    dts_ref = []
    for _ in range(nsteps):
        dt = orc.time_step(p_ref, G, *dx)
        dts_ref.append(dt)
        orc.godunov(p_ref, c_ref, G, *dx, cfl * dt)
        orc.cons_to_prim(c_ref, G, out=p_ref)
    dts = orc.simulate(prim, cons, G, *dx, cfl, nsteps)
    assert np.array_equal(dts, np.array(dts_ref))
    assert np.array_equal(prim, p_ref) and np.array_equal(cons, c_ref)
    assert np.array_equal(cons[:, ghosts], cons_ghosts0)  # frozen ghost layer
    assert not np.array_equal(prim[:, 1:-1, 1:-1, 1:-1], prim0[:, 1:-1, 1:-1, 1:-1])  # the interior evolves


@pytest.mark.parametrize("perm", [(1, 0, 2), (2, 1, 0), (0, 2, 1)])
def test_godunov_axis_permutation(orc, perm):
    """Relabelling the axes (fields and velocity components together) commutes
    with the update to round-off."""
    n = 8
    dx = (0.1, 0.1, 0.1)
    prim = random_prim((n, n, n), seed=6)
    cons = orc.prim_to_cons(prim, G)
    out = cons.copy()
    dt = 0.4 * orc.time_step(prim, G, *dx)
    orc.godunov(prim, out, G, *dx, dt)

    # spatial axes of a field array are (k, j, i) = (2, 1, 0) in array order
    ax = [2 - perm[2], 2 - perm[1], 2 - perm[0]]

    def permute(s):
        t = np.stack([np.transpose(s[f], ax) for f in range(5)])
        v = t[2:5].copy()
        t[2:5] = v[list(perm)]
        return np.ascontiguousarray(t)

    p2, c2 = permute(prim), permute(cons)
    o2 = c2.copy()
    orc.godunov(p2, o2, G, *dx, dt)
    inc, inc2 = permute(out) - permute(cons), o2 - c2
    assert np.allclose(inc2, inc, rtol=1e-10, atol=1e-13 * np.max(np.abs(inc)))


# ------------------------------------------------------------ Riemann problems


def run_1d(orc, axis, n, left, right, t_end, cfl=0.5):
    """1-D problem along `axis` through the 3-D oracle, zero-gradient ghosts
    refreshed every step so transverse face fluxes cancel exactly."""
    shape = [3, 3, 3]
    shape[axis] = n + 2
    n0, n1, n2 = shape
    dx = 1.0 / n
    prim = np.zeros((5, n2, n1, n0))
    x = ((np.arange(n + 2) - 1) + 0.5) * dx
    for f in range(5):
        line = np.where(x < 0.5, left[f], right[f])
        view = np.moveaxis(prim[f], 2 - axis, 0)
        view[...] = line[:, None, None]
    # left/right velocity is the normal one: move it to component `axis`
    vel = prim[2:5].copy()
    prim[2:5] = 0.0
    prim[2 + axis] = vel[0]
    cons = orc.prim_to_cons(prim, G)
    t = 0.0
    while t < t_end:
        prim = orc.cons_to_prim(cons, G)
        for f in range(5):
            v = np.moveaxis(prim[f], 2 - axis, 0)  # (n+2, a, b)
            v[0] = v[1]
            v[-1] = v[-2]
            interior = v[:, 1, 1].copy()
            v[...] = interior[:, None, None]
        dt = cfl * orc.time_step(prim, G, dx, dx, dx)
        dt = min(dt, t_end - t)
        orc.godunov(prim, cons, G, dx, dx, dx, dt)
        t += dt
    prim = orc.cons_to_prim(cons, G)
    rho = np.moveaxis(prim[0], 2 - axis, 0)[1:-1, 1, 1]
    u = np.moveaxis(prim[2 + axis], 2 - axis, 0)[1:-1, 1, 1]
    p = np.moveaxis(prim[1], 2 - axis, 0)[1:-1, 1, 1]
    return x[1:-1], rho, u, p


PROBLEMS = {
    # name: (left (rho, p, u), right, t_end)
    "sod": ((1.0, 1.0, 0.0), (0.125, 0.1, 0.0), 0.2),
    "toro1": ((1.0, 1.0, 0.75), (0.125, 0.1, 0.0), 0.2),
    "lax": ((0.445, 3.528, 0.698), (0.5, 0.571, 0.0), 0.13),
}


@pytest.mark.parametrize("name", sorted(PROBLEMS))
def test_riemann_problems_converge_to_exact(orc, name):
    left, right, t_end = PROBLEMS[name]
    L = (left[0], left[1], left[2], 0.0, 0.0)
    R = (right[0], right[1], right[2], 0.0, 0.0)
    errs = []
    for n in (100, 200, 400):
        x, rho, u, p = run_1d(orc, 0, n, L, R, t_end)
        ex = sample((x - 0.5) / t_end, left[0], left[2], left[1], right[0], right[2], right[1], G)
        errs.append(np.mean(np.abs(rho - ex[0])))
        assert np.mean(np.abs(p - ex[2])) < 0.06 * max(left[1], right[1])
    assert errs[0] < 0.08 and errs[2] < 0.03, errs
    assert errs[1] < 0.8 * errs[0] and errs[2] < 0.8 * errs[1], errs


def test_riemann_problem_identical_along_each_axis(orc):
    """Same 1-D problem along x, y and z. Only the CFL sum's operand order differs
    (time_step.hpp:33), so the results agree to round-off, not bitwise; a stride or
    component mix-up in one direction would break this."""
    L, R = (1.0, 1.0, 0.0, 0.0, 0.0), (0.125, 0.1, 0.0, 0.0, 0.0)
    res = [run_1d(orc, a, 64, L, R, 0.1) for a in (0, 1, 2)]
    for a in (1, 2):
        for f in (1, 2, 3):
            assert np.allclose(res[a][f], res[0][f], rtol=1e-12, atol=1e-13)


# ----------------------------------------------------------------------- init


def test_init_implode_brute_force(orc):
    """Cell-by-cell in exact rational arithmetic: inside iff r^2 < 1/16."""
    n = 12
    dx = 1.0 / n
    prim = orc.init_implode(n + 2, n + 2, n + 2, dx, dx, dx)
    inside = 0
    for k in range(n + 2):
        for j in range(n + 2):
            for i in range(n + 2):
                c = [Fraction(2 * (m - 1) + 1, 2 * n) - Fraction(1, 2) for m in (i, j, k)]
                r2 = sum(v * v for v in c)
                want = r2 < Fraction(1, 16)
                inside += want
                assert prim[0, k, j, i] == (1.0 if want else 0.125)
                assert prim[1, k, j, i] == (1.0 if want else 0.14)
    assert inside == int((prim[0] == 1.0).sum()) > 0
    assert not prim[2:].any()


def test_init_diagonal_brute_force(orc):
    n0, n1, n2 = 10, 8, 6
    dx = (1.0 / 8, 1.0 / 6, 1.0 / 4)
    prim = orc.init_diagonal(n0, n1, n2, *dx, offset=0.45)
    for k in range(n2):
        for j in range(n1):
            for i in range(n0):
                s = Fraction(2 * i - 1, 16) + Fraction(2 * j - 1, 12) + Fraction(2 * k - 1, 8)
                # exact sum is never within rounding distance of 0.45 on this grid
                assert abs(s - Fraction(45, 100)) > Fraction(1, 10**6)
                hi = s > Fraction(45, 100)
                assert prim[0, k, j, i] == (1.0 if hi else 0.125)
                assert prim[1, k, j, i] == (1.0 if hi else 0.14)


# ---------------------------------------------------------------------- FP32


def test_fp32_instantiation_tracks_fp64(orc):
    prim = random_prim((12, 10, 8), seed=8)
    cons = orc.prim_to_cons(prim, G)
    dx = (0.1, 0.1, 0.1)
    dt = 0.4 * orc.time_step(prim, G, *dx)
    out64 = cons.copy()
    orc.godunov(prim, out64, G, *dx, dt)
    p32, c32 = prim.astype(np.float32), cons.astype(np.float32)
    orc.godunov(p32, c32, G, *dx, dt)
    for f in range(5):  # normwise per field: momenta cross zero
        assert np.max(np.abs(c32[f] - out64[f])) < 2e-6 * np.max(np.abs(out64[f]))
