// Adjoint kernels of one Godunov step with the time step frozen.
//
// Forward step n (euler_simulation.cpp:56-67): U_{n+1} = U_n - dtodv_n * R(P_n) on
// interior cells (ghosts unchanged), P_{n+1} = c2p(U_{n+1}) on all cells, where
// R(P)_c = sum_d ds_d (F_d(P_c, P_{c+d}) - F_d(P_{c-d}, P_c)) (godunov.hpp:48-90).
// With mu_{n+1} = dL/dU_{n+1}, the step's adjoint is
//   mu_n = mu_{n+1} + c2p^T( -dtodv_n * R^T(P_n) mu_{n+1}|interior ),
// and for the first step (P_0 given, U_0 = p2c(P_0)):
//   dL/dP_0 = -dtodv_0 * R^T(P_0) mu_1|interior + p2c^T mu_1.
// R^T is evaluated per cell by gathering the vector-Jacobian products of the six
// faces of the cell: a face between a (low side) and b (high side) carries the
// weight g = ds_d (mu_a [a interior] - mu_b [b interior]); a receives
// (dF/dqL)^T g, b receives (dF/dqR)^T g. Each thread owns one cell, so the result
// is race-free and deterministic.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "device_math.hpp"
#include "kernels.hpp"

namespace eb {

template <class T, class I>
__device__ __forceinline__ Flux<T> cons_as_flux(ConsPtrs<T const> const& a, I idx)
{
    return {a.d[idx], a.e[idx], a.mx0[idx], a.mx1[idx], a.mx2[idx]};
}

template <class T>
__device__ __forceinline__ Flux<T> face_weight(T ds, Flux<T> const& wa, bool ia, Flux<T> const& wb, bool ib)
{
    T const z = T(0);
    return {ds * ((ia ? wa.d : z) - (ib ? wb.d : z)), ds * ((ia ? wa.e : z) - (ib ? wb.e : z)),
            ds * ((ia ? wa.m0 : z) - (ib ? wb.m0 : z)), ds * ((ia ? wa.m1 : z) - (ib ? wb.m1 : z)),
            ds * ((ia ? wa.m2 : z) - (ib ? wb.m2 : z))};
}

template <class T>
__device__ __forceinline__ bool nonzero(Flux<T> const& g)
{
    return g.d != T(0) || g.e != T(0) || g.m0 != T(0) || g.m1 != T(0) || g.m2 != T(0);
}

// Adds R^T contributions of the two faces of cell c along Dir to gq.
template <int Dir, class T, class I>
__device__ __forceinline__ void gather_faces(
        PrimPtrs<T const> const& prim,
        ConsPtrs<T const> const& mu,
        I c,
        I stride,
        I pos,
        I n,
        bool int_c,
        bool int_lo,
        bool int_hi,
        Prim<T> const& qc,
        Flux<T> const& wc,
        T ds,
        T gamma,
        T gm1,
        PrimGrad<T>& gq)
{
    PrimGrad<T> dummy {T(0), T(0), T(0), T(0), T(0)};
    if (pos >= 1 && (int_lo || int_c)) {
        I const cl = c - stride;
        Flux<T> const g = face_weight(ds, cons_as_flux(mu, cl), int_lo, wc, int_c);
        if (nonzero(g)) {
            hllc_vjp<Dir>(load_prim(prim, cl), qc, gamma, gm1, g, dummy, gq);
        }
    }
    if (pos <= n - 2 && (int_c || int_hi)) {
        I const ch = c + stride;
        Flux<T> const g = face_weight(ds, wc, int_c, cons_as_flux(mu, ch), int_hi);
        if (nonzero(g)) {
            hllc_vjp<Dir>(qc, load_prim(prim, ch), gamma, gm1, g, gq, dummy);
        }
    }
}

// Mode 0: mu_out = mu_in + c2p^T(-dtodv * R^T mu_in)           (steps n >= 1)
// Mode 1: gprim  = -dtodv * R^T mu_in + p2c^T mu_in            (step 0)
template <class T, class I, int Mode>
__global__ __launch_bounds__(256) void k_godunov_vjp(
        PrimPtrs<T const> prim,
        ConsPtrs<T const> mu_in,
        ConsPtrs<T> mu_out,
        PrimPtrs<T> gprim,
        I n0,
        I n1,
        I n2,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        DtArg<T> dta)
{
    I const n = n0 * n1 * n2;
    I const stride = static_cast<I>(gridDim.x) * 256;
    T const dtodv = resolve_dtodv(dta);
    for (I c = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); c < n; c += stride) {
        I const i = c % n0;
        I const j = (c / n0) % n1;
        I const k = c / (n0 * n1);
        auto interior = [&](I a, I b, I e) {
            return a >= 1 && a <= n0 - 2 && b >= 1 && b <= n1 - 2 && e >= 1 && e <= n2 - 2;
        };
        bool const ic = interior(i, j, k);
        Prim<T> const qc = load_prim(prim, c);
        Flux<T> const wc = cons_as_flux(mu_in, c);
        PrimGrad<T> gq {T(0), T(0), T(0), T(0), T(0)};
        gather_faces<0>(prim, mu_in, c, I(1), i, n0, ic, interior(i - 1, j, k), interior(i + 1, j, k), qc, wc, ds0,
                        gamma, gm1, gq);
        gather_faces<1>(prim, mu_in, c, n0, j, n1, ic, interior(i, j - 1, k), interior(i, j + 1, k), qc, wc, ds1,
                        gamma, gm1, gq);
        gather_faces<2>(prim, mu_in, c, n0 * n1, k, n2, ic, interior(i, j, k - 1), interior(i, j, k + 1), qc, wc,
                        ds2, gamma, gm1, gq);
        PrimGrad<T> const gp {-dtodv * gq.d, -dtodv * gq.p, -dtodv * gq.u0, -dtodv * gq.u1, -dtodv * gq.u2};
        if constexpr (Mode == 0) {
            Cons<T> const r = c2p_vjp(qc, gp, gm1);
            mu_out.d[c] = wc.d + r.d;
            mu_out.e[c] = wc.e + r.e;
            mu_out.mx0[c] = wc.m0 + r.m0;
            mu_out.mx1[c] = wc.m1 + r.m1;
            mu_out.mx2[c] = wc.m2 + r.m2;
        } else {
            PrimGrad<T> const r = p2c_vjp(qc, Cons<T> {wc.d, wc.e, wc.m0, wc.m1, wc.m2}, gm1);
            gprim.d[c] = gp.d + r.d;
            gprim.p[c] = gp.p + r.p;
            gprim.ux0[c] = gp.u0 + r.u0;
            gprim.ux1[c] = gp.u1 + r.u1;
            gprim.ux2[c] = gp.u2 + r.u2;
        }
    }
}

// gcons = c2p^T(gprim) at the states with primitives `prim` (all cells).
template <class T, class I>
__global__ __launch_bounds__(256) void k_c2p_vjp(PrimPtrs<T const> prim, PrimPtrs<T const> gprim, ConsPtrs<T> gcons,
                                                 I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I c = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); c < n; c += stride) {
        Prim<T> const g = load_prim(gprim, c);
        Cons<T> const r = c2p_vjp(load_prim(prim, c), PrimGrad<T> {g.d, g.p, g.u0, g.u1, g.u2}, gm1);
        store_cons(gcons, c, r);
    }
}

// gprim = p2c^T(gcons) at primitives `prim` (all cells).
template <class T, class I>
__global__ __launch_bounds__(256) void k_p2c_vjp(PrimPtrs<T const> prim, ConsPtrs<T const> gcons, PrimPtrs<T> gprim,
                                                 I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I c = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); c < n; c += stride) {
        PrimGrad<T> const r = p2c_vjp(load_prim(prim, c), load_cons(gcons, c), gm1);
        store_prim(gprim, c, Prim<T> {r.d, r.p, r.u0, r.u1, r.u2});
    }
}

// Batched HLLC VJP (validation entry): face f has states qL[f], qR[f] (SoA, five
// arrays each) and weight g[f]; writes (dF/dqL)^T g and (dF/dqR)^T g.
template <class T, int Dir>
__global__ __launch_bounds__(256) void k_hllc_vjp(PrimPtrs<T const> qL, PrimPtrs<T const> qR, ConsPtrs<T const> g,
                                                  PrimPtrs<T> gL, PrimPtrs<T> gR, int64_t n, T gamma)
{
    int64_t const f = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (f >= n) {
        return;
    }
    PrimGrad<T> a {T(0), T(0), T(0), T(0), T(0)};
    PrimGrad<T> b {T(0), T(0), T(0), T(0), T(0)};
    Cons<T> const w = load_cons(g, f);
    hllc_vjp<Dir>(load_prim(qL, f), load_prim(qR, f), gamma, gamma - T(1), Flux<T> {w.d, w.e, w.m0, w.m1, w.m2}, a,
                  b);
    store_prim(gL, f, Prim<T> {a.d, a.p, a.u0, a.u1, a.u2});
    store_prim(gR, f, Prim<T> {b.d, b.p, b.u0, b.u1, b.u2});
}

// Batched HLLC flux (validation entry), same layout.
template <class T, int Dir>
__global__ __launch_bounds__(256) void k_hllc_flux(PrimPtrs<T const> qL, PrimPtrs<T const> qR, ConsPtrs<T> F,
                                                   int64_t n, T gamma)
{
    int64_t const f = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (f >= n) {
        return;
    }
    Prim<T> const a = load_prim(qL, f);
    Prim<T> const b = load_prim(qR, f);
    Flux<T> const r = hllc_flux<Dir>(a, sound_speed(gamma, a.d, a.p), b, sound_speed(gamma, b.d, b.p), gamma - T(1));
    store_cons(F, f, Cons<T> {r.d, r.e, r.m0, r.m1, r.m2});
}

} // namespace eb
