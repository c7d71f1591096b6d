// Device-side point physics for the Godunov step: SoA pointer bundles, the
// perfect-gas relations, conservative<->primitive conversions and the HLLC
// Riemann flux. Expression trees keep the reference's operation order so that,
// compiled with contraction off, every result is the IEEE result of the
// reference's own arithmetic:
//   conversions   euler_arrays.hpp:38-75, perfect_gas.hpp:15-28
//   HLLC          hllc.hpp:52-104
// The HLLC here takes the sound speeds as inputs: a cell's sound speed is
// computed once (sqrt(gamma*p/rho), perfect_gas.hpp:17) and shared by every
// face it touches, which yields the same bits as recomputing it per call.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace eb {

template <class T>
struct PrimPtrs
{
    T* d;
    T* p;
    T* ux0;
    T* ux1;
    T* ux2;
};

template <class T>
struct ConsPtrs
{
    T* d;
    T* e;
    T* mx0;
    T* mx1;
    T* mx2;
};

template <class T>
struct Prim
{
    T d, p, u0, u1, u2;
};

template <class T>
struct Cons
{
    T d, e, m0, m1, m2;
};

template <class T>
struct Flux
{
    T d, e, m0, m1, m2;
};

// Next index of a grid-stride loop, or n once the stride would reach past n.
// idx + stride is formed only when it stays below n, so a 32-bit index cannot
// overflow however large the grid is (n itself fits the index type).
template <class I>
__device__ __forceinline__ I grid_next(I idx, I n, I stride)
{
    return n - idx > stride ? idx + stride : n;
}

// Ternary min/max with the std:: semantics Kokkos::min/max follow.
template <class T>
__device__ __forceinline__ T tmin(T a, T b)
{
    return (b < a) ? b : a;
}

template <class T>
__device__ __forceinline__ T tmax(T a, T b)
{
    return (a < b) ? b : a;
}

template <class T, class I>
__device__ __forceinline__ Prim<T> load_prim(PrimPtrs<T const> const& a, I idx)
{
    return {a.d[idx], a.p[idx], a.ux0[idx], a.ux1[idx], a.ux2[idx]};
}

template <class T, class I>
__device__ __forceinline__ Cons<T> load_cons(ConsPtrs<T const> const& a, I idx)
{
    return {a.d[idx], a.e[idx], a.mx0[idx], a.mx1[idx], a.mx2[idx]};
}

template <class T, class I>
__device__ __forceinline__ Cons<T> load_cons(ConsPtrs<T> const& a, I idx)
{
    return {a.d[idx], a.e[idx], a.mx0[idx], a.mx1[idx], a.mx2[idx]};
}

template <class T, class I>
__device__ __forceinline__ void store_prim(PrimPtrs<T> const& a, I idx, Prim<T> const& q)
{
    a.d[idx] = q.d;
    a.p[idx] = q.p;
    a.ux0[idx] = q.u0;
    a.ux1[idx] = q.u1;
    a.ux2[idx] = q.u2;
}

template <class T, class I>
__device__ __forceinline__ void store_cons(ConsPtrs<T> const& a, I idx, Cons<T> const& u)
{
    a.d[idx] = u.d;
    a.e[idx] = u.e;
    a.mx0[idx] = u.m0;
    a.mx1[idx] = u.m1;
    a.mx2[idx] = u.m2;
}

// rho*|u|^2/2 on primitives (euler_arrays.hpp:39-42).
template <class T>
__device__ __forceinline__ T ke_prim(Prim<T> const& q)
{
    return q.d * ((q.u0 * q.u0) + (q.u1 * q.u1) + (q.u2 * q.u2)) / T(2);
}

// |m|^2/rho/2 on conservatives (euler_arrays.hpp:45-48).
template <class T>
__device__ __forceinline__ T ke_cons(Cons<T> const& u)
{
    return ((u.m0 * u.m0) + (u.m1 * u.m1) + (u.m2 * u.m2)) / u.d / T(2);
}

// cons -> prim: p = (gamma-1)*(e - ke), u = m * (1/rho)
// (cons_to_prim.hpp:28-31, euler_arrays.hpp:50-75, perfect_gas.hpp:27).
template <class T>
__device__ __forceinline__ Prim<T> cons_to_prim_pt(Cons<T> const& u, T gm1)
{
    T const p = gm1 * (u.e - ke_cons(u));
    T const vol = T(1) / u.d;
    return {u.d, p, u.m0 * vol, u.m1 * vol, u.m2 * vol};
}

// prim -> cons: e = ke + p/(gamma-1), m = rho*u
// (prim_to_cons.hpp:28-31, euler_arrays.hpp:56-64, perfect_gas.hpp:22).
template <class T>
__device__ __forceinline__ Cons<T> prim_to_cons_pt(Prim<T> const& q, T gm1)
{
    T const eint = q.p / gm1;
    return {q.d, ke_prim(q) + eint, q.d * q.u0, q.d * q.u1, q.d * q.u2};
}

// sqrt(gamma*p/rho) (perfect_gas.hpp:17).
template <class T>
__device__ __forceinline__ T sound_speed(T gamma, T d, T p)
{
    return sqrt(gamma * p / d);
}

template <int Dir, class T>
__device__ __forceinline__ T normal(Prim<T> const& q)
{
    if constexpr (Dir == 0) {
        return q.u0;
    } else if constexpr (Dir == 1) {
        return q.u1;
    } else {
        return q.u2;
    }
}

// Total energy per volume p/(gamma-1) + rho|u|^2/2 in the order of hllc.hpp:82.
template <class T>
__device__ __forceinline__ T total_energy(Prim<T> const& q, T gm1)
{
    return q.p / gm1 + ke_prim(q);
}

// HLLC flux across a face normal to Dir, left state qL (sound speed cL, total
// energy eL), right state qR (cR, eR). Davis wave speeds, acoustic star state,
// upwind side chosen by u* > 0 (ties to the right), supersonic test
// S_L*S_R > 0; the energy star term divides by (S - u*). Follows hllc.hpp:61-103
// line by line; cL, cR, eL, eR are the per-cell values hllc.hpp:64-65,82 computes,
// passed in so that a cell computes them once for all of its faces.
template <int Dir, class T>
__device__ __forceinline__ Flux<T> hllc_flux(Prim<T> const& qL, T cL, T eL, Prim<T> const& qR, T cR, T eR)
{
    T const unL = normal<Dir>(qL);
    T const unR = normal<Dir>(qR);
    T const cmax = tmax(cL, cR);
    T const SL = tmin(unL, unR) - cmax;
    T const SR = tmax(unL, unR) + cmax;
    T const rcL = qL.d * (SL - unL);
    T const rcR = qR.d * (SR - unR);
    T const ustar = (qR.p - qL.p + rcL * unL - rcR * unR) / (rcL - rcR);
    T const pstar = T(0.5) * (qL.p + qR.p + rcL * (ustar - unL) + rcR * (ustar - unR));
    bool const left = ustar > T(0);
    T const S = left ? SL : SR;
    T const qd = left ? qL.d : qR.d;
    T const qp = left ? qL.p : qR.p;
    T const un = left ? unL : unR;
    T const etot = left ? eL : eR;
    bool const supersonic = SL * SR > T(0);
    T const un_o = supersonic ? un : ustar;
    T const ptot_o = supersonic ? qp : pstar;
    T const ratio = (S - un) / (S - un_o);
    T const d_o = ratio * qd;
    T const etot_o = (ratio * etot) + ((ptot_o * un_o - qp * un) / (S - ustar));
    T const mass = d_o * un_o;
    Flux<T> f;
    f.d = mass;
    f.e = (etot_o + ptot_o) * un_o;
    f.m0 = mass * (left ? qL.u0 : qR.u0);
    f.m1 = mass * (left ? qL.u1 : qR.u1);
    f.m2 = mass * (left ? qL.u2 : qR.u2);
    if constexpr (Dir == 0) {
        f.m0 = (mass * un_o) + ptot_o;
    } else if constexpr (Dir == 1) {
        f.m1 = (mass * un_o) + ptot_o;
    } else {
        f.m2 = (mass * un_o) + ptot_o;
    }
    return f;
}

// Same flux with the per-cell quantities computed here (reference-order kernel).
template <int Dir, class T>
__device__ __forceinline__ Flux<T> hllc_flux(Prim<T> const& qL, T cL, Prim<T> const& qR, T cR, T gm1)
{
    return hllc_flux<Dir>(qL, cL, total_energy(qL, gm1), qR, cR, total_energy(qR, gm1));
}

// flux += ds * (fR - fL), component-wise (godunov.hpp:58-62 and its y/z twins).
template <class T>
__device__ __forceinline__ void accumulate(Flux<T>& acc, T ds, Flux<T> const& fR, Flux<T> const& fL)
{
    acc.d += ds * (fR.d - fL.d);
    acc.e += ds * (fR.e - fL.e);
    acc.m0 += ds * (fR.m0 - fL.m0);
    acc.m1 += ds * (fR.m1 - fL.m1);
    acc.m2 += ds * (fR.m2 - fL.m2);
}

// cons -= dtodv * flux (godunov.hpp:84-90).
template <class T>
__device__ __forceinline__ void apply_update(Cons<T>& u, T dtodv, Flux<T> const& f)
{
    u.d -= dtodv * f.d;
    u.e -= dtodv * f.e;
    u.m0 -= dtodv * f.m0;
    u.m1 -= dtodv * f.m1;
    u.m2 -= dtodv * f.m2;
}

// 1/dt candidate of one cell (time_step.hpp:28-34).
template <class T>
__device__ __forceinline__ T inv_dt_cell(Prim<T> const& q, T c, T invdx0, T invdx1, T invdx2)
{
    T const cx0 = c + fabs(q.u0);
    T const cx1 = c + fabs(q.u1);
    T const cx2 = c + fabs(q.u2);
    return (cx0 * invdx0) + (cx1 * invdx1) + (cx2 * invdx2);
}

} // namespace eb
