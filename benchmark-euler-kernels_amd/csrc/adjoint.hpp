// Hand-written reverse-mode derivatives (vector-Jacobian products) of the point
// physics of the Godunov step, for the adjoint of a rollout with the time step
// frozen per step.
//
//   hllc_vjp      : (dF/dqL)^T g and (dF/dqR)^T g of the HLLC flux (hllc.hpp:53-104),
//                   through every branch: Davis min/max speed selection, upwind side
//                   (u* > 0), supersonic test (S_L*S_R > 0). At a switch the
//                   derivative of the branch the forward took is used.
//   c2p_vjp       : J^T g of cons -> prim (cons_to_prim.hpp:28-31) at a state given
//                   by its primitives.
//   p2c_vjp       : J^T g of prim -> cons (prim_to_cons.hpp:28-31).
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace eb {

template <class T>
struct PrimGrad
{
    T d, p, u0, u1, u2;
};

template <int Dir, class T>
__device__ __forceinline__ T& comp(PrimGrad<T>& g)
{
    if constexpr (Dir == 0) {
        return g.u0;
    } else if constexpr (Dir == 1) {
        return g.u1;
    } else {
        return g.u2;
    }
}

template <int Dir, class T>
__device__ __forceinline__ T flux_normal(Flux<T> const& f)
{
    if constexpr (Dir == 0) {
        return f.m0;
    } else if constexpr (Dir == 1) {
        return f.m1;
    } else {
        return f.m2;
    }
}

// Reciprocal for the derivative arithmetic, which need not round like the forward
// step (only the branch decisions must match, and those depend on signs and on
// comparisons of values computed identically on both sides): the hardware
// estimate refined by EB_VJP_NEWTON Newton steps. Each step squares the relative
// error; one step leaves about 1e-14, two about 1 ulp.
#ifndef EB_VJP_NEWTON
#define EB_VJP_NEWTON 1
#endif
__device__ __forceinline__ double fast_rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    for (int it = 0; it < EB_VJP_NEWTON; ++it) {
        r = fma(r, fma(-x, r, 1.0), r);
    }
    return r;
#else
    return 1.0 / x;
#endif
}

__device__ __forceinline__ float fast_rcp(float x)
{
    return 1.0f / x;
}

// Per-cell quantity the face derivatives share: the sound speed c.
template <class T>
struct VjpCell
{
    Prim<T> q;
    T c;
};

template <class T>
__device__ __forceinline__ VjpCell<T> vjp_cell(Prim<T> const& q, T gamma)
{
    return {q, sound_speed(gamma, q.d, q.p)};
}

// Adds (dF/dqL)^T g to gl and (dF/dqR)^T g to gr. Forward values are recomputed
// in the order of hllc.hpp:61-103, with the same branch decisions as the flux
// (the quotients are formed as products with reciprocals, which keeps their
// signs, so u* > 0 and S_L*S_R > 0 select the same branches).
template <int Dir, class T>
__device__ __forceinline__ void hllc_vjp(
        VjpCell<T> const& L,
        VjpCell<T> const& R,
        T gm1,
        Flux<T> const& g,
        PrimGrad<T>& gl,
        PrimGrad<T>& gr)
{
    Prim<T> const& qL = L.q;
    Prim<T> const& qR = R.q;
    // ---- forward
    T const unL = normal<Dir>(qL);
    T const unR = normal<Dir>(qR);
    T const cL = L.c;
    T const cR = R.c;
    bool const sel_c = cL < cR; // max(cL, cR) picks cR
    T const cmax = sel_c ? cR : cL;
    bool const sel_min = unR < unL; // min picks unR
    bool const sel_max = unL < unR; // max picks unR
    T const umin = sel_min ? unR : unL;
    T const umax = sel_max ? unR : unL;
    T const SL = umin - cmax;
    T const SR = umax + cmax;
    T const aL = SL - unL;
    T const aR = SR - unR;
    T const rcL = qL.d * aL;
    T const rcR = qR.d * aR;
    T const num = qR.p - qL.p + rcL * unL - rcR * unR;
    T const rden = fast_rcp(rcL - rcR);
    T const us = num * rden;
    T const bL = us - unL;
    T const bR = us - unR;
    T const ps = T(0.5) * (qL.p + qR.p + rcL * bL + rcR * bR);
    bool const left = us > T(0);
    Prim<T> const& q = left ? qL : qR;
    T const S = left ? SL : SR;
    T const un = left ? unL : unR;
    // 1/(gamma-1) is loop-invariant in the kernels, so this is hoisted
    T const igm1 = fast_rcp(gm1);
    T const etot = q.p * igm1 + ke_prim(q);
    bool const sup = SL * SR > T(0);
    T const uo = sup ? un : us;
    T const po = sup ? q.p : ps;
    T const A = S - un;
    T const rB = fast_rcp(S - uo);
    T const ratio = A * rB;
    T const d_o = ratio * q.d;
    T const C = po * uo - q.p * un;
    T const rD = fast_rcp(S - us);
    T const X = C * rD;
    T const eo = ratio * etot + X;
    T const mass = d_o * uo;

    // Reverse sweep. Its products may contract into FMAs (they need not round like
    // the forward step); the branch decisions above are taken without contraction.
    {
#pragma clang fp contract(fast)
        // ---- reverse: outputs F = (mass, (eo+po)*uo, mass*u_t, mass*uo + po)
        T const g_mn = flux_normal<Dir>(g);
        T g_mass = g.d + g_mn * uo;
        T g_uo = g.e * (eo + po) + g_mn * mass;
        T g_po = g.e * uo + g_mn;
        T const g_eo = g.e * uo;
        // transverse momentum fluxes mass * q.u_t
        PrimGrad<T> gq {T(0), T(0), T(0), T(0), T(0)};
        if constexpr (Dir != 0) {
            g_mass += g.m0 * q.u0;
            gq.u0 += g.m0 * mass;
        }
        if constexpr (Dir != 1) {
            g_mass += g.m1 * q.u1;
            gq.u1 += g.m1 * mass;
        }
        if constexpr (Dir != 2) {
            g_mass += g.m2 * q.u2;
            gq.u2 += g.m2 * mass;
        }
        // mass = d_o * uo
        T const g_do = g_mass * uo;
        g_uo += g_mass * d_o;
        // eo = ratio * etot + X
        T g_ratio = g_eo * etot;
        T const g_etot = g_eo * ratio;
        // X = C / D
        T const g_C = g_eo * rD;
        T const g_D = -g_C * X;
        // C = po*uo - q.p*un
        g_po += g_C * uo;
        g_uo += g_C * po;
        gq.p += -g_C * un;
        T g_un = -g_C * q.p;
        // D = S - us
        T g_S = g_D;
        T g_us = -g_D;
        // d_o = ratio * q.d
        g_ratio += g_do * q.d;
        gq.d += g_do * ratio;
        // ratio = A / B; A = S - un; B = S - uo
        T const g_A = g_ratio * rB;
        T const g_B = -g_A * ratio;
        g_S += g_A + g_B;
        g_un += -g_A;
        g_uo += -g_B;
        // uo, po selections
        T g_ps = T(0);
        if (sup) {
            g_un += g_uo;
            gq.p += g_po;
        } else {
            g_us += g_uo;
            g_ps = g_po;
        }
        // etot = p/gm1 + d*|u|^2/2 of the upwind state
        {
            T const u2 = (q.u0 * q.u0) + (q.u1 * q.u1) + (q.u2 * q.u2);
            gq.p += g_etot * igm1;
            gq.d += g_etot * u2 * T(0.5);
            gq.u0 += g_etot * q.d * q.u0;
            gq.u1 += g_etot * q.d * q.u1;
            gq.u2 += g_etot * q.d * q.u2;
        }
        comp<Dir>(gq) += g_un;
        // route the upwind state's and S's adjoints to the chosen side
        T g_SL = T(0);
        T g_SR = T(0);
        if (left) {
            g_SL += g_S;
            gl.d += gq.d;
            gl.p += gq.p;
            gl.u0 += gq.u0;
            gl.u1 += gq.u1;
            gl.u2 += gq.u2;
        } else {
            g_SR += g_S;
            gr.d += gq.d;
            gr.p += gq.p;
            gr.u0 += gq.u0;
            gr.u1 += gq.u1;
            gr.u2 += gq.u2;
        }
        // ps = 0.5*((pL + pR) + rcL*bL + rcR*bR)
        T const h = T(0.5) * g_ps;
        T g_pL = h;
        T g_pR = h;
        T g_rcL = h * bL;
        T g_rcR = h * bR;
        T const g_bL = h * rcL;
        T const g_bR = h * rcR;
        g_us += g_bL + g_bR;
        T g_unL = -g_bL;
        T g_unR = -g_bR;
        // us = num / den
        T const g_num = g_us * rden;
        T const g_den = -g_num * us;
        // num = (pR - pL) + rcL*unL - rcR*unR; den = rcL - rcR
        g_pR += g_num;
        g_pL -= g_num;
        g_rcL += g_num * unL + g_den;
        g_unL += g_num * rcL;
        g_rcR += -g_num * unR - g_den;
        g_unR += -g_num * rcR;
        // rcL = dL*aL; rcR = dR*aR
        T g_dL = g_rcL * aL;
        T g_dR = g_rcR * aR;
        T const g_aL = g_rcL * qL.d;
        T const g_aR = g_rcR * qR.d;
        // aL = SL - unL; aR = SR - unR
        g_SL += g_aL;
        g_unL -= g_aL;
        g_SR += g_aR;
        g_unR -= g_aR;
        // SL = umin - cmax; SR = umax + cmax
        T const g_cmax = g_SR - g_SL;
        if (sel_min) {
            g_unR += g_SL;
        } else {
            g_unL += g_SL;
        }
        if (sel_max) {
            g_unR += g_SR;
        } else {
            g_unL += g_SR;
        }
        // c = sqrt(gamma*p/d): dc = c/2 * (dp/p - dd/d), on the side that set cmax
        {
            Prim<T> const& qc = sel_c ? qR : qL;
            T const gcc = g_cmax * cmax;
            T const gp = gcc * (T(0.5) * fast_rcp(qc.p));
            T const gd = gcc * (T(0.5) * fast_rcp(qc.d));
            if (sel_c) {
                g_pR += gp;
                g_dR -= gd;
            } else {
                g_pL += gp;
                g_dL -= gd;
            }
        }
        gl.d += g_dL;
        gl.p += g_pL;
        comp<Dir>(gl) += g_unL;
        gr.d += g_dR;
        gr.p += g_pR;
        comp<Dir>(gr) += g_unR;
    }
}

// Same, from the primitives alone.
template <int Dir, class T>
__device__ __forceinline__ void hllc_vjp(
        Prim<T> const& qL,
        Prim<T> const& qR,
        T gamma,
        T gm1,
        Flux<T> const& g,
        PrimGrad<T>& gl,
        PrimGrad<T>& gr)
{
    hllc_vjp<Dir>(vjp_cell(qL, gamma), vjp_cell(qR, gamma), gm1, g, gl, gr);
}

// J^T g of cons -> prim at the state with primitives q (u = m/rho, p = (gamma-1)(e - |m|^2/(2rho))).
template <class T>
__device__ __forceinline__ Cons<T> c2p_vjp(Prim<T> const& q, PrimGrad<T> const& g, T gm1)
{
#pragma clang fp contract(fast)
    T const inv = fast_rcp(q.d);
    T const ke = q.d * ((q.u0 * q.u0) + (q.u1 * q.u1) + (q.u2 * q.u2)) * T(0.5);
    T const g_ke = -gm1 * g.p;
    Cons<T> r;
    r.d = g.d - g_ke * ke * inv - (g.u0 * q.u0 + g.u1 * q.u1 + g.u2 * q.u2) * inv;
    r.e = gm1 * g.p;
    r.m0 = g_ke * q.u0 + g.u0 * inv;
    r.m1 = g_ke * q.u1 + g.u1 * inv;
    r.m2 = g_ke * q.u2 + g.u2 * inv;
    return r;
}

// J^T g of prim -> cons (e = p/(gamma-1) + rho|u|^2/2, m = rho*u).
template <class T>
__device__ __forceinline__ PrimGrad<T> p2c_vjp(Prim<T> const& q, Cons<T> const& g, T gm1)
{
#pragma clang fp contract(fast)
    T const u2 = (q.u0 * q.u0) + (q.u1 * q.u1) + (q.u2 * q.u2);
    PrimGrad<T> r;
    r.d = g.d + g.e * u2 * T(0.5) + g.m0 * q.u0 + g.m1 * q.u1 + g.m2 * q.u2;
    r.p = g.e * fast_rcp(gm1);
    r.u0 = g.e * q.d * q.u0 + g.m0 * q.d;
    r.u1 = g.e * q.d * q.u1 + g.m1 * q.d;
    r.u2 = g.e * q.d * q.u2 + g.m2 * q.d;
    return r;
}

} // namespace eb
