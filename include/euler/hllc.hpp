// HLLC Riemann solver (euler_operators/hllc.hpp:11-105): the normal component
// get<Dir> of a point struct, and the flux functor godunov takes. The library's
// kernels evaluate the same flux internally; this functor is the reference's
// public form of it, callable from host code and from device code built with
// hipcc. Its arithmetic keeps the reference's expression trees operation for
// operation (Davis wave speeds, acoustic star state, upwind side chosen by u* > 0
// with ties to the right, energy star term over (S - u*)), so with contraction
// off it returns the reference's bits. The functor and the point formulas it
// calls switch contraction off themselves under clang and hipcc (EB_IEEE_ORDER,
// honoured unless the caller builds with -ffp-contract=fast); host code built
// with GCC needs -ffp-contract=off for the same guarantee.
#pragma once

#include <cstddef>
#include <type_traits>

#include "euler_arrays.hpp"
#include "macros.hpp"
#include "perfect_gas.hpp"

namespace eb {

// Normal velocity / momentum / momentum flux along Dir (hllc.hpp:11-48).
template <std::size_t Dir, class T>
EB_FUNCTION T get(std::integral_constant<std::size_t, Dir> /*dir*/, EulerPrim<T> const& prim) noexcept
{
    static_assert(Dir < 3, "direction 0, 1 or 2");
    if constexpr (Dir == 0) {
        return prim.ux0;
    } else if constexpr (Dir == 1) {
        return prim.ux1;
    } else {
        return prim.ux2;
    }
}

template <std::size_t Dir, class T>
EB_FUNCTION T get(std::integral_constant<std::size_t, Dir> /*dir*/, EulerCons<T> const& cons) noexcept
{
    static_assert(Dir < 3, "direction 0, 1 or 2");
    if constexpr (Dir == 0) {
        return cons.mx0;
    } else if constexpr (Dir == 1) {
        return cons.mx1;
    } else {
        return cons.mx2;
    }
}

template <std::size_t Dir, class T>
EB_FUNCTION T get(std::integral_constant<std::size_t, Dir> /*dir*/, EulerFlux<T> const& flux) noexcept
{
    static_assert(Dir < 3, "direction 0, 1 or 2");
    if constexpr (Dir == 0) {
        return flux.mx0;
    } else if constexpr (Dir == 1) {
        return flux.mx1;
    } else {
        return flux.mx2;
    }
}

namespace detail {

// std::min / std::max semantics (the second operand only when strictly
// smaller / larger), which Kokkos::min / Kokkos::max follow.
template <class T>
EB_FUNCTION constexpr T tmin(T const a, T const b) noexcept
{
    return b < a ? b : a;
}

template <class T>
EB_FUNCTION constexpr T tmax(T const a, T const b) noexcept
{
    return a < b ? b : a;
}

} // namespace detail

// Direction tag of the Riemann solver call (godunov.hpp:12-13).
template <std::size_t N>
using dir_t = std::integral_constant<std::size_t, N>;

struct hllc
{
    // Flux across a face normal to Dir between the left state qL and the right
    // state qR (hllc.hpp:52-104).
    template <std::size_t Dir, class T, class U>
    EB_FUNCTION EulerFlux<T> operator()(std::integral_constant<std::size_t, Dir> dir, PerfectGas<U> const& eos,
                                        EulerPrim<T> const& qL, EulerPrim<T> const& qR) const noexcept
    {
        EB_IEEE_ORDER
        static_assert(Dir < 3, "direction 0, 1 or 2");
        // wave speeds from the normal velocities and the larger sound speed (:61-68)
        T const unL = get(dir, qL);
        T const unR = get(dir, qR);
        T const cmax = detail::tmax(eos.speed_of_sound(qL.d, qL.p), eos.speed_of_sound(qR.d, qR.p));
        T const sL = detail::tmin(unL, unR) - cmax;
        T const sR = detail::tmax(unL, unR) + cmax;
        // acoustic star state (:70-76)
        T const mL = qL.d * (sL - unL);
        T const mR = qR.d * (sR - unR);
        T const u_star = (qR.p - qL.p + mL * unL - mR * unR) / (mL - mR);
        T const p_star = static_cast<U>(0.5) * (qL.p + qR.p + mL * (u_star - unL) + mR * (u_star - unR));
        // upwind side; u* = 0 goes right (:78-82)
        bool const from_left = u_star > 0;
        T const s = from_left ? sL : sR;
        EulerPrim<T> const& q = from_left ? qL : qR;
        T const un = get(dir, q);
        T const e_tot = eos.internal_energy(q.d, q.p) + kinetic_energy(q);
        // state on the upwind side of the contact; supersonic: the upwind state (:84-88)
        bool const supersonic = sL * sR > 0;
        T const un_o = supersonic ? un : u_star;
        T const p_o = supersonic ? q.p : p_star;
        T const d_o = (s - un) / (s - un_o) * q.d;
        T const e_o = ((s - un) / (s - un_o) * e_tot) + ((p_o * un_o - q.p * un) / (s - u_star));
        // fluxes (:90-102)
        EulerFlux<T> f {};
        f.d = d_o * un_o;
        f.e = (e_o + p_o) * un_o;
        f.mx0 = d_o * un_o * q.ux0;
        f.mx1 = d_o * un_o * q.ux1;
        f.mx2 = d_o * un_o * q.ux2;
        T const normal = (d_o * un_o * un_o) + p_o;
        if constexpr (Dir == 0) {
            f.mx0 = normal;
        } else if constexpr (Dir == 1) {
            f.mx1 = normal;
        } else {
            f.mx2 = normal;
        }
        return f;
    }
};

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/hllc.hpp and godunov.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::dir_t;
using eb::get;
using eb::hllc;
#endif
