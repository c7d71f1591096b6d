// CPU oracle: a plain C++ restatement of the reference's euler_operators, kept as
// TEST INFRASTRUCTURE ONLY. The HIP library never calls it; tests, the smoke check
// and the CPU-baseline leg of the benchmark use it as the checker.
//
// Every function follows the operation order of the reference header it cites, so
// that with floating-point contraction off (-ffp-contract=off) its results are the
// IEEE results of the reference's own expression trees. Index math is int64.
// Layout: SoA, layout_left, linear index i + n0*(j + n1*k).
//
// Parity pins (what checks each function, see tests/test_oracle_*.py):
//   kinetic_energy / to_cons / to_prim / cons_to_prim / prim_to_cons :
//       hand-derived closed forms, p2c(c2p(x)) round trip to a few ulp.
//   speed_of_sound / time_step : closed form 1/(3*sqrt(1.4)) on uniform states,
//       hand-derived anisotropic values, NaN cells skipped, DBL_MAX identity.
//   hllc : consistency F(q,q) = physical flux, supersonic upwinding, exact
//       stationary/moving contact, mirror symmetry, and agreement with an
//       independently written Toro-form HLLC (F*_K = F_K + S_K (U*_K - U_K)).
//   godunov : uniform state is the identity, telescoped conservation balance,
//       axis-permutation symmetry, Sod/Toro 1-D problems converging to the
//       exact Riemann solution along each axis.
//   init_implode / init_diagonal : hand-computed cells and cell counts.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <limits>

namespace oracle {

template <class T>
struct Prim
{
    T d, p, ux0, ux1, ux2;
};

template <class T>
struct Cons
{
    T d, e, mx0, mx1, mx2;
};

template <class T>
struct Flux
{
    T d, e, mx0, mx1, mx2;
};

template <class T>
struct PrimArrays
{
    T* d;
    T* p;
    T* ux0;
    T* ux1;
    T* ux2;
};

template <class T>
struct ConsArrays
{
    T* d;
    T* e;
    T* mx0;
    T* mx1;
    T* mx2;
};

// std::min / std::max semantics (ternary), as used by Kokkos::min/max.
template <class T>
inline T kmin(T a, T b)
{
    return (b < a) ? b : a;
}

template <class T>
inline T kmax(T a, T b)
{
    return (a < b) ? b : a;
}

// euler_arrays.hpp:38-42
template <class T>
inline T kinetic_energy(Prim<T> const& q)
{
    return q.d * ((q.ux0 * q.ux0) + (q.ux1 * q.ux1) + (q.ux2 * q.ux2)) / 2;
}

// euler_arrays.hpp:44-48
template <class T>
inline T kinetic_energy(Cons<T> const& u)
{
    return ((u.mx0 * u.mx0) + (u.mx1 * u.mx1) + (u.mx2 * u.mx2)) / u.d / 2;
}

// euler_arrays.hpp:50-54
template <class T>
inline T internal_energy(Cons<T> const& u)
{
    return u.e - kinetic_energy(u);
}

// euler_arrays.hpp:56-64
template <class T>
inline Cons<T> to_cons(Prim<T> const& q, T const int_e)
{
    return {q.d, kinetic_energy(q) + int_e, q.d * q.ux0, q.d * q.ux1, q.d * q.ux2};
}

// euler_arrays.hpp:66-75 (reciprocal, then multiply)
template <class T>
inline Prim<T> to_prim(Cons<T> const& u, T const p)
{
    T const vol_spe = 1 / u.d;
    return {u.d, p, u.mx0 * vol_spe, u.mx1 * vol_spe, u.mx2 * vol_spe};
}

// perfect_gas.hpp:6-29
template <class T>
struct Gas
{
    T gamma;
    T speed_of_sound(T d, T p) const { return std::sqrt(gamma * p / d); }
    T internal_energy(T /*d*/, T p) const { return p / (gamma - 1); }
    T pressure(T /*d*/, T eint) const { return (gamma - 1) * eint; }
};

// hllc.hpp:11-22
template <int Dir, class T>
inline T normal_velocity(Prim<T> const& q)
{
    if constexpr (Dir == 0) {
        return q.ux0;
    } else if constexpr (Dir == 1) {
        return q.ux1;
    } else {
        return q.ux2;
    }
}

// hllc.hpp:50-105 — RAMSES-style direct star-state HLLC with Davis wave speeds.
template <int Dir, class T>
inline Flux<T> hllc(Gas<T> const& eos, Prim<T> const& qL, Prim<T> const& qR)
{
    T const unL = normal_velocity<Dir>(qL);
    T const unR = normal_velocity<Dir>(qR);
    T const cL = eos.speed_of_sound(qL.d, qL.p);
    T const cR = eos.speed_of_sound(qR.d, qR.p);
    T const SL = kmin(unL, unR) - kmax(cL, cR);
    T const SR = kmax(unL, unR) + kmax(cL, cR);
    T const rcL = qL.d * (SL - unL);
    T const rcR = qR.d * (SR - unR);
    T const ustar = (qR.p - qL.p + rcL * unL - rcR * unR) / (rcL - rcR);
    T const pstar = static_cast<T>(0.5) * (qL.p + qR.p + rcL * (ustar - unL) + rcR * (ustar - unR));
    T const S = ustar > 0 ? SL : SR;
    Prim<T> const q = ustar > 0 ? qL : qR;
    T const un = normal_velocity<Dir>(q);
    T const etot = eos.internal_energy(q.d, q.p) + kinetic_energy(q);
    T const un_o = SL * SR > 0 ? un : ustar;
    T const ptot_o = SL * SR > 0 ? q.p : pstar;
    T const d_o = (S - un) / (S - un_o) * q.d;
    T const etot_o = ((S - un) / (S - un_o) * etot) + ((ptot_o * un_o - q.p * un) / (S - ustar));
    Flux<T> f {};
    f.d = d_o * un_o;
    f.e = (etot_o + ptot_o) * un_o;
    f.mx0 = d_o * un_o * q.ux0;
    f.mx1 = d_o * un_o * q.ux1;
    f.mx2 = d_o * un_o * q.ux2;
    if constexpr (Dir == 0) {
        f.mx0 = (d_o * un_o * un_o) + ptot_o;
    } else if constexpr (Dir == 1) {
        f.mx1 = (d_o * un_o * un_o) + ptot_o;
    } else {
        f.mx2 = (d_o * un_o * un_o) + ptot_o;
    }
    return f;
}

template <class T>
inline Prim<T> load(PrimArrays<T> const& a, int64_t idx)
{
    return {a.d[idx], a.p[idx], a.ux0[idx], a.ux1[idx], a.ux2[idx]};
}

template <class T>
inline Cons<T> load(ConsArrays<T> const& a, int64_t idx)
{
    return {a.d[idx], a.e[idx], a.mx0[idx], a.mx1[idx], a.mx2[idx]};
}

template <class T>
inline void store(PrimArrays<T> const& a, int64_t idx, Prim<T> const& q)
{
    a.d[idx] = q.d;
    a.p[idx] = q.p;
    a.ux0[idx] = q.ux0;
    a.ux1[idx] = q.ux1;
    a.ux2[idx] = q.ux2;
}

template <class T>
inline void store(ConsArrays<T> const& a, int64_t idx, Cons<T> const& u)
{
    a.d[idx] = u.d;
    a.e[idx] = u.e;
    a.mx0[idx] = u.mx0;
    a.mx1[idx] = u.mx1;
    a.mx2[idx] = u.mx2;
}

// cons_to_prim.hpp:20-32 — all cells, ghosts included.
template <class T>
void cons_to_prim(ConsArrays<T> const& cons, PrimArrays<T> const& prim, Gas<T> const& eos, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        Cons<T> const u = load(cons, idx);
        store(prim, idx, to_prim(u, eos.pressure(u.d, internal_energy(u))));
    }
}

// prim_to_cons.hpp:20-32 — all cells, ghosts included.
template <class T>
void prim_to_cons(PrimArrays<T> const& prim, ConsArrays<T> const& cons, Gas<T> const& eos, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < n; ++idx) {
        Prim<T> const q = load(prim, idx);
        store(cons, idx, to_cons(q, eos.internal_energy(q.d, q.p)));
    }
}

// time_step.hpp:18-39 — all cells, ghosts included; Min reducer identity is the
// largest finite value and a NaN candidate never replaces the running minimum.
template <class T>
T time_step(
        PrimArrays<T> const& prim,
        Gas<T> const& eos,
        T dx0,
        T dx1,
        T dx2,
        int64_t n)
{
    T const invdx0 = 1 / dx0;
    T const invdx1 = 1 / dx1;
    T const invdx2 = 1 / dx2;
    T dt = std::numeric_limits<T>::max();
#pragma omp parallel
    {
        T dt_loc = std::numeric_limits<T>::max();
#pragma omp for schedule(static) nowait
        for (int64_t idx = 0; idx < n; ++idx) {
            Prim<T> const q = load(prim, idx);
            T const cs = eos.speed_of_sound(q.d, q.p);
            T const cx0 = cs + std::abs(q.ux0);
            T const cx1 = cs + std::abs(q.ux1);
            T const cx2 = cs + std::abs(q.ux2);
            T const invdt = (cx0 * invdx0) + (cx1 * invdx1) + (cx2 * invdx2);
            dt_loc = kmin(dt_loc, 1 / invdt);
        }
#pragma omp critical
        {
            if (dt_loc < dt) {
                dt = dt_loc;
            }
        }
    }
    return dt;
}

// godunov.hpp:31-90 — interior cells only, six Riemann solves per cell (each face
// solved from both sides, as in the reference), then cons -= (dt/dv) * flux.
template <class T>
void godunov(
        PrimArrays<T> const& prim,
        ConsArrays<T> const& cons,
        Gas<T> const& eos,
        T dx0,
        T dx1,
        T dx2,
        T dt,
        int64_t n0,
        int64_t n1,
        int64_t n2)
{
    // uniform_mesh.hpp:33-51
    T const ds0 = dx1 * dx2;
    T const ds1 = dx2 * dx0;
    T const ds2 = dx0 * dx1;
    T const dv = dx0 * dx1 * dx2;
    T const dtodv = dt / dv;
    int64_t const s1 = n0;
    int64_t const s2 = n0 * n1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k < n2 - 1; ++k) {
        for (int64_t j = 1; j < n1 - 1; ++j) {
            for (int64_t i = 1; i < n0 - 1; ++i) {
                int64_t const c = i + s1 * j + s2 * k;
                Prim<T> const q = load(prim, c);
                Flux<T> flux {};
                {
                    Flux<T> const fL = hllc<0>(eos, load(prim, c - 1), q);
                    Flux<T> const fR = hllc<0>(eos, q, load(prim, c + 1));
                    flux.d += ds0 * (fR.d - fL.d);
                    flux.e += ds0 * (fR.e - fL.e);
                    flux.mx0 += ds0 * (fR.mx0 - fL.mx0);
                    flux.mx1 += ds0 * (fR.mx1 - fL.mx1);
                    flux.mx2 += ds0 * (fR.mx2 - fL.mx2);
                }
                {
                    Flux<T> const fL = hllc<1>(eos, load(prim, c - s1), q);
                    Flux<T> const fR = hllc<1>(eos, q, load(prim, c + s1));
                    flux.d += ds1 * (fR.d - fL.d);
                    flux.e += ds1 * (fR.e - fL.e);
                    flux.mx0 += ds1 * (fR.mx0 - fL.mx0);
                    flux.mx1 += ds1 * (fR.mx1 - fL.mx1);
                    flux.mx2 += ds1 * (fR.mx2 - fL.mx2);
                }
                {
                    Flux<T> const fL = hllc<2>(eos, load(prim, c - s2), q);
                    Flux<T> const fR = hllc<2>(eos, q, load(prim, c + s2));
                    flux.d += ds2 * (fR.d - fL.d);
                    flux.e += ds2 * (fR.e - fL.e);
                    flux.mx0 += ds2 * (fR.mx0 - fL.mx0);
                    flux.mx1 += ds2 * (fR.mx1 - fL.mx1);
                    flux.mx2 += ds2 * (fR.mx2 - fL.mx2);
                }
                Cons<T> u = load(cons, c);
                u.d -= dtodv * flux.d;
                u.e -= dtodv * flux.e;
                u.mx0 -= dtodv * flux.mx0;
                u.mx1 -= dtodv * flux.mx1;
                u.mx2 -= dtodv * flux.mx2;
                store(cons, c, u);
            }
        }
    }
}

// init_implode.hpp:16-43 — sphere r^2 < 0.25^2 about the cube centre, ghosts included.
// The centre coordinate is formed in double ((i-1)+0.5 promotes to double) and then
// rounded to T.
template <class T>
void init_implode(PrimArrays<T> const& prim, T dx0, T dx1, T dx2, int64_t n0, int64_t n1, int64_t n2)
{
    T const density_in = 1.;
    T const density_out = 0.125;
    T const pressure_in = 1.;
    T const pressure_out = 0.14;
    T const radius2 = 0.25 * 0.25;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < n2; ++k) {
        for (int64_t j = 0; j < n1; ++j) {
            for (int64_t i = 0; i < n0; ++i) {
                T const c0 = ((static_cast<double>(i - 1) + 0.5) * dx0) - 0.5;
                T const c1 = ((static_cast<double>(j - 1) + 0.5) * dx1) - 0.5;
                T const c2 = ((static_cast<double>(k - 1) + 0.5) * dx2) - 0.5;
                T const r2 = (c0 * c0) + (c1 * c1) + (c2 * c2);
                Prim<T> const q {r2 < radius2 ? density_in : density_out,
                                 r2 < radius2 ? pressure_in : pressure_out,
                                 T(0),
                                 T(0),
                                 T(0)};
                store(prim, i + n0 * (j + n1 * k), q);
            }
        }
    }
}

// Diagonal implosion (no function in the reference; the measured configuration asks
// for it): cell centres x_d = ((i_d-1)+0.5)*dx_d on the unit cube, formed in double
// and rounded to T, summed left to right in T:
//   x0+x1+x2 > offset  ->  (d_hi, p_hi)   else (d_lo, p_lo);  u = 0.
template <class T>
void init_diagonal(
        PrimArrays<T> const& prim,
        T dx0,
        T dx1,
        T dx2,
        T offset,
        T d_hi,
        T p_hi,
        T d_lo,
        T p_lo,
        int64_t n0,
        int64_t n1,
        int64_t n2)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < n2; ++k) {
        for (int64_t j = 0; j < n1; ++j) {
            for (int64_t i = 0; i < n0; ++i) {
                T const x0 = (static_cast<double>(i - 1) + 0.5) * dx0;
                T const x1 = (static_cast<double>(j - 1) + 0.5) * dx1;
                T const x2 = (static_cast<double>(k - 1) + 0.5) * dx2;
                bool const hi = (x0 + x1 + x2) > offset;
                Prim<T> const q {hi ? d_hi : d_lo, hi ? p_hi : p_lo, T(0), T(0), T(0)};
                store(prim, i + n0 * (j + n1 * k), q);
            }
        }
    }
}

} // namespace oracle
