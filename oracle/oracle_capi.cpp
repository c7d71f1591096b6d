// extern "C" face of the CPU oracle (test infrastructure only; see euler_oracle.hpp).
// Arrays are passed as a pointer to five field pointers in the reference's field
// order: prim {d, p, ux0, ux1, ux2}, cons {d, e, mx0, mx1, mx2}.
#include <cstdint>

#include <omp.h>

#include "euler_oracle.hpp"

namespace {

template <class T>
oracle::PrimArrays<T> prim_of(T* const* f)
{
    return {f[0], f[1], f[2], f[3], f[4]};
}

template <class T>
oracle::ConsArrays<T> cons_of(T* const* f)
{
    return {f[0], f[1], f[2], f[3], f[4]};
}

template <class T>
void hllc_any(int dir, T gamma, T const* qL, T const* qR, T* out)
{
    oracle::Gas<T> const eos {gamma};
    oracle::Prim<T> const L {qL[0], qL[1], qL[2], qL[3], qL[4]};
    oracle::Prim<T> const R {qR[0], qR[1], qR[2], qR[3], qR[4]};
    oracle::Flux<T> f {};
    if (dir == 0) {
        f = oracle::hllc<0>(eos, L, R);
    } else if (dir == 1) {
        f = oracle::hllc<1>(eos, L, R);
    } else {
        f = oracle::hllc<2>(eos, L, R);
    }
    out[0] = f.d;
    out[1] = f.e;
    out[2] = f.mx0;
    out[3] = f.mx1;
    out[4] = f.mx2;
}

// euler_simulation.cpp:56-68 — dt = time_step(prim); godunov(prim, cons, cfl*dt);
// cons_to_prim(cons -> prim), repeated nsteps times. dts[it] receives the raw dt.
template <class T>
void simulate_any(
        T* const* prim,
        T* const* cons,
        T gamma,
        T dx0,
        T dx1,
        T dx2,
        T cfl,
        int64_t nsteps,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        T* dts)
{
    oracle::Gas<T> const eos {gamma};
    auto const P = prim_of(prim);
    auto const U = cons_of(cons);
    int64_t const n = n0 * n1 * n2;
    for (int64_t it = 0; it < nsteps; ++it) {
        T const dt = oracle::time_step(P, eos, dx0, dx1, dx2, n);
        if (dts != nullptr) {
            dts[it] = dt;
        }
        oracle::godunov(P, U, eos, dx0, dx1, dx2, cfl * dt, n0, n1, n2);
        oracle::cons_to_prim(U, P, eos, n);
    }
}

} // namespace

extern "C" {

void oracle_set_threads(int n)
{
    omp_set_num_threads(n);
}

int oracle_max_threads()
{
    return omp_get_max_threads();
}

#define ORACLE_DEFINE(SUF, T)                                                                       \
    void oracle_hllc_##SUF(int dir, T gamma, T const* qL, T const* qR, T* out)                     \
    {                                                                                               \
        hllc_any<T>(dir, gamma, qL, qR, out);                                                       \
    }                                                                                               \
    void oracle_cons_to_prim_##SUF(T* const* cons, T* const* prim, T gamma, int64_t n)             \
    {                                                                                               \
        oracle::cons_to_prim(cons_of(cons), prim_of(prim), oracle::Gas<T> {gamma}, n);              \
    }                                                                                               \
    void oracle_prim_to_cons_##SUF(T* const* prim, T* const* cons, T gamma, int64_t n)             \
    {                                                                                               \
        oracle::prim_to_cons(prim_of(prim), cons_of(cons), oracle::Gas<T> {gamma}, n);              \
    }                                                                                               \
    T oracle_time_step_##SUF(T* const* prim, T gamma, T dx0, T dx1, T dx2, int64_t n)              \
    {                                                                                               \
        return oracle::time_step(prim_of(prim), oracle::Gas<T> {gamma}, dx0, dx1, dx2, n);         \
    }                                                                                               \
    void oracle_godunov_##SUF(                                                                      \
            T* const* prim,                                                                         \
            T* const* cons,                                                                         \
            T gamma,                                                                                \
            T dx0,                                                                                  \
            T dx1,                                                                                  \
            T dx2,                                                                                  \
            T dt,                                                                                   \
            int64_t n0,                                                                             \
            int64_t n1,                                                                             \
            int64_t n2)                                                                             \
    {                                                                                               \
        oracle::godunov(                                                                            \
                prim_of(prim), cons_of(cons), oracle::Gas<T> {gamma}, dx0, dx1, dx2, dt, n0, n1, n2); \
    }                                                                                               \
    void oracle_init_implode_##SUF(T* const* prim, T dx0, T dx1, T dx2, int64_t n0, int64_t n1, int64_t n2) \
    {                                                                                               \
        oracle::init_implode(prim_of(prim), dx0, dx1, dx2, n0, n1, n2);                             \
    }                                                                                               \
    void oracle_init_diagonal_##SUF(                                                                \
            T* const* prim,                                                                         \
            T dx0,                                                                                  \
            T dx1,                                                                                  \
            T dx2,                                                                                  \
            T offset,                                                                               \
            T d_hi,                                                                                 \
            T p_hi,                                                                                 \
            T d_lo,                                                                                 \
            T p_lo,                                                                                 \
            int64_t n0,                                                                             \
            int64_t n1,                                                                             \
            int64_t n2)                                                                             \
    {                                                                                               \
        oracle::init_diagonal(prim_of(prim), dx0, dx1, dx2, offset, d_hi, p_hi, d_lo, p_lo, n0, n1, n2); \
    }                                                                                               \
    void oracle_simulate_##SUF(                                                                     \
            T* const* prim,                                                                         \
            T* const* cons,                                                                         \
            T gamma,                                                                                \
            T dx0,                                                                                  \
            T dx1,                                                                                  \
            T dx2,                                                                                  \
            T cfl,                                                                                  \
            int64_t nsteps,                                                                         \
            int64_t n0,                                                                             \
            int64_t n1,                                                                             \
            int64_t n2,                                                                             \
            T* dts)                                                                                 \
    {                                                                                               \
        simulate_any<T>(prim, cons, gamma, dx0, dx1, dx2, cfl, nsteps, n0, n1, n2, dts);            \
    }

ORACLE_DEFINE(f64, double)
ORACLE_DEFINE(f32, float)

#undef ORACLE_DEFINE

} // extern "C"
