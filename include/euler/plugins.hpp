// Plug points the reference fixes (euler_operators/godunov.hpp:26-28 takes exactly
// `hllc`, `PerfectGas<T>` and `UniformMesh3d<T>`): the Riemann solver, the
// equation of state and the mesh. This header adds overloads of godunov,
// cons_to_prim, prim_to_cons and time_step for any types with the reference's
// member signatures; their kernels are compiled by the caller's hipcc, so the
// caller's types run on the GPU. The reference types keep dispatching to the
// library's tiled kernels (the overloads in godunov.hpp etc. are more specialised
// and win overload resolution); the generic kernels follow the reference's loops
// (one thread per cell, six Riemann solves per interior cell, godunov.hpp:40-90)
// and give the reference's results for any solver, EOS and mesh.
//
// Requirements (T the real type; members callable from device code, i.e.
// EB_FUNCTION / __host__ __device__):
//   Riemann solver: EulerFlux<T> operator()(std::integral_constant<std::size_t, D>,
//                   EOS const&, EulerPrim<T> const& qL, EulerPrim<T> const& qR) const
//                   for D = 0, 1, 2 (hllc.hpp:50-105);
//   EOS:            T speed_of_sound(T d, T p) const, T internal_energy(T d, T p) const,
//                   T pressure(T d, T eint) const (perfect_gas.hpp:15-28);
//   mesh:           T dx0() .. dx2(), ds0() .. ds2(), dv() const (uniform_mesh.hpp:33-51).
// Every solver, EOS and mesh object is passed to the kernels by value.
#pragma once

#if !defined(__HIPCC__)
#error "euler/plugins.hpp holds device code: build the including file with hipcc"
#endif

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>

#include <hip/hip_runtime.h>

#include "euler.hpp"

namespace eb {

namespace plugin_detail {

template <std::size_t D>
using dir = std::integral_constant<std::size_t, D>;

// Thread -> cell (i, j, k) over an n0 x n1 x n2 box starting at (o, o, o), i fastest.
template <class I>
struct Box
{
    I n0, n1, n2, o;
};

template <class I>
__device__ __forceinline__ bool cell_of(Box<I> const& b, long long g, I& i, I& j, I& k)
{
    long long const n01 = static_cast<long long>(b.n0) * b.n1;
    if (g >= n01 * b.n2) {
        return false;
    }
    k = static_cast<I>(g / n01) + b.o;
    long long const r = g % n01;
    j = static_cast<I>(r / b.n0) + b.o;
    i = static_cast<I>(r % b.n0) + b.o;
    return true;
}

template <class Solver, class EOS, std::size_t D, class PV, class T, class I>
__device__ __forceinline__ void add_direction(Solver const& rs, EOS const& eos, PV const& prim, EulerPrim<T> const& q,
                                              T ds, I i, I j, I k, EulerFlux<T>& flux)
{
    EB_IEEE_ORDER
    I const im = D == 0 ? i - 1 : i, jm = D == 1 ? j - 1 : j, km = D == 2 ? k - 1 : k;
    I const ip = D == 0 ? i + 1 : i, jp = D == 1 ? j + 1 : j, kp = D == 2 ? k + 1 : k;
    EulerPrim<T> const qL = load(prim, im, jm, km);
    EulerPrim<T> const qR = load(prim, ip, jp, kp);
    EulerFlux<T> const fL = rs(dir<D> {}, eos, qL, q);
    EulerFlux<T> const fR = rs(dir<D> {}, eos, q, qR);
    flux.d += ds * (fR.d - fL.d);
    flux.e += ds * (fR.e - fL.e);
    flux.mx0 += ds * (fR.mx0 - fL.mx0);
    flux.mx1 += ds * (fR.mx1 - fL.mx1);
    flux.mx2 += ds * (fR.mx2 - fL.mx2);
}

// godunov.hpp:40-90 for any solver, EOS and mesh; dtodv formed on the host (:32).
template <class Solver, class EOS, class PV, class CV, class T, class I>
__global__ __launch_bounds__(256) void k_godunov(PV prim, CV cons, EOS eos, Solver rs, T ds0, T ds1, T ds2, T dtodv,
                                                 Box<I> box)
{
    EB_IEEE_ORDER
    I i, j, k;
    if (!cell_of(box, static_cast<long long>(blockIdx.x) * 256 + threadIdx.x, i, j, k)) {
        return;
    }
    EulerPrim<T> const q = load(prim, i, j, k);
    EulerFlux<T> flux {};
    add_direction<Solver, EOS, 0>(rs, eos, prim, q, ds0, i, j, k, flux);
    add_direction<Solver, EOS, 1>(rs, eos, prim, q, ds1, i, j, k, flux);
    add_direction<Solver, EOS, 2>(rs, eos, prim, q, ds2, i, j, k, flux);
    EulerCons<T> u = load(cons, i, j, k);
    u.d -= dtodv * flux.d;
    u.e -= dtodv * flux.e;
    u.mx0 -= dtodv * flux.mx0;
    u.mx1 -= dtodv * flux.mx1;
    u.mx2 -= dtodv * flux.mx2;
    store(u, cons, i, j, k);
}

// cons_to_prim.hpp:28-31
template <class EOS, class CV, class PV, class T, class I>
__global__ __launch_bounds__(256) void k_cons_to_prim(CV cons, PV prim, EOS eos, Box<I> box)
{
    I i, j, k;
    if (cell_of(box, static_cast<long long>(blockIdx.x) * 256 + threadIdx.x, i, j, k)) {
        EulerCons<T> const u = load(cons, i, j, k);
        store(to_prim(u, eos.pressure(u.d, internal_energy(u))), prim, i, j, k);
    }
}

// prim_to_cons.hpp:28-31
template <class EOS, class PV, class CV, class T, class I>
__global__ __launch_bounds__(256) void k_prim_to_cons(PV prim, CV cons, EOS eos, Box<I> box)
{
    I i, j, k;
    if (cell_of(box, static_cast<long long>(blockIdx.x) * 256 + threadIdx.x, i, j, k)) {
        EulerPrim<T> const q = load(prim, i, j, k);
        store(to_cons(q, eos.internal_energy(q.d, q.p)), cons, i, j, k);
    }
}

// time_step.hpp:28-38: per-block minima of 1/invdt (Kokkos::Min semantics: a
// candidate replaces the incumbent only when strictly smaller, so NaN never
// enters), then one block combines them. The minimum is exact, so the combine
// order does not matter.
template <class T>
__device__ __forceinline__ T min_keep(T incumbent, T candidate)
{
    return candidate < incumbent ? candidate : incumbent;
}

template <class T>
__device__ __forceinline__ T block_min_256(T v, T* shared)
{
    shared[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) {
            shared[threadIdx.x] = min_keep(shared[threadIdx.x], shared[threadIdx.x + s]);
        }
        __syncthreads();
    }
    return shared[0];
}

template <class EOS, class Mesh, class PV, class T, class I>
__global__ __launch_bounds__(256) void k_time_step_partial(PV prim, EOS eos, T invdx0, T invdx1, T invdx2, Box<I> box,
                                                           T* partial)
{
    EB_IEEE_ORDER
    __shared__ T sh[256];
    T local = std::numeric_limits<T>::max();
    long long const total = static_cast<long long>(box.n0) * box.n1 * box.n2;
    for (long long g = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; g < total;
         g += static_cast<long long>(gridDim.x) * 256) {
        I i, j, k;
        cell_of(box, g, i, j, k);
        EulerPrim<T> const q = load(prim, i, j, k);
        T const cs = eos.speed_of_sound(q.d, q.p);
        T const cx0 = cs + fabs(q.ux0);
        T const cx1 = cs + fabs(q.ux1);
        T const cx2 = cs + fabs(q.ux2);
        T const invdt = (cx0 * invdx0) + (cx1 * invdx1) + (cx2 * invdx2);
        local = min_keep(local, T(1) / invdt);
    }
    T const m = block_min_256(local, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = m;
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_time_step_final(T const* partial, int n, T* out)
{
    __shared__ T sh[256];
    T local = std::numeric_limits<T>::max();
    for (int b = static_cast<int>(threadIdx.x); b < n; b += 256) {
        local = min_keep(local, partial[b]);
    }
    T const m = block_min_256(local, sh);
    if (threadIdx.x == 0) {
        out[0] = m;
    }
}

template <class I>
unsigned int blocks_for(Box<I> const& b)
{
    long long const n = static_cast<long long>(b.n0) * b.n1 * b.n2;
    return static_cast<unsigned int>((n + 255) / 256);
}

inline void launched(char const* what)
{
    detail::check_hip(hipGetLastError(), what);
}

template <class A>
Box<typename decltype(A::d)::index_type> all_cells(A const& a)
{
    using I = typename decltype(A::d)::index_type;
    return {static_cast<I>(a.d.extent(0)), static_cast<I>(a.d.extent(1)), static_cast<I>(a.d.extent(2)), I(0)};
}

} // namespace plugin_detail

// godunov with any Riemann solver, EOS and mesh (godunov.hpp:15-92).
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2, class EOS, class Mesh,
          class RiemannSolver>
void godunov(HipExecSpace const& exec_space,
             EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
             EulerConsArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
             EOS const& eos, Mesh const& mesh, RiemannSolver const& riemann_solver, T const dt)
{
    detail::check_pair(prim_arrays, cons_arrays);
    auto const n = plugin_detail::all_cells(prim_arrays);
    if (n.n0 < 3 || n.n1 < 3 || n.n2 < 3) {
        return;
    }
    plugin_detail::Box<IndexType> const inner {static_cast<IndexType>(n.n0 - 2), static_cast<IndexType>(n.n1 - 2),
                                              static_cast<IndexType>(n.n2 - 2), IndexType(1)};
    T const dtodv = dt / mesh.dv(); // godunov.hpp:32, on the host
    using PV = std::remove_cv_t<std::remove_reference_t<decltype(prim_arrays)>>;
    using CV = std::remove_cv_t<std::remove_reference_t<decltype(cons_arrays)>>;
    hipLaunchKernelGGL((plugin_detail::k_godunov<RiemannSolver, EOS, PV, CV, T, IndexType>),
                       dim3(plugin_detail::blocks_for(inner)), dim3(256), 0, exec_space.stream(), prim_arrays,
                       cons_arrays, eos, riemann_solver, mesh.ds0(), mesh.ds1(), mesh.ds2(), dtodv, inner);
    plugin_detail::launched("godunov (plug-in kernel)");
}

// cons_to_prim with any EOS (cons_to_prim.hpp:7-33).
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2, class EOS>
void cons_to_prim(HipExecSpace const& exec_space,
                  EulerConsArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
                  EulerPrimArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
                  EOS const& eos)
{
    detail::check_pair(cons_arrays, prim_arrays);
    auto const box = plugin_detail::all_cells(prim_arrays);
    if (prim_arrays.d.size() == 0) {
        return;
    }
    using CV = std::remove_cv_t<std::remove_reference_t<decltype(cons_arrays)>>;
    using PV = std::remove_cv_t<std::remove_reference_t<decltype(prim_arrays)>>;
    hipLaunchKernelGGL((plugin_detail::k_cons_to_prim<EOS, CV, PV, T, IndexType>), dim3(plugin_detail::blocks_for(box)),
                       dim3(256), 0, exec_space.stream(), cons_arrays, prim_arrays, eos, box);
    plugin_detail::launched("cons_to_prim (plug-in kernel)");
}

// prim_to_cons with any EOS (prim_to_cons.hpp:7-33).
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2, class EOS>
void prim_to_cons(HipExecSpace const& exec_space,
                  EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
                  EulerConsArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
                  EOS const& eos)
{
    detail::check_pair(prim_arrays, cons_arrays);
    auto const box = plugin_detail::all_cells(prim_arrays);
    if (prim_arrays.d.size() == 0) {
        return;
    }
    using PV = std::remove_cv_t<std::remove_reference_t<decltype(prim_arrays)>>;
    using CV = std::remove_cv_t<std::remove_reference_t<decltype(cons_arrays)>>;
    hipLaunchKernelGGL((plugin_detail::k_prim_to_cons<EOS, PV, CV, T, IndexType>), dim3(plugin_detail::blocks_for(box)),
                       dim3(256), 0, exec_space.stream(), prim_arrays, cons_arrays, eos, box);
    plugin_detail::launched("prim_to_cons (plug-in kernel)");
}

// time_step with any EOS and mesh (time_step.hpp:8-41), returned to the host.
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2, class EOS, class Mesh>
T time_step(HipExecSpace const& exec_space,
            EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
            EOS const& eos, Mesh const& mesh)
{
    detail::check_fields(prim_arrays);
    auto const box = plugin_detail::all_cells(prim_arrays);
    // the inverse spacings are formed on the host (time_step.hpp:18-20)
    T const invdx0 = 1 / mesh.dx0();
    T const invdx1 = 1 / mesh.dx1();
    T const invdx2 = 1 / mesh.dx2();
    unsigned int const blocks = std::min(1024u, std::max(1u, plugin_detail::blocks_for(box)));
    View<T> scratch(exec_space, "time_step partial minima", blocks + 1);
    using PV = std::remove_cv_t<std::remove_reference_t<decltype(prim_arrays)>>;
    hipLaunchKernelGGL((plugin_detail::k_time_step_partial<EOS, Mesh, PV, T, IndexType>), dim3(blocks), dim3(256), 0,
                       exec_space.stream(), prim_arrays, eos, invdx0, invdx1, invdx2, box, scratch.data());
    plugin_detail::launched("time_step (plug-in kernel)");
    hipLaunchKernelGGL(plugin_detail::k_time_step_final<T>, dim3(1), dim3(256), 0, exec_space.stream(), scratch.data(),
                       static_cast<int>(blocks), scratch.data() + blocks);
    plugin_detail::launched("time_step (plug-in final)");
    T dt {};
    detail::check_hip(hipMemcpyAsync(&dt, scratch.data() + blocks, sizeof(T), hipMemcpyDeviceToHost, exec_space.stream()),
                      "hipMemcpyAsync");
    exec_space.fence();
    return dt;
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The overloads above under the reference's global spellings as well (a
// using-declaration names only the overloads declared before it).
using eb::godunov;
using eb::cons_to_prim;
using eb::prim_to_cons;
using eb::time_step;
#endif
