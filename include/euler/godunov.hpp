// godunov (euler_operators/godunov.hpp:15-92): cons -= dt/dv * sum_d ds_d
// (F_{d+1/2} - F_{d-1/2}) on interior cells, F the HLLC flux of the primitives.
// The reference's form takes dt on the host; the device-dt form reads
// cfl * dt_dev[0] on the GPU.
#pragma once

#include "detail_capi.hpp"
#include "hllc.hpp"
#include "perfect_gas.hpp"
#include "uniform_mesh.hpp"

namespace eb {

template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void godunov(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        EulerConsArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
        PerfectGas<T> const& eos,
        UniformMesh3d<T> const& mesh,
        hllc const& /*riemann_solver*/,
        T const dt)
{
    detail::check_pair(prim_arrays, cons_arrays);
    auto const p = detail::ptrs(prim_arrays);
    auto const c = detail::ptrs(cons_arrays);
    detail::check(detail::capi<T>::godunov(exec_space.stream(), &p, &c, detail::ext(prim_arrays, 0),
                                           detail::ext(prim_arrays, 1), detail::ext(prim_arrays, 2),
                                           detail::index_bits<IndexType>(prim_arrays), eos.gamma(), mesh.dx0(), mesh.dx1(),
                                           mesh.dx2(), dt));
}

template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void godunov(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        EulerConsArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
        PerfectGas<T> const& eos,
        UniformMesh3d<T> const& mesh,
        hllc const& /*riemann_solver*/,
        T const* dt_dev,
        T const cfl)
{
    detail::check_pair(prim_arrays, cons_arrays);
    auto const p = detail::ptrs(prim_arrays);
    auto const c = detail::ptrs(cons_arrays);
    detail::check(detail::capi<T>::godunov_dev(exec_space.stream(), &p, &c, detail::ext(prim_arrays, 0),
                                               detail::ext(prim_arrays, 1), detail::ext(prim_arrays, 2),
                                               detail::index_bits<IndexType>(prim_arrays), eos.gamma(), mesh.dx0(),
                                               mesh.dx1(), mesh.dx2(), dt_dev, cfl));
}

// Six Riemann solves per cell in the reference's order (validation kernel).
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void godunov_reference_order(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        EulerConsArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
        PerfectGas<T> const& eos,
        UniformMesh3d<T> const& mesh,
        hllc const& /*riemann_solver*/,
        T const dt)
{
    detail::check_pair(prim_arrays, cons_arrays);
    auto const p = detail::ptrs(prim_arrays);
    auto const c = detail::ptrs(cons_arrays);
    detail::check(detail::capi<T>::godunov_reference_order(
            exec_space.stream(), &p, &c, detail::ext(prim_arrays, 0), detail::ext(prim_arrays, 1),
            detail::ext(prim_arrays, 2), detail::index_bits<IndexType>(prim_arrays), eos.gamma(), mesh.dx0(), mesh.dx1(),
            mesh.dx2(), dt));
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/godunov.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::godunov;
#endif
