// time_step (euler_operators/time_step.hpp:8-41): min over all cells, ghosts
// included, of 1 / sum_d (c + |u_d|) / dx_d. The reference returns it to the host
// (blocking); the second overload leaves it in device memory instead, so a step
// loop never waits for the host.
#pragma once

#include "detail_capi.hpp"
#include "perfect_gas.hpp"
#include "uniform_mesh.hpp"

namespace eb {

template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
T time_step(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        PerfectGas<T> const& eos,
        UniformMesh3d<T> const& mesh)
{
    detail::check_fields(prim_arrays);
    auto const p = detail::ptrs(prim_arrays);
    T dt {};
    detail::check(detail::capi<T>::time_step_sync(exec_space.context(), exec_space.stream(), &p,
                                                  detail::ext(prim_arrays, 0), detail::ext(prim_arrays, 1),
                                                  detail::ext(prim_arrays, 2), detail::index_bits<IndexType>(prim_arrays),
                                                  eos.gamma(), mesh.dx0(), mesh.dx1(), mesh.dx2(), &dt));
    return dt;
}

// Asynchronous form: the raw dt is written to dt_dev[0] (device memory).
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void time_step(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        PerfectGas<T> const& eos,
        UniformMesh3d<T> const& mesh,
        T* dt_dev)
{
    detail::check_fields(prim_arrays);
    auto const p = detail::ptrs(prim_arrays);
    detail::check(detail::capi<T>::time_step(exec_space.context(), exec_space.stream(), &p,
                                             detail::ext(prim_arrays, 0), detail::ext(prim_arrays, 1),
                                             detail::ext(prim_arrays, 2), detail::index_bits<IndexType>(prim_arrays),
                                             eos.gamma(), mesh.dx0(), mesh.dx1(), mesh.dx2(), dt_dev));
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/time_step.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::time_step;
#endif
