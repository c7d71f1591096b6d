// Initial states. init_implode follows euler_operators/init_implode.hpp:7-45
// (sphere of radius 0.25 about the centre, ghosts included); init_diagonal and
// init_random have no reference counterpart (measured configurations).
#pragma once

#include <cstdint>

#include "detail_capi.hpp"
#include "uniform_mesh.hpp"

namespace eb {

template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void init_implode(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        UniformMesh3d<T> const& mesh)
{
    detail::check_fields(prim_arrays);
    auto const p = detail::ptrs(prim_arrays);
    detail::check(detail::capi<T>::init_implode(exec_space.stream(), &p, detail::ext(prim_arrays, 0),
                                                detail::ext(prim_arrays, 1), detail::ext(prim_arrays, 2),
                                                detail::index_bits<IndexType>(prim_arrays), mesh.dx0(), mesh.dx1(),
                                                mesh.dx2()));
}

// x0 + x1 + x2 > offset -> (d_hi, p_hi), else (d_lo, p_lo); u = 0; cell centres
// ((i-1)+0.5)*dx.
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void init_diagonal(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        UniformMesh3d<T> const& mesh,
        T offset = T(0.45),
        T d_hi = T(1),
        T p_hi = T(1),
        T d_lo = T(0.125),
        T p_lo = T(0.14))
{
    detail::check_fields(prim_arrays);
    auto const p = detail::ptrs(prim_arrays);
    detail::check(detail::capi<T>::init_diagonal(exec_space.stream(), &p, detail::ext(prim_arrays, 0),
                                                 detail::ext(prim_arrays, 1), detail::ext(prim_arrays, 2),
                                                 detail::index_bits<IndexType>(prim_arrays), mesh.dx0(), mesh.dx1(),
                                                 mesh.dx2(), offset, d_hi, p_hi, d_lo, p_lo));
}

// Seeded state rho ~ U[0.5,2), p ~ U[0.5,2), u_k ~ N(0,0.3), counter-based.
template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void init_random(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        std::uint64_t seed = 1234)
{
    detail::check_fields(prim_arrays);
    auto const p = detail::ptrs(prim_arrays);
    detail::check(detail::capi<T>::init_random(exec_space.stream(), &p, static_cast<int64_t>(size(prim_arrays)),
                                               detail::index_bits<IndexType>(prim_arrays), seed, 0.5, 2.0, 0.5, 2.0, 0.3));
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/init_implode.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::init_implode;
#endif
