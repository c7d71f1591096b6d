// cons_to_prim (euler_operators/cons_to_prim.hpp:7-33): primitives of every
// cell, ghosts included, from the conservatives.
#pragma once

#include "detail_capi.hpp"
#include "perfect_gas.hpp"

namespace eb {

template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void cons_to_prim(
        HipExecSpace const& exec_space,
        EulerConsArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
        EulerPrimArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        PerfectGas<T> const& eos)
{
    detail::check_pair(cons_arrays, prim_arrays);
    auto const c = detail::ptrs(cons_arrays);
    auto const p = detail::ptrs(prim_arrays);
    detail::check(detail::capi<T>::cons_to_prim(exec_space.stream(), &c, &p, detail::ext(prim_arrays, 0),
                                                detail::ext(prim_arrays, 1), detail::ext(prim_arrays, 2),
                                                detail::index_bits<IndexType>(prim_arrays), eos.gamma()));
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/cons_to_prim.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::cons_to_prim;
#endif
