// prim_to_cons (euler_operators/prim_to_cons.hpp:7-33): conservatives of every
// cell, ghosts included, from the primitives.
#pragma once

#include "detail_capi.hpp"
#include "perfect_gas.hpp"

namespace eb {

template <class T, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void prim_to_cons(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<T const, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        EulerConsArrays<mdspan<T, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
        PerfectGas<T> const& eos)
{
    detail::check_pair(prim_arrays, cons_arrays);
    auto const p = detail::ptrs(prim_arrays);
    auto const c = detail::ptrs(cons_arrays);
    detail::check(detail::capi<T>::prim_to_cons(exec_space.stream(), &p, &c, detail::ext(prim_arrays, 0),
                                                detail::ext(prim_arrays, 1), detail::ext(prim_arrays, 2),
                                                detail::index_bits<IndexType>(prim_arrays), eos.gamma()));
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/prim_to_cons.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::prim_to_cons;
#endif
