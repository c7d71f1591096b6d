// Benchmark harness helpers of the reference (benchmarks/benchmark_utils.hpp):
// int_cast with its range check (:10-17), init_from_value (:19-44) and the two
// init_from_state overloads (:46-87), over the C ABI's device fill.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "detail_capi.hpp"

namespace eb {

// Conversion that throws when the value does not fit R (benchmark_utils.hpp:10-17).
template <class R, class T>
R int_cast(T t)
{
    if (std::in_range<R>(t)) {
        return static_cast<R>(t);
    }
    throw std::runtime_error("Conversion cannot preserve value representation");
}

// Every element of the array (ghosts included) set to value.
template <class ElementType, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void init_from_value(HipExecSpace const& exec_space,
                     mdspan<ElementType, extents<IndexType, E0, E1, E2>, layout_left> const& array,
                     std::type_identity_t<ElementType> const& value)
{
    if (array.size() != 0 && array.data_handle() == nullptr) {
        throw Error(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    detail::check(detail::capi<ElementType>::fill(exec_space.stream(), array.data_handle(),
                                                  static_cast<std::int64_t>(array.size()), value));
}

template <class ElementType, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void init_from_state(
        HipExecSpace const& exec_space,
        EulerPrimArrays<mdspan<ElementType, extents<IndexType, E0, E1, E2>, layout_left>> const& prim_arrays,
        EulerPrim<ElementType> const& prim)
{
    init_from_value(exec_space, prim_arrays.d, prim.d);
    init_from_value(exec_space, prim_arrays.p, prim.p);
    init_from_value(exec_space, prim_arrays.ux0, prim.ux0);
    init_from_value(exec_space, prim_arrays.ux1, prim.ux1);
    init_from_value(exec_space, prim_arrays.ux2, prim.ux2);
}

template <class ElementType, class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
void init_from_state(
        HipExecSpace const& exec_space,
        EulerConsArrays<mdspan<ElementType, extents<IndexType, E0, E1, E2>, layout_left>> const& cons_arrays,
        EulerCons<ElementType> const& cons)
{
    init_from_value(exec_space, cons_arrays.d, cons.d);
    init_from_value(exec_space, cons_arrays.e, cons.e);
    init_from_value(exec_space, cons_arrays.mx0, cons.mx0);
    init_from_value(exec_space, cons_arrays.mx1, cons.mx1);
    init_from_value(exec_space, cons_arrays.mx2, cons.mx2);
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (benchmarks/benchmark_utils.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::int_cast;
using eb::init_from_value;
using eb::init_from_state;
#endif
