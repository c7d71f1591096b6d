// Mapping of the C++ operator templates onto the C ABI (euler_capi.h): the real
// type picks the _f64 / _f32 entry points, the view's IndexType picks the index
// width, and every status other than EB_OK becomes an eb::Error.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>

#include "../euler_capi.h"
#include "euler_arrays.hpp"
#include "exec_space.hpp"
#include "mdspan.hpp"

namespace eb::detail {

template <class T>
struct capi;

#define EB_CAPI_TRAITS(T, SUF)                                                  \
    template <>                                                                 \
    struct capi<T>                                                              \
    {                                                                           \
        using prim = eb_prim_##SUF;                                             \
        using cons = eb_cons_##SUF;                                             \
        static constexpr auto cons_to_prim = &eb_cons_to_prim_##SUF;            \
        static constexpr auto prim_to_cons = &eb_prim_to_cons_##SUF;            \
        static constexpr auto time_step = &eb_time_step_##SUF;                  \
        static constexpr auto time_step_sync = &eb_time_step_sync_##SUF;        \
        static constexpr auto godunov = &eb_godunov_##SUF;                      \
        static constexpr auto godunov_dev = &eb_godunov_dev_##SUF;              \
        static constexpr auto godunov_reference_order = &eb_godunov_reference_order_##SUF; \
        static constexpr auto ghost_refresh = &eb_ghost_refresh_##SUF;          \
        static constexpr auto step_fused = &eb_step_fused_##SUF;                \
        static constexpr auto init_implode = &eb_init_implode_##SUF;            \
        static constexpr auto init_diagonal = &eb_init_diagonal_##SUF;          \
        static constexpr auto init_random = &eb_init_random_##SUF;              \
        static constexpr auto fill = &eb_fill_##SUF;                            \
    };
EB_CAPI_TRAITS(double, f64)
EB_CAPI_TRAITS(float, f32)
#undef EB_CAPI_TRAITS

// Index width of the kernels for arrays of the given extents. 64-bit index types
// use the int64 kernels. Narrower ones use the 32-bit kernels up to
// EB_MAX_CELLS_INDEX32 cells, and the int64 kernels (same results) above it while
// the cell count still fits IndexType: `int` reaches 2^31 - 1 cells and `unsigned`
// 2^32 - 1, as with the reference's index_t (benchmarks/index_type.hpp:3-4). A
// count past IndexType is sent to the 32-bit kernels, which reject it with
// EB_ERR_INDEX_OVERFLOW.
template <class IndexType>
int index_bits(int64_t n0, int64_t n1, int64_t n2) noexcept
{
    static_assert(std::is_integral_v<IndexType>, "IndexType must be an integer type");
    if constexpr (sizeof(IndexType) >= 8) {
        return 64;
    } else {
        if (n0 < 0 || n1 < 0 || n2 < 0) {
            return 32; // rejected by the C ABI as an invalid argument
        }
        unsigned __int128 const total = static_cast<unsigned __int128>(n0) * static_cast<unsigned __int128>(n1)
                                        * static_cast<unsigned __int128>(n2);
        if (total <= static_cast<unsigned __int128>(EB_MAX_CELLS_INDEX32)) {
            return 32;
        }
        return total <= static_cast<unsigned __int128>(std::numeric_limits<IndexType>::max()) ? 64 : 32;
    }
}

template <class IndexType, class A>
int index_bits(A const& arrays) noexcept
{
    return index_bits<IndexType>(static_cast<int64_t>(arrays.d.extent(0)), static_cast<int64_t>(arrays.d.extent(1)),
                                 static_cast<int64_t>(arrays.d.extent(2)));
}

template <class E, class X>
void same_extents(mdspan<E, X, layout_left> const& ref, mdspan<E, X, layout_left> const& f)
{
    for (std::size_t r = 0; r < 3; ++r) {
        if (ref.extent(r) != f.extent(r)) {
            throw Error(EB_ERR_INVALID_ARGUMENT, "fields of one state differ in extents");
        }
    }
}

template <class E, class X>
void check_fields(EulerPrimArrays<mdspan<E, X, layout_left>> const& a)
{
    same_extents(a.d, a.p);
    same_extents(a.d, a.ux0);
    same_extents(a.d, a.ux1);
    same_extents(a.d, a.ux2);
}

template <class E, class X>
void check_fields(EulerConsArrays<mdspan<E, X, layout_left>> const& a)
{
    same_extents(a.d, a.e);
    same_extents(a.d, a.mx0);
    same_extents(a.d, a.mx1);
    same_extents(a.d, a.mx2);
}

template <class A, class B>
void check_pair(A const& a, B const& b)
{
    check_fields(a);
    check_fields(b);
    for (std::size_t r = 0; r < 3; ++r) {
        if (static_cast<std::size_t>(a.d.extent(r)) != static_cast<std::size_t>(b.d.extent(r))) {
            throw Error(EB_ERR_INVALID_ARGUMENT, "prim and cons arrays differ in extents");
        }
    }
}

template <class E, class X>
auto ptrs(EulerPrimArrays<mdspan<E, X, layout_left>> const& a)
{
    using T = std::remove_const_t<E>;
    return typename capi<T>::prim {const_cast<T*>(a.d.data_handle()), const_cast<T*>(a.p.data_handle()),
                                   const_cast<T*>(a.ux0.data_handle()), const_cast<T*>(a.ux1.data_handle()),
                                   const_cast<T*>(a.ux2.data_handle())};
}

template <class E, class X>
auto ptrs(EulerConsArrays<mdspan<E, X, layout_left>> const& a)
{
    using T = std::remove_const_t<E>;
    return typename capi<T>::cons {const_cast<T*>(a.d.data_handle()), const_cast<T*>(a.e.data_handle()),
                                   const_cast<T*>(a.mx0.data_handle()), const_cast<T*>(a.mx1.data_handle()),
                                   const_cast<T*>(a.mx2.data_handle())};
}

template <class A>
int64_t ext(A const& a, std::size_t r)
{
    return static_cast<int64_t>(a.d.extent(r));
}

} // namespace eb::detail
