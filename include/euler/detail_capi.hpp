// Mapping of the C++ operator templates onto the C ABI (euler_capi.h): the real
// type picks the _f64 / _f32 entry points, the view's IndexType picks the index
// width, and every status other than EB_OK becomes an eb::Error.
#pragma once

#include <cstddef>
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

// 64-bit index types use the int64 kernels; narrower ones the 32-bit kernels,
// which reject extents whose cell count does not fit a signed 32-bit index.
template <class IndexType>
constexpr int index_bits() noexcept
{
    static_assert(std::is_integral_v<IndexType>, "IndexType must be an integer type");
    return sizeof(IndexType) >= 8 ? 64 : 32;
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
