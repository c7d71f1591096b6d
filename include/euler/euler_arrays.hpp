// SoA state arrays (euler_operators/euler_arrays.hpp:8-411): per-cell structs
// EulerPrim / EulerCons / EulerFlux, the five-field containers EulerPrimArrays /
// EulerConsArrays over any view type, allocation of zero-initialised 1-D device
// fields, their rank-3 layout_left views and the size helpers.
#pragma once

#include <cstddef>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../euler_capi.h"
#include "exec_space.hpp"
#include "macros.hpp"
#include "mdspan.hpp"

namespace eb {

template <class T>
struct EulerPrim
{
    T d;
    T p;
    T ux0;
    T ux1;
    T ux2;
};

template <class T>
struct EulerCons
{
    T d;
    T e;
    T mx0;
    T mx1;
    T mx2;
};

template <class T>
struct EulerFlux
{
    T d;
    T e;
    T mx0;
    T mx1;
    T mx2;
};

// Point formulas in the reference's operation order (euler_arrays.hpp:38-75),
// callable from host and device code.
template <class T>
EB_FUNCTION T kinetic_energy(EulerPrim<T> const& prim) noexcept
{
    EB_IEEE_ORDER
    return prim.d * ((prim.ux0 * prim.ux0) + (prim.ux1 * prim.ux1) + (prim.ux2 * prim.ux2)) / 2;
}

template <class T>
EB_FUNCTION T kinetic_energy(EulerCons<T> const& cons) noexcept
{
    EB_IEEE_ORDER
    return ((cons.mx0 * cons.mx0) + (cons.mx1 * cons.mx1) + (cons.mx2 * cons.mx2)) / cons.d / 2;
}

template <class T>
EB_FUNCTION T internal_energy(EulerCons<T> const& cons) noexcept
{
    EB_IEEE_ORDER
    return cons.e - kinetic_energy(cons);
}

template <class T>
EB_FUNCTION EulerCons<T> to_cons(EulerPrim<T> const& prim, T const int_e) noexcept
{
    EB_IEEE_ORDER
    return {prim.d, kinetic_energy(prim) + int_e, prim.d * prim.ux0, prim.d * prim.ux1, prim.d * prim.ux2};
}

template <class T>
EB_FUNCTION EulerPrim<T> to_prim(EulerCons<T> const& cons, T const p) noexcept
{
    EB_IEEE_ORDER
    T const vol_spe = 1 / cons.d;
    return {cons.d, p, cons.mx0 * vol_spe, cons.mx1 * vol_spe, cons.mx2 * vol_spe};
}

// Owning 1-D device allocation with shared ownership (the role of Kokkos::View<T*>).
template <class T>
class View
{
public:
    View() = default;

    View(HipExecSpace const& exec_space, std::string label, std::size_t n) : m_label(std::move(label)), m_size(n)
    {
        T* p = nullptr;
        if (n > 0) {
            detail::check_hip(hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)), "hipMalloc");
            m_data = std::shared_ptr<T>(p, [](T* q) { (void)hipFree(q); });
            detail::check_hip(hipMemsetAsync(p, 0, n * sizeof(T), exec_space.stream()), "hipMemsetAsync");
        }
    }

    T* data() const noexcept
    {
        return m_data.get();
    }

    std::size_t size() const noexcept
    {
        return m_size;
    }

    std::string const& label() const noexcept
    {
        return m_label;
    }

private:
    std::shared_ptr<T> m_data;
    std::string m_label;
    std::size_t m_size = 0;
};

template <class ViewT>
struct EulerPrimArrays
{
    ViewT d;
    ViewT p;
    ViewT ux0;
    ViewT ux1;
    ViewT ux2;
};

template <class ViewT>
struct EulerConsArrays
{
    ViewT d;
    ViewT e;
    ViewT mx0;
    ViewT mx1;
    ViewT mx2;
};

template <class T>
EulerPrimArrays<View<T>> create_prim_arrays_1d(HipExecSpace const& exec_space, std::size_t const n)
{
    return {View<T>(exec_space, "density", n), View<T>(exec_space, "pressure", n), View<T>(exec_space, "ux0", n),
            View<T>(exec_space, "ux1", n), View<T>(exec_space, "ux2", n)};
}

template <class T>
EulerConsArrays<View<T>> create_cons_arrays_1d(HipExecSpace const& exec_space, std::size_t const n)
{
    return {View<T>(exec_space, "density", n), View<T>(exec_space, "energy", n), View<T>(exec_space, "mx0", n),
            View<T>(exec_space, "mx1", n), View<T>(exec_space, "mx2", n)};
}

namespace detail {

template <class MdspanOut, class V>
MdspanOut view_as(V const& v, typename MdspanOut::index_type nx, typename MdspanOut::index_type ny,
                  typename MdspanOut::index_type nz)
{
    if (static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny) * static_cast<std::size_t>(nz) > v.size()) {
        throw Error(EB_ERR_INVALID_ARGUMENT, "to_mdspan: extents exceed the allocation");
    }
    return MdspanOut(v.data(), nx, ny, nz);
}

} // namespace detail

// 3-D layout_left views of the 1-D fields (euler_arrays.hpp:130-142, 281-293).
template <class MdspanOut, class T>
EulerPrimArrays<MdspanOut> to_mdspan(EulerPrimArrays<View<T>> const& a, typename MdspanOut::index_type nx,
                                     typename MdspanOut::index_type ny, typename MdspanOut::index_type nz)
{
    using detail::view_as;
    return {view_as<MdspanOut>(a.d, nx, ny, nz), view_as<MdspanOut>(a.p, nx, ny, nz),
            view_as<MdspanOut>(a.ux0, nx, ny, nz), view_as<MdspanOut>(a.ux1, nx, ny, nz),
            view_as<MdspanOut>(a.ux2, nx, ny, nz)};
}

template <class MdspanOut, class T>
EulerConsArrays<MdspanOut> to_mdspan(EulerConsArrays<View<T>> const& a, typename MdspanOut::index_type nx,
                                     typename MdspanOut::index_type ny, typename MdspanOut::index_type nz)
{
    using detail::view_as;
    return {view_as<MdspanOut>(a.d, nx, ny, nz), view_as<MdspanOut>(a.e, nx, ny, nz),
            view_as<MdspanOut>(a.mx0, nx, ny, nz), view_as<MdspanOut>(a.mx1, nx, ny, nz),
            view_as<MdspanOut>(a.mx2, nx, ny, nz)};
}

// Element-type conversion of the views (euler_arrays.hpp:87-104, 238-255).
template <class MdspanOut, class E, class X, class L>
EulerPrimArrays<MdspanOut> as(EulerPrimArrays<mdspan<E, X, L>> const& a) noexcept
{
    return {a.d, a.p, a.ux0, a.ux1, a.ux2};
}

template <class MdspanOut, class E, class X, class L>
EulerConsArrays<MdspanOut> as(EulerConsArrays<mdspan<E, X, L>> const& a) noexcept
{
    return {a.d, a.e, a.mx0, a.mx1, a.mx2};
}

template <class T, class X, class L>
EulerPrimArrays<mdspan<T const, X, L>> as_const(EulerPrimArrays<mdspan<T, X, L>> const& a) noexcept
{
    return as<mdspan<T const, X, L>>(a);
}

template <class T, class X, class L>
EulerConsArrays<mdspan<T const, X, L>> as_const(EulerConsArrays<mdspan<T, X, L>> const& a) noexcept
{
    return as<mdspan<T const, X, L>>(a);
}

// Raw field pointers (euler_arrays.hpp:119-128, 270-279).
template <class E, class X, class L>
EulerPrimArrays<E*> data_handle(EulerPrimArrays<mdspan<E, X, L>> const& a) noexcept
{
    return {a.d.data_handle(), a.p.data_handle(), a.ux0.data_handle(), a.ux1.data_handle(), a.ux2.data_handle()};
}

template <class E, class X, class L>
EulerConsArrays<E*> data_handle(EulerConsArrays<mdspan<E, X, L>> const& a) noexcept
{
    return {a.d.data_handle(), a.e.data_handle(), a.mx0.data_handle(), a.mx1.data_handle(), a.mx2.data_handle()};
}

// One cell of the five fields as a point struct, and back (euler_arrays.hpp:144-214,
// 295-365): through rank-3 views at (i, j, k), or through raw field pointers at a
// linear index. Callable from host code on host memory and from device code.
template <class E, class I, std::size_t E0, std::size_t E1, std::size_t E2, class L>
EB_FUNCTION EulerPrim<std::remove_const_t<E>> load(EulerPrimArrays<mdspan<E, extents<I, E0, E1, E2>, L>> const& a,
                                                   I const i, I const j, I const k) noexcept
{
    return {a.d(i, j, k), a.p(i, j, k), a.ux0(i, j, k), a.ux1(i, j, k), a.ux2(i, j, k)};
}

template <class E, class I, std::size_t E0, std::size_t E1, std::size_t E2, class L>
EB_FUNCTION EulerCons<std::remove_const_t<E>> load(EulerConsArrays<mdspan<E, extents<I, E0, E1, E2>, L>> const& a,
                                                   I const i, I const j, I const k) noexcept
{
    return {a.d(i, j, k), a.e(i, j, k), a.mx0(i, j, k), a.mx1(i, j, k), a.mx2(i, j, k)};
}

template <class E, class I>
EB_FUNCTION EulerPrim<std::remove_const_t<E>> load(EulerPrimArrays<E*> const& a, I const i) noexcept
{
    return {a.d[i], a.p[i], a.ux0[i], a.ux1[i], a.ux2[i]};
}

template <class E, class I>
EB_FUNCTION EulerCons<std::remove_const_t<E>> load(EulerConsArrays<E*> const& a, I const i) noexcept
{
    return {a.d[i], a.e[i], a.mx0[i], a.mx1[i], a.mx2[i]};
}

template <class T, class I, std::size_t E0, std::size_t E1, std::size_t E2, class L>
EB_FUNCTION void store(EulerPrim<T> const& q, EulerPrimArrays<mdspan<T, extents<I, E0, E1, E2>, L>> const& a,
                       I const i, I const j, I const k) noexcept
{
    a.d(i, j, k) = q.d;
    a.p(i, j, k) = q.p;
    a.ux0(i, j, k) = q.ux0;
    a.ux1(i, j, k) = q.ux1;
    a.ux2(i, j, k) = q.ux2;
}

template <class T, class I, std::size_t E0, std::size_t E1, std::size_t E2, class L>
EB_FUNCTION void store(EulerCons<T> const& u, EulerConsArrays<mdspan<T, extents<I, E0, E1, E2>, L>> const& a,
                       I const i, I const j, I const k) noexcept
{
    a.d(i, j, k) = u.d;
    a.e(i, j, k) = u.e;
    a.mx0(i, j, k) = u.mx0;
    a.mx1(i, j, k) = u.mx1;
    a.mx2(i, j, k) = u.mx2;
}

template <class T, class I>
EB_FUNCTION void store(EulerPrim<T> const& q, EulerPrimArrays<T*> const& a, I const i) noexcept
{
    a.d[i] = q.d;
    a.p[i] = q.p;
    a.ux0[i] = q.ux0;
    a.ux1[i] = q.ux1;
    a.ux2[i] = q.ux2;
}

template <class T, class I>
EB_FUNCTION void store(EulerCons<T> const& u, EulerConsArrays<T*> const& a, I const i) noexcept
{
    a.d[i] = u.d;
    a.e[i] = u.e;
    a.mx0[i] = u.mx0;
    a.mx1[i] = u.mx1;
    a.mx2[i] = u.mx2;
}

// The same slices of the five fields (euler_arrays.hpp:106-117, 257-268): one
// slice per rank, as eb::submdspan takes them.
template <class E, class X, class L, class... S>
auto subspan(EulerPrimArrays<mdspan<E, X, L>> const& a, S const&... slices)
{
    using Sub = decltype(submdspan(a.d, slices...));
    return EulerPrimArrays<Sub> {submdspan(a.d, slices...), submdspan(a.p, slices...), submdspan(a.ux0, slices...),
                                 submdspan(a.ux1, slices...), submdspan(a.ux2, slices...)};
}

template <class E, class X, class L, class... S>
auto subspan(EulerConsArrays<mdspan<E, X, L>> const& a, S const&... slices)
{
    using Sub = decltype(submdspan(a.d, slices...));
    return EulerConsArrays<Sub> {submdspan(a.d, slices...), submdspan(a.e, slices...), submdspan(a.mx0, slices...),
                                 submdspan(a.mx1, slices...), submdspan(a.mx2, slices...)};
}

// Elements per field and bytes of the five fields (euler_arrays.hpp:367-399).
template <class E, class X, class L>
std::size_t size(EulerPrimArrays<mdspan<E, X, L>> const& a) noexcept
{
    return a.d.size();
}

template <class E, class X, class L>
std::size_t size(EulerConsArrays<mdspan<E, X, L>> const& a) noexcept
{
    return a.d.size();
}

template <class E, class X, class L>
std::size_t size_bytes(EulerPrimArrays<mdspan<E, X, L>> const& a) noexcept
{
    return 5 * sizeof(E) * size(a);
}

template <class E, class X, class L>
std::size_t size_bytes(EulerConsArrays<mdspan<E, X, L>> const& a) noexcept
{
    return 5 * sizeof(E) * size(a);
}

// Device field -> host vector (blocking on the exec space), for output and checks.
template <class E, class X, class L>
std::vector<std::remove_const_t<E>> to_host(HipExecSpace const& exec_space, mdspan<E, X, L> const& field)
{
    std::vector<std::remove_const_t<E>> out(field.size());
    if (!out.empty()) {
        detail::check_hip(hipMemcpyAsync(out.data(), field.data_handle(), out.size() * sizeof(E),
                                         hipMemcpyDeviceToHost, exec_space.stream()),
                          "hipMemcpyAsync");
    }
    exec_space.fence();
    return out;
}

// Host vector -> device field (blocking).
template <class T, class X, class L>
void from_host(HipExecSpace const& exec_space, std::vector<T> const& src, mdspan<T, X, L> const& field)
{
    if (src.size() != field.size()) {
        throw Error(EB_ERR_INVALID_ARGUMENT, "from_host: size mismatch");
    }
    if (!src.empty()) {
        detail::check_hip(hipMemcpyAsync(field.data_handle(), src.data(), src.size() * sizeof(T),
                                         hipMemcpyHostToDevice, exec_space.stream()),
                          "hipMemcpyAsync");
    }
    exec_space.fence();
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/euler_arrays.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::EulerPrim;
using eb::EulerCons;
using eb::EulerFlux;
using eb::kinetic_energy;
using eb::internal_energy;
using eb::to_cons;
using eb::to_prim;
using eb::EulerPrimArrays;
using eb::EulerConsArrays;
using eb::create_prim_arrays_1d;
using eb::create_cons_arrays_1d;
using eb::to_mdspan;
using eb::as;
using eb::as_const;
using eb::data_handle;
using eb::size;
using eb::size_bytes;
using eb::load;
using eb::store;
using eb::subspan;
#endif
