// Rank-3 layout_left view: the subset of std::mdspan the Euler operators use
// (the reference spells it Kokkos::mdspan<T, Kokkos::extents<I, E0, E1, E2>,
// Kokkos::layout_left>, euler_operators/godunov.hpp:15-24). The view does not
// own its memory; the data handle is usually device memory, which host code
// must not dereference (operator() is for host buffers).
#pragma once

#include <cstddef>
#include <limits>
#include <stdexcept>
#include <type_traits>

namespace eb {

inline constexpr std::size_t dynamic_extent = std::numeric_limits<std::size_t>::max();

template <class IndexType, std::size_t E0, std::size_t E1, std::size_t E2>
class extents
{
public:
    using index_type = IndexType;
    using size_type = std::make_unsigned_t<IndexType>;
    using rank_type = std::size_t;

    static constexpr rank_type rank() noexcept
    {
        return 3;
    }

    static constexpr std::size_t static_extent(rank_type r) noexcept
    {
        return r == 0 ? E0 : (r == 1 ? E1 : E2);
    }

    constexpr extents() noexcept
        : m_ext {E0 == dynamic_extent ? IndexType(0) : static_cast<IndexType>(E0),
                 E1 == dynamic_extent ? IndexType(0) : static_cast<IndexType>(E1),
                 E2 == dynamic_extent ? IndexType(0) : static_cast<IndexType>(E2)}
    {
    }

    // All three extents; a static extent must equal the value given for it.
    constexpr extents(IndexType n0, IndexType n1, IndexType n2) : m_ext {n0, n1, n2}
    {
        for (rank_type r = 0; r < 3; ++r) {
            if (static_extent(r) != dynamic_extent && static_cast<std::size_t>(m_ext[r]) != static_extent(r)) {
                throw std::invalid_argument("extent differs from the static extent");
            }
        }
    }

    constexpr IndexType extent(rank_type r) const noexcept
    {
        return m_ext[r];
    }

private:
    IndexType m_ext[3];
};

template <class IndexType, std::size_t Rank>
struct dextents_of
{
    static_assert(Rank == 3, "the Euler operators use rank-3 views");
    using type = extents<IndexType, dynamic_extent, dynamic_extent, dynamic_extent>;
};

template <class IndexType, std::size_t Rank>
using dextents = typename dextents_of<IndexType, Rank>::type;

// Column-major: i + n0 * (j + n1 * k).
struct layout_left
{
};

template <class ElementType, class Extents, class LayoutPolicy = layout_left>
class mdspan
{
    static_assert(std::is_same_v<LayoutPolicy, layout_left>, "only layout_left is supported");

public:
    using element_type = ElementType;
    using value_type = std::remove_cv_t<ElementType>;
    using extents_type = Extents;
    using index_type = typename Extents::index_type;
    using layout_type = LayoutPolicy;
    using data_handle_type = ElementType*;
    using reference = ElementType&;

    constexpr mdspan() noexcept : m_ptr(nullptr), m_ext() {}

    constexpr mdspan(data_handle_type p, index_type n0, index_type n1, index_type n2)
        : m_ptr(p), m_ext(n0, n1, n2)
    {
    }

    constexpr mdspan(data_handle_type p, extents_type const& e) noexcept : m_ptr(p), m_ext(e) {}

    // mdspan<T> -> mdspan<T const>
    template <class Other, std::enable_if_t<std::is_same_v<ElementType, Other const>, int> = 0>
    constexpr mdspan(mdspan<Other, Extents, LayoutPolicy> const& o) noexcept
        : m_ptr(o.data_handle()), m_ext(o.extents())
    {
    }

    constexpr data_handle_type data_handle() const noexcept
    {
        return m_ptr;
    }

    constexpr extents_type const& extents() const noexcept
    {
        return m_ext;
    }

    constexpr index_type extent(std::size_t r) const noexcept
    {
        return m_ext.extent(r);
    }

    static constexpr std::size_t rank() noexcept
    {
        return 3;
    }

    constexpr std::size_t size() const noexcept
    {
        return static_cast<std::size_t>(extent(0)) * static_cast<std::size_t>(extent(1))
               * static_cast<std::size_t>(extent(2));
    }

    constexpr index_type mapping(index_type i, index_type j, index_type k) const noexcept
    {
        return i + extent(0) * (j + extent(1) * k);
    }

    // Host memory only.
    constexpr reference operator()(index_type i, index_type j, index_type k) const noexcept
    {
        return m_ptr[mapping(i, j, k)];
    }

private:
    data_handle_type m_ptr;
    extents_type m_ext;
};

} // namespace eb
