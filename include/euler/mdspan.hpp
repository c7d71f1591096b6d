// Multidimensional views: the subset of std::mdspan (C++23) and submdspan (C++26)
// the Euler operators and their callers use. The reference spells these
// Kokkos::mdspan<T, Kokkos::extents<I, E0, E1, E2>, Kokkos::layout_left> and
// Kokkos::submdspan (euler_operators/godunov.hpp:15-24, euler_arrays.hpp:106-117).
// Views do not own their memory; the data handle is usually device memory, which
// host code must not dereference. Element access, index mapping and extents are
// callable from host and device code (EB_FUNCTION); constructors that check
// static extents throw and are host-only.
#pragma once

#include <cstddef>
#include <limits>
#include <stdexcept>
#include <tuple>
#include <type_traits>
#include <utility>

#include "macros.hpp"

namespace eb {

inline constexpr std::size_t dynamic_extent = std::numeric_limits<std::size_t>::max();

template <class IndexType, std::size_t... Es>
class extents
{
public:
    using index_type = IndexType;
    using size_type = std::make_unsigned_t<IndexType>;
    using rank_type = std::size_t;

    static_assert(std::is_integral_v<IndexType>, "IndexType must be an integer type");

    static constexpr rank_type rank() noexcept
    {
        return sizeof...(Es);
    }

    static constexpr rank_type rank_dynamic() noexcept
    {
        return ((Es == dynamic_extent ? 1 : 0) + ... + 0);
    }

    static constexpr std::size_t static_extent(rank_type r) noexcept
    {
        constexpr std::size_t e[] = {Es..., 0};
        return e[r];
    }

    // Static extents at their value, dynamic ones 0.
    constexpr extents() noexcept : m_ext {static_cast<IndexType>(Es == dynamic_extent ? 0 : Es)...} {}

    // Either every extent (a static one must equal the value given for it) or
    // only the dynamic ones, in order.
    template <class... Is,
              std::enable_if_t<(sizeof...(Is) > 0) && (sizeof...(Is) == rank() || sizeof...(Is) == rank_dynamic())
                                       && (std::is_convertible_v<Is, IndexType> && ...),
                               int> = 0>
    constexpr explicit extents(Is... is) : m_ext {}
    {
        IndexType const given[] = {static_cast<IndexType>(is)...};
        if constexpr (sizeof...(Is) == rank()) {
            for (rank_type r = 0; r < rank(); ++r) {
                if (static_extent(r) != dynamic_extent && static_cast<std::size_t>(given[r]) != static_extent(r)) {
                    throw std::invalid_argument("extent differs from the static extent");
                }
                m_ext[r] = given[r];
            }
        } else {
            rank_type d = 0;
            for (rank_type r = 0; r < rank(); ++r) {
                m_ext[r] = static_extent(r) == dynamic_extent ? given[d++] : static_cast<IndexType>(static_extent(r));
            }
        }
    }

    EB_FUNCTION constexpr IndexType extent(rank_type r) const noexcept
    {
        return m_ext[r];
    }

private:
    IndexType m_ext[rank() > 0 ? rank() : 1];
};

namespace detail {

template <class IndexType, class Seq>
struct dextents_impl;

template <class IndexType, std::size_t... R>
struct dextents_impl<IndexType, std::index_sequence<R...>>
{
    using type = extents<IndexType, ((void)R, dynamic_extent)...>;
};

} // namespace detail

template <class IndexType, std::size_t Rank>
using dextents = typename detail::dextents_impl<IndexType, std::make_index_sequence<Rank>>::type;

// Column-major: offset i0 + e0 * (i1 + e1 * (i2 + ...)).
struct layout_left
{
    template <class Extents>
    class mapping
    {
    public:
        using extents_type = Extents;
        using index_type = typename Extents::index_type;
        using rank_type = std::size_t;
        using layout_type = layout_left;

        constexpr mapping() noexcept = default;

        constexpr mapping(extents_type const& e) noexcept : m_ext(e) {}

        EB_FUNCTION constexpr extents_type const& extents() const noexcept
        {
            return m_ext;
        }

        template <class... Is>
        EB_FUNCTION constexpr index_type operator()(Is... is) const noexcept
        {
            static_assert(sizeof...(Is) == extents_type::rank(), "one index per rank");
            index_type const idx[] = {static_cast<index_type>(is)..., index_type(0)};
            index_type off = 0;
            for (rank_type r = extents_type::rank(); r-- > 0;) {
                off = idx[r] + m_ext.extent(r) * off;
            }
            return off;
        }

        EB_FUNCTION constexpr index_type stride(rank_type r) const noexcept
        {
            index_type s = 1;
            for (rank_type q = 0; q < r; ++q) {
                s *= m_ext.extent(q);
            }
            return s;
        }

        EB_FUNCTION constexpr index_type required_span_size() const noexcept
        {
            index_type s = 1;
            for (rank_type r = 0; r < extents_type::rank(); ++r) {
                s *= m_ext.extent(r);
            }
            return s;
        }

        static constexpr bool is_always_exhaustive() noexcept
        {
            return true;
        }

    private:
        extents_type m_ext;
    };
};

// Any strides: offset sum_r i_r * stride_r (the layout of most submdspan results).
struct layout_stride
{
    template <class Extents>
    class mapping
    {
    public:
        using extents_type = Extents;
        using index_type = typename Extents::index_type;
        using rank_type = std::size_t;
        using layout_type = layout_stride;
        static constexpr rank_type kRank = extents_type::rank();

        constexpr mapping() noexcept = default;

        template <class S>
        constexpr mapping(extents_type const& e, S const& strides) noexcept : m_ext(e), m_str {}
        {
            for (rank_type r = 0; r < kRank; ++r) {
                m_str[r] = static_cast<index_type>(strides[r]);
            }
        }

        EB_FUNCTION constexpr extents_type const& extents() const noexcept
        {
            return m_ext;
        }

        template <class... Is>
        EB_FUNCTION constexpr index_type operator()(Is... is) const noexcept
        {
            static_assert(sizeof...(Is) == kRank, "one index per rank");
            index_type const idx[] = {static_cast<index_type>(is)..., index_type(0)};
            index_type off = 0;
            for (rank_type r = 0; r < kRank; ++r) {
                off += idx[r] * m_str[r];
            }
            return off;
        }

        EB_FUNCTION constexpr index_type stride(rank_type r) const noexcept
        {
            return m_str[r];
        }

        EB_FUNCTION constexpr index_type required_span_size() const noexcept
        {
            index_type s = 1;
            for (rank_type r = 0; r < kRank; ++r) {
                if (m_ext.extent(r) == 0) {
                    return 0;
                }
                s += (m_ext.extent(r) - 1) * m_str[r];
            }
            return s;
        }

        static constexpr bool is_always_exhaustive() noexcept
        {
            return false;
        }

    private:
        extents_type m_ext;
        index_type m_str[kRank > 0 ? kRank : 1];
    };
};

template <class ElementType, class Extents, class LayoutPolicy = layout_left>
class mdspan
{
public:
    using element_type = ElementType;
    using value_type = std::remove_cv_t<ElementType>;
    using extents_type = Extents;
    using index_type = typename Extents::index_type;
    using size_type = typename Extents::size_type;
    using rank_type = std::size_t;
    using layout_type = LayoutPolicy;
    using mapping_type = typename LayoutPolicy::template mapping<Extents>;
    using data_handle_type = ElementType*;
    using reference = ElementType&;

    constexpr mdspan() noexcept = default;

    // Data handle and extents (all of them, or only the dynamic ones).
    template <class... Is,
              std::enable_if_t<(sizeof...(Is) > 0) && (std::is_convertible_v<Is, index_type> && ...), int> = 0>
    constexpr mdspan(data_handle_type p, Is... is) : m_ptr(p), m_map(extents_type(static_cast<index_type>(is)...))
    {
    }

    constexpr mdspan(data_handle_type p, extents_type const& e) noexcept : m_ptr(p), m_map(e) {}

    constexpr mdspan(data_handle_type p, mapping_type const& m) noexcept : m_ptr(p), m_map(m) {}

    // mdspan<T> -> mdspan<T const>
    template <class Other, std::enable_if_t<std::is_same_v<ElementType, Other const>, int> = 0>
    constexpr mdspan(mdspan<Other, Extents, LayoutPolicy> const& o) noexcept
        : m_ptr(o.data_handle()), m_map(o.mapping())
    {
    }

    EB_FUNCTION constexpr data_handle_type data_handle() const noexcept
    {
        return m_ptr;
    }

    EB_FUNCTION constexpr mapping_type const& mapping() const noexcept
    {
        return m_map;
    }

    EB_FUNCTION constexpr extents_type const& extents() const noexcept
    {
        return m_map.extents();
    }

    EB_FUNCTION constexpr index_type extent(rank_type r) const noexcept
    {
        return m_map.extents().extent(r);
    }

    EB_FUNCTION constexpr index_type stride(rank_type r) const noexcept
    {
        return m_map.stride(r);
    }

    static constexpr rank_type rank() noexcept
    {
        return extents_type::rank();
    }

    static constexpr rank_type rank_dynamic() noexcept
    {
        return extents_type::rank_dynamic();
    }

    static constexpr std::size_t static_extent(rank_type r) noexcept
    {
        return extents_type::static_extent(r);
    }

    // Product of the extents.
    EB_FUNCTION constexpr std::size_t size() const noexcept
    {
        std::size_t s = 1;
        for (rank_type r = 0; r < rank(); ++r) {
            s *= static_cast<std::size_t>(extent(r));
        }
        return s;
    }

    EB_FUNCTION constexpr bool empty() const noexcept
    {
        return size() == 0;
    }

    // The element at the given indices (memory the caller can dereference).
    template <class... Is>
    EB_FUNCTION constexpr reference operator()(Is... is) const noexcept
    {
        return m_ptr[m_map(static_cast<index_type>(is)...)];
    }

private:
    data_handle_type m_ptr = nullptr;
    mapping_type m_map;
};

// ---------------------------------------------------------------- submdspan ---

// Slice specifier keeping a whole extent.
struct full_extent_t
{
    explicit full_extent_t() = default;
};
inline constexpr full_extent_t full_extent {};

namespace detail {

// Slice kinds: an index (drops the rank), the full extent, or a range
// [first, last) given as a pair or two-element tuple.
enum class slice_kind { index, full, range };

template <class S>
constexpr slice_kind kind_of() noexcept
{
    using D = std::remove_cv_t<std::remove_reference_t<S>>;
    if constexpr (std::is_same_v<D, full_extent_t>) {
        return slice_kind::full;
    } else if constexpr (std::is_integral_v<D>) {
        return slice_kind::index;
    } else {
        static_assert(std::tuple_size_v<D> == 2, "a range slice is a pair or tuple {first, last}");
        return slice_kind::range;
    }
}

template <class I, class S>
constexpr I slice_first(S const& s) noexcept
{
    if constexpr (kind_of<S>() == slice_kind::full) {
        return 0;
    } else if constexpr (kind_of<S>() == slice_kind::index) {
        return static_cast<I>(s);
    } else {
        return static_cast<I>(std::get<0>(s));
    }
}

template <class I, class S>
constexpr I slice_extent(S const& s, I ext) noexcept
{
    if constexpr (kind_of<S>() == slice_kind::full) {
        return ext;
    } else if constexpr (kind_of<S>() == slice_kind::index) {
        return 0;
    } else {
        return static_cast<I>(std::get<1>(s)) - static_cast<I>(std::get<0>(s));
    }
}

template <class... S>
constexpr std::size_t sub_rank() noexcept
{
    return ((kind_of<S>() == slice_kind::index ? 0 : 1) + ... + 0);
}

// layout_left survives when the first k-1 slices keep their whole extent, slice k-1
// is a full extent or a range and every later slice is an index (k = sub rank), as
// std::submdspan specifies for layout_left.
template <class... S>
constexpr bool keeps_layout_left() noexcept
{
    constexpr slice_kind kinds[] = {kind_of<S>()..., slice_kind::index};
    constexpr std::size_t k = sub_rank<S...>();
    if constexpr (k == 0) {
        return true;
    } else {
        for (std::size_t r = 0; r < sizeof...(S); ++r) {
            slice_kind const want = r + 1 < k ? slice_kind::full : (r + 1 == k ? kinds[r] : slice_kind::index);
            if (kinds[r] != want || (r + 1 == k && kinds[r] == slice_kind::index)) {
                return false;
            }
        }
        return true;
    }
}

template <class I, std::size_t K>
struct small_array
{
    I v[K > 0 ? K : 1];

    constexpr I operator[](std::size_t r) const noexcept
    {
        return v[r];
    }
};

template <class I, std::size_t K, std::size_t... R>
constexpr dextents<I, K> make_dextents(small_array<I, K> const& e, std::index_sequence<R...>)
{
    return dextents<I, K>(e.v[R]...);
}

} // namespace detail

// View of the part of m selected by one slice per rank (Kokkos::submdspan as
// called by euler_arrays.hpp:106-117): an index drops the rank, full_extent keeps
// it, {first, last} keeps [first, last). The result has dynamic extents and
// layout_left where std::submdspan keeps it, else layout_stride.
template <class E, class X, class L, class... S>
auto submdspan(mdspan<E, X, L> const& m, S const&... slices)
{
    static_assert(sizeof...(S) == X::rank(), "one slice per rank");
    using I = typename X::index_type;
    constexpr std::size_t K = detail::sub_rank<S...>();
    I const first[] = {detail::slice_first<I>(slices)..., I(0)};
    detail::slice_kind const kinds[] = {detail::kind_of<S>()..., detail::slice_kind::index};
    detail::small_array<I, K> ext {};
    detail::small_array<I, K> str {};
    I offset = 0;
    {
        std::size_t r = 0;
        std::size_t k = 0;
        auto take = [&](auto const& s) {
            offset += first[r] * m.stride(r);
            if (kinds[r] != detail::slice_kind::index) {
                ext.v[k] = detail::slice_extent<I>(s, m.extent(r));
                str.v[k] = m.stride(r);
                ++k;
            }
            ++r;
        };
        (take(slices), ...);
    }
    auto const sub_ext = detail::make_dextents(ext, std::make_index_sequence<K> {});
    if constexpr (std::is_same_v<L, layout_left> && detail::keeps_layout_left<S...>()) {
        return mdspan<E, dextents<I, K>, layout_left>(m.data_handle() + offset, sub_ext);
    } else {
        using M = typename layout_stride::template mapping<dextents<I, K>>;
        return mdspan<E, dextents<I, K>, layout_stride>(m.data_handle() + offset, M(sub_ext, str));
    }
}

} // namespace eb
