// .npy output of one field (the role of simulations/save_npy.hpp:90-140): NPY
// format 1.0, little-endian, fortran_order True for the layout_left extents
// (n0, n1, n2), so numpy.load returns an array indexed [i, j, k]. Device fields
// are copied to the host first.
#pragma once

#include <array>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <ostream>
#include <string>
#include <type_traits>

#include "euler_arrays.hpp"

namespace eb {

namespace detail {

template <class T>
char const* npy_descr()
{
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "save_npy writes f8 or f4 fields");
    return std::is_same_v<T, double> ? "<f8" : "<f4";
}

} // namespace detail

template <class T>
void save_npy(std::ostream& os, T const* data, std::array<std::size_t, 3> const& shape)
{
    // Every extent is followed by ", " inside the parentheses, as the reference's
    // writer emits it (simulations/save_npy.cpp:37-46), so the files are byte-identical.
    std::string dims;
    for (std::size_t e : shape) {
        dims += std::to_string(e) + ", ";
    }
    std::string dict = std::string("{'descr': '") + detail::npy_descr<T>() + "', 'fortran_order': True, 'shape': ("
                       + dims + "), }";
    // magic (6) + version (2) + header length (2) + dict + padding + '\n' is a multiple of 64
    std::size_t const unpadded = 10 + dict.size() + 1;
    dict.append((64 - unpadded % 64) % 64, ' ');
    dict.push_back('\n');
    std::uint16_t const hlen = static_cast<std::uint16_t>(dict.size());
    char const head[10] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0, static_cast<char>(hlen & 0xff),
                           static_cast<char>(hlen >> 8)};
    os.write(head, 10);
    os.write(dict.data(), static_cast<std::streamsize>(dict.size()));
    os.write(reinterpret_cast<char const*>(data),
             static_cast<std::streamsize>(shape[0] * shape[1] * shape[2] * sizeof(T)));
}

template <class E, class X>
void save_npy(std::ostream& os, HipExecSpace const& exec_space, mdspan<E, X, layout_left> const& field)
{
    auto const host = to_host(exec_space, field);
    save_npy(os, host.data(),
             {static_cast<std::size_t>(field.extent(0)), static_cast<std::size_t>(field.extent(1)),
              static_cast<std::size_t>(field.extent(2))});
}

template <class E, class X>
void save_npy(std::filesystem::path const& filename, HipExecSpace const& exec_space,
              mdspan<E, X, layout_left> const& field)
{
    std::ofstream os(filename, std::ios::binary);
    if (!os) {
        throw Error(EB_ERR_INVALID_ARGUMENT, "save_npy: cannot open " + filename.string());
    }
    save_npy(os, exec_space, field);
}

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (simulations/save_npy.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::save_npy;
#endif
