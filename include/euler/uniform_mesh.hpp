// Uniform Cartesian mesh (euler_operators/uniform_mesh.hpp:5-52): cell widths,
// face areas ds_d and cell volume dv, callable from host and device code.
#pragma once

#include "macros.hpp"

namespace eb {

template <class T>
class UniformMesh3d
{
public:
    EB_FUNCTION UniformMesh3d(T dx0, T dx1, T dx2) noexcept : m_dx0(dx0), m_dx1(dx1), m_dx2(dx2) {}

    EB_FUNCTION T dx0() const noexcept { return m_dx0; }
    EB_FUNCTION T dx1() const noexcept { return m_dx1; }
    EB_FUNCTION T dx2() const noexcept { return m_dx2; }
    EB_FUNCTION T ds0() const noexcept
    {
        EB_IEEE_ORDER
        return m_dx1 * m_dx2;
    }
    EB_FUNCTION T ds1() const noexcept
    {
        EB_IEEE_ORDER
        return m_dx2 * m_dx0;
    }
    EB_FUNCTION T ds2() const noexcept
    {
        EB_IEEE_ORDER
        return m_dx0 * m_dx1;
    }
    EB_FUNCTION T dv() const noexcept
    {
        EB_IEEE_ORDER
        return m_dx0 * m_dx1 * m_dx2;
    }

private:
    T m_dx0;
    T m_dx1;
    T m_dx2;
};

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/uniform_mesh.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::UniformMesh3d;
#endif
