// Perfect-gas equation of state (euler_operators/perfect_gas.hpp:6-29). The
// library's kernels take gamma by value; these members are the same formulas for
// host code and for device code built with hipcc.
#pragma once

#include <cmath>

#include "macros.hpp"

namespace eb {

template <class T>
class PerfectGas
{
public:
    EB_FUNCTION explicit PerfectGas(T const gamma) : m_gamma(gamma) {}

    EB_FUNCTION T gamma() const noexcept
    {
        return m_gamma;
    }

    EB_FUNCTION T speed_of_sound(T const density, T const pressure) const noexcept
    {
        EB_IEEE_ORDER
        return std::sqrt(m_gamma * pressure / density);
    }

    EB_FUNCTION T internal_energy(T const /*density*/, T const pressure) const noexcept
    {
        EB_IEEE_ORDER
        return pressure / (m_gamma - 1);
    }

    EB_FUNCTION T pressure(T const /*density*/, T const int_e) const noexcept
    {
        EB_IEEE_ORDER
        return (m_gamma - 1) * int_e;
    }

private:
    T m_gamma;
};

} // namespace eb

#ifndef EB_NO_GLOBAL_NAMES
// The reference declares these at global scope (euler_operators/perfect_gas.hpp); the same
// unqualified spellings name them here. Define EB_NO_GLOBAL_NAMES to keep them
// in namespace eb only.
using eb::PerfectGas;
#endif
