// Perfect-gas equation of state (euler_operators/perfect_gas.hpp:6-29). The GPU
// kernels take gamma by value; these members are the host-side formulas.
#pragma once

#include <cmath>

namespace eb {

template <class T>
class PerfectGas
{
public:
    explicit PerfectGas(T const gamma) : m_gamma(gamma) {}

    T gamma() const noexcept
    {
        return m_gamma;
    }

    T speed_of_sound(T const density, T const pressure) const noexcept
    {
        return std::sqrt(m_gamma * pressure / density);
    }

    T internal_energy(T const /*density*/, T const pressure) const noexcept
    {
        return pressure / (m_gamma - 1);
    }

    T pressure(T const /*density*/, T const int_e) const noexcept
    {
        return (m_gamma - 1) * int_e;
    }

private:
    T m_gamma;
};

} // namespace eb
