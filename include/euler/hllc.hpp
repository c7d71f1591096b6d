// HLLC Riemann solver tag (euler_operators/hllc.hpp:50-105). The flux itself is
// evaluated inside the GPU kernels; on the host the solver is only selected.
#pragma once

namespace eb {

struct hllc
{
};

} // namespace eb
