// C++ API of the MI355X Godunov step: the reference's euler_operators interface
// (euler_arrays, perfect_gas, uniform_mesh, hllc, cons_to_prim, prim_to_cons,
// time_step, godunov, init_implode) over the C ABI of libeuler_hip.so, plus the
// fused step loop, .npy output and the benchmark harness helpers.
#pragma once

#include "benchmark_utils.hpp"
#include "cons_to_prim.hpp"
#include "euler_arrays.hpp"
#include "exec_space.hpp"
#include "godunov.hpp"
#include "hllc.hpp"
#include "init_implode.hpp"
#include "mdspan.hpp"
#include "perfect_gas.hpp"
#include "prim_to_cons.hpp"
#include "save_npy.hpp"
#include "step_loop.hpp"
#include "time_step.hpp"
#include "uniform_mesh.hpp"
