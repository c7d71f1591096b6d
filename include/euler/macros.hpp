// Function attributes of the C++ API's point functions.
//
// EB_FUNCTION: callable from host code and from device code built with hipcc.
// EB_IEEE_ORDER: first statement of a point formula's body; with clang or hipcc
// it switches floating-point contraction off for that body, so the formula
// returns the IEEE result of the reference's operation order (the reference's
// bits) under -ffp-contract=on, off or fast-honor-pragmas (hipcc's default);
// -ffp-contract=fast ignores the pragma. Host code built with GCC needs
// -ffp-contract=off for the same guarantee (GCC has no per-function form).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define EB_FUNCTION __host__ __device__
#else
#define EB_FUNCTION
#endif

#if defined(__clang__)
#define EB_IEEE_ORDER _Pragma("clang fp contract(off)")
#else
#define EB_IEEE_ORDER
#endif
