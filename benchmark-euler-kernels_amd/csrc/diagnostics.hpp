// Conservation diagnostics in FP64 with Neumaier-compensated sums (no reference
// counterpart; the reference computes no totals). Deterministic: every sum is
// taken in a fixed order (thread-strided per block, then blocks in index order),
// so repeated runs give identical bits.
//
//   k_conserved_partials + k_compensated_final: totals over the interior cells of
//       the five conservative fields (d, e, mx0, mx1, mx2).
//   k_boundary_partials + k_compensated_final: the amount B by which one Godunov
//       step lowers those totals, B = dtodv * sum_d ds_d (sum of F over the high
//       boundary faces - sum over the low boundary faces) (godunov.hpp:48-90
//       summed over the interior: interior faces telescope), accumulated across
//       steps as a compensated (sum, compensation) pair per field.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "kernels.hpp"

namespace eb {

// Neumaier's compensated addition: s + c carries the running sum.
__device__ __forceinline__ void neumaier_add(double& s, double& c, double x)
{
    double const t = s + x;
    if (fabs(s) >= fabs(x)) {
        c += (s - t) + x;
    } else {
        c += (x - t) + s;
    }
    s = t;
}

constexpr int kDiagThreads = 256;

// Block-level combine of per-thread (s, c) pairs for 5 fields, in thread order;
// writes the block's pairs to partials[10*block .. 10*block+9].
__device__ __forceinline__ void block_pairs_out(double const (&s)[5], double const (&c)[5], double* partials)
{
    __shared__ double ls[5][kDiagThreads];
    __shared__ double lc[5][kDiagThreads];
    int const t = threadIdx.x;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        ls[f][t] = s[f];
        lc[f][t] = c[f];
    }
    __syncthreads();
    if (t < 5) {
        double S = 0.0;
        double C = 0.0;
        for (int u = 0; u < kDiagThreads; ++u) {
            neumaier_add(S, C, ls[t][u]);
            C += lc[t][u];
        }
        partials[10 * blockIdx.x + 2 * t] = S;
        partials[10 * blockIdx.x + 2 * t + 1] = C;
    }
}

// Interior totals: block b takes interior rows (j, k) b, b + G, ...; threads stride i.
template <class T, class I>
__global__ __launch_bounds__(kDiagThreads) void k_conserved_partials(ConsPtrs<T const> cons, I n0, I n1, I n2,
                                                                     double* partials)
{
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    I const ny = n1 - 2;
    I const rows = ny * (n2 - 2);
    for (I r = blockIdx.x; r < rows; r += gridDim.x) {
        I const j = 1 + r % ny;
        I const k = 1 + r / ny;
        I const base = n0 * (j + n1 * k);
        for (I i = 1 + static_cast<I>(threadIdx.x); i <= n0 - 2; i += kDiagThreads) {
            I const idx = base + i;
            neumaier_add(s[0], c[0], static_cast<double>(cons.d[idx]));
            neumaier_add(s[1], c[1], static_cast<double>(cons.e[idx]));
            neumaier_add(s[2], c[2], static_cast<double>(cons.mx0[idx]));
            neumaier_add(s[3], c[3], static_cast<double>(cons.mx1[idx]));
            neumaier_add(s[4], c[4], static_cast<double>(cons.mx2[idx]));
        }
    }
    block_pairs_out(s, c, partials);
}

template <int Dir, class T, class I>
__device__ __forceinline__ Flux<T> boundary_face(PrimPtrs<T const> const& prim, I a, I b, T gamma, T gm1)
{
    Prim<T> const qa = load_prim(prim, a);
    Prim<T> const qb = load_prim(prim, b);
    return hllc_flux<Dir>(qa, sound_speed(gamma, qa.d, qa.p), qb, sound_speed(gamma, qb.d, qb.p), gm1);
}

// One step's boundary outflow B. Faces: x (per interior (j,k)), y (per interior
// (i,k)), z (per interior (i,j)); low side first, then high side, in each group.
template <class T, class I>
__global__ __launch_bounds__(kDiagThreads) void k_boundary_partials(PrimPtrs<T const> prim, I n0, I n1, I n2,
                                                                    T gamma, T gm1, T ds0, T ds1, T ds2,
                                                                    DtArg<T> dta, double* partials)
{
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    I const nx = n0 - 2;
    I const ny = n1 - 2;
    I const nz = n2 - 2;
    I const fx = ny * nz;
    I const fy = nx * nz;
    I const fz = nx * ny;
    I const total = 2 * (fx + fy + fz);
    double const dtodv = static_cast<double>(resolve_dtodv(dta));
    I const stride = static_cast<I>(gridDim.x) * kDiagThreads;
    for (I g = static_cast<I>(blockIdx.x) * kDiagThreads + threadIdx.x; g < total; g = grid_next(g, total, stride)) {
        Flux<T> f;
        double w;
        if (g < 2 * fx) {
            I const h = g % fx;
            bool const high = g >= fx;
            I const j = 1 + h % ny;
            I const k = 1 + h / ny;
            I const a = (high ? n0 - 2 : 0) + n0 * (j + n1 * k);
            f = boundary_face<0>(prim, a, a + 1, gamma, gm1);
            w = static_cast<double>(ds0) * (high ? 1.0 : -1.0);
        } else if (g < 2 * (fx + fy)) {
            I const g2 = g - 2 * fx;
            I const h = g2 % fy;
            bool const high = g2 >= fy;
            I const i = 1 + h % nx;
            I const k = 1 + h / nx;
            I const a = i + n0 * ((high ? n1 - 2 : 0) + n1 * k);
            f = boundary_face<1>(prim, a, a + n0, gamma, gm1);
            w = static_cast<double>(ds1) * (high ? 1.0 : -1.0);
        } else {
            I const g3 = g - 2 * (fx + fy);
            I const h = g3 % fz;
            bool const high = g3 >= fz;
            I const i = 1 + h % nx;
            I const j = 1 + h / nx;
            I const a = i + n0 * (j + n1 * (high ? n2 - 2 : 0));
            f = boundary_face<2>(prim, a, a + n0 * n1, gamma, gm1);
            w = static_cast<double>(ds2) * (high ? 1.0 : -1.0);
        }
        double const k2 = dtodv * w;
        neumaier_add(s[0], c[0], k2 * static_cast<double>(f.d));
        neumaier_add(s[1], c[1], k2 * static_cast<double>(f.e));
        neumaier_add(s[2], c[2], k2 * static_cast<double>(f.m0));
        neumaier_add(s[3], c[3], k2 * static_cast<double>(f.m1));
        neumaier_add(s[4], c[4], k2 * static_cast<double>(f.m2));
    }
    block_pairs_out(s, c, partials);
}

// Combines nblocks partial pairs in block order. accumulate = false: out[f] =
// total; accumulate = true: (out[2f], out[2f+1]) += total as a compensated pair.
__global__ void k_compensated_final(double const* partials, int nblocks, double* out, int accumulate)
{
    int const f = threadIdx.x;
    if (f >= 5) {
        return;
    }
    double S = 0.0;
    double C = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        neumaier_add(S, C, partials[10 * b + 2 * f]);
        C += partials[10 * b + 2 * f + 1];
    }
    if (accumulate != 0) {
        double s = out[2 * f];
        double c = out[2 * f + 1];
        neumaier_add(s, c, S);
        c += C;
        out[2 * f] = s;
        out[2 * f + 1] = c;
    } else {
        out[f] = S + C;
    }
}

} // namespace eb
