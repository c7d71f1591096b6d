// HIP kernels of the Godunov step for gfx950 (CDNA4, wave64).
//
// Arrays are the reference's SoA layout (euler_arrays.hpp:77-85, 228-236): five
// separate fields, layout_left, linear index i + n0*(j + n1*k), extents including
// the one-cell ghost layer. I is the index type (int32_t or int64_t).
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "reduce.hpp"

namespace eb {

// dt/dv for the update: either a host value, or cfl * dt_dev[0] / dv evaluated on
// the device in the same order the reference's host code uses
// (euler_simulation.cpp:65 then godunov.hpp:32).
template <class T>
struct DtArg
{
    T dtodv;
    T const* dt_dev;
    T cfl;
    T dv;
};

template <class T>
__device__ __forceinline__ T resolve_dtodv(DtArg<T> const& a)
{
    if (a.dt_dev != nullptr) {
        T const dt = a.cfl * a.dt_dev[0];
        return dt / a.dv;
    }
    return a.dtodv;
}

// ---------------------------------------------------------------- pointwise ---

// cons_to_prim.hpp:20-32, all cells (grid-stride so any extent fits the grid).
template <class T, class I>
__global__ __launch_bounds__(256) void k_cons_to_prim(ConsPtrs<T const> cons, PrimPtrs<T> prim, I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx += stride) {
        store_prim(prim, idx, cons_to_prim_pt(load_cons(cons, idx), gm1));
    }
}

// prim_to_cons.hpp:20-32, all cells.
template <class T, class I>
__global__ __launch_bounds__(256) void k_prim_to_cons(PrimPtrs<T const> prim, ConsPtrs<T> cons, I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx += stride) {
        store_cons(cons, idx, prim_to_cons_pt(load_prim(prim, idx), gm1));
    }
}

// time_step.hpp:18-39, all cells: dt = min over cells of 1/invdt, written to
// sink.dt_out on the device by the last block to finish.
template <class T, class I>
__global__ __launch_bounds__(256) void k_time_step(
        PrimPtrs<T const> prim,
        I n,
        T gamma,
        T invdx0,
        T invdx1,
        T invdx2,
        DtSink<T> sink)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x);
    T local = Limits<T>::max();
    // Four cells in flight per thread to keep enough loads outstanding.
    for (; idx + 3 * stride < n; idx += 4 * stride) {
        Prim<T> q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            q[u] = load_prim(prim, idx + u * stride);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            T const c = sound_speed(gamma, q[u].d, q[u].p);
            local = min_join(local, T(1) / inv_dt_cell(q[u], c, invdx0, invdx1, invdx2));
        }
    }
    for (; idx < n; idx += stride) {
        Prim<T> const q = load_prim(prim, idx);
        T const c = sound_speed(gamma, q.d, q.p);
        local = min_join(local, T(1) / inv_dt_cell(q, c, invdx0, invdx1, invdx2));
    }
    T const bm = block_min<T, 4>(local);
    if (threadIdx.x == 0) {
        grid_min_commit(sink, bm, gridDim.x);
    }
}

// ------------------------------------------------------------------- init ---

// init_implode.hpp:16-43: sphere r^2 < 0.25^2 about the cube centre; the centre
// coordinate is formed in double and rounded to T; ghosts included.
template <class T, class I>
__global__ __launch_bounds__(256) void k_init_implode(PrimPtrs<T> prim, I n0, I n1, I n2, T dx0, T dx1, T dx2)
{
    I const n = n0 * n1 * n2;
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx += stride) {
    I const i = idx % n0;
    I const j = (idx / n0) % n1;
    I const k = idx / (n0 * n1);
    T const c0 = ((static_cast<double>(i - 1) + 0.5) * dx0) - 0.5;
    T const c1 = ((static_cast<double>(j - 1) + 0.5) * dx1) - 0.5;
    T const c2 = ((static_cast<double>(k - 1) + 0.5) * dx2) - 0.5;
    T const r2 = (c0 * c0) + (c1 * c1) + (c2 * c2);
    T const radius2 = T(0.25 * 0.25);
    bool const in = r2 < radius2;
    store_prim(prim, idx, Prim<T> {in ? T(1.) : T(0.125), in ? T(1.) : T(0.14), T(0), T(0), T(0)});
    }
}

// Diagonal interface on the unit cube: x0+x1+x2 > offset -> (d_hi, p_hi), else
// (d_lo, p_lo), with cell centres x_d = ((i_d-1)+0.5)*dx_d formed in double and
// rounded to T; u = 0; ghosts included.
template <class T, class I>
__global__ __launch_bounds__(256) void k_init_diagonal(
        PrimPtrs<T> prim,
        I n0,
        I n1,
        I n2,
        T dx0,
        T dx1,
        T dx2,
        T offset,
        T d_hi,
        T p_hi,
        T d_lo,
        T p_lo)
{
    I const n = n0 * n1 * n2;
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx += stride) {
    I const i = idx % n0;
    I const j = (idx / n0) % n1;
    I const k = idx / (n0 * n1);
    T const x0 = (static_cast<double>(i - 1) + 0.5) * dx0;
    T const x1 = (static_cast<double>(j - 1) + 0.5) * dx1;
    T const x2 = (static_cast<double>(k - 1) + 0.5) * dx2;
    bool const hi = (x0 + x1 + x2) > offset;
    store_prim(prim, idx, Prim<T> {hi ? d_hi : d_lo, hi ? p_hi : p_lo, T(0), T(0), T(0)});
    }
}

// Counter-based generator (SplitMix64 finaliser) keyed by (seed, field, index), so
// the synthetic state does not depend on the launch shape.
__device__ __forceinline__ unsigned long long splitmix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform01(unsigned long long seed, int field, unsigned long long idx)
{
    unsigned long long const h = splitmix(splitmix(seed * 16ull + static_cast<unsigned long long>(field)) ^ idx);
    return static_cast<double>(h >> 11) * 0x1.0p-53;
}

// Seeded random primitive state: rho ~ U[d_lo, d_hi), p ~ U[p_lo, p_hi),
// u_k ~ N(0, u_sigma) (Box-Muller in FP64), drawn in FP64 and rounded to T.
template <class T, class I>
__global__ __launch_bounds__(256) void k_init_random(
        PrimPtrs<T> prim,
        I n,
        unsigned long long seed,
        double d_lo,
        double d_hi,
        double p_lo,
        double p_hi,
        double u_sigma)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx += stride) {
    unsigned long long const g = static_cast<unsigned long long>(idx);
    double const d = d_lo + (d_hi - d_lo) * uniform01(seed, 0, g);
    double const p = p_lo + (p_hi - p_lo) * uniform01(seed, 1, g);
    double u[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double const a = 1.0 - uniform01(seed, 2 + 2 * c, g);
        double const b = uniform01(seed, 3 + 2 * c, g);
        u[c] = u_sigma * sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
    }
    store_prim(prim, idx, Prim<T> {T(d), T(p), T(u[0]), T(u[1]), T(u[2])});
    }
}

// --------------------------------------------------------------- godunov ---

// Reference-order Godunov update (godunov.hpp:40-90): one thread per interior
// cell, six Riemann solves per cell. Kept as the parity baseline for the tiled
// kernel below.
template <class T, class I>
__global__ __launch_bounds__(256) void k_godunov_naive(
        PrimPtrs<T const> prim,
        ConsPtrs<T> cons,
        I n0,
        I n1,
        I n2,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        DtArg<T> dta)
{
    // Flat grid: block b covers a 64 x 4 patch of plane k.
    I const tiles_x = (n0 - 2 + 63) / 64;
    I const tiles_y = (n1 - 2 + 3) / 4;
    I const b = static_cast<I>(blockIdx.x);
    I const i = 1 + (b % tiles_x) * 64 + static_cast<I>(threadIdx.x);
    I const j = 1 + ((b / tiles_x) % tiles_y) * 4 + static_cast<I>(threadIdx.y);
    I const k = 1 + b / (tiles_x * tiles_y);
    if (i > n0 - 2 || j > n1 - 2 || k > n2 - 2) {
        return;
    }
    I const s1 = n0;
    I const s2 = n0 * n1;
    I const c = i + s1 * j + s2 * k;
    Prim<T> const q = load_prim(prim, c);
    T const cq = sound_speed(gamma, q.d, q.p);
    Flux<T> acc {T(0), T(0), T(0), T(0), T(0)};
    {
        Prim<T> const qL = load_prim(prim, c - 1);
        Prim<T> const qR = load_prim(prim, c + 1);
        Flux<T> const fL = hllc_flux<0>(qL, sound_speed(gamma, qL.d, qL.p), q, cq, gm1);
        Flux<T> const fR = hllc_flux<0>(q, cq, qR, sound_speed(gamma, qR.d, qR.p), gm1);
        accumulate(acc, ds0, fR, fL);
    }
    {
        Prim<T> const qL = load_prim(prim, c - s1);
        Prim<T> const qR = load_prim(prim, c + s1);
        Flux<T> const fL = hllc_flux<1>(qL, sound_speed(gamma, qL.d, qL.p), q, cq, gm1);
        Flux<T> const fR = hllc_flux<1>(q, cq, qR, sound_speed(gamma, qR.d, qR.p), gm1);
        accumulate(acc, ds1, fR, fL);
    }
    {
        Prim<T> const qL = load_prim(prim, c - s2);
        Prim<T> const qR = load_prim(prim, c + s2);
        Flux<T> const fL = hllc_flux<2>(qL, sound_speed(gamma, qL.d, qL.p), q, cq, gm1);
        Flux<T> const fR = hllc_flux<2>(q, cq, qR, sound_speed(gamma, qR.d, qR.p), gm1);
        accumulate(acc, ds2, fR, fL);
    }
    Cons<T> u = load_cons(cons, c);
    apply_update(u, resolve_dtodv(dta), acc);
    store_cons(cons, c, u);
}

// Block-to-tile map that keeps neighbouring tiles on one XCD: the hardware deals
// workgroups round-robin over the 8 XCDs, so block b runs on XCD b%8; giving each
// XCD a contiguous range of tiles lets halo re-reads hit that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblk)
{
    if ((nblk & 7) != 0) {
        return b;
    }
    return (b & 7) * (nblk >> 3) + (b >> 3);
}

// Tiled, face-once Godunov update (2.5-D march).
//
// A block of 64 x BY threads owns a 64 x BY tile of interior cells in (i, j) and
// marches along k over `kchunk` planes. Per plane:
//   A. every thread stages its own cell (primitive + sound speed) in LDS, and the
//      first 128 + 2*BY threads stage the halo ring (rows j0-1, j0+BY; columns
//      i0-1, i0+64);
//   B. each thread solves the x face on its low side, the y face on its low side,
//      and the z face on its high side; the tile's high-side x faces (BY of them)
//      and high-side y faces (64) are solved by two of the waves;
//   C. each thread combines low/high faces in the reference's order
//      (x, then y, then z; godunov.hpp:48-82), updates cons in place.
// The z face on the low side comes from the previous plane (registers), the x face
// on the high side from the next lane (wave shuffle), the y face on the high side
// from LDS. Every face is solved once, with the same operands as the reference's
// second solve of it, so results are bitwise those of the reference order.
//
// Fused = true additionally converts the updated cell to primitives into
// prim_out (cons_to_prim.hpp:28-31) and reduces the next dt over the new
// primitives (time_step.hpp:28-38), publishing it through `sink`.
template <class T, class I, int BY, bool Fused>
__global__ __launch_bounds__(64 * BY) void k_godunov_tiled(
        PrimPtrs<T const> prim,
        ConsPtrs<T> cons,
        PrimPtrs<T> prim_out,
        I n0,
        I n1,
        I n2,
        int kchunk,
        int tiles_x,
        int tiles_y,
        int tiles_z,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        DtArg<T> dta,
        T invdx0,
        T invdx1,
        T invdx2,
        DtSink<T> sink)
{
    constexpr int BX = 64;
    constexpr int LX = BX + 2;
    constexpr int LY = BY + 2;
    constexpr int NHALO = 2 * BX + 2 * BY;
    __shared__ T sq[6][LY][LX];   // staged plane: d, p, u0, u1, u2, c
    __shared__ T sfy[5][BY + 1][BX]; // y faces: row r sits between tile rows r-1 and r
    __shared__ T sfx[5][BY];      // x faces on the tile's high-i edge

    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    int const bx = t % tiles_x;
    int const by = (t / tiles_x) % tiles_y;
    int const bz = t / (tiles_x * tiles_y);
    int const tx = threadIdx.x;
    int const ty = threadIdx.y;
    int const lid = tx + BX * ty;

    I const i0 = 1 + static_cast<I>(bx) * BX;
    I const j0 = 1 + static_cast<I>(by) * BY;
    I const kbeg = 1 + static_cast<I>(bz) * kchunk;
    I const kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    I const s1 = n0;
    I const s2 = n0 * n1;

    // Own cell (clamped into the array for loads; only valid cells are written).
    I const i = i0 + tx;
    I const j = j0 + ty;
    bool const valid = (i <= n0 - 2) && (j <= n1 - 2);
    I const ic = (i < n0 - 1) ? i : n0 - 1;
    I const jc = (j < n1 - 1) ? j : n1 - 1;
    I const col = ic + s1 * jc;

    // Halo cell handled by this thread (if any): LDS (row, column) and clamped
    // global (i, j).
    int hrow = 0;
    int hcol = 0;
    I hcolg = 0;
    bool const has_halo = lid < NHALO;
    if (has_halo) {
        if (lid < BX) {
            hrow = 0;
            hcol = lid + 1;
        } else if (lid < 2 * BX) {
            hrow = BY + 1;
            hcol = lid - BX + 1;
        } else if (lid < 2 * BX + BY) {
            hrow = lid - 2 * BX + 1;
            hcol = 0;
        } else {
            hrow = lid - 2 * BX - BY + 1;
            hcol = LX - 1;
        }
        I hi = i0 - 1 + hcol;
        I hj = j0 - 1 + hrow;
        hi = (hi < n0 - 1) ? hi : n0 - 1;
        hj = (hj < n1 - 1) ? hj : n1 - 1;
        hcolg = hi + s1 * hj;
    }

    T const dtodv = resolve_dtodv(dta);
    T local_min = Limits<T>::max();

    // Prologue: z face below the first plane.
    Prim<T> qc = load_prim(prim, col + s2 * kbeg);
    T cc = sound_speed(gamma, qc.d, qc.p);
    Flux<T> fz_lo;
    {
        Prim<T> const qm = load_prim(prim, col + s2 * (kbeg - 1));
        T const cm = sound_speed(gamma, qm.d, qm.p);
        fz_lo = hllc_flux<2>(qm, cm, qc, cc, gm1);
    }
    Prim<T> qn = load_prim(prim, col + s2 * (kbeg + 1));
    T cn = sound_speed(gamma, qn.d, qn.p);

    for (I k = kbeg; k < kend; ++k) {
        // A. stage plane k.
        sq[0][ty + 1][tx + 1] = qc.d;
        sq[1][ty + 1][tx + 1] = qc.p;
        sq[2][ty + 1][tx + 1] = qc.u0;
        sq[3][ty + 1][tx + 1] = qc.u1;
        sq[4][ty + 1][tx + 1] = qc.u2;
        sq[5][ty + 1][tx + 1] = cc;
        if (has_halo) {
            Prim<T> const h = load_prim(prim, hcolg + s2 * k);
            sq[0][hrow][hcol] = h.d;
            sq[1][hrow][hcol] = h.p;
            sq[2][hrow][hcol] = h.u0;
            sq[3][hrow][hcol] = h.u1;
            sq[4][hrow][hcol] = h.u2;
            sq[5][hrow][hcol] = sound_speed(gamma, h.d, h.p);
        }
        Cons<T> u = load_cons(cons, col + s2 * k);
        __syncthreads();

        // B. faces.
        Flux<T> fx_lo;
        Flux<T> fy_lo;
        {
            Prim<T> const qw {sq[0][ty + 1][tx], sq[1][ty + 1][tx], sq[2][ty + 1][tx], sq[3][ty + 1][tx], sq[4][ty + 1][tx]};
            fx_lo = hllc_flux<0>(qw, sq[5][ty + 1][tx], qc, cc, gm1);
        }
        {
            Prim<T> const qs {sq[0][ty][tx + 1], sq[1][ty][tx + 1], sq[2][ty][tx + 1], sq[3][ty][tx + 1], sq[4][ty][tx + 1]};
            fy_lo = hllc_flux<1>(qs, sq[5][ty][tx + 1], qc, cc, gm1);
        }
        Flux<T> const fz_hi = hllc_flux<2>(qc, cc, qn, cn, gm1);
        sfy[0][ty][tx] = fy_lo.d;
        sfy[1][ty][tx] = fy_lo.e;
        sfy[2][ty][tx] = fy_lo.m0;
        sfy[3][ty][tx] = fy_lo.m1;
        sfy[4][ty][tx] = fy_lo.m2;
        if (ty == BY - 1) {
            // High-side y faces of the tile: rows BY and BY+1 of the staged plane.
            Prim<T> const a {sq[0][BY][tx + 1], sq[1][BY][tx + 1], sq[2][BY][tx + 1], sq[3][BY][tx + 1], sq[4][BY][tx + 1]};
            Prim<T> const b {sq[0][BY + 1][tx + 1], sq[1][BY + 1][tx + 1], sq[2][BY + 1][tx + 1], sq[3][BY + 1][tx + 1], sq[4][BY + 1][tx + 1]};
            Flux<T> const f = hllc_flux<1>(a, sq[5][BY][tx + 1], b, sq[5][BY + 1][tx + 1], gm1);
            sfy[0][BY][tx] = f.d;
            sfy[1][BY][tx] = f.e;
            sfy[2][BY][tx] = f.m0;
            sfy[3][BY][tx] = f.m1;
            sfy[4][BY][tx] = f.m2;
        }
        if (ty == (BY > 1 ? BY - 2 : 0) && tx < BY) {
            // High-side x faces of the tile: columns 64 and 65 of row tx.
            int const r = tx + 1;
            Prim<T> const a {sq[0][r][BX], sq[1][r][BX], sq[2][r][BX], sq[3][r][BX], sq[4][r][BX]};
            Prim<T> const b {sq[0][r][BX + 1], sq[1][r][BX + 1], sq[2][r][BX + 1], sq[3][r][BX + 1], sq[4][r][BX + 1]};
            Flux<T> const f = hllc_flux<0>(a, sq[5][r][BX], b, sq[5][r][BX + 1], gm1);
            sfx[0][tx] = f.d;
            sfx[1][tx] = f.e;
            sfx[2][tx] = f.m0;
            sfx[3][tx] = f.m1;
            sfx[4][tx] = f.m2;
        }
        __syncthreads();

        // C. combine and update.
        Flux<T> fx_hi;
        fx_hi.d = __shfl_down(fx_lo.d, 1, 64);
        fx_hi.e = __shfl_down(fx_lo.e, 1, 64);
        fx_hi.m0 = __shfl_down(fx_lo.m0, 1, 64);
        fx_hi.m1 = __shfl_down(fx_lo.m1, 1, 64);
        fx_hi.m2 = __shfl_down(fx_lo.m2, 1, 64);
        if (tx == BX - 1) {
            fx_hi = Flux<T> {sfx[0][ty], sfx[1][ty], sfx[2][ty], sfx[3][ty], sfx[4][ty]};
        }
        Flux<T> const fy_hi {sfy[0][ty + 1][tx], sfy[1][ty + 1][tx], sfy[2][ty + 1][tx], sfy[3][ty + 1][tx], sfy[4][ty + 1][tx]};
        Flux<T> acc {T(0), T(0), T(0), T(0), T(0)};
        accumulate(acc, ds0, fx_hi, fx_lo);
        accumulate(acc, ds1, fy_hi, fy_lo);
        accumulate(acc, ds2, fz_hi, fz_lo);
        apply_update(u, dtodv, acc);
        if (valid) {
            I const c = col + s2 * k;
            store_cons(cons, c, u);
            if constexpr (Fused) {
                Prim<T> const q1 = cons_to_prim_pt(u, gm1);
                store_prim(prim_out, c, q1);
                T const c1 = sound_speed(gamma, q1.d, q1.p);
                local_min = min_join(local_min, T(1) / inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
            }
        }

        // Advance the march.
        fz_lo = fz_hi;
        qc = qn;
        cc = cn;
        if (k + 1 < kend) {
            qn = load_prim(prim, col + s2 * (k + 2));
            cn = sound_speed(gamma, qn.d, qn.p);
        }
    }

    if constexpr (Fused) {
        T const bm = block_min<T, BY>(local_min);
        if (lid == 0) {
            grid_min_commit(sink, bm, static_cast<unsigned int>(nblk));
        }
    }
}

// Ghost-cell refresh for the fused step: prim_out = cons_to_prim(cons) on the
// ghost layer only (the ghost conservatives never change, godunov.hpp:40-43), and
// the time-step minimum over those cells published to sink (time_step.hpp:28-38).
// One thread per ghost cell, enumerated face by face.
template <class T, class I>
__global__ __launch_bounds__(256) void k_ghost_refresh(
        ConsPtrs<T const> cons,
        PrimPtrs<T> prim_out,
        I n0,
        I n1,
        I n2,
        T gamma,
        T gm1,
        T invdx0,
        T invdx1,
        T invdx2,
        DtSink<T> sink)
{
    // Ghost cells: planes k = 0 and k = n2-1 (n0*n1 each, when n2 > 1), then rows
    // j = 0 and j = n1-1 of the inner planes (n0 each), then columns i = 0 and
    // i = n0-1 of the inner rows.
    I const nkz = (n2 > 1) ? 2 : 1;
    I const planes = nkz * n0 * n1;
    I const inner_k = (n2 > 2) ? n2 - 2 : 0;
    I const nky = (n1 > 1) ? 2 : 1;
    I const rows = inner_k * nky * n0;
    I const inner_j = (n1 > 2) ? n1 - 2 : 0;
    I const nkx = (n0 > 1) ? 2 : 1;
    I const cols = inner_k * inner_j * nkx;
    I const total = planes + rows + cols;
    I const stride = static_cast<I>(gridDim.x) * 256;
    T local = Limits<T>::max();
    for (I g = static_cast<I>(blockIdx.x) * 256 + threadIdx.x; g < total; g += stride) {
        I i;
        I j;
        I k;
        if (g < planes) {
            I const side = g / (n0 * n1);
            I const r = g % (n0 * n1);
            i = r % n0;
            j = r / n0;
            k = side == 0 ? 0 : n2 - 1;
        } else if (g < planes + rows) {
            I const h = g - planes;
            I const side = h / (inner_k * n0);
            I const r = h % (inner_k * n0);
            i = r % n0;
            k = 1 + r / n0;
            j = side == 0 ? 0 : n1 - 1;
        } else {
            I const h = g - planes - rows;
            I const side = h / (inner_k * inner_j);
            I const r = h % (inner_k * inner_j);
            j = 1 + r % inner_j;
            k = 1 + r / inner_j;
            i = side == 0 ? 0 : n0 - 1;
        }
        I const c = i + n0 * (j + n1 * k);
        Prim<T> const q = cons_to_prim_pt(load_cons(cons, c), gm1);
        store_prim(prim_out, c, q);
        T const cs = sound_speed(gamma, q.d, q.p);
        local = min_join(local, T(1) / inv_dt_cell(q, cs, invdx0, invdx1, invdx2));
    }
    T const bm = block_min<T, 4>(local);
    if (threadIdx.x == 0) {
        grid_min_commit(sink, bm, gridDim.x);
    }
}

} // namespace eb
