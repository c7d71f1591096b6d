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

#ifndef EB_HALO_PREFETCH
#define EB_HALO_PREFETCH 1 // tiled kernel loads its halo ring one plane ahead
#endif

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
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx = grid_next(idx, n, stride)) {
        store_prim(prim, idx, cons_to_prim_pt(load_cons(cons, idx), gm1));
    }
}

// prim_to_cons.hpp:20-32, all cells.
template <class T, class I>
__global__ __launch_bounds__(256) void k_prim_to_cons(PrimPtrs<T const> prim, ConsPtrs<T> cons, I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx = grid_next(idx, n, stride)) {
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
    for (; n - idx > 3 * stride; idx = grid_next(idx, n, 4 * stride)) {
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
    for (; idx < n; idx = grid_next(idx, n, stride)) {
        Prim<T> const q = load_prim(prim, idx);
        T const c = sound_speed(gamma, q.d, q.p);
        local = min_join(local, T(1) / inv_dt_cell(q, c, invdx0, invdx1, invdx2));
    }
    T const bm = block_min<T, 4>(local);
    if (threadIdx.x == 0) {
        grid_min_commit(sink, bm, blockIdx.x, gridDim.x);
    }
}

// The same reduction with 16-byte non-temporal loads (the fields are read once):
// each lane reads V = 16/sizeof(T) consecutive cells of every field per load, four
// such groups in flight. Needs the five field pointers 16-byte aligned; the cells
// past the last whole group are taken one per thread. The minimum is exact, so the
// result is that of k_time_step.
template <class T, class I>
__global__ __launch_bounds__(256) void k_time_step_vec(
        PrimPtrs<T const> prim,
        I n,
        T gamma,
        T invdx0,
        T invdx1,
        T invdx2,
        DtSink<T> sink)
{
    constexpr int V = 16 / static_cast<int>(sizeof(T));
    typedef T Vec __attribute__((ext_vector_type(V)));
    I const ng = n / V; // whole groups
    I const stride = static_cast<I>(gridDim.x) * 256;
    I g = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x);
    T local = Limits<T>::max();
    auto group = [&](I gi) {
        Vec const d = __builtin_nontemporal_load(reinterpret_cast<Vec const*>(prim.d) + gi);
        Vec const p = __builtin_nontemporal_load(reinterpret_cast<Vec const*>(prim.p) + gi);
        Vec const u0 = __builtin_nontemporal_load(reinterpret_cast<Vec const*>(prim.ux0) + gi);
        Vec const u1 = __builtin_nontemporal_load(reinterpret_cast<Vec const*>(prim.ux1) + gi);
        Vec const u2 = __builtin_nontemporal_load(reinterpret_cast<Vec const*>(prim.ux2) + gi);
        Prim<T> q[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            q[v] = Prim<T> {d[v], p[v], u0[v], u1[v], u2[v]};
        }
        return [=](T& acc) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                T const c = sound_speed(gamma, q[v].d, q[v].p);
                acc = min_join(acc, T(1) / inv_dt_cell(q[v], c, invdx0, invdx1, invdx2));
            }
        };
    };
    for (; ng - g > 3 * stride; g = grid_next(g, ng, 4 * stride)) {
        auto const a0 = group(g);
        auto const a1 = group(g + stride);
        auto const a2 = group(g + 2 * stride);
        auto const a3 = group(g + 3 * stride);
        a0(local);
        a1(local);
        a2(local);
        a3(local);
    }
    for (; g < ng; g = grid_next(g, ng, stride)) {
        group(g)(local);
    }
    // the last n % V cells
    I const rem = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x);
    if (rem < n - ng * V) {
        Prim<T> const q = load_prim(prim, ng * V + rem);
        T const c = sound_speed(gamma, q.d, q.p);
        local = min_join(local, T(1) / inv_dt_cell(q, c, invdx0, invdx1, invdx2));
    }
    T const bm = block_min<T, 4>(local);
    if (threadIdx.x == 0) {
        grid_min_commit(sink, bm, blockIdx.x, gridDim.x);
    }
}

// ------------------------------------------------------------------- init ---

// Every element of one field set to a value (the init_from_value role of
// benchmarks/benchmark_utils.hpp:19-44), one grid-stride store per element.
template <class T>
__global__ __launch_bounds__(256) void k_fill(T* a, int64_t n, T v)
{
    int64_t const stride = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; idx < n; idx = grid_next(idx, n, stride)) {
        a[idx] = v;
    }
}

// Bandwidth probe, not an operator of the reference: per cell the plain flux
// kernel's HBM traffic (five fields of `a` read, five fields of `b` read and
// written in place, one access of sizeof(T) per field) with one add per field and
// no tiles, halos or barriers, so its rate is what plain streaming of that traffic
// reaches on the part.
template <class T>
__global__ __launch_bounds__(256) void k_stream_probe(PrimPtrs<T const> a, ConsPtrs<T> b, int64_t n)
{
    int64_t const stride = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; idx < n; idx = grid_next(idx, n, stride)) {
        T const x0 = a.d[idx];
        T const x1 = a.p[idx];
        T const x2 = a.ux0[idx];
        T const x3 = a.ux1[idx];
        T const x4 = a.ux2[idx];
        T const y0 = b.d[idx];
        T const y1 = b.e[idx];
        T const y2 = b.mx0[idx];
        T const y3 = b.mx1[idx];
        T const y4 = b.mx2[idx];
        b.d[idx] = y0 + x0;
        b.e[idx] = y1 + x1;
        b.mx0[idx] = y2 + x2;
        b.mx1[idx] = y3 + x3;
        b.mx2[idx] = y4 + x4;
    }
}

// init_implode.hpp:16-43: sphere r^2 < 0.25^2 about the cube centre; the centre
// coordinate is formed in double and rounded to T; ghosts included.
template <class T, class I>
__global__ __launch_bounds__(256) void k_init_implode(PrimPtrs<T> prim, I n0, I n1, I n2, T dx0, T dx1, T dx2)
{
    I const n = n0 * n1 * n2;
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx = grid_next(idx, n, stride)) {
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
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx = grid_next(idx, n, stride)) {
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
    for (I idx = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); idx < n; idx = grid_next(idx, n, stride)) {
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
// XCD a contiguous range of tiles lets halo re-reads hit that XCD's L2. With
// nblk = 8q + r, XCD x takes the q + (x < r) tiles from x*q + min(x, r) on, in block
// order, so the map is a bijection of [0, nblk) for every block count.
__device__ __forceinline__ int xcd_tile(int b, int nblk)
{
    int const q = nblk >> 3;
    int const r = nblk & 7;
    int const x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Hides how a wave-uniform value was computed from the optimiser (an empty asm
// that keeps it in scalar registers). Used on the plane index of a march so that
// loop strength reduction does not turn the scalar plane base plus 32-bit lane
// offset into per-lane 64-bit pointers carried across iterations (two VGPRs per
// field and cell).
template <class I>
__device__ __forceinline__ void uniform_opaque(I& v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(v));
#endif
}

// The same for a per-lane 32-bit byte offset of a march, redefined at each use
// site (no instruction): the offset's zero extension then stays next to the
// accesses, where instruction selection folds it into the saddr + voffset form of
// each global load and store, instead of a loop-invariant 64-bit copy of the
// offset added to every field's plane base by a 64-bit VALU add per access (and
// 64-bit addresses held in VGPRs from a load to the store that reuses them).
__device__ __forceinline__ void lane_opaque(uint32_t& v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(v));
#endif
}

// Pins the five values of a conservative update where they are formed. Without
// it the compiler may sink the update, and with it the wait for the conservative
// loads, into the store's validity branch; a wave that skips the branch then
// reaches the next plane with those loads still counted in flight, and every
// plane starts with a wait that also covers the previous plane's stores.
template <class T>
__device__ __forceinline__ void keep_here(Cons<T>& u)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(u.d), "+v"(u.e), "+v"(u.m0), "+v"(u.m1), "+v"(u.m2));
#endif
}

// Pointer into the global address space. Every field array is device memory, so
// accesses through it are global loads and stores even where the pointer was read
// from memory (the batched ensemble's instance descriptors), where the compiler
// would otherwise emit flat accesses, which also count against the LDS wait
// counter and so make every LDS barrier wait for the global traffic in flight.
template <class T>
using global_ptr = __attribute__((address_space(1))) T*;

// Plane-relative addressing: a wave-uniform plane base pointer (scalar registers,
// index type I, so 64-bit extents are honoured) plus a per-lane 32-bit byte offset
// inside the plane. This is the saddr + voffset form of global loads/stores: one
// VGPR of address per lane for all five fields.
template <class T>
__device__ __forceinline__ T ld(T const* plane, uint32_t byte_off)
{
    return *(global_ptr<T const>)(reinterpret_cast<char const*>(plane) + byte_off);
}

// FP64 stores are non-temporal (streamed past the L2 rather than left dirty in
// it): a fused step's per-block device-scope release then has little to write
// back, 128^3 FP64 85 -> 74 us, 512^3 4.89 -> 4.86 ms, the plain flux kernel
// unchanged. FP32 keeps ordinary stores: non-temporal ones made its fused step
// 9% slower at 512^3.
template <class T>
__device__ __forceinline__ void st(T* plane, uint32_t byte_off, T v)
{
    global_ptr<T> const p = (global_ptr<T>)(reinterpret_cast<char*>(plane) + byte_off);
    if constexpr (sizeof(T) == 8) {
        __builtin_nontemporal_store(v, p);
    } else {
        *p = v;
    }
}

template <class T, class I>
__device__ __forceinline__ Prim<T> ld_prim(PrimPtrs<T const> const& a, I plane, uint32_t off)
{
    return {ld(a.d + plane, off), ld(a.p + plane, off), ld(a.ux0 + plane, off), ld(a.ux1 + plane, off),
            ld(a.ux2 + plane, off)};
}

template <class T, class I>
__device__ __forceinline__ Cons<T> ld_cons(ConsPtrs<T> const& a, I plane, uint32_t off)
{
    return {ld<T>(a.d + plane, off), ld<T>(a.e + plane, off), ld<T>(a.mx0 + plane, off), ld<T>(a.mx1 + plane, off),
            ld<T>(a.mx2 + plane, off)};
}

template <class T, class I>
__device__ __forceinline__ void st_cons(ConsPtrs<T> const& a, I plane, uint32_t off, Cons<T> const& u)
{
    st(a.d + plane, off, u.d);
    st(a.e + plane, off, u.e);
    st(a.mx0 + plane, off, u.m0);
    st(a.mx1 + plane, off, u.m1);
    st(a.mx2 + plane, off, u.m2);
}

template <class T, class I>
__device__ __forceinline__ void st_prim(PrimPtrs<T> const& a, I plane, uint32_t off, Prim<T> const& q)
{
    st(a.d + plane, off, q.d);
    st(a.p + plane, off, q.p);
    st(a.ux0 + plane, off, q.u0);
    st(a.ux1 + plane, off, q.u1);
    st(a.ux2 + plane, off, q.u2);
}

// s_waitcnt vmcnt(0) through the builtin, so the compiler's own wait tracking
// knows that every earlier vector-memory operation (LDS transfers included) has
// completed. Encoding for gfx9-family targets: vmcnt 0, expcnt 7, lgkmcnt 15.
// (Device-only builtins are kept out of the host pass, where a __global__
// template's body is parsed but not compiled.)
__device__ __forceinline__ void wait_vmem_all()
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(0x0F70);
#endif
}

// One 16-byte transfer per lane from global memory straight into LDS. The LDS
// destination is wave-uniform: lane l writes lds + 16*l bytes.
__device__ __forceinline__ void transfer16(void const* global, void* lds)
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(global, lds, 16, 0, 0);
#endif
}

// Workgroup barrier that orders LDS traffic only: it waits for this wave's LDS
// operations, not for its outstanding global loads, so prefetches stay in flight
// across it. No kernel here communicates through global memory inside a launch.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Tiled, face-once Godunov update (2.5-D march along k).
//
// A block of 64 x BY threads owns a 64 x BY tile of interior cells in (i, j) and
// marches along k over `kchunk` planes. Per plane k:
//   A. each thread stages its own cell (primitives, sound speed c and total energy
//      E, hllc.hpp:64-65,82) in LDS, and the first 128 + 2*BY threads stage the halo
//      ring (rows j0-1 and j0+BY, columns i0-1 and i0+64);
//   B. each thread solves the x face and the y face on its low side and the z face
//      on its high side; the tile's high-side x faces (BY of them) and y faces (64)
//      are solved by two of the waves;
//   C. each thread combines low and high faces in the reference's order (x, then y,
//      then z; godunov.hpp:48-82) and updates its cell in place.
// The low z face is carried from the previous plane in registers, the high x face
// comes from the next lane (wave shuffle), the high y face from LDS. Each face is
// solved once with the operands of the reference's second solve of it, so the
// result is bitwise that of the reference order.
//
// Loads of plane k (own cell of plane k+1, halo and conservatives of plane k) are
// issued at the top of the plane; the barriers only order LDS, so the
// conservatives stay in flight through the flux work.
//
// Fused = true additionally converts the updated cell to primitives into prim_out
// (cons_to_prim.hpp:28-31) and reduces the next dt over the new primitives
// (time_step.hpp:28-38), publishing it through `sink`.
//
// Requires n0, n1, n2 >= 3 and n0*n1*sizeof(T) < 2^32 (32-bit in-plane offsets).
//
// The tile body is a device function so that the single-grid kernel below and the
// batched ensemble kernel (one launch over the tiles of many grids) share it. It
// returns the thread's contribution to the next dt (largest finite value unless
// Fused). `apply` = false leaves the conservatives' values unchanged while doing
// the same loads, solves and stores (used to time launch shapes on live state).
// Inclusive box of interior cells for the boxed form of the paired-row tile body.
template <class I>
struct CellBox
{
    I lo0;
    I hi0;
    I lo1;
    I hi1;
    I lo2;
    I hi2;
};

template <class T, class I, int BY, bool Fused>
__device__ __forceinline__ T godunov_tile(
        PrimPtrs<T const> const& prim,
        ConsPtrs<T> const& cons,
        PrimPtrs<T> const& prim_out,
        I n0,
        I n1,
        I n2,
        int kchunk,
        int bx,
        int by,
        int bz,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        T dtodv,
        T invdx0,
        T invdx1,
        T invdx2,
        bool apply)
{
    constexpr int BX = 64;
    constexpr int LX = BX + 2;
    constexpr int LY = BY + 2;
    constexpr int NHALO = 2 * BX + 2 * BY;
    constexpr int NF = 7; // staged per cell: d, p, u0, u1, u2, c, E
    __shared__ T sq[NF][LY][LX];
    __shared__ T sfy[5][BY + 1][BX]; // y faces: row r sits between tile rows r-1 and r
    __shared__ T sfx[5][BY];         // x faces on the tile's high-i edge

    int const tx = threadIdx.x;
    int const ty = threadIdx.y;
    int const lid = tx + BX * ty;

    I const i0 = 1 + static_cast<I>(bx) * BX;
    I const j0 = 1 + static_cast<I>(by) * BY;
    I const kbeg = 1 + static_cast<I>(bz) * kchunk;
    I const kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    I const s2 = n0 * n1;

    // Own cell, clamped into the array for loads (only valid cells are written).
    I const i = i0 + tx;
    I const j = j0 + ty;
    bool const valid = (i <= n0 - 2) && (j <= n1 - 2);
    I const ic = (i < n0 - 1) ? i : n0 - 1;
    I const jc = (j < n1 - 1) ? j : n1 - 1;
    uint32_t const own = static_cast<uint32_t>(ic + n0 * jc) * static_cast<uint32_t>(sizeof(T));

    // Halo cell staged by this thread (if any): LDS (row, column), clamped offset.
    int hrow = 0;
    int hcol = 0;
    uint32_t halo = 0;
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
        halo = static_cast<uint32_t>(hi + n0 * hj) * static_cast<uint32_t>(sizeof(T));
    }

    DtFold<T> local_dt(dt_foldable(invdx0, invdx1, invdx2));

    auto stage = [&](int r, int c, Prim<T> const& q, T cs, T e) {
        sq[0][r][c] = q.d;
        sq[1][r][c] = q.p;
        sq[2][r][c] = q.u0;
        sq[3][r][c] = q.u1;
        sq[4][r][c] = q.u2;
        sq[5][r][c] = cs;
        sq[6][r][c] = e;
    };
    auto cell = [&](int r, int c) { return Prim<T> {sq[0][r][c], sq[1][r][c], sq[2][r][c], sq[3][r][c], sq[4][r][c]}; };

    // Prologue: z face below the first plane, own plane kbeg in registers.
    Flux<T> fz_lo;
    Prim<T> q_cur = ld_prim(prim, s2 * kbeg, own);
    T c_cur = sound_speed(gamma, q_cur.d, q_cur.p);
    T e_cur = total_energy(q_cur, gm1);
    {
        Prim<T> const qm = ld_prim(prim, s2 * (kbeg - 1), own);
        fz_lo = hllc_flux<2>(qm, sound_speed(gamma, qm.d, qm.p), total_energy(qm, gm1), q_cur, c_cur, e_cur);
    }
    Prim<T> q_next {};
    Prim<T> h_cur {};
#if EB_HALO_PREFETCH
    if (has_halo) {
        h_cur = ld_prim(prim, s2 * kbeg, halo);
    }
#endif

    for (I k = kbeg; k < kend; ++k) {
        I plane = s2 * k;
        uniform_opaque(plane);
        q_next = ld_prim(prim, plane + s2, own);
#if EB_HALO_PREFETCH
        // halo of the next plane, one plane ahead like the own cell
        Prim<T> h_next {};
        if (has_halo) {
            h_next = ld_prim(prim, plane + s2, halo);
        }
#else
        if (has_halo) {
            h_cur = ld_prim(prim, plane, halo);
        }
#endif
        Cons<T> u = ld_cons(cons, plane, own);

        // A. stage plane k.
        stage(ty + 1, tx + 1, q_cur, c_cur, e_cur);
        if (has_halo) {
            stage(hrow, hcol, h_cur, sound_speed(gamma, h_cur.d, h_cur.p), total_energy(h_cur, gm1));
        }
        lds_barrier();

        // B. faces, in an order that keeps few face results live while another
        // face is solved: y (kept in LDS only), the tile's high-side y and x faces,
        // z, and x last (exchanged by shuffle right away).
        {
            Flux<T> const f = hllc_flux<1>(cell(ty, tx + 1), sq[5][ty][tx + 1], sq[6][ty][tx + 1], q_cur, c_cur, e_cur);
            sfy[0][ty][tx] = f.d;
            sfy[1][ty][tx] = f.e;
            sfy[2][ty][tx] = f.m0;
            sfy[3][ty][tx] = f.m1;
            sfy[4][ty][tx] = f.m2;
        }
        if (ty == BY - 1) {
            // High-side y faces of the tile: rows BY and BY+1 of the staged plane.
            Flux<T> const f = hllc_flux<1>(cell(BY, tx + 1), sq[5][BY][tx + 1], sq[6][BY][tx + 1],
                                           cell(BY + 1, tx + 1), sq[5][BY + 1][tx + 1], sq[6][BY + 1][tx + 1]);
            sfy[0][BY][tx] = f.d;
            sfy[1][BY][tx] = f.e;
            sfy[2][BY][tx] = f.m0;
            sfy[3][BY][tx] = f.m1;
            sfy[4][BY][tx] = f.m2;
        }
        if (ty == (BY > 1 ? BY - 2 : 0) && tx < BY) {
            // High-side x faces of the tile: columns 64 and 65 of row tx+1.
            int const r = tx + 1;
            Flux<T> const f = hllc_flux<0>(cell(r, BX), sq[5][r][BX], sq[6][r][BX], cell(r, BX + 1),
                                           sq[5][r][BX + 1], sq[6][r][BX + 1]);
            sfx[0][tx] = f.d;
            sfx[1][tx] = f.e;
            sfx[2][tx] = f.m0;
            sfx[3][tx] = f.m1;
            sfx[4][tx] = f.m2;
        }
        T const c_next = sound_speed(gamma, q_next.d, q_next.p);
        T const e_next = total_energy(q_next, gm1);
        Flux<T> const fz_hi = hllc_flux<2>(q_cur, c_cur, e_cur, q_next, c_next, e_next);
        Flux<T> const fx_lo
                = hllc_flux<0>(cell(ty + 1, tx), sq[5][ty + 1][tx], sq[6][ty + 1][tx], q_cur, c_cur, e_cur);
        Flux<T> fx_hi;
        fx_hi.d = __shfl_down(fx_lo.d, 1, 64);
        fx_hi.e = __shfl_down(fx_lo.e, 1, 64);
        fx_hi.m0 = __shfl_down(fx_lo.m0, 1, 64);
        fx_hi.m1 = __shfl_down(fx_lo.m1, 1, 64);
        fx_hi.m2 = __shfl_down(fx_lo.m2, 1, 64);
        lds_barrier();

        // C. combine (x, y, z in the reference's order) and update.
        if (tx == BX - 1) {
            fx_hi = Flux<T> {sfx[0][ty], sfx[1][ty], sfx[2][ty], sfx[3][ty], sfx[4][ty]};
        }
        Flux<T> const fy_lo {sfy[0][ty][tx], sfy[1][ty][tx], sfy[2][ty][tx], sfy[3][ty][tx], sfy[4][ty][tx]};
        Flux<T> const fy_hi {sfy[0][ty + 1][tx], sfy[1][ty + 1][tx], sfy[2][ty + 1][tx], sfy[3][ty + 1][tx],
                             sfy[4][ty + 1][tx]};
        Flux<T> acc {T(0), T(0), T(0), T(0), T(0)};
        accumulate(acc, ds0, fx_hi, fx_lo);
        accumulate(acc, ds1, fy_hi, fy_lo);
        accumulate(acc, ds2, fz_hi, fz_lo);
        if (apply) {
            apply_update(u, dtodv, acc);
        }
        if (valid) {
            st_cons(cons, plane, own, u);
            if constexpr (Fused) {
                Prim<T> const q1 = cons_to_prim_pt(u, gm1);
                st_prim(prim_out, plane, own, q1);
                T const c1 = sound_speed(gamma, q1.d, q1.p);
                local_dt.add(inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
            }
        }

        // Advance the march.
        fz_lo = fz_hi;
        q_cur = q_next;
        c_cur = c_next;
        e_cur = e_next;
#if EB_HALO_PREFETCH
        h_cur = h_next;
#endif
    }
    return local_dt.result();
}

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
    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    int const bx = t % tiles_x;
    int const by = (t / tiles_x) % tiles_y;
    int const bz = t / (tiles_x * tiles_y);
    T const local_min = godunov_tile<T, I, BY, Fused>(prim, cons, prim_out, n0, n1, n2, kchunk, bx, by, bz, gamma,
                                                       gm1, ds0, ds1, ds2, resolve_dtodv(dta), invdx0, invdx1,
                                                       invdx2, true);
    if constexpr (Fused) {
        T const bm = block_min<T, BY>(local_min);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            grid_min_commit(sink, bm, blockIdx.x, static_cast<unsigned int>(nblk));
        }
    }
}

// Paired-row Godunov update (2.5-D march along k), configuration variant 4.
//
// A block of BY/2 waves owns a 64 x BY column of interior cells; each lane owns two
// cells of one column, rows 2w and 2w+1 of the tile (cells A and B of wave w). The
// face between A and B is solved once in registers and serves both cells; the face
// below A goes through LDS to the wave below (whose B it tops); x faces come from
// the next lane by shuffle; z faces are carried to the next plane. Compared with one
// cell per lane, a tile twice as tall costs the same number of waves: half the halo
// rows and half the tile-edge face work per cell, two independent cells per lane,
// and two barriers per plane for 64*BY cells. Same per-cell sums as the reference
// (x, y, z order, godunov.hpp:48-82), so bitwise equal to the other variants.
// `apply` = false leaves the conservatives' values unchanged (launch-shape timing).
// WithDt = false (fused form) skips the per-cell time-step candidates: for callers
// that replay recorded time steps; the returned minimum is then meaningless.
template <class T, class I, int BY, bool Fused, bool Boxed = false, bool WithDt = true>
__device__ __forceinline__ T godunov_tile_pair(
        PrimPtrs<T const> const& prim,
        ConsPtrs<T> const& cons,
        PrimPtrs<T> const& prim_out,
        I n0,
        I n1,
        I n2,
        int kchunk,
        int bx,
        int by,
        int bz,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        T dtodv,
        T invdx0,
        T invdx1,
        T invdx2,
        bool apply = true,
        CellBox<I> const& box = CellBox<I> {})
{
    static_assert(BY % 2 == 0, "two rows per wave");
    static_assert(2 * 64 + 2 * BY <= 32 * BY, "one halo cell per thread at most (BY >= 8)");
    constexpr int BX = 64;
    constexpr int NW = BY / 2;
    constexpr int LX = BX + 2;
    constexpr int LY = BY + 2;
    constexpr int NHALO = 2 * BX + 2 * BY;
    constexpr int NF = 7; // staged per cell: d, p, u0, u1, u2, c, E
    __shared__ T sq[NF][LY][LX];
    __shared__ T sfy[5][NW + 1][BX]; // [w]: face between staged rows 2w and 2w+1 (below wave w's A)
    __shared__ T sfx[5][BY];         // x faces on the tile's high-i edge, per row

    int const tx = threadIdx.x;
    int const ty = threadIdx.y;
    int const lid = tx + BX * ty;
    int const rA = 2 * ty + 1; // staged rows of A and B
    int const rB = 2 * ty + 2;

    I const i0 = 1 + static_cast<I>(bx) * BX;
    I const j0 = 1 + static_cast<I>(by) * BY;
    I kbeg = 1 + static_cast<I>(bz) * kchunk;
    I kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    if constexpr (Boxed) {
        // the tile's planes inside the box (block-uniform; never empty for a tile
        // of the box's tile range)
        kbeg = kbeg > box.lo2 ? kbeg : box.lo2;
        kend = kend < box.hi2 + 1 ? kend : box.hi2 + 1;
    }
    I const s2 = n0 * n1;

    I const i = i0 + tx;
    I const jA = j0 + 2 * ty;
    I const jB = jA + 1;
    bool validA = (i <= n0 - 2) && (jA <= n1 - 2);
    bool validB = (i <= n0 - 2) && (jB <= n1 - 2);
    if constexpr (Boxed) {
        bool const in_i = i >= box.lo0 && i <= box.hi0;
        validA = validA && in_i && jA >= box.lo1 && jA <= box.hi1;
        validB = validB && in_i && jB >= box.lo1 && jB <= box.hi1;
    }
    I const ic = (i < n0 - 1) ? i : n0 - 1;
    I const jAc = (jA < n1 - 1) ? jA : n1 - 1;
    I const jBc = (jB < n1 - 1) ? jB : n1 - 1;
    uint32_t offA = static_cast<uint32_t>(ic + n0 * jAc) * static_cast<uint32_t>(sizeof(T));
    uint32_t offB = static_cast<uint32_t>(ic + n0 * jBc) * static_cast<uint32_t>(sizeof(T));
    // Plane-relative addressing pinned at every access site (lane_opaque), in the
    // flux kernel: 4% (FP64) and 13% (FP32) fewer VGPRs, 1.5% faster at 512^3.
    // The fused step keeps the compiler's own addressing: with fewer VGPRs its FP32
    // form would fit two 8-wave blocks per CU, which this march runs 10% slower.
    constexpr bool pin_offsets = !Fused;

    int hrow = 0;
    int hcol = 0;
    uint32_t halo = 0;
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
        halo = static_cast<uint32_t>(hi + n0 * hj) * static_cast<uint32_t>(sizeof(T));
    }

    DtFold<T> local_dt(dt_foldable(invdx0, invdx1, invdx2));

    auto stage = [&](int r, int c, Prim<T> const& q, T cs, T e) {
        sq[0][r][c] = q.d;
        sq[1][r][c] = q.p;
        sq[2][r][c] = q.u0;
        sq[3][r][c] = q.u1;
        sq[4][r][c] = q.u2;
        sq[5][r][c] = cs;
        sq[6][r][c] = e;
    };
    auto cell = [&](int r, int c) { return Prim<T> {sq[0][r][c], sq[1][r][c], sq[2][r][c], sq[3][r][c], sq[4][r][c]}; };
    auto put = [](T (&dst)[5], Flux<T> const& f) {
        dst[0] = f.d;
        dst[1] = f.e;
        dst[2] = f.m0;
        dst[3] = f.m1;
        dst[4] = f.m2;
    };
    auto shfl_flux = [](Flux<T> const& f) {
        return Flux<T> {__shfl_down(f.d, 1, 64), __shfl_down(f.e, 1, 64), __shfl_down(f.m0, 1, 64),
                        __shfl_down(f.m1, 1, 64), __shfl_down(f.m2, 1, 64)};
    };

    // Prologue: z faces below the first plane; own cells of plane kbeg in registers.
    Prim<T> qA = ld_prim(prim, s2 * kbeg, offA);
    Prim<T> qB = ld_prim(prim, s2 * kbeg, offB);
    T cA = sound_speed(gamma, qA.d, qA.p);
    T cB = sound_speed(gamma, qB.d, qB.p);
    T eA = total_energy(qA, gm1);
    T eB = total_energy(qB, gm1);
    Flux<T> fzA_lo;
    Flux<T> fzB_lo;
    {
        Prim<T> const m = ld_prim(prim, s2 * (kbeg - 1), offA);
        fzA_lo = hllc_flux<2>(m, sound_speed(gamma, m.d, m.p), total_energy(m, gm1), qA, cA, eA);
    }
    {
        Prim<T> const m = ld_prim(prim, s2 * (kbeg - 1), offB);
        fzB_lo = hllc_flux<2>(m, sound_speed(gamma, m.d, m.p), total_energy(m, gm1), qB, cB, eB);
    }
    Prim<T> h_cur {};
    if (has_halo) {
        h_cur = ld_prim(prim, s2 * kbeg, halo);
    }
    if constexpr (pin_offsets) {
        // the prologue's loads land before the march, so that no register is
        // still pending at the loop head (a wait there would run at every plane)
        wait_vmem_all();
    }

    for (I k = kbeg; k < kend; ++k) {
        I plane = s2 * k;
        uniform_opaque(plane);
        if constexpr (pin_offsets) {
            lane_opaque(offA);
            lane_opaque(offB);
            lane_opaque(halo);
        }
        Prim<T> const qA_next = ld_prim(prim, plane + s2, offA);
        Prim<T> const qB_next = ld_prim(prim, plane + s2, offB);
        Prim<T> h_next {};
        if (has_halo) {
            h_next = ld_prim(prim, plane + s2, halo);
        }

        // A. stage plane k.
        stage(rA, tx + 1, qA, cA, eA);
        stage(rB, tx + 1, qB, cB, eB);
        if (has_halo) {
            stage(hrow, hcol, h_cur, sound_speed(gamma, h_cur.d, h_cur.p), total_energy(h_cur, gm1));
        }
        lds_barrier();

        // B. faces: y below A (to LDS), y between A and B (registers), the tile's
        // high-side y and x faces, z of A and B, then x of A and B (shuffled at once).
        // The own cells are read back from LDS, so no copy of them stays live here.
        {
            Flux<T> const f = hllc_flux<1>(cell(rA - 1, tx + 1), sq[5][rA - 1][tx + 1], sq[6][rA - 1][tx + 1],
                                           cell(rA, tx + 1), sq[5][rA][tx + 1], sq[6][rA][tx + 1]);
            T tmp[5];
            put(tmp, f);
            for (int c = 0; c < 5; ++c) {
                sfy[c][ty][tx] = tmp[c];
            }
        }
        Flux<T> const fAB = hllc_flux<1>(cell(rA, tx + 1), sq[5][rA][tx + 1], sq[6][rA][tx + 1], cell(rB, tx + 1),
                                         sq[5][rB][tx + 1], sq[6][rB][tx + 1]);
        if (ty == NW - 1) {
            Flux<T> const f = hllc_flux<1>(cell(BY, tx + 1), sq[5][BY][tx + 1], sq[6][BY][tx + 1],
                                           cell(BY + 1, tx + 1), sq[5][BY + 1][tx + 1], sq[6][BY + 1][tx + 1]);
            T tmp[5];
            put(tmp, f);
            for (int c = 0; c < 5; ++c) {
                sfy[c][NW][tx] = tmp[c];
            }
        }
        if (ty == (NW > 1 ? NW - 2 : 0) && tx < BY) {
            int const r = tx + 1;
            Flux<T> const f = hllc_flux<0>(cell(r, BX), sq[5][r][BX], sq[6][r][BX], cell(r, BX + 1),
                                           sq[5][r][BX + 1], sq[6][r][BX + 1]);
            T tmp[5];
            put(tmp, f);
            for (int c = 0; c < 5; ++c) {
                sfx[c][tx] = tmp[c];
            }
        }
        T const cA_next = sound_speed(gamma, qA_next.d, qA_next.p);
        T const eA_next = total_energy(qA_next, gm1);
        Flux<T> const fzA_hi
                = hllc_flux<2>(cell(rA, tx + 1), sq[5][rA][tx + 1], sq[6][rA][tx + 1], qA_next, cA_next, eA_next);
        T const cB_next = sound_speed(gamma, qB_next.d, qB_next.p);
        T const eB_next = total_energy(qB_next, gm1);
        Flux<T> const fzB_hi
                = hllc_flux<2>(cell(rB, tx + 1), sq[5][rB][tx + 1], sq[6][rB][tx + 1], qB_next, cB_next, eB_next);
        Flux<T> const fxA_lo = hllc_flux<0>(cell(rA, tx), sq[5][rA][tx], sq[6][rA][tx], cell(rA, tx + 1),
                                            sq[5][rA][tx + 1], sq[6][rA][tx + 1]);
        Flux<T> fxA_hi = shfl_flux(fxA_lo);
        Flux<T> const fxB_lo = hllc_flux<0>(cell(rB, tx), sq[5][rB][tx], sq[6][rB][tx], cell(rB, tx + 1),
                                            sq[5][rB][tx + 1], sq[6][rB][tx + 1]);
        Flux<T> fxB_hi = shfl_flux(fxB_lo);
        if constexpr (pin_offsets) {
            lane_opaque(offA);
            lane_opaque(offB);
        }
        Cons<T> uA = ld_cons(cons, plane, offA);
        Cons<T> uB = ld_cons(cons, plane, offB);
        lds_barrier();

        // C. combine (x, y, z in the reference's order) and update both cells.
        if (tx == BX - 1) {
            fxA_hi = Flux<T> {sfx[0][rA - 1], sfx[1][rA - 1], sfx[2][rA - 1], sfx[3][rA - 1], sfx[4][rA - 1]};
            fxB_hi = Flux<T> {sfx[0][rB - 1], sfx[1][rB - 1], sfx[2][rB - 1], sfx[3][rB - 1], sfx[4][rB - 1]};
        }
        Flux<T> const fyA_lo {sfy[0][ty][tx], sfy[1][ty][tx], sfy[2][ty][tx], sfy[3][ty][tx], sfy[4][ty][tx]};
        Flux<T> const fyB_hi {sfy[0][ty + 1][tx], sfy[1][ty + 1][tx], sfy[2][ty + 1][tx], sfy[3][ty + 1][tx],
                              sfy[4][ty + 1][tx]};
        Flux<T> accA {T(0), T(0), T(0), T(0), T(0)};
        accumulate(accA, ds0, fxA_hi, fxA_lo);
        accumulate(accA, ds1, fAB, fyA_lo);
        accumulate(accA, ds2, fzA_hi, fzA_lo);
        Flux<T> accB {T(0), T(0), T(0), T(0), T(0)};
        accumulate(accB, ds0, fxB_hi, fxB_lo);
        accumulate(accB, ds1, fyB_hi, fAB);
        accumulate(accB, ds2, fzB_hi, fzB_lo);
        if (apply) {
            apply_update(uA, dtodv, accA);
            apply_update(uB, dtodv, accB);
        }
        if constexpr (pin_offsets) {
            keep_here(uA);
            keep_here(uB);
            lane_opaque(offA);
            lane_opaque(offB);
        }
        if (validA) {
            st_cons(cons, plane, offA, uA);
            if constexpr (Fused) {
                Prim<T> const q1 = cons_to_prim_pt(uA, gm1);
                st_prim(prim_out, plane, offA, q1);
                if constexpr (WithDt) {
                    T const c1 = sound_speed(gamma, q1.d, q1.p);
                    local_dt.add(inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
                }
            }
        }
        if (validB) {
            st_cons(cons, plane, offB, uB);
            if constexpr (Fused) {
                Prim<T> const q1 = cons_to_prim_pt(uB, gm1);
                st_prim(prim_out, plane, offB, q1);
                if constexpr (WithDt) {
                    T const c1 = sound_speed(gamma, q1.d, q1.p);
                    local_dt.add(inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
                }
            }
        }

        // Advance the march.
        fzA_lo = fzA_hi;
        fzB_lo = fzB_hi;
        qA = qA_next;
        qB = qB_next;
        cA = cA_next;
        cB = cB_next;
        eA = eA_next;
        eB = eB_next;
        h_cur = h_next;
    }
    return local_dt.result();
}

template <class T, class I, int BY, bool Fused>
__global__ __launch_bounds__(32 * BY) __attribute__((amdgpu_waves_per_eu(2))) void k_godunov_pair(
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
    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    int const bx = t % tiles_x;
    int const by = (t / tiles_x) % tiles_y;
    int const bz = t / (tiles_x * tiles_y);
    T const local_min = godunov_tile_pair<T, I, BY, Fused>(prim, cons, prim_out, n0, n1, n2, kchunk, bx, by, bz,
                                                            gamma, gm1, ds0, ds1, ds2, resolve_dtodv(dta), invdx0,
                                                            invdx1, invdx2);
    if constexpr (Fused) {
        T const bm = block_min<T, BY / 2>(local_min);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            grid_min_commit(sink, bm, blockIdx.x, static_cast<unsigned int>(nblk));
        }
    }
}

// The fused step of the paired-row tiles restricted to a box of interior cells:
// tiles_x x tiles_y x tiles_z tiles starting at tile (tile0_x, tile0_y, tile0_z)
// are marched over the box's planes, and only cells inside the box are written
// (the others keep their values). With WithDt the next dt is the minimum over the
// box; without, no time step is computed and the sink is not touched.
template <class T, class I, int BY, bool WithDt>
__global__ __launch_bounds__(32 * BY) __attribute__((amdgpu_waves_per_eu(2))) void k_godunov_box_pair(
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
        int tile0_x,
        int tile0_y,
        int tile0_z,
        CellBox<I> box,
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
    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    int const bx = tile0_x + t % tiles_x;
    int const by = tile0_y + (t / tiles_x) % tiles_y;
    int const bz = tile0_z + t / (tiles_x * tiles_y);
    T const local_min = godunov_tile_pair<T, I, BY, true, true, WithDt>(
            prim, cons, prim_out, n0, n1, n2, kchunk, bx, by, bz, gamma, gm1, ds0, ds1, ds2, resolve_dtodv(dta),
            invdx0, invdx1, invdx2, true, box);
    if constexpr (WithDt) {
        T const bm = block_min<T, BY / 2>(local_min);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            grid_min_commit(sink, bm, blockIdx.x, static_cast<unsigned int>(nblk));
        }
    }
}

// Linear-tile Godunov update (2.5-D march along k), configuration variant 3.
//
// A block of NW waves owns 64*NW consecutive cells of a k-plane in array order,
// a = i + n0*j, taken from the interior rows j = 1 .. n1-2 (ghost columns i = 0
// and n0-1 ride along and produce nothing). Rows wrap inside the tile, so no lane
// is lost to a ragged x extent. Per plane k:
//   A. each thread stages its cell (primitives, c, E); the block also stages the
//      row below its first cell and the row above its last (2*n0 halo cells).
//      Halo cell t of each thread (t < 2*n0) is loaded one plane ahead, after
//      the staging of the plane before, so its latency overlaps a whole plane and
//      waiting for it never waits for the plane's own loads, which are issued
//      later; the rare second halo cell of a thread (2*n0 > 64*NW) is loaded in
//      place;
//   B. each thread solves the x face and the y face on its low side (from the
//      staged neighbours a-1 and a-n0) and the z face on its high side; the y
//      faces above the last n0 cells and the x face after the last cell are
//      solved once more by the first threads;
//   C. each thread combines its faces in the reference's order (x, y, z;
//      godunov.hpp:48-82) and updates its cell in place.
// As in the 64 x BY tiles, every face is solved once with the operands of the
// reference's second solve of it, so results are bitwise the reference order's.
// Requires n0 <= 64*NW, n0, n1, n2 >= 3 and n0*n1*sizeof(T) < 2^32.
template <class T, class I, int NW, bool Fused>
__device__ __forceinline__ T godunov_tile_lin(
        PrimPtrs<T const> const& prim,
        ConsPtrs<T> const& cons,
        PrimPtrs<T> const& prim_out,
        I n0,
        I n1,
        I n2,
        int kchunk,
        int tile,
        int bz,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        T dtodv,
        T invdx0,
        T invdx1,
        T invdx2,
        bool apply)
{
    constexpr int NT = 64 * NW;     // cells per tile
    constexpr int LS = 3 * NT;      // staged cells: n0 below + NT + n0 above, n0 <= NT
    constexpr int NF = 7;           // d, p, u0, u1, u2, c, E
    __shared__ T sq[NF][LS];
    __shared__ T sfx[5][NT + 1];    // x face on the low side of tile cell t; [NT]: after the last cell
    __shared__ T sfy[5][NT];        // y face on the low side of tile cell t
    __shared__ T sfe[5][NT];        // y faces above tile cells NT-n0 .. NT-1

    int const t = threadIdx.x + 64 * threadIdx.y;
    I const s2 = n0 * n1;
    I const a_begin = n0;            // first cell of row 1
    I const a_end = n0 * (n1 - 1);   // one past the last cell of row n1-2
    I const A = a_begin + static_cast<I>(tile) * NT;
    I const a = A + t;
    I const kbeg = 1 + static_cast<I>(bz) * kchunk;
    I const kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    int const w0 = static_cast<int>(n0); // staged offset of tile cell 0

    I const i = a % n0;
    bool const valid = a < a_end && i >= 1 && i <= n0 - 2;
    auto clampa = [&](I x) { return x < s2 - 1 ? x : s2 - 1; };
    uint32_t const own = static_cast<uint32_t>(clampa(a)) * static_cast<uint32_t>(sizeof(T));

    DtFold<T> local_dt(dt_foldable(invdx0, invdx1, invdx2));
    auto stage = [&](int r, Prim<T> const& q, T cs, T e) {
        sq[0][r] = q.d;
        sq[1][r] = q.p;
        sq[2][r] = q.u0;
        sq[3][r] = q.u1;
        sq[4][r] = q.u2;
        sq[5][r] = cs;
        sq[6][r] = e;
    };
    auto cell = [&](int r) { return Prim<T> {sq[0][r], sq[1][r], sq[2][r], sq[3][r], sq[4][r]}; };
    auto put = [&](auto& arr, int r, Flux<T> const& f) {
        arr[0][r] = f.d;
        arr[1][r] = f.e;
        arr[2][r] = f.m0;
        arr[3][r] = f.m1;
        arr[4][r] = f.m2;
    };

    // halo cell h: the row below the tile for h < n0, the row above it after that
    auto halo_off = [&](int h) {
        I const b = h < w0 ? A - n0 + h : A + NT + (h - w0);
        return static_cast<uint32_t>(clampa(b)) * static_cast<uint32_t>(sizeof(T));
    };
    auto halo_slot = [&](int h) { return h < w0 ? h : w0 + NT + (h - w0); };
    bool const has_h0 = t < 2 * w0;
    uint32_t const h0_off = has_h0 ? halo_off(t) : 0u;

    Flux<T> fz_lo;
    Prim<T> q_cur = ld_prim(prim, s2 * kbeg, own);
    T c_cur = sound_speed(gamma, q_cur.d, q_cur.p);
    T e_cur = total_energy(q_cur, gm1);
    {
        Prim<T> const qm = ld_prim(prim, s2 * (kbeg - 1), own);
        fz_lo = hllc_flux<2>(qm, sound_speed(gamma, qm.d, qm.p), total_energy(qm, gm1), q_cur, c_cur, e_cur);
    }
    Prim<T> h0 {};
    if (has_h0) {
        h0 = ld_prim(prim, s2 * kbeg, h0_off);
    }

    for (I k = kbeg; k < kend; ++k) {
        I const plane = s2 * k;
        Prim<T> const q_next = ld_prim(prim, plane + s2, own);
        Cons<T> u = ld_cons(cons, plane, own);

        // A. stage plane k: own cell, then the halo rows (2*n0 cells, h < n0 below);
        // then the first halo cell of the next plane is requested (k + 1 <= n2 - 1).
        stage(w0 + t, q_cur, c_cur, e_cur);
        if (has_h0) {
            stage(halo_slot(t), h0, sound_speed(gamma, h0.d, h0.p), total_energy(h0, gm1));
            h0 = ld_prim(prim, plane + s2, h0_off);
        }
        for (int h = t + NT; h < 2 * w0; h += NT) {
            Prim<T> const q = ld_prim(prim, plane, halo_off(h));
            stage(halo_slot(h), q, sound_speed(gamma, q.d, q.p), total_energy(q, gm1));
        }
        lds_barrier();

        // B. faces.
        T const c_next = sound_speed(gamma, q_next.d, q_next.p);
        T const e_next = total_energy(q_next, gm1);
        int const r = w0 + t;
        Flux<T> const fx_lo = hllc_flux<0>(cell(r - 1), sq[5][r - 1], sq[6][r - 1], q_cur, c_cur, e_cur);
        Flux<T> const fy_lo = hllc_flux<1>(cell(r - w0), sq[5][r - w0], sq[6][r - w0], q_cur, c_cur, e_cur);
        Flux<T> const fz_hi = hllc_flux<2>(q_cur, c_cur, e_cur, q_next, c_next, e_next);
        put(sfx, t, fx_lo);
        put(sfy, t, fy_lo);
        if (t < w0) {
            // y face above tile cell NT - n0 + t
            int const lo = w0 + NT - w0 + t; // staged index of that cell
            Flux<T> const f
                    = hllc_flux<1>(cell(lo), sq[5][lo], sq[6][lo], cell(lo + w0), sq[5][lo + w0], sq[6][lo + w0]);
            put(sfe, t, f);
        }
        if (t == NT - 1) {
            Flux<T> const f = hllc_flux<0>(q_cur, c_cur, e_cur, cell(r + 1), sq[5][r + 1], sq[6][r + 1]);
            sfx[0][NT] = f.d;
            sfx[1][NT] = f.e;
            sfx[2][NT] = f.m0;
            sfx[3][NT] = f.m1;
            sfx[4][NT] = f.m2;
        }
        lds_barrier();

        // C. combine and update.
        Flux<T> const fx_hi {sfx[0][t + 1], sfx[1][t + 1], sfx[2][t + 1], sfx[3][t + 1], sfx[4][t + 1]};
        int const up = t + w0;
        Flux<T> const fy_hi = up < NT ? Flux<T> {sfy[0][up], sfy[1][up], sfy[2][up], sfy[3][up], sfy[4][up]}
                                      : Flux<T> {sfe[0][up - NT], sfe[1][up - NT], sfe[2][up - NT],
                                                 sfe[3][up - NT], sfe[4][up - NT]};
        Flux<T> acc {T(0), T(0), T(0), T(0), T(0)};
        accumulate(acc, ds0, fx_hi, fx_lo);
        accumulate(acc, ds1, fy_hi, fy_lo);
        accumulate(acc, ds2, fz_hi, fz_lo);
        if (apply) {
            apply_update(u, dtodv, acc);
        }
        if (valid) {
            st_cons(cons, plane, own, u);
            if constexpr (Fused) {
                Prim<T> const q1 = cons_to_prim_pt(u, gm1);
                st_prim(prim_out, plane, own, q1);
                T const c1 = sound_speed(gamma, q1.d, q1.p);
                local_dt.add(inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
            }
        }

        fz_lo = fz_hi;
        q_cur = q_next;
        c_cur = c_next;
        e_cur = e_next;
    }
    return local_dt.result();
}

template <class T, class I, int NW, bool Fused>
__global__ __launch_bounds__(64 * NW) void k_godunov_linear(
        PrimPtrs<T const> prim,
        ConsPtrs<T> cons,
        PrimPtrs<T> prim_out,
        I n0,
        I n1,
        I n2,
        int kchunk,
        int tiles_plane,
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
    int const nblk = tiles_plane * tiles_z;
    int const tb = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    T const local_min = godunov_tile_lin<T, I, NW, Fused>(prim, cons, prim_out, n0, n1, n2, kchunk, tb % tiles_plane,
                                                           tb / tiles_plane, gamma, gm1, ds0, ds1, ds2,
                                                           resolve_dtodv(dta), invdx0, invdx1, invdx2, true);
    if constexpr (Fused) {
        T const bm = block_min<T, NW>(local_min);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            grid_min_commit(sink, bm, blockIdx.x, static_cast<unsigned int>(nblk));
        }
    }
}

// Streamed Godunov update (2.5-D march along k) with a producer wave.
//
// A block owns a 64 x BY tile of interior cells in (i, j), marches along k over
// `kchunk` planes, and has BY compute waves (one per tile row, one thread per
// cell) plus one producer wave. The producer streams the primitive fields of
// plane k+2, halo ring included, straight from HBM into LDS with 16-byte
// global_load_lds transfers into a three-plane ring, while the compute waves work
// on plane k; the compute waves never touch global memory for primitives, so they
// never wait on those transfers, and the producer waits for one plane's transfers
// a full plane after issuing them.
//
// Per plane k (three LDS barriers):
//   0. plane k+1 has landed (the producer waited before arriving);
//   A. compute waves load the conservatives of plane k into registers and compute
//      the sound speed c and total energy E (hllc.hpp:64-65,82) of every tile cell
//      of plane k+1; the producer issues plane k+2;
//   B. faces of plane k: x and y faces on each cell's low side, z face on its high
//      side; the tile's high-side x faces (BY) and y faces (64) are solved by two of
//      the waves;
//   C. each thread combines faces in the reference's order (x, y, z;
//      godunov.hpp:48-82) and updates its cell in place.
// Every face is solved once with the operands of the reference's second solve of
// it, so results are bitwise those of the reference order.
//
// LDS rows are 16-byte aligned copies of global rows: row r of a plane holds
// elements starting at the aligned-down linear index of cell (i0-1, j0-1+r), so a
// cell's column is shifted by that row's misalignment `shift`.
//
// Requires n0, n1, n2 >= 3, n0*n1*sizeof(T) < 2^32 and 16-byte aligned fields.
template <class T, class I, int BY, bool Fused>
__global__ __launch_bounds__(64 * (BY + 1)) void k_godunov_stream(
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
    constexpr int E = 16 / static_cast<int>(sizeof(T)); // elements per 16-byte transfer
    constexpr int P = ((LX + E - 1) + E - 1) / E * E;   // LDS row pitch with room for the shift
    constexpr int L = P / E;                            // transfers per row
    constexpr int NXFER = LY * L;                       // transfers per field and plane
    constexpr int NCH = (NXFER + 63) / 64;              // wave instructions per field and plane
    constexpr int NCELL = LY * LX;
    __shared__ __align__(16) T sp[3][5][LY * P]; // primitive ring: planes k-1+slot
    __shared__ T sd[2][2][LY][LX];               // c and E of planes k and k+1
    __shared__ T sfy[5][BY + 1][BX];
    __shared__ T sfx[5][BY];

    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    int const bx = t % tiles_x;
    int const by = (t / tiles_x) % tiles_y;
    int const bz = t / (tiles_x * tiles_y);
    int const tx = threadIdx.x;
    int const ty = threadIdx.y;
    bool const producer = ty == BY;

    I const i0 = 1 + static_cast<I>(bx) * BX;
    I const j0 = 1 + static_cast<I>(by) * BY;
    I const kbeg = 1 + static_cast<I>(bz) * kchunk;
    I const kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    I const s2 = n0 * n1;
    I const ntot = s2 * n2;

    // Clamped global row of LDS row r, and the 32-bit pieces of its linear index
    // used to find the row's misalignment in plane k.
    auto row_j = [&](int r) {
        I const jr = j0 - 1 + r;
        return (jr < n1 - 1) ? jr : n1 - 1;
    };
    auto shift = [&](int r, I k) {
        uint32_t const lin = static_cast<uint32_t>(i0 - 1) + static_cast<uint32_t>(n0) * static_cast<uint32_t>(row_j(r))
                             + static_cast<uint32_t>(s2) * static_cast<uint32_t>(k);
        return static_cast<int>(lin & static_cast<uint32_t>(E - 1));
    };
    auto slot = [&](I k) { return static_cast<int>((k - kbeg + 1) % 3); };
    auto dslot = [&](I k) { return static_cast<int>((k - kbeg + 1) & 1); };

    // Producer: all transfers of plane k into ring slot b.
    auto transfer = [&](I k, int b) {
        T const* const fields[5] = {prim.d, prim.p, prim.ux0, prim.ux1, prim.ux2};
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            int const x = ch * 64 + tx;
            if (x < NXFER) {
                int const r = x / L;
                int const seg = x % L;
                I const rowlin = (i0 - 1) + n0 * (row_j(r) + n1 * k);
                I const g = rowlin - (rowlin & static_cast<I>(E - 1)) + static_cast<I>(seg * E);
                if (g + E <= ntot) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) {
                        transfer16(fields[f] + g, &sp[b][f][ch * 64 * E]);
                    }
                }
            }
        }
    };

    // Compute-wave views of the ring.
    auto prim_at = [&](int b, int r, int c, int sh) {
        int const o = r * P + sh + c;
        return Prim<T> {sp[b][0][o], sp[b][1][o], sp[b][2][o], sp[b][3][o], sp[b][4][o]};
    };
    // c and E of every tile cell of plane k (ring slot b) into sd[db].
    auto derive = [&](I k, int b, int db) {
        for (int n = tx + BX * ty; n < NCELL; n += BX * BY) {
            int const r = n / LX;
            int const c = n % LX;
            Prim<T> const q = prim_at(b, r, c, shift(r, k));
            sd[db][0][r][c] = sound_speed(gamma, q.d, q.p);
            sd[db][1][r][c] = total_energy(q, gm1);
        }
    };

    T const dtodv = resolve_dtodv(dta);
    T local_min = Limits<T>::max();

    // Own cell (compute waves), clamped into the array for loads.
    I const i = i0 + tx;
    I const j = j0 + ty;
    bool const valid = !producer && (i <= n0 - 2) && (j <= n1 - 2);
    I const ic = (i < n0 - 1) ? i : n0 - 1;
    I const jc = (j < n1 - 1) ? j : n1 - 1;
    uint32_t const own = static_cast<uint32_t>(ic + n0 * jc) * static_cast<uint32_t>(sizeof(T));

    // Prologue: planes kbeg-1, kbeg, kbeg+1 in the ring; c, E of plane kbeg for the
    // tile; the low z face of plane kbeg in registers.
    if (producer) {
        transfer(kbeg - 1, slot(kbeg - 1));
        transfer(kbeg, slot(kbeg));
        transfer(kbeg + 1, slot(kbeg + 1));
        wait_vmem_all();
    }
    lds_barrier();
    Flux<T> fz_lo {};
    if (!producer) {
        derive(kbeg, slot(kbeg), dslot(kbeg));
        Prim<T> const qm = prim_at(slot(kbeg - 1), ty + 1, tx + 1, shift(ty + 1, kbeg - 1));
        Prim<T> const q0 = prim_at(slot(kbeg), ty + 1, tx + 1, shift(ty + 1, kbeg));
        fz_lo = hllc_flux<2>(qm, sound_speed(gamma, qm.d, qm.p), total_energy(qm, gm1), q0,
                             sound_speed(gamma, q0.d, q0.p), total_energy(q0, gm1));
    }
    lds_barrier();

    if (producer) {
        for (I k = kbeg; k < kend; ++k) {
            // 0. plane k+1 landed (issued a plane ago); then stream plane k+2 into
            // the slot of plane k-1, which no wave reads any more.
            wait_vmem_all();
            lds_barrier();
            if (k + 2 <= kend) {
                transfer(k + 2, slot(k + 2));
            }
            lds_barrier(); // A | B
            lds_barrier(); // B | C
        }
    } else {
        for (I k = kbeg; k < kend; ++k) {
            I const plane = s2 * k;
            int const b = slot(k);
            int const bn = slot(k + 1);
            int const db = dslot(k);
            int const dbn = dslot(k + 1);
            lds_barrier(); // 0

            // A. conservatives of plane k; c, E of plane k+1.
            Cons<T> u = ld_cons(cons, plane, own);
            derive(k + 1, bn, dbn);
            lds_barrier();

            // B. faces.
            int const sh_own = shift(ty + 1, k);
            int const sh_dn = shift(ty, k);
            Prim<T> const qc = prim_at(b, ty + 1, tx + 1, sh_own);
            T const cc = sd[db][0][ty + 1][tx + 1];
            T const ec = sd[db][1][ty + 1][tx + 1];
            Flux<T> const fx_lo = hllc_flux<0>(prim_at(b, ty + 1, tx, sh_own), sd[db][0][ty + 1][tx],
                                               sd[db][1][ty + 1][tx], qc, cc, ec);
            Flux<T> const fy_lo = hllc_flux<1>(prim_at(b, ty, tx + 1, sh_dn), sd[db][0][ty][tx + 1],
                                               sd[db][1][ty][tx + 1], qc, cc, ec);
            Flux<T> const fz_hi = hllc_flux<2>(qc, cc, ec, prim_at(bn, ty + 1, tx + 1, shift(ty + 1, k + 1)),
                                               sd[dbn][0][ty + 1][tx + 1], sd[dbn][1][ty + 1][tx + 1]);
            sfy[0][ty][tx] = fy_lo.d;
            sfy[1][ty][tx] = fy_lo.e;
            sfy[2][ty][tx] = fy_lo.m0;
            sfy[3][ty][tx] = fy_lo.m1;
            sfy[4][ty][tx] = fy_lo.m2;
            if (ty == BY - 1) {
                // High-side y faces of the tile: rows BY and BY+1.
                int const s0 = shift(BY, k);
                int const s1 = shift(BY + 1, k);
                Flux<T> const f = hllc_flux<1>(prim_at(b, BY, tx + 1, s0), sd[db][0][BY][tx + 1], sd[db][1][BY][tx + 1],
                                               prim_at(b, BY + 1, tx + 1, s1), sd[db][0][BY + 1][tx + 1],
                                               sd[db][1][BY + 1][tx + 1]);
                sfy[0][BY][tx] = f.d;
                sfy[1][BY][tx] = f.e;
                sfy[2][BY][tx] = f.m0;
                sfy[3][BY][tx] = f.m1;
                sfy[4][BY][tx] = f.m2;
            }
            if (ty == (BY > 1 ? BY - 2 : 0) && tx < BY) {
                // High-side x faces of the tile: columns 64 and 65 of row tx+1.
                int const r = tx + 1;
                int const sr = shift(r, k);
                Flux<T> const f = hllc_flux<0>(prim_at(b, r, BX, sr), sd[db][0][r][BX], sd[db][1][r][BX],
                                               prim_at(b, r, BX + 1, sr), sd[db][0][r][BX + 1], sd[db][1][r][BX + 1]);
                sfx[0][tx] = f.d;
                sfx[1][tx] = f.e;
                sfx[2][tx] = f.m0;
                sfx[3][tx] = f.m1;
                sfx[4][tx] = f.m2;
            }
            lds_barrier();

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
            Flux<T> const fy_hi {sfy[0][ty + 1][tx], sfy[1][ty + 1][tx], sfy[2][ty + 1][tx], sfy[3][ty + 1][tx],
                                 sfy[4][ty + 1][tx]};
            Flux<T> acc {T(0), T(0), T(0), T(0), T(0)};
            accumulate(acc, ds0, fx_hi, fx_lo);
            accumulate(acc, ds1, fy_hi, fy_lo);
            accumulate(acc, ds2, fz_hi, fz_lo);
            apply_update(u, dtodv, acc);
            if (valid) {
                st_cons(cons, plane, own, u);
                if constexpr (Fused) {
                    Prim<T> const q1 = cons_to_prim_pt(u, gm1);
                    st_prim(prim_out, plane, own, q1);
                    T const c1 = sound_speed(gamma, q1.d, q1.p);
                    local_min = min_join(local_min, T(1) / inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
                }
            }
            fz_lo = fz_hi;
        }
    }

    if constexpr (Fused) {
        T const bm = block_min<T, BY + 1>(local_min);
        if (tx == 0 && ty == 0) {
            grid_min_commit(sink, bm, blockIdx.x, static_cast<unsigned int>(nblk));
        }
    }
}

// Face-parallel Godunov update (2.5-D march along k), configuration variant 0.
//
// A block owns a 64 x BY tile of interior cells in (i, j) and marches along k over
// `kchunk` planes. It has 3*BY waves in three roles, one wave per tile row each:
//   X waves solve the x face on the low side of every cell of their row,
//   Y waves solve the y face on the low side,
//   Z waves solve the z face on the high side and update the cell.
// The tile's high-side x faces (BY) are solved by X wave 0 and its high-side y
// faces (64) by Y wave BY-1, in a second pass. Each thread therefore carries the
// state of about one Riemann solve, which keeps registers low enough for two
// blocks (6*BY waves) per CU; face fluxes meet in LDS.
//
// Per plane k (two LDS-only barriers):
//   S. the first (BY+2)*66 threads stage their cell of plane k+1 (loaded one plane
//      earlier) with its sound speed and total energy (hllc.hpp:64-65,82) into the
//      other half of a double-buffered tile, then load their cell of plane k+2;
//      Z threads load the conservatives of plane k;
//   F. faces of plane k: x and y faces from tile k, z faces from tiles k and k+1;
//   U. Z threads combine faces in the reference's order (x, y, z; godunov.hpp:48-82)
//      and update their cell in place.
// Every face is solved once with the operands of the reference's second solve of
// it, so results are bitwise those of the reference order.
//
// Fused = true additionally converts the updated cell to primitives into prim_out
// (cons_to_prim.hpp:28-31) and reduces the next dt over the new primitives
// (time_step.hpp:28-38), publishing it through `sink`.
//
// Requires n0, n1, n2 >= 3 and n0*n1*sizeof(T) < 2^32 (32-bit in-plane offsets).
template <class T, class I, int BY, bool Fused>
__global__ __launch_bounds__(192 * BY) void k_godunov_faces(
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
    constexpr int NSTAGE = LX * LY;
    constexpr int NF = 7; // d, p, u0, u1, u2, c, E
    static_assert(NSTAGE <= 192 * BY, "one staged cell per thread");
    __shared__ T tile[2][NF][LY][LX];
    __shared__ T fx[5][BY][BX + 1]; // x face between columns c-1 and c of row r
    __shared__ T fy[5][BY + 1][BX]; // y face between rows r-1 and r of column c

    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x), nblk);
    int const bx = t % tiles_x;
    int const by = (t / tiles_x) % tiles_y;
    int const bz = t / (tiles_x * tiles_y);
    int const lane = threadIdx.x;
    int const wave = threadIdx.y; // 0..3*BY-1
    int const role = wave / BY;   // 0: X, 1: Y, 2: Z
    int const row = wave % BY;
    int const lid = lane + BX * wave;

    I const i0 = 1 + static_cast<I>(bx) * BX;
    I const j0 = 1 + static_cast<I>(by) * BY;
    I const kbeg = 1 + static_cast<I>(bz) * kchunk;
    I const kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    I const s2 = n0 * n1;

    // Staged cell of this thread: tile (sr, sc), global column clamped into the array.
    bool const stager = lid < NSTAGE;
    int const sr = lid / LX;
    int const sc = lid % LX;
    uint32_t soff = 0;
    if (stager) {
        I si = i0 - 1 + sc;
        I sj = j0 - 1 + sr;
        si = (si < n0 - 1) ? si : n0 - 1;
        sj = (sj < n1 - 1) ? sj : n1 - 1;
        soff = static_cast<uint32_t>(si + n0 * sj) * static_cast<uint32_t>(sizeof(T));
    }
    // Cell updated by a Z thread.
    I const ci = i0 + lane;
    I const cj = j0 + row;
    bool const valid = role == 2 && ci <= n0 - 2 && cj <= n1 - 2;
    I const cic = (ci < n0 - 1) ? ci : n0 - 1;
    I const cjc = (cj < n1 - 1) ? cj : n1 - 1;
    uint32_t const coff = static_cast<uint32_t>(cic + n0 * cjc) * static_cast<uint32_t>(sizeof(T));

    auto stage = [&](int b, Prim<T> const& q) {
        tile[b][0][sr][sc] = q.d;
        tile[b][1][sr][sc] = q.p;
        tile[b][2][sr][sc] = q.u0;
        tile[b][3][sr][sc] = q.u1;
        tile[b][4][sr][sc] = q.u2;
        tile[b][5][sr][sc] = sound_speed(gamma, q.d, q.p);
        tile[b][6][sr][sc] = total_energy(q, gm1);
    };
    auto cell = [&](int b, int r, int c) {
        return Prim<T> {tile[b][0][r][c], tile[b][1][r][c], tile[b][2][r][c], tile[b][3][r][c], tile[b][4][r][c]};
    };
    T const dtodv = resolve_dtodv(dta);
    T local_min = Limits<T>::max();

    // Prologue: plane kbeg staged, plane kbeg+1 in the stagers' registers, low z
    // face of the first plane in the Z threads' registers.
    Prim<T> q_stage {};
    if (stager) {
        stage(static_cast<int>(kbeg & 1), ld_prim(prim, s2 * kbeg, soff));
        q_stage = ld_prim(prim, s2 * (kbeg + 1), soff);
    }
    Flux<T> fz_lo {};
    if (role == 2) {
        Prim<T> const qm = ld_prim(prim, s2 * (kbeg - 1), coff);
        Prim<T> const q0 = ld_prim(prim, s2 * kbeg, coff);
        fz_lo = hllc_flux<2>(qm, sound_speed(gamma, qm.d, qm.p), total_energy(qm, gm1), q0,
                             sound_speed(gamma, q0.d, q0.p), total_energy(q0, gm1));
    }

    for (I k = kbeg; k < kend; ++k) {
        int const b = static_cast<int>(k & 1);
        int const bn = b ^ 1;
        I const plane = s2 * k;
        // S. stage plane k+1, prefetch plane k+2; Z threads fetch cons of plane k.
        if (stager) {
            stage(bn, q_stage);
            if (k + 2 <= kend) {
                q_stage = ld_prim(prim, plane + 2 * s2, soff);
            }
        }
        Cons<T> u {};
        if (role == 2) {
            u = ld_cons(cons, plane, coff);
        }
        lds_barrier();

        // F. faces of plane k.
        Flux<T> fz_hi {};
        if (role == 0) {
            int const r = row + 1;
            Flux<T> const f = hllc_flux<0>(cell(b, r, lane), tile[b][5][r][lane], tile[b][6][r][lane],
                                           cell(b, r, lane + 1), tile[b][5][r][lane + 1], tile[b][6][r][lane + 1]);
            fx[0][row][lane] = f.d;
            fx[1][row][lane] = f.e;
            fx[2][row][lane] = f.m0;
            fx[3][row][lane] = f.m1;
            fx[4][row][lane] = f.m2;
            if (row == 0 && lane < BY) {
                // High-side x faces of the tile: columns 64 | 65 of every row.
                int const rr = lane + 1;
                Flux<T> const g = hllc_flux<0>(cell(b, rr, BX), tile[b][5][rr][BX], tile[b][6][rr][BX],
                                               cell(b, rr, BX + 1), tile[b][5][rr][BX + 1], tile[b][6][rr][BX + 1]);
                fx[0][lane][BX] = g.d;
                fx[1][lane][BX] = g.e;
                fx[2][lane][BX] = g.m0;
                fx[3][lane][BX] = g.m1;
                fx[4][lane][BX] = g.m2;
            }
        } else if (role == 1) {
            int const c = lane + 1;
            Flux<T> const f = hllc_flux<1>(cell(b, row, c), tile[b][5][row][c], tile[b][6][row][c],
                                           cell(b, row + 1, c), tile[b][5][row + 1][c], tile[b][6][row + 1][c]);
            fy[0][row][lane] = f.d;
            fy[1][row][lane] = f.e;
            fy[2][row][lane] = f.m0;
            fy[3][row][lane] = f.m1;
            fy[4][row][lane] = f.m2;
            if (row == BY - 1) {
                // High-side y faces of the tile: rows BY | BY+1.
                Flux<T> const g = hllc_flux<1>(cell(b, BY, c), tile[b][5][BY][c], tile[b][6][BY][c],
                                               cell(b, BY + 1, c), tile[b][5][BY + 1][c], tile[b][6][BY + 1][c]);
                fy[0][BY][lane] = g.d;
                fy[1][BY][lane] = g.e;
                fy[2][BY][lane] = g.m0;
                fy[3][BY][lane] = g.m1;
                fy[4][BY][lane] = g.m2;
            }
        } else {
            int const r = row + 1;
            int const c = lane + 1;
            fz_hi = hllc_flux<2>(cell(b, r, c), tile[b][5][r][c], tile[b][6][r][c], cell(bn, r, c),
                                 tile[bn][5][r][c], tile[bn][6][r][c]);
        }
        lds_barrier();

        // U. combine and update (Z threads).
        if (role == 2) {
            Flux<T> const fxl {fx[0][row][lane], fx[1][row][lane], fx[2][row][lane], fx[3][row][lane],
                               fx[4][row][lane]};
            Flux<T> const fxh {fx[0][row][lane + 1], fx[1][row][lane + 1], fx[2][row][lane + 1],
                               fx[3][row][lane + 1], fx[4][row][lane + 1]};
            Flux<T> const fyl {fy[0][row][lane], fy[1][row][lane], fy[2][row][lane], fy[3][row][lane],
                               fy[4][row][lane]};
            Flux<T> const fyh {fy[0][row + 1][lane], fy[1][row + 1][lane], fy[2][row + 1][lane],
                               fy[3][row + 1][lane], fy[4][row + 1][lane]};
            Flux<T> acc {T(0), T(0), T(0), T(0), T(0)};
            accumulate(acc, ds0, fxh, fxl);
            accumulate(acc, ds1, fyh, fyl);
            accumulate(acc, ds2, fz_hi, fz_lo);
            apply_update(u, dtodv, acc);
            if (valid) {
                st_cons(cons, plane, coff, u);
                if constexpr (Fused) {
                    Prim<T> const q1 = cons_to_prim_pt(u, gm1);
                    st_prim(prim_out, plane, coff, q1);
                    T const c1 = sound_speed(gamma, q1.d, q1.p);
                    local_min = min_join(local_min, T(1) / inv_dt_cell(q1, c1, invdx0, invdx1, invdx2));
                }
            }
            fz_lo = fz_hi;
        }
    }

    if constexpr (Fused) {
        T const bm = block_min<T, 3 * BY>(local_min);
        if (lid == 0) {
            grid_min_commit(sink, bm, blockIdx.x, static_cast<unsigned int>(nblk));
        }
    }
}

// Ghost-cell refresh for the fused step: prim_out = cons_to_prim(cons) on the
// ghost layer only (the ghost conservatives never change, godunov.hpp:40-43), and
// the time-step minimum over those cells published to sink (time_step.hpp:28-38).
// One thread per ghost cell, enumerated face by face. Reduce = false writes the
// ghost primitives only (buffers whose ghost dt is not needed).
template <class T, class I, bool Reduce = true>
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
    for (I g = static_cast<I>(blockIdx.x) * 256 + threadIdx.x; g < total; g = grid_next(g, total, stride)) {
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
        if constexpr (Reduce) {
            T const cs = sound_speed(gamma, q.d, q.p);
            local = min_join(local, T(1) / inv_dt_cell(q, cs, invdx0, invdx1, invdx2));
        }
    }
    if constexpr (Reduce) {
        T const bm = block_min<T, 4>(local);
        if (threadIdx.x == 0) {
            grid_min_commit(sink, bm, blockIdx.x, gridDim.x);
        }
    }
}

} // namespace eb
