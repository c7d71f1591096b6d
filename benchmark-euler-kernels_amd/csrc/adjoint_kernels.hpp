// Adjoint kernels of one Godunov step with the time step frozen.
//
// Forward step n (euler_simulation.cpp:56-67): U_{n+1} = U_n - dtodv_n * R(P_n) on
// interior cells (ghosts unchanged), P_{n+1} = c2p(U_{n+1}) on all cells, where
// R(P)_c = sum_d ds_d (F_d(P_c, P_{c+d}) - F_d(P_{c-d}, P_c)) (godunov.hpp:48-90).
// With mu_{n+1} = dL/dU_{n+1}, the step's adjoint is
//   mu_n = mu_{n+1} + c2p^T( -dtodv_n * R^T(P_n) mu_{n+1}|interior ),
// and for the first step (P_0 given, U_0 = p2c(P_0)):
//   dL/dP_0 = -dtodv_0 * R^T(P_0) mu_1|interior + p2c^T mu_1.
// A face between a (low side) and b (high side) carries the weight
// g = ds_d (mu_a [a interior] - mu_b [b interior]); a receives (dF/dqL)^T g and b
// receives (dF/dqR)^T g. Each face is differentiated once and each cell's sum is
// formed by the one thread that owns the cell, so the result is race-free and
// deterministic.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "device_math.hpp"
#include "kernels.hpp"

#ifndef EB_VJP_MIN_WAVES
#define EB_VJP_MIN_WAVES 2 // waves per SIMD the tiled adjoint kernel is compiled for
#endif

namespace eb {

// Keeps the scheduler from interleaving consecutive face derivatives: each one
// holds ~100 live values, and overlapping two of them would spill.
__device__ __forceinline__ void sched_fence()
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}

template <class T>
__device__ __forceinline__ bool nonzero(Flux<T> const& w)
{
    return w.d != T(0) || w.e != T(0) || w.m0 != T(0) || w.m1 != T(0) || w.m2 != T(0);
}

template <class T>
__device__ __forceinline__ bool nonzero(PrimGrad<T> const& g)
{
    return g.d != T(0) || g.p != T(0) || g.u0 != T(0) || g.u1 != T(0) || g.u2 != T(0);
}

// Face weight ds * (mu_a [a interior] - mu_b [b interior]), the flags as 0/1 factors.
template <class T>
__device__ __forceinline__ Flux<T> face_weight(T ds, Flux<T> const& wa, bool ia, Flux<T> const& wb, bool ib)
{
    T const fa = ia ? T(1) : T(0);
    T const fb = ib ? T(1) : T(0);
    return {ds * fma(-fb, wb.d, fa * wa.d), ds * fma(-fb, wb.e, fa * wa.e), ds * fma(-fb, wb.m0, fa * wa.m0),
            ds * fma(-fb, wb.m1, fa * wa.m1), ds * fma(-fb, wb.m2, fa * wa.m2)};
}

template <class T>
__device__ __forceinline__ void add_to(PrimGrad<T>& a, PrimGrad<T> const& b)
{
    a.d += b.d;
    a.p += b.p;
    a.u0 += b.u0;
    a.u1 += b.u1;
    a.u2 += b.u2;
}

// Output of one cell from its gathered face adjoints gq (see the header comment):
// the five values, in the order of the destination's fields (mu_out's d, e, mx0,
// mx1, mx2 in mode 0, gprim's d, p, ux0, ux1, ux2 in mode 1), handed to `put`.
template <class T, int Mode, class Put>
__device__ __forceinline__ void vjp_emit(Prim<T> const& q, Flux<T> const& w, PrimGrad<T> const& gq, T dtodv, T gm1,
                                         Put&& put)
{
#pragma clang fp contract(fast)
    PrimGrad<T> const gp {-dtodv * gq.d, -dtodv * gq.p, -dtodv * gq.u0, -dtodv * gq.u1, -dtodv * gq.u2};
    if constexpr (Mode == 0) {
        Cons<T> const r = c2p_vjp(q, gp, gm1);
        put(w.d + r.d, w.e + r.e, w.m0 + r.m0, w.m1 + r.m1, w.m2 + r.m2);
    } else {
        PrimGrad<T> const r = p2c_vjp(q, Cons<T> {w.d, w.e, w.m0, w.m1, w.m2}, gm1);
        put(gp.d + r.d, gp.p + r.p, gp.u0 + r.u0, gp.u1 + r.u1, gp.u2 + r.u2);
    }
}

// The same output stored at linear cell index c.
template <class T, class I, int Mode>
__device__ __forceinline__ void vjp_output(ConsPtrs<T> const& mu_out, PrimPtrs<T> const& gprim, I c, Prim<T> const& q,
                                           Flux<T> const& w, PrimGrad<T> const& gq, T dtodv, T gm1)
{
    vjp_emit<T, Mode>(q, w, gq, dtodv, gm1, [&](T v0, T v1, T v2, T v3, T v4) {
        if constexpr (Mode == 0) {
            mu_out.d[c] = v0;
            mu_out.e[c] = v1;
            mu_out.mx0[c] = v2;
            mu_out.mx1[c] = v3;
            mu_out.mx2[c] = v4;
        } else {
            gprim.d[c] = v0;
            gprim.p[c] = v1;
            gprim.ux0[c] = v2;
            gprim.ux1[c] = v3;
            gprim.ux2[c] = v4;
        }
    });
}

// The same output stored plane-relative: scalar plane base, 32-bit lane byte offset.
template <class T, class I, int Mode>
__device__ __forceinline__ void vjp_output_at(ConsPtrs<T> const& mu_out, PrimPtrs<T> const& gprim, I plane,
                                              uint32_t off, Prim<T> const& q, Flux<T> const& w,
                                              PrimGrad<T> const& gq, T dtodv, T gm1)
{
    vjp_emit<T, Mode>(q, w, gq, dtodv, gm1, [&](T v0, T v1, T v2, T v3, T v4) {
        if constexpr (Mode == 0) {
            st(mu_out.d + plane, off, v0);
            st(mu_out.e + plane, off, v1);
            st(mu_out.mx0 + plane, off, v2);
            st(mu_out.mx1 + plane, off, v3);
            st(mu_out.mx2 + plane, off, v4);
        } else {
            st(gprim.d + plane, off, v0);
            st(gprim.p + plane, off, v1);
            st(gprim.ux0 + plane, off, v2);
            st(gprim.ux1 + plane, off, v3);
            st(gprim.ux2 + plane, off, v4);
        }
    });
}

// Ghost cells: each ghost cell on a face of the array (not an edge or corner)
// shares exactly one face with the interior; its share of that face's derivative
// is its whole gathered adjoint. Edge and corner ghosts touch no interior cell.
// Threads take ghost cells first, first + stride, ...
template <class T, class I, int Mode>
__device__ __forceinline__ void vjp_ghost_cells(
        PrimPtrs<T const> const& prim,
        ConsPtrs<T const> const& mu,
        ConsPtrs<T> const& mu_out,
        PrimPtrs<T> const& gprim,
        I n0,
        I n1,
        I n2,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        T dtodv,
        I first,
        I stride)
{
    I const nkz = (n2 > 1) ? 2 : 1;
    I const planes = nkz * n0 * n1;
    I const inner_k = (n2 > 2) ? n2 - 2 : 0;
    I const nky = (n1 > 1) ? 2 : 1;
    I const rows = inner_k * nky * n0;
    I const inner_j = (n1 > 2) ? n1 - 2 : 0;
    I const nkx = (n0 > 1) ? 2 : 1;
    I const cols = inner_k * inner_j * nkx;
    I const total = planes + rows + cols;
    auto inside = [](I a, I n) { return a >= 1 && a <= n - 2; };
    for (I g = first; g < total; g = grid_next(g, total, stride)) {
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
        Prim<T> const q = load_prim(prim, c);
        Cons<T> const u = load_cons(mu, c);
        Flux<T> const w {u.d, u.e, u.m0, u.m1, u.m2};
        PrimGrad<T> gq {T(0), T(0), T(0), T(0), T(0)};
        PrimGrad<T> dummy {T(0), T(0), T(0), T(0), T(0)};
        bool const ii = inside(i, n0);
        bool const jj = inside(j, n1);
        bool const kk = inside(k, n2);
        // the one interior neighbour, if any, and which side of the face the ghost is on
        int dir = -1;
        bool ghost_low = false;
        I other = c;
        if (jj && kk && (i == 0 || i == n0 - 1) && n0 >= 3) {
            dir = 0;
            ghost_low = i == 0;
            other = ghost_low ? c + 1 : c - 1;
        } else if (ii && kk && (j == 0 || j == n1 - 1) && n1 >= 3) {
            dir = 1;
            ghost_low = j == 0;
            other = ghost_low ? c + n0 : c - n0;
        } else if (ii && jj && (k == 0 || k == n2 - 1) && n2 >= 3) {
            dir = 2;
            ghost_low = k == 0;
            other = ghost_low ? c + n0 * n1 : c - n0 * n1;
        }
        if (dir >= 0) {
            Cons<T> const uo = load_cons(mu, other);
            Flux<T> const wo {uo.d, uo.e, uo.m0, uo.m1, uo.m2};
            T const ds = dir == 0 ? ds0 : (dir == 1 ? ds1 : ds2);
            Flux<T> const g = ghost_low ? face_weight(ds, w, false, wo, true) : face_weight(ds, wo, true, w, false);
            if (nonzero(g)) { // the face carries no adjoint otherwise
                VjpCell<T> const vg = vjp_cell(q, gamma);
                VjpCell<T> const vo = vjp_cell(load_prim(prim, other), gamma);
                if (dir == 0) {
                    ghost_low ? hllc_vjp<0>(vg, vo, gm1, g, gq, dummy) : hllc_vjp<0>(vo, vg, gm1, g, dummy, gq);
                } else if (dir == 1) {
                    ghost_low ? hllc_vjp<1>(vg, vo, gm1, g, gq, dummy) : hllc_vjp<1>(vo, vg, gm1, g, dummy, gq);
                } else {
                    ghost_low ? hllc_vjp<2>(vg, vo, gm1, g, gq, dummy) : hllc_vjp<2>(vo, vg, gm1, g, dummy, gq);
                }
            }
        }
        vjp_output<T, I, Mode>(mu_out, gprim, c, q, w, gq, dtodv, gm1);
    }
}

template <class T, class I, int Mode>
__global__ __launch_bounds__(256) void k_godunov_vjp_ghost(
        PrimPtrs<T const> prim,
        ConsPtrs<T const> mu,
        ConsPtrs<T> mu_out,
        PrimPtrs<T> gprim,
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
    vjp_ghost_cells<T, I, Mode>(prim, mu, mu_out, gprim, n0, n1, n2, gamma, gm1, ds0, ds1, ds2, resolve_dtodv(dta),
                                static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x),
                                static_cast<I>(gridDim.x) * 256);
}

// Tiled, face-once adjoint step over the interior cells (the forward kernel's
// geometry: 64 x BY tiles of interior cells, a march along k over `kchunk` planes).
// Per plane k each thread stages its cell (primitives, c and the incoming adjoint
// mu) in LDS and differentiates the x and y faces on its low side
// and the z face on its high side; two of the waves also take the tile's high-side
// x and y faces. Each face's two halves go to its two cells: the low cell's half
// by wave shuffle (x), LDS (y) or registers carried to the next plane (z). Faces
// between an interior and a ghost cell are included; the ghost cells' own outputs
// come from the first `ghost_blocks` blocks (vjp_ghost_cells), which recompute the
// one interior face of each.
template <class T, class I, int BY, int Mode>
__global__ __launch_bounds__(64 * BY, EB_VJP_MIN_WAVES) void k_godunov_vjp_tiled(
        PrimPtrs<T const> prim,
        ConsPtrs<T const> mu,
        ConsPtrs<T> mu_out,
        PrimPtrs<T> gprim,
        I n0,
        I n1,
        I n2,
        int kchunk,
        int tile0_x,
        int tile0_y,
        int tile0_z,
        int tiles_x,
        int tiles_y,
        int tiles_z,
        int ghost_blocks,
        T gamma,
        T gm1,
        T ds0,
        T ds1,
        T ds2,
        DtArg<T> dta)
{
    if (static_cast<int>(blockIdx.x) < ghost_blocks) {
        // The first blocks take the ghost cells, so that work overlaps the tiles.
        I const lt = static_cast<I>(threadIdx.x + 64 * threadIdx.y);
        vjp_ghost_cells<T, I, Mode>(prim, mu, mu_out, gprim, n0, n1, n2, gamma, gm1, ds0, ds1, ds2,
                                    resolve_dtodv(dta), static_cast<I>(blockIdx.x) * (64 * BY) + lt,
                                    static_cast<I>(ghost_blocks) * (64 * BY));
        return;
    }
    constexpr int BX = 64;
    constexpr int LX = BX + 2;
    constexpr int LY = BY + 2;
    constexpr int NHALO = 2 * BX + 2 * BY;
    constexpr int NF = 11; // d, p, u0, u1, u2, c, mu (5)
    __shared__ T sq[NF][LY][LX];
    __shared__ T sgy[5][BY + 1][BX]; // low-cell halves of the y faces: row r = face between tile rows r-1, r
    __shared__ T sgx[5][BY];         // low-cell halves of the tile's high-side x faces
    __shared__ T sgl[5][BY][BX];     // low-cell halves of the x faces: [r][c] belongs to cell c-1 of row r

    int const nblk = tiles_x * tiles_y * tiles_z;
    int const t = xcd_tile(static_cast<int>(blockIdx.x) - ghost_blocks, nblk);
    int const bx = tile0_x + t % tiles_x;
    int const by = tile0_y + (t / tiles_x) % tiles_y;
    int const bz = tile0_z + t / (tiles_x * tiles_y);
    int const tx = threadIdx.x;
    int const ty = threadIdx.y;
    int const lid = tx + BX * ty;

    I const i0 = 1 + static_cast<I>(bx) * BX;
    I const j0 = 1 + static_cast<I>(by) * BY;
    I const kbeg = 1 + static_cast<I>(bz) * kchunk;
    I const kend = (kbeg + kchunk < n2 - 1) ? kbeg + kchunk : n2 - 1;
    I const s2 = n0 * n1;

    I const i = i0 + tx;
    I const j = j0 + ty;
    auto col_int = [&](I ii) { return ii >= 1 && ii <= n0 - 2; };
    bool const row_ok = j <= n1 - 2;
    bool const own_int = col_int(i) && row_ok;
    I const ic = (i < n0 - 1) ? i : n0 - 1;
    I const jc = (j < n1 - 1) ? j : n1 - 1;
    uint32_t own = static_cast<uint32_t>(ic + n0 * jc) * static_cast<uint32_t>(sizeof(T));

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

    T const dtodv = resolve_dtodv(dta);
    ConsPtrs<T> const muw {const_cast<T*>(mu.d), const_cast<T*>(mu.e), const_cast<T*>(mu.mx0),
                           const_cast<T*>(mu.mx1), const_cast<T*>(mu.mx2)};
    auto load_mu = [&](I plane, uint32_t off) {
        Cons<T> const u = ld_cons(muw, plane, off);
        return Flux<T> {u.d, u.e, u.m0, u.m1, u.m2};
    };
    auto stage = [&](int r, int c, VjpCell<T> const& v, Flux<T> const& w) {
        sq[0][r][c] = v.q.d;
        sq[1][r][c] = v.q.p;
        sq[2][r][c] = v.q.u0;
        sq[3][r][c] = v.q.u1;
        sq[4][r][c] = v.q.u2;
        sq[5][r][c] = v.c;
        sq[6][r][c] = w.d;
        sq[7][r][c] = w.e;
        sq[8][r][c] = w.m0;
        sq[9][r][c] = w.m1;
        sq[10][r][c] = w.m2;
    };
    auto cell = [&](int r, int c) {
        return VjpCell<T> {Prim<T> {sq[0][r][c], sq[1][r][c], sq[2][r][c], sq[3][r][c], sq[4][r][c]}, sq[5][r][c]};
    };
    auto weight = [&](int r, int c) { return Flux<T> {sq[6][r][c], sq[7][r][c], sq[8][r][c], sq[9][r][c], sq[10][r][c]}; };
    PrimGrad<T> const zero {T(0), T(0), T(0), T(0), T(0)};

    // Prologue: the z face below the first plane gives the own cell's first part.
    VjpCell<T> v_cur = vjp_cell(ld_prim(prim, s2 * kbeg, own), gamma);
    Flux<T> w_cur = load_mu(s2 * kbeg, own);
    PrimGrad<T> carry = zero;
    {
        VjpCell<T> const vm = vjp_cell(ld_prim(prim, s2 * (kbeg - 1), own), gamma);
        Flux<T> const wm = load_mu(s2 * (kbeg - 1), own);
        PrimGrad<T> dummy = zero;
        hllc_vjp<2>(vm, v_cur, gm1, face_weight(ds2, wm, own_int && kbeg - 1 >= 1, w_cur, own_int), dummy, carry);
    }
    // Loads run one plane ahead of their use: the own cell of plane k+1 and the halo
    // of plane k are in registers when plane k starts, so the test of whether the
    // plane is active (which needs the adjoint of plane k+1 and of the halo) does not
    // wait for memory. Planes past the chunk are never loaded: the prefetch plane is
    // clamped to kend (<= n2-1), whose values are then not used.
    Prim<T> q_nx {};
    Flux<T> w_nx {};
    Prim<T> h_q {};
    Flux<T> h_w {};
    if (kbeg < kend) {
        q_nx = ld_prim(prim, s2 * (kbeg + 1), own);
        w_nx = load_mu(s2 * (kbeg + 1), own);
        if (has_halo) {
            h_q = ld_prim(prim, s2 * kbeg, halo);
            h_w = load_mu(s2 * kbeg, halo);
        }
    }

    for (I k = kbeg; k < kend; ++k) {
        I plane = s2 * k;
        uniform_opaque(plane);
        lane_opaque(own);
        lane_opaque(halo);
        Prim<T> const q_next = q_nx;
        Flux<T> const w_next = w_nx;
        {
            I const kp = k + 2 <= kend ? k + 2 : kend;
            q_nx = ld_prim(prim, s2 * kp, own);
            w_nx = load_mu(s2 * kp, own);
        }
        Prim<T> h_qn {};
        Flux<T> h_wn {};
        if (has_halo) {
            h_qn = ld_prim(prim, plane + s2, halo);
            h_wn = load_mu(plane + s2, halo);
        }

        // A. stage plane k. Where the incoming adjoint and the carried z-face part
        // vanish over the whole tile, its ring and the plane above, every face
        // weight of the plane is zero: the face derivatives are skipped (the
        // adjoint of a localised loss is zero outside its domain of dependence).
        stage(ty + 1, tx + 1, v_cur, w_cur);
        bool active = nonzero(w_cur) || nonzero(w_next) || nonzero(carry);
        if (has_halo) {
            stage(hrow, hcol, vjp_cell(h_q, gamma), h_w);
            active = active || nonzero(h_w);
        }
        if (__syncthreads_or(active ? 1 : 0) == 0) {
            if (own_int) {
                vjp_output_at<T, I, Mode>(mu_out, gprim, plane, own, cell(ty + 1, tx + 1).q, weight(ty + 1, tx + 1),
                                          carry, dtodv, gm1);
            }
            carry = zero;
            v_cur = vjp_cell(q_next, gamma);
            w_cur = w_next;
            h_q = h_qn;
            h_w = h_wn;
            continue;
        }

        // B. face derivatives. The own cell is read back from LDS where it is needed,
        // so that no copy of it stays in registers across the derivatives.
        PrimGrad<T> gq = carry;
        {
            PrimGrad<T> gl_x = zero;
            hllc_vjp<0>(cell(ty + 1, tx), cell(ty + 1, tx + 1), gm1,
                        face_weight(ds0, weight(ty + 1, tx), col_int(i - 1) && row_ok, weight(ty + 1, tx + 1),
                                    own_int),
                        gl_x, gq);
            sgl[0][ty][tx] = gl_x.d;
            sgl[1][ty][tx] = gl_x.p;
            sgl[2][ty][tx] = gl_x.u0;
            sgl[3][ty][tx] = gl_x.u1;
            sgl[4][ty][tx] = gl_x.u2;
        }
        sched_fence();
        {
            PrimGrad<T> gl_y = zero;
            hllc_vjp<1>(cell(ty, tx + 1), cell(ty + 1, tx + 1), gm1,
                        face_weight(ds1, weight(ty, tx + 1), col_int(i) && j - 1 >= 1 && j - 1 <= n1 - 2,
                                    weight(ty + 1, tx + 1), own_int),
                        gl_y, gq);
            sgy[0][ty][tx] = gl_y.d;
            sgy[1][ty][tx] = gl_y.p;
            sgy[2][ty][tx] = gl_y.u0;
            sgy[3][ty][tx] = gl_y.u1;
            sgy[4][ty][tx] = gl_y.u2;
        }
        sched_fence();
        if (ty == BY - 1) {
            // High-side y faces of the tile: own row and the halo row above it.
            PrimGrad<T> gl = zero;
            PrimGrad<T> dummy = zero;
            hllc_vjp<1>(cell(BY, tx + 1), cell(BY + 1, tx + 1), gm1,
                        face_weight(ds1, weight(BY, tx + 1), own_int, weight(BY + 1, tx + 1),
                                    col_int(i) && j + 1 <= n1 - 2),
                        gl, dummy);
            sgy[0][BY][tx] = gl.d;
            sgy[1][BY][tx] = gl.p;
            sgy[2][BY][tx] = gl.u0;
            sgy[3][BY][tx] = gl.u1;
            sgy[4][BY][tx] = gl.u2;
        }
        sched_fence();
        if (ty == (BY > 1 ? BY - 2 : 0) && tx < BY) {
            // High-side x faces of the tile: columns 64 and 65 of row tx+1.
            int const r = tx + 1;
            bool const rok = j0 + tx <= n1 - 2;
            PrimGrad<T> gl = zero;
            PrimGrad<T> dummy = zero;
            hllc_vjp<0>(cell(r, BX), cell(r, BX + 1), gm1,
                        face_weight(ds0, weight(r, BX), col_int(i0 + BX - 1) && rok, weight(r, BX + 1),
                                    col_int(i0 + BX) && rok),
                        gl, dummy);
            sgx[0][tx] = gl.d;
            sgx[1][tx] = gl.p;
            sgx[2][tx] = gl.u0;
            sgx[3][tx] = gl.u1;
            sgx[4][tx] = gl.u2;
        }
        sched_fence();
        VjpCell<T> const v_next = vjp_cell(q_next, gamma);
        PrimGrad<T> gr_z = zero;
        hllc_vjp<2>(cell(ty + 1, tx + 1), v_next, gm1,
                    face_weight(ds2, weight(ty + 1, tx + 1), own_int, w_next, own_int && k + 1 <= n2 - 2), gq, gr_z);
        lds_barrier();

        // C. the low-cell halves from the next lane (x) and the next row (y).
        PrimGrad<T> const hx = tx < BX - 1
                                       ? PrimGrad<T> {sgl[0][ty][tx + 1], sgl[1][ty][tx + 1], sgl[2][ty][tx + 1],
                                                      sgl[3][ty][tx + 1], sgl[4][ty][tx + 1]}
                                       : PrimGrad<T> {sgx[0][ty], sgx[1][ty], sgx[2][ty], sgx[3][ty], sgx[4][ty]};
        add_to(gq, hx);
        add_to(gq, PrimGrad<T> {sgy[0][ty + 1][tx], sgy[1][ty + 1][tx], sgy[2][ty + 1][tx], sgy[3][ty + 1][tx],
                                sgy[4][ty + 1][tx]});
        lane_opaque(own);
        if (own_int) {
            vjp_output_at<T, I, Mode>(mu_out, gprim, plane, own, cell(ty + 1, tx + 1).q, weight(ty + 1, tx + 1), gq,
                                      dtodv, gm1);
        }

        carry = gr_z;
        v_cur = v_next;
        w_cur = w_next;
        h_q = h_qn;
        h_w = h_wn;
    }
}

// gcons = c2p^T(gprim) at the states with primitives `prim` (all cells).
template <class T, class I>
__global__ __launch_bounds__(256) void k_c2p_vjp(PrimPtrs<T const> prim, PrimPtrs<T const> gprim, ConsPtrs<T> gcons,
                                                 I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I c = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); c < n; c = grid_next(c, n, stride)) {
        Prim<T> const g = load_prim(gprim, c);
        Cons<T> const r = c2p_vjp(load_prim(prim, c), PrimGrad<T> {g.d, g.p, g.u0, g.u1, g.u2}, gm1);
        store_cons(gcons, c, r);
    }
}

// gprim = p2c^T(gcons) at primitives `prim` (all cells).
template <class T, class I>
__global__ __launch_bounds__(256) void k_p2c_vjp(PrimPtrs<T const> prim, ConsPtrs<T const> gcons, PrimPtrs<T> gprim,
                                                 I n, T gm1)
{
    I const stride = static_cast<I>(gridDim.x) * 256;
    for (I c = static_cast<I>(blockIdx.x) * 256 + static_cast<I>(threadIdx.x); c < n; c = grid_next(c, n, stride)) {
        PrimGrad<T> const r = p2c_vjp(load_prim(prim, c), load_cons(gcons, c), gm1);
        store_prim(gprim, c, Prim<T> {r.d, r.p, r.u0, r.u1, r.u2});
    }
}

// Batched HLLC VJP (validation entry): face f has states qL[f], qR[f] (SoA, five
// arrays each) and weight g[f]; writes (dF/dqL)^T g and (dF/dqR)^T g.
template <class T, int Dir>
__global__ __launch_bounds__(256) void k_hllc_vjp(PrimPtrs<T const> qL, PrimPtrs<T const> qR, ConsPtrs<T const> g,
                                                  PrimPtrs<T> gL, PrimPtrs<T> gR, int64_t n, T gamma)
{
    int64_t const f = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (f >= n) {
        return;
    }
    PrimGrad<T> a {T(0), T(0), T(0), T(0), T(0)};
    PrimGrad<T> b {T(0), T(0), T(0), T(0), T(0)};
    Cons<T> const w = load_cons(g, f);
    hllc_vjp<Dir>(load_prim(qL, f), load_prim(qR, f), gamma, gamma - T(1), Flux<T> {w.d, w.e, w.m0, w.m1, w.m2}, a,
                  b);
    store_prim(gL, f, Prim<T> {a.d, a.p, a.u0, a.u1, a.u2});
    store_prim(gR, f, Prim<T> {b.d, b.p, b.u0, b.u1, b.u2});
}

// Batched HLLC flux (validation entry), same layout.
template <class T, int Dir>
__global__ __launch_bounds__(256) void k_hllc_flux(PrimPtrs<T const> qL, PrimPtrs<T const> qR, ConsPtrs<T> F,
                                                   int64_t n, T gamma)
{
    int64_t const f = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (f >= n) {
        return;
    }
    Prim<T> const a = load_prim(qL, f);
    Prim<T> const b = load_prim(qR, f);
    Flux<T> const r = hllc_flux<Dir>(a, sound_speed(gamma, a.d, a.p), b, sound_speed(gamma, b.d, b.p), gamma - T(1));
    store_cons(F, f, Cons<T> {r.d, r.e, r.m0, r.m1, r.m2});
}

} // namespace eb
