/* C ABI of the MI355X (gfx950) Godunov step for the 3-D compressible Euler
 * equations: HLLC Riemann solver, perfect-gas EOS, uniform mesh, first order.
 *
 * Every entry point takes plain device pointers and sizes, enqueues work on the
 * given HIP stream, never throws, and returns an eb_status code. Arrays follow the
 * reference's SoA layout (euler_operators/euler_arrays.hpp:77-85, 228-236): five
 * separate fields per state, layout_left, linear index i + n0*(j + n1*k), extents
 * n0 x n1 x n2 INCLUDING the one-cell ghost layer. The caller owns every array.
 *
 * index_bits selects the kernels' index type: 32 (signed 32-bit) or 64 (int64_t).
 * With 32 the total cell count must not exceed EB_MAX_CELLS_INDEX32, else
 * EB_ERR_INDEX_OVERFLOW. The index width never changes a result, so a caller whose
 * index type reaches further (the reference's `int` up to 2^31-1 cells, `unsigned`
 * up to 2^32-1, benchmarks/index_type.hpp:3-4 and the int_cast rule of
 * benchmarks/benchmark_utils.hpp:10-17) passes 64 above that limit; the C++ API
 * in include/euler/ does so.
 *
 * Each function exists for f64 (real_t = double) and f32 (real_t = float), the
 * reference's real_type switch (benchmarks/real_type.hpp:3-4).
 */
#ifndef EULER_CAPI_H
#define EULER_CAPI_H

#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest total cell count the 32-bit index kernels accept: 2^31 - 1 less 2^22
 * cells kept free so that no index formed inside a kernel can pass 2^31 - 1. */
#define EB_MAX_CELLS_INDEX32 ((int64_t)2147483647 - ((int64_t)1 << 22))

typedef enum eb_status {
    EB_OK = 0,
    EB_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative extent, bad option */
    EB_ERR_INDEX_OVERFLOW = 2,   /* extents do not fit the requested index type */
    EB_ERR_HIP = 3,              /* a HIP call or kernel launch failed */
    EB_ERR_OUT_OF_MEMORY = 4,    /* device allocation of context scratch failed */
    EB_ERR_UNSUPPORTED = 5       /* option outside the supported set */
} eb_status;

/* EulerPrim / EulerCons field order (euler_arrays.hpp:8-26). */
typedef struct eb_prim_f64 { double* d; double* p; double* ux0; double* ux1; double* ux2; } eb_prim_f64;
typedef struct eb_cons_f64 { double* d; double* e; double* mx0; double* mx1; double* mx2; } eb_cons_f64;
typedef struct eb_prim_f32 { float* d; float* p; float* ux0; float* ux1; float* ux2; } eb_prim_f32;
typedef struct eb_cons_f32 { float* d; float* e; float* mx0; float* mx1; float* mx2; } eb_cons_f32;

/* Device-side scratch for the single-pass CFL reduction (replaces the Kokkos
 * reduction buffer behind time_step.hpp:22-39). One context serves one device;
 * reductions enqueued through one context must not overlap in time (use one
 * context per concurrently running stream). */
typedef struct eb_context eb_context;

int eb_context_create(eb_context** ctx);
int eb_context_destroy(eb_context* ctx);

const char* eb_status_string(int status);
/* Message of the last failure on the calling thread ("" if none). */
const char* eb_last_error(void);

/* Godunov kernel configuration, process-wide (used by eb_godunov*, eb_step_fused*):
 *   variant 0: face-parallel waves, tile height `by` in {2, 4};
 *   variant 1: cell-parallel waves, one row of cells per wave, `by` in {4, 8, 16};
 *   variant 2: streamed (producer wave + global->LDS transfers), `by` in {4, 8};
 *              needs 16-byte aligned primitive fields (else EB_ERR_UNSUPPORTED);
 *   variant 3: linear tiles of 64*by consecutive cells of a plane (rows wrap, no
 *              lane lost to ragged x extents), `by` in {4, 8}; needs n0 <= 64*by
 *              (else EB_ERR_UNSUPPORTED);
 *   variant 4: cell-parallel waves, two rows of cells per wave (the face between
 *              them solved once in registers), `by` in {8, 16} rows (default);
 *   `by` 0 (any variant): the tile height is chosen per kernel (variant 4: 16;
 *   variant 1: 4 for the FP64 flux kernel, 8 for its fused form and for FP32);
 *   `kchunk` >= 1 planes marched per block along k, or 0: chosen per launch from
 *   the grid (long chunks, shortened until the launch has ~512 blocks).
 *   Defaults (4, 0, 0).
 * All configurations give bitwise identical results. The triple is replaced as one
 * atomic value: each launch reads a consistent configuration even while another
 * thread sets a new one, and such a change can alter speed, never results. */
int eb_set_godunov_config(int variant, int by, int kchunk);
int eb_get_godunov_config(int* variant, int* by, int* kchunk);
/* The tile height the current configuration launches for reals of real_bytes
 * (4 or 8) in the flux kernel (fused = 0) or the fused step (fused = 1). */
int eb_get_godunov_tile_height(int real_bytes, int fused, int* by);

/* ---- cons_to_prim.hpp:7-33 --------------------------------------------- */
int eb_cons_to_prim_f64(hipStream_t stream, const eb_cons_f64* cons, const eb_prim_f64* prim,
                        int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma);
int eb_cons_to_prim_f32(hipStream_t stream, const eb_cons_f32* cons, const eb_prim_f32* prim,
                        int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma);

/* ---- prim_to_cons.hpp:7-33 --------------------------------------------- */
int eb_prim_to_cons_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* cons,
                        int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma);
int eb_prim_to_cons_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* cons,
                        int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma);

/* ---- time_step.hpp:8-41 -------------------------------------------------
 * Asynchronous form: the raw dt (min over all cells, ghosts included, of
 * 1/sum_d((c+|u_d|)/dx_d); no CFL factor) is written to dt_dev[0] on the device.
 * Blocking form: same value returned to *dt_host, as the reference does. */
int eb_time_step_f64(eb_context* ctx, hipStream_t stream, const eb_prim_f64* prim,
                     int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                     double dx0, double dx1, double dx2, double* dt_dev);
int eb_time_step_f32(eb_context* ctx, hipStream_t stream, const eb_prim_f32* prim,
                     int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                     float dx0, float dx1, float dx2, float* dt_dev);
int eb_time_step_sync_f64(eb_context* ctx, hipStream_t stream, const eb_prim_f64* prim,
                          int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                          double dx0, double dx1, double dx2, double* dt_host);
int eb_time_step_sync_f32(eb_context* ctx, hipStream_t stream, const eb_prim_f32* prim,
                          int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                          float dx0, float dx1, float dx2, float* dt_host);
/* eb_time_step_f64 without a context, for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): no
 * reference counterpart. It allocates nothing. Its reduction record (4224 bytes:
 * 33 accumulators of a running minimum and a count) lives in a static device array that every
 * GPU the library runs on holds until the process ends; the record's size does
 * not grow with the extents, so these are limited only as for eb_time_step_f64.
 * The record is reset by a small kernel at the start of every call. Calls that
 * overlap in time on one GPU (two streams, or two host threads) share the record
 * and race: serialise them. Same result as eb_time_step_f64, raw dt in dt_dev[0]. */
int eb_time_step_noalloc_f64(hipStream_t stream, const eb_prim_f64* prim,
                             int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                             double dx0, double dx1, double dx2, double* dt_dev);

/* ---- godunov.hpp:15-92 (hllc, PerfectGas, UniformMesh3d) ----------------
 * Updates cons in place on interior cells from the (const) primitives with a
 * host time step dt, like the reference. prim and cons must not alias. Extents
 * below 3 in any direction leave cons unchanged. Tiled, face-once kernel. */
int eb_godunov_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* cons,
                   int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                   double dx0, double dx1, double dx2, double dt);
int eb_godunov_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* cons,
                   int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                   float dx0, float dx1, float dx2, float dt);
/* Same update with dt = cfl * dt_dev[0] read on the device (no host sync). */
int eb_godunov_dev_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* cons,
                       int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                       double dx0, double dx1, double dx2, const double* dt_dev, double cfl);
int eb_godunov_dev_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* cons,
                       int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                       float dx0, float dx1, float dx2, const float* dt_dev, float cfl);
/* Reference-order kernel (six Riemann solves per cell), for validation. */
int eb_godunov_reference_order_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* cons,
                                   int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                                   double dx0, double dx1, double dx2, double dt);
int eb_godunov_reference_order_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* cons,
                                   int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                                   float dx0, float dx1, float dx2, float dt);

/* ---- fused step (extension of euler_simulation.cpp:56-68) ---------------
 * One launch performing godunov(prim_in, cons, cfl*dt_in[0]) followed by
 * cons_to_prim(cons -> prim_out) on interior cells and the next time_step over
 * prim_out, whose raw dt is written to dt_out[0]. Ghost cells of prim_out and
 * their dt contribution come from eb_ghost_refresh (ghost_min = its output;
 * with ghost_min_dev NULL it writes the ghost primitives only, no reduction).
 * time_acc (nullable) accumulates cfl*dt_in[0] in FP64. prim_out must differ
 * from prim_in; dt_out must differ from dt_in. */
int eb_ghost_refresh_f64(eb_context* ctx, hipStream_t stream, const eb_cons_f64* cons, const eb_prim_f64* prim_out,
                         int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                         double dx0, double dx1, double dx2, double* ghost_min_dev);
int eb_ghost_refresh_f32(eb_context* ctx, hipStream_t stream, const eb_cons_f32* cons, const eb_prim_f32* prim_out,
                         int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                         float dx0, float dx1, float dx2, float* ghost_min_dev);
int eb_step_fused_f64(eb_context* ctx, hipStream_t stream, const eb_prim_f64* prim_in, const eb_cons_f64* cons,
                      const eb_prim_f64* prim_out, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                      double gamma, double dx0, double dx1, double dx2, double cfl,
                      const double* dt_in, const double* ghost_min, double* dt_out, double* time_acc);
int eb_step_fused_f32(eb_context* ctx, hipStream_t stream, const eb_prim_f32* prim_in, const eb_cons_f32* cons,
                      const eb_prim_f32* prim_out, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                      float gamma, float dx0, float dx1, float dx2, float cfl,
                      const float* dt_in, const float* ghost_min, float* dt_out, double* time_acc);
/* The same step on the box of interior cells region = {i_lo, i_hi, j_lo, j_hi,
 * k_lo, k_hi} (inclusive, clipped to the interior): only the box's cells are
 * written, the others keep their values, and dt_out is the minimum over the box.
 * dt_out may be null (then time_acc must be null too): no time step is computed,
 * for callers that replay recorded time steps.
 * Exact on the box when prim_in is exact on the box grown by one cell and cons on
 * the box. For callers that only need part of the grid advanced, such as the
 * adjoint's recompute of a localised adjoint. No reference counterpart. */
int eb_step_fused_region_f64(eb_context* ctx, hipStream_t stream, const eb_prim_f64* prim_in,
                             const eb_cons_f64* cons, const eb_prim_f64* prim_out, int64_t n0, int64_t n1,
                             int64_t n2, int index_bits, double gamma, double dx0, double dx1, double dx2,
                             double cfl, const double* dt_in, const double* ghost_min, double* dt_out,
                             double* time_acc, const int64_t* region);
int eb_step_fused_region_f32(eb_context* ctx, hipStream_t stream, const eb_prim_f32* prim_in,
                             const eb_cons_f32* cons, const eb_prim_f32* prim_out, int64_t n0, int64_t n1,
                             int64_t n2, int index_bits, float gamma, float dx0, float dx1, float dx2, float cfl,
                             const float* dt_in, const float* ghost_min, float* dt_out, double* time_acc,
                             const int64_t* region);

/* ---- init_implode.hpp:7-45 (sphere r < 0.25 about the centre) ---------- */
int eb_init_implode_f64(hipStream_t stream, const eb_prim_f64* prim, int64_t n0, int64_t n1, int64_t n2,
                        int index_bits, double dx0, double dx1, double dx2);
int eb_init_implode_f32(hipStream_t stream, const eb_prim_f32* prim, int64_t n0, int64_t n1, int64_t n2,
                        int index_bits, float dx0, float dx1, float dx2);
/* Diagonal interface x0+x1+x2 > offset -> (d_hi, p_hi) else (d_lo, p_lo), u = 0,
 * cell centres ((i-1)+0.5)*dx on the unit cube. No reference counterpart: the
 * measured implosion configuration asks for it. */
int eb_init_diagonal_f64(hipStream_t stream, const eb_prim_f64* prim, int64_t n0, int64_t n1, int64_t n2,
                         int index_bits, double dx0, double dx1, double dx2, double offset,
                         double d_hi, double p_hi, double d_lo, double p_lo);
int eb_init_diagonal_f32(hipStream_t stream, const eb_prim_f32* prim, int64_t n0, int64_t n1, int64_t n2,
                         int index_bits, float dx0, float dx1, float dx2, float offset,
                         float d_hi, float p_hi, float d_lo, float p_lo);
/* Seeded synthetic state (counter-based, launch-shape independent):
 * rho ~ U[d_lo,d_hi), p ~ U[p_lo,p_hi), u_k ~ N(0,u_sigma). No reference
 * counterpart (the reference benchmarks use uniform states,
 * benchmarks/benchmark_utils.hpp:53-87). */
int eb_init_random_f64(hipStream_t stream, const eb_prim_f64* prim, int64_t n, int index_bits, uint64_t seed,
                       double d_lo, double d_hi, double p_lo, double p_hi, double u_sigma);
int eb_init_random_f32(hipStream_t stream, const eb_prim_f32* prim, int64_t n, int index_bits, uint64_t seed,
                       double d_lo, double d_hi, double p_lo, double p_hi, double u_sigma);
/* ---- benchmarks/benchmark_utils.hpp:19-44 (init_from_value) ---------------
 * a[0..n) = value on the device, for one field of an array set. */
int eb_fill_f64(hipStream_t stream, double* a, int64_t n, double value);
int eb_fill_f32(hipStream_t stream, float* a, int64_t n, float value);
/* ---- bandwidth probe (no counterpart in the reference) -------------------
 * For measurement only: b.f[i] += a.f[i] for the five fields and i in [0, n),
 * one access of the element size per field and cell, no tiles or barriers. Per
 * cell it moves the plain godunov kernel's HBM traffic (five fields read from
 * one array set, five read and written in place in another), so its rate is
 * what plain streaming of that traffic reaches on the device. */
int eb_stream_probe_f64(hipStream_t stream, const eb_prim_f64* a, const eb_cons_f64* b, int64_t n);
int eb_stream_probe_f32(hipStream_t stream, const eb_prim_f32* a, const eb_cons_f32* b, int64_t n);


/* ---- adjoint (no reference counterpart; the reference has no derivatives) ---
 * Reverse mode of the step U_{n+1} = U_n - dtodv*R(P_n), P_{n+1} = c2p(U_{n+1}),
 * with the time step frozen (dtodv = cfl*dt_dev[0]/dv as in eb_godunov_dev).
 * mu_* are adjoints of the conservatives (cons layout), all cells.
 *   eb_godunov_vjp:       mu_out = mu_in + c2p^T(-dtodv * R^T(P) mu_in|interior)
 *   eb_godunov_vjp_first: gprim  = -dtodv * R^T(P_0) mu_in|interior + p2c^T mu_in
 * (the first step reads the given P_0, later steps read c2p(U_n)).
 * region (nullable): interior cells {i_lo, i_hi, j_lo, j_hi, k_lo, k_hi}
 * (inclusive) whose outputs are computed; interior cells outside it are not
 * written (ghost cells always are). Exact for a localised adjoint: when mu_in
 * vanishes outside the region shrunk by one cell, so does the result outside the
 * region, and zero-filled outputs stay correct. */
int eb_godunov_vjp_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* mu_in,
                       const eb_cons_f64* mu_out, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                       double gamma, double dx0, double dx1, double dx2, const double* dt_dev, double cfl,
                       const int64_t* region);
int eb_godunov_vjp_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* mu_in,
                       const eb_cons_f32* mu_out, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                       float gamma, float dx0, float dx1, float dx2, const float* dt_dev, float cfl,
                       const int64_t* region);
int eb_godunov_vjp_first_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* mu_in,
                             const eb_prim_f64* gprim, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                             double gamma, double dx0, double dx1, double dx2, const double* dt_dev, double cfl,
                             const int64_t* region);
int eb_godunov_vjp_first_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* mu_in,
                             const eb_prim_f32* gprim, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                             float gamma, float dx0, float dx1, float dx2, const float* dt_dev, float cfl,
                             const int64_t* region);
/* gcons = (d prim / d cons)^T gprim at the states with primitives prim. */
int eb_cons_to_prim_vjp_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_prim_f64* gprim,
                            const eb_cons_f64* gcons, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                            double gamma);
int eb_cons_to_prim_vjp_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_prim_f32* gprim,
                            const eb_cons_f32* gcons, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                            float gamma);
/* gprim = (d cons / d prim)^T gcons. */
int eb_prim_to_cons_vjp_f64(hipStream_t stream, const eb_prim_f64* prim, const eb_cons_f64* gcons,
                            const eb_prim_f64* gprim, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                            double gamma);
int eb_prim_to_cons_vjp_f32(hipStream_t stream, const eb_prim_f32* prim, const eb_cons_f32* gcons,
                            const eb_prim_f32* gprim, int64_t n0, int64_t n1, int64_t n2, int index_bits,
                            float gamma);
/* Batched HLLC flux (hllc.hpp:53-104) and its VJP over nfaces independent faces
 * normal to dir (0, 1, 2): validation entry points. */
int eb_hllc_flux_f64(hipStream_t stream, int dir, const eb_prim_f64* qL, const eb_prim_f64* qR,
                     const eb_cons_f64* flux, int64_t nfaces, double gamma);
int eb_hllc_flux_f32(hipStream_t stream, int dir, const eb_prim_f32* qL, const eb_prim_f32* qR,
                     const eb_cons_f32* flux, int64_t nfaces, float gamma);
int eb_hllc_vjp_f64(hipStream_t stream, int dir, const eb_prim_f64* qL, const eb_prim_f64* qR,
                    const eb_cons_f64* g, const eb_prim_f64* gL, const eb_prim_f64* gR, int64_t nfaces,
                    double gamma);
int eb_hllc_vjp_f32(hipStream_t stream, int dir, const eb_prim_f32* qL, const eb_prim_f32* qR,
                    const eb_cons_f32* g, const eb_prim_f32* gL, const eb_prim_f32* gR, int64_t nfaces,
                    float gamma);


/* ---- conservation diagnostics (FP64, compensated; no reference counterpart) --
 * Deterministic Neumaier-compensated sums in FP64, for either real type.
 * eb_conserved_totals: totals_dev[0..4] = sums over the interior cells of
 *   (d, e, mx0, mx1, mx2).
 * eb_boundary_flux: the amount B by which one godunov step with
 *   dt = cfl * dt_dev[0] (as eb_godunov_dev) lowers those totals, i.e.
 *   dt/dv * sum_d ds_d * (sum of the flux over the high boundary faces - sum over
 *   the low boundary faces), added to acc_dev[0..9] as (sum, compensation) pairs
 *   per field. With the frozen ghost layer (godunov.hpp:40-43), totals(U_n) +
 *   sum of B over the steps equals totals(U_0) up to the update's rounding. */
int eb_conserved_totals_f64(eb_context* ctx, hipStream_t stream, const eb_cons_f64* cons,
                            int64_t n0, int64_t n1, int64_t n2, int index_bits, double* totals_dev);
int eb_conserved_totals_f32(eb_context* ctx, hipStream_t stream, const eb_cons_f32* cons,
                            int64_t n0, int64_t n1, int64_t n2, int index_bits, double* totals_dev);
int eb_boundary_flux_f64(eb_context* ctx, hipStream_t stream, const eb_prim_f64* prim,
                         int64_t n0, int64_t n1, int64_t n2, int index_bits, double gamma,
                         double dx0, double dx1, double dx2, const double* dt_dev, double cfl, double* acc_dev);
int eb_boundary_flux_f32(eb_context* ctx, hipStream_t stream, const eb_prim_f32* prim,
                         int64_t n0, int64_t n1, int64_t n2, int index_bits, float gamma,
                         float dx0, float dx1, float dx2, const float* dt_dev, float cfl, double* acc_dev);

/* ---- batched ensemble (no reference counterpart) ---------------------------
 * Many independent problems, each with its own extents, gamma, cell widths and
 * CFL factor, advanced together: one persistent launch per step over a queue of
 * tiles drawn from every instance. Each instance keeps its own raw time step on
 * the device (time_step.hpp:8-41 semantics) and its FP64 simulated time; results
 * are bitwise those of the single-grid fused step (eb_step_fused_*). The caller
 * owns every array: prim_a holds the initial primitives, prim_b is a second
 * primitive buffer of the same extents, cons holds prim_to_cons(prim_a). Step
 * parity p reads buffer p (a = 0, b = 1) and writes the other.
 * Protocol: eb_batch_start; eb_batch_step(parity 0); eb_batch_refresh_ghosts(0);
 * then eb_batch_step with parity 1, 0, 1, ... (steps are graph-capturable). */
typedef struct eb_batch eb_batch;
typedef struct eb_batch_instance_f64 {
    eb_prim_f64 prim_a; eb_prim_f64 prim_b; eb_cons_f64 cons;
    int64_t n0, n1, n2;
    double gamma, dx0, dx1, dx2, cfl;
} eb_batch_instance_f64;
typedef struct eb_batch_instance_f32 {
    eb_prim_f32 prim_a; eb_prim_f32 prim_b; eb_cons_f32 cons;
    int64_t n0, n1, n2;
    float gamma, dx0, dx1, dx2, cfl;
} eb_batch_instance_f32;
int eb_batch_create_f64(eb_batch** batch, const eb_batch_instance_f64* instances, int ninstances, int index_bits);
int eb_batch_create_f32(eb_batch** batch, const eb_batch_instance_f32* instances, int ninstances, int index_bits);
int eb_batch_destroy(eb_batch* batch);
/* Tile layout 0: 64 x by columns of cells, by in {4, 8, 16}; layout 1: runs of
 * 64*by consecutive cells of a plane, by in {4, 8} (also 16 in FP32), every
 * instance with n0 <= 64*by; layout 2: 64 x by columns with two rows of cells per lane, by in
 * {8, 16}. kchunk planes per tile (0: whole columns). Default (0, 8, 32).
 * eb_batch_autotune picks all three by timing. On failure the previous layout
 * stays in place. Both wait for all work on the device before replacing the
 * tile descriptors (configuration time; not for graph capture), and a graph
 * captured before them must be captured again. */
int eb_batch_set_tiles(eb_batch* batch, int layout, int by, int kchunk);
int eb_batch_get_tiles(const eb_batch* batch, int* layout, int* by, int* kchunk, int* ntiles, int* grid);
/* Device arrays: dt = 3 reals per instance ([0], [1] ping-pong raw dt, [2] the
 * ghost cells' dt minimum), time = 1 double per instance. */
int eb_batch_device_state(const eb_batch* batch, void** dt, double** time);
/* Copies those arrays (3*ninstances reals, ninstances doubles) to caller memory,
 * host or device, enqueued on stream; either destination may be null. */
int eb_batch_copy_state(const eb_batch* batch, hipStream_t stream, void* dt_dst, double* time_dst);
/* First raw dt of every instance from prim_a (ghosts included), ghost cells of
 * prim_b from the ghost conservatives. */
int eb_batch_start(eb_batch* batch, hipStream_t stream);
/* Ghost cells of buffer (0: prim_a, 1: prim_b) from the ghost conservatives. */
int eb_batch_refresh_ghosts(eb_batch* batch, hipStream_t stream, int buffer);
/* One step of every instance in one launch. */
int eb_batch_step(eb_batch* batch, hipStream_t stream, int parity);
/* The same step as one eb_step_fused_* launch per instance, with the
 * process-wide single-grid configuration (for comparison). */
int eb_batch_step_per_instance(eb_batch* batch, hipStream_t stream, int parity);
/* Times every (tile height, kchunk) candidate on the live ensemble with a zero
 * time step (the state is left bitwise unchanged; buffer 1-parity is scratch),
 * keeps the fastest and returns its time per launch. Blocking. */
int eb_batch_autotune(eb_batch* batch, hipStream_t stream, int parity, int reps, float* best_ms);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* EULER_CAPI_H */
