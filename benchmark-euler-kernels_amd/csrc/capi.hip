// extern "C" layer of the Godunov-step library: argument checking, index-type
// dispatch (the reference's index_t switch), launch shapes, status codes.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <mutex>
#include <string>

#include <hip/hip_runtime.h>

#include "euler_capi.h"
#include "adjoint_kernels.hpp"
#include "diagnostics.hpp"
#include "kernels.hpp"

struct eb_context
{
    int device;
    eb::MinScratch* scratch; // one slot: reductions through one context are serialised
    double* dt_slot;         // device landing slot for the blocking time step
    double* partials;        // per-block (sum, compensation) pairs of the FP64 diagnostics
};

// Reduction words of eb_time_step_noalloc_f64, which allocates nothing: one
// record of the single-pass grid minimum (32 slot accumulators and the top one,
// each a running minimum's ordered encoding and a count of contributions on its
// own 128-byte line: 4224 bytes), in a device array that
// every GPU the library runs on holds until the process ends. On the host the
// name `scratch` is not a device address: host code uses the per-device address
// from hipGetSymbolAddress (noalloc_scratch below) for kernel arguments.
static __device__ __align__(256) eb::MinScratch scratch[1];

namespace {

thread_local std::string g_last_error;

// Godunov launch configuration (process-wide): kernel variant (4 = paired rows,
// the default), tile height (0: chosen per kernel by auto_by) and planes per block
// (0: sized per launch). The three live in one atomic word, so every launch reads
// a consistent triple even while another thread reconfigures; all configurations
// give bitwise identical results, so a concurrent change affects speed only.
struct GodunovConfig
{
    int variant;
    int by;
    int kchunk;
};

std::atomic<uint64_t> g_config {uint64_t(4) << 56};

uint64_t pack_config(int variant, int by, int kchunk)
{
    return (uint64_t(uint8_t(variant)) << 56) | (uint64_t(uint8_t(by)) << 48) | uint64_t(uint32_t(kchunk));
}

GodunovConfig load_config()
{
    uint64_t const w = g_config.load();
    return GodunovConfig {int((w >> 56) & 0xff), int((w >> 48) & 0xff), int(uint32_t(w & 0xffffffffu))};
}

bool config_supported(int variant, int by)
{
    return (variant >= 0 && variant <= 4 && by == 0) || (variant == 0 && (by == 2 || by == 4))
           || (variant == 4 && (by == 8 || by == 16))
           || (variant == 1 && (by == 4 || by == 8 || by == 16)) || (variant == 2 && (by == 4 || by == 8))
           || (variant == 3 && (by == 4 || by == 8));
}

// Tile height when the configuration leaves it open (by = 0), from MI355X timings
// at 512^3. Paired rows (variant 4, the default): 64 x 16 tiles, eight waves of two
// rows each, for the flux kernel and the fused step in FP64 and FP32. One row per
// lane (variant 1): the FP64 flux kernel needs ~148 VGPRs, three waves per SIMD,
// which 4-wave blocks (three per CU) can use and 8-wave blocks (one per CU) cannot,
// so it takes by = 4; its fused form and the FP32 kernels run best with by = 8.
int auto_by(int variant, std::size_t real_bytes, bool fused)
{
    if (variant == 0) {
        return 4;
    }
    if (variant == 4) {
        return 16;
    }
    if (variant == 1 && real_bytes == 8 && !fused) {
        return 4;
    }
    return 8;
}

int fail(int status, std::string const& msg)
{
    g_last_error = msg;
    return status;
}

int check_launch(char const* what)
{
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        return fail(EB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
    return EB_OK;
}

// Blocks of the FP64 diagnostic reductions (partials held by the context).
constexpr int kDiagBlocks = 2048;

// Largest total cell count accepted with 32-bit indices. Grid-stride loops step
// with grid_next, which never forms an index past n; the 2^22 cells below 2^31 - 1
// are a margin for any other index arithmetic near the end of the arrays.
constexpr int64_t kMax32 = EB_MAX_CELLS_INDEX32;
static_assert(kMax32 <= std::numeric_limits<int32_t>::max() - (int64_t(1) << 22));

int check_extents(int64_t n0, int64_t n1, int64_t n2, int index_bits)
{
    if (n0 < 0 || n1 < 0 || n2 < 0) {
        return fail(EB_ERR_INVALID_ARGUMENT, "negative extent");
    }
    if (index_bits != 32 && index_bits != 64) {
        return fail(EB_ERR_INVALID_ARGUMENT, "index_bits must be 32 or 64");
    }
    // Overflow-safe product check.
    __int128 const total = static_cast<__int128>(n0) * n1 * n2;
    __int128 const limit = index_bits == 32 ? kMax32 : std::numeric_limits<int64_t>::max() / 4;
    if (total > limit) {
        return fail(EB_ERR_INDEX_OVERFLOW, "extents do not fit the index type");
    }
    return EB_OK;
}

template <class T, class S>
eb::PrimPtrs<T> prim_ptrs(S const* s)
{
    return {s->d, s->p, s->ux0, s->ux1, s->ux2};
}

template <class T, class S>
eb::ConsPtrs<T> cons_ptrs(S const* s)
{
    return {s->d, s->e, s->mx0, s->mx1, s->mx2};
}

template <class S>
bool prim_null(S const* s)
{
    return s == nullptr || s->d == nullptr || s->p == nullptr || s->ux0 == nullptr || s->ux1 == nullptr
           || s->ux2 == nullptr;
}

template <class S>
bool cons_null(S const* s)
{
    return s == nullptr || s->d == nullptr || s->e == nullptr || s->mx0 == nullptr || s->mx1 == nullptr
           || s->mx2 == nullptr;
}

unsigned int pointwise_grid(int64_t n)
{
    int64_t const blocks = (n + 255) / 256;
    return static_cast<unsigned int>(blocks < (int64_t(1) << 20) ? (blocks > 0 ? blocks : 1) : (int64_t(1) << 20));
}

unsigned int reduction_grid(int64_t n)
{
    // About eight resident blocks per CU on a 256-CU part; fewer for small inputs.
    int64_t const blocks = (n + 1023) / 1024;
    return static_cast<unsigned int>(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
}

// Blocks of a grid-stride minimum over n items whose blocks each commit to the
// device-wide record (grid_min_commit: a device-scope release, which writes back
// the XCD's L2, and two device-scope atomics; tens of ns per block, measured
// ~40-50 ns when all blocks shared one word). The streaming side wants many blocks, the
// commit few: about 2 n^(1/3) blocks, at most 768, is within a few percent of the
// fastest block count measured at every size from 128^3 to 512^3 cells in FP64
// and FP32 (time_step 512^3 FP64 1.03 -> 0.84 ms, 256^3 0.20 -> 0.12 ms, 128^3
// 0.061 -> 0.026 ms against eight blocks per CU), and no block gets fewer than
// 256 items.
unsigned int commit_grid(int64_t n)
{
    int64_t const by_commit = static_cast<int64_t>(2.0 * std::cbrt(static_cast<double>(n > 0 ? n : 0)));
    int64_t const by_work = (n + 255) / 256;
    int64_t const blocks = std::min<int64_t>(768, std::min(by_commit, by_work));
    return static_cast<unsigned int>(blocks > 0 ? blocks : 1);
}

// ---------------------------------------------------------------- launchers

template <class T, class I, class CS, class PS>
int run_cons_to_prim(hipStream_t s, CS const* cons, PS const* prim, int64_t n0, int64_t n1, int64_t n2, T gamma)
{
    int64_t const n = n0 * n1 * n2;
    if (n == 0) {
        return EB_OK;
    }
    auto const c = cons_ptrs<T>(cons);
    eb::ConsPtrs<T const> const cc {c.d, c.e, c.mx0, c.mx1, c.mx2};
    hipLaunchKernelGGL((eb::k_cons_to_prim<T, I>), dim3(pointwise_grid(n)), dim3(256), 0, s, cc,
                       prim_ptrs<T>(prim), static_cast<I>(n), static_cast<T>(gamma - 1));
    return check_launch("cons_to_prim");
}

template <class T, class I, class PS, class CS>
int run_prim_to_cons(hipStream_t s, PS const* prim, CS const* cons, int64_t n0, int64_t n1, int64_t n2, T gamma)
{
    int64_t const n = n0 * n1 * n2;
    if (n == 0) {
        return EB_OK;
    }
    auto const p = prim_ptrs<T>(prim);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    hipLaunchKernelGGL((eb::k_prim_to_cons<T, I>), dim3(pointwise_grid(n)), dim3(256), 0, s, pc,
                       cons_ptrs<T>(cons), static_cast<I>(n), static_cast<T>(gamma - 1));
    return check_launch("prim_to_cons");
}

template <class T>
eb::DtSink<T> plain_sink(eb::MinScratch* scratch, T* dt_out)
{
    return eb::DtSink<T> {scratch, dt_out, nullptr, nullptr, nullptr, T(0)};
}

template <class T>
eb::DtSink<T> plain_sink(eb_context* ctx, T* dt_out)
{
    return plain_sink<T>(ctx->scratch, dt_out);
}

template <class T>
bool aligned16(eb::PrimPtrs<T const> const& p)
{
    auto a = [](void const* q) { return (reinterpret_cast<uintptr_t>(q) & 15u) == 0; };
    return a(p.d) && a(p.p) && a(p.ux0) && a(p.ux1) && a(p.ux2);
}

template <class T, class I, class PS>
int run_time_step(
        eb::MinScratch* scratch,
        hipStream_t s,
        PS const* prim,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        T gamma,
        T dx0,
        T dx1,
        T dx2,
        T* dt_dev)
{
    int64_t const n = n0 * n1 * n2;
    auto const p = prim_ptrs<T>(prim);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    // time_step.hpp:18-20: the inverse spacings are formed on the host.
    T const invdx0 = T(1) / dx0;
    T const invdx1 = T(1) / dx1;
    T const invdx2 = T(1) / dx2;
    if (aligned16(pc)) {
        // sized in groups of 16 bytes, so that each thread has four groups in flight
        hipLaunchKernelGGL((eb::k_time_step_vec<T, I>), dim3(commit_grid(n / (16 / int64_t(sizeof(T))))), dim3(256),
                           0, s, pc, static_cast<I>(n), gamma, invdx0, invdx1, invdx2, plain_sink<T>(scratch, dt_dev));
    } else {
        hipLaunchKernelGGL((eb::k_time_step<T, I>), dim3(commit_grid(n)), dim3(256), 0, s, pc, static_cast<I>(n),
                           gamma, invdx0, invdx1, invdx2, plain_sink<T>(scratch, dt_dev));
    }
    return check_launch("time_step");
}

struct Mesh
{
    template <class T>
    static void faces(T dx0, T dx1, T dx2, T& ds0, T& ds1, T& ds2, T& dv)
    {
        // uniform_mesh.hpp:33-51
        ds0 = dx1 * dx2;
        ds1 = dx2 * dx0;
        ds2 = dx0 * dx1;
        dv = dx0 * dx1 * dx2;
    }
};

template <int Variant, class T, class I, int BY, bool Fused>
void launch_tiled(
        hipStream_t s,
        eb::PrimPtrs<T const> const& pc,
        eb::ConsPtrs<T> const& cu,
        eb::PrimPtrs<T> const& pout,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        int kchunk,
        T gamma,
        T ds0,
        T ds1,
        T ds2,
        eb::DtArg<T> const& dta,
        T invdx0,
        T invdx1,
        T invdx2,
        eb::DtSink<T> const& sink)
{
    int const tiles_x = static_cast<int>((n0 - 2 + 63) / 64);
    int const tiles_y = static_cast<int>((n1 - 2 + BY - 1) / BY);
    int const tiles_z = static_cast<int>((n2 - 2 + kchunk - 1) / kchunk);
    int const nblk = tiles_x * tiles_y * tiles_z;
    if constexpr (Variant == 3) {
        int const tiles_plane = static_cast<int>((n0 * (n1 - 2) + 64 * BY - 1) / (64 * BY));
        hipLaunchKernelGGL((eb::k_godunov_linear<T, I, BY, Fused>), dim3(tiles_plane * tiles_z), dim3(64, BY), 0, s,
                           pc, cu, pout, static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk,
                           tiles_plane, tiles_z, gamma, static_cast<T>(gamma - 1), ds0, ds1, ds2, dta, invdx0, invdx1,
                           invdx2, sink);
    } else if constexpr (Variant == 2) {
        hipLaunchKernelGGL((eb::k_godunov_stream<T, I, BY, Fused>), dim3(nblk), dim3(64, BY + 1), 0, s, pc, cu, pout,
                           static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk, tiles_x, tiles_y,
                           tiles_z, gamma, static_cast<T>(gamma - 1), ds0, ds1, ds2, dta, invdx0, invdx1, invdx2,
                           sink);
    } else if constexpr (Variant == 4) {
        hipLaunchKernelGGL((eb::k_godunov_pair<T, I, BY, Fused>), dim3(nblk), dim3(64, BY / 2), 0, s, pc, cu, pout,
                           static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk, tiles_x, tiles_y,
                           tiles_z, gamma, static_cast<T>(gamma - 1), ds0, ds1, ds2, dta, invdx0, invdx1, invdx2,
                           sink);
    } else if constexpr (Variant == 0) {
        hipLaunchKernelGGL((eb::k_godunov_faces<T, I, BY, Fused>), dim3(nblk), dim3(64, 3 * BY), 0, s, pc, cu, pout,
                           static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk, tiles_x, tiles_y,
                           tiles_z, gamma, static_cast<T>(gamma - 1), ds0, ds1, ds2, dta, invdx0, invdx1, invdx2,
                           sink);
    } else {
        hipLaunchKernelGGL((eb::k_godunov_tiled<T, I, BY, Fused>), dim3(nblk), dim3(64, BY), 0, s, pc, cu, pout,
                           static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk, tiles_x, tiles_y,
                           tiles_z, gamma, static_cast<T>(gamma - 1), ds0, ds1, ds2, dta, invdx0, invdx1, invdx2,
                           sink);
    }
}

// Planes marched per block by the paired-row kernels, which run one block per CU:
// the fewest rounds of 256 blocks (one per CU) whose chunks are at most `lmax`
// planes long, with the chunks as long as that round count allows. Small grids
// get one round of short chunks (64^3: 1-plane chunks, 256 blocks, 12 us against
// 45 us with 8-plane chunks and 32 blocks); long chunks save the plane a block
// re-reads below its chunk. Measured lmax: 128 planes in FP64 (128^3: 8 planes;
// 320^3: 64; 512^3: 128), 32 in the FP32 flux kernel (faster at 400^3 and 448^3,
// level at 256^3, 320^3 and 512^3), 128 in the fused step in both precisions (every
// block commits to the one time-step word, so fewer blocks pay: FP32 at 512^3
// 2.70 against 2.91 ms with 32-plane chunks, 256^3 0.353 against 0.383 ms).
int64_t pair_lmax(std::size_t real_bytes, bool fused)
{
    return real_bytes == 8 || fused ? 128 : 32;
}

int64_t pair_kchunk(int64_t columns, int64_t nz, int64_t lmax, int64_t slots = 256)
{
    if (nz < 1) {
        return 1;
    }
    for (int64_t rounds = 1;; ++rounds) {
        int64_t const per = std::max<int64_t>(1, rounds * slots / columns); // chunks per column
        int64_t const kc = (nz + per - 1) / per;
        if (kc <= lmax || per >= nz) {
            return kc;
        }
    }
}

// The fused step of the paired-row kernel (64 x 16 tiles) on the box of interior
// cells [lo, hi] (inclusive, per axis): the tiles that meet the box are marched
// over the box's planes and only the box's cells are written. Used where only part
// of the grid has to advance (the adjoint's recompute of a localised adjoint).
template <class T, class I>
int launch_fused_region(hipStream_t s, eb::PrimPtrs<T const> const& pc, eb::ConsPtrs<T> const& cu,
                        eb::PrimPtrs<T> const& pout, int64_t n0, int64_t n1, int64_t n2, int64_t const* lo,
                        int64_t const* hi, T gamma, T ds0, T ds1, T ds2, eb::DtArg<T> const& dta, T invdx0, T invdx1,
                        T invdx2, eb::DtSink<T> const& sink)
{
    constexpr int BY = 16;
    int64_t const x0 = (lo[0] - 1) / 64;
    int64_t const x1 = (hi[0] - 1) / 64 + 1;
    int64_t const y0 = (lo[1] - 1) / BY;
    int64_t const y1 = (hi[1] - 1) / BY + 1;
    int64_t const columns = (x1 - x0) * (y1 - y0);
    int const kchunk = static_cast<int>(pair_kchunk(columns, hi[2] - lo[2] + 1, pair_lmax(sizeof(T), true)));
    int64_t const z0 = (lo[2] - 1) / kchunk;
    int64_t const z1 = (hi[2] - 1) / kchunk + 1;
    int64_t const nblk = columns * (z1 - z0);
    if (nblk > std::numeric_limits<int32_t>::max() / 2) {
        return fail(EB_ERR_INDEX_OVERFLOW, "tile count exceeds the grid limit");
    }
    eb::CellBox<I> const box {static_cast<I>(lo[0]), static_cast<I>(hi[0]), static_cast<I>(lo[1]),
                              static_cast<I>(hi[1]), static_cast<I>(lo[2]), static_cast<I>(hi[2])};
    // Without dt_out the step computes no time step (callers that replay recorded ones).
    auto const kernel = sink.dt_out != nullptr ? eb::k_godunov_box_pair<T, I, BY, true>
                                               : eb::k_godunov_box_pair<T, I, BY, false>;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned int>(nblk)), dim3(64, BY / 2), 0, s, pc, cu, pout,
                       static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk, static_cast<int>(x1 - x0),
                       static_cast<int>(y1 - y0), static_cast<int>(z1 - z0), static_cast<int>(x0),
                       static_cast<int>(y0), static_cast<int>(z0), box, gamma, static_cast<T>(gamma - 1), ds0, ds1,
                       ds2, dta, invdx0, invdx1, invdx2, sink);
    return check_launch("step_fused_region");
}

// Planes marched per block when the configured kchunk is 0. A block re-reads one
// primitive plane below its chunk, so long chunks save traffic; the chunk is
// shortened until the launch has about 512 blocks (two per CU) so that small
// grids still fill the GPU: 128 planes at 512^3 (2048 blocks), 8 at 128^3.
int auto_kchunk(int64_t n0, int64_t n1, int64_t n2, int by)
{
    int64_t const nz = n2 - 2;
    int64_t const columns = ((n0 - 2 + 63) / 64) * ((n1 - 2 + by - 1) / by);
    int64_t kc = nz * columns / 512;
    kc = kc < 8 ? 8 : (kc > 128 ? 128 : kc);
    return static_cast<int>(kc < nz ? kc : (nz > 0 ? nz : 1));
}

template <class T, class I, bool Fused>
int dispatch_tiled(
        hipStream_t s,
        eb::PrimPtrs<T const> const& pc,
        eb::ConsPtrs<T> const& cu,
        eb::PrimPtrs<T> const& pout,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        T gamma,
        T ds0,
        T ds1,
        T ds2,
        eb::DtArg<T> const& dta,
        T invdx0,
        T invdx1,
        T invdx2,
        eb::DtSink<T> const& sink)
{
    GodunovConfig const cfg = load_config();
    int const variant = cfg.variant;
    int const by_set = cfg.by;
    int const by = by_set > 0 ? by_set : auto_by(variant, sizeof(T), Fused);
    int const kc_set = cfg.kchunk;
    int const kchunk = kc_set > 0 ? kc_set
                       : variant == 4 ? static_cast<int>(pair_kchunk(((n0 - 2 + 63) / 64) * ((n1 - 2 + by - 1) / by),
                                                                     n2 - 2, pair_lmax(sizeof(T), Fused)))
                                      : auto_kchunk(n0, n1, n2, by);
    // The block count must fit a 32-bit grid and a plane must fit 32-bit offsets.
    int64_t const columns = variant == 3 ? (n0 * (n1 - 2) + 64 * by - 1) / (64 * by)
                                         : ((n0 - 2 + 63) / 64) * ((n1 - 2 + by - 1) / by);
    int64_t const tiles = columns * ((n2 - 2 + kchunk - 1) / kchunk);
    if (tiles > std::numeric_limits<int32_t>::max()) {
        return fail(EB_ERR_INDEX_OVERFLOW, "tile count exceeds the grid limit");
    }
    if (static_cast<__int128>(n0) * n1 * static_cast<int64_t>(sizeof(T)) >= (static_cast<__int128>(1) << 32)) {
        return fail(EB_ERR_UNSUPPORTED, "a k-plane must hold fewer than 2^32 bytes per field");
    }
#define EB_LAUNCH(V, B)                                                                                          \
    launch_tiled<V, T, I, B, Fused>(s, pc, cu, pout, n0, n1, n2, kchunk, gamma, ds0, ds1, ds2, dta, invdx0, invdx1, \
                                    invdx2, sink)
    if (variant == 2 && !aligned16(pc)) {
        return fail(EB_ERR_UNSUPPORTED, "the streamed variant needs 16-byte aligned primitive fields");
    }
    if (variant == 3 && n0 > 64 * by) {
        return fail(EB_ERR_UNSUPPORTED, "linear tiles need rows of at most 64*by cells (n0 <= 64*by)");
    }
    if (variant == 4 && by == 8) {
        EB_LAUNCH(4, 8);
    } else if (variant == 4 && by == 16) {
        EB_LAUNCH(4, 16);
    } else if (variant == 3 && by == 4) {
        EB_LAUNCH(3, 4);
    } else if (variant == 3 && by == 8) {
        EB_LAUNCH(3, 8);
    } else if (variant == 2 && by == 4) {
        EB_LAUNCH(2, 4);
    } else if (variant == 2 && by == 8) {
        EB_LAUNCH(2, 8);
    } else if (variant == 0 && by == 2) {
        EB_LAUNCH(0, 2);
    } else if (variant == 0 && by == 4) {
        EB_LAUNCH(0, 4);
    } else if (variant == 1 && by == 4) {
        EB_LAUNCH(1, 4);
    } else if (variant == 1 && by == 8) {
        EB_LAUNCH(1, 8);
    } else if (variant == 1 && by == 16) {
        EB_LAUNCH(1, 16);
    } else {
        return fail(EB_ERR_UNSUPPORTED, "unsupported godunov configuration");
    }
#undef EB_LAUNCH
    return check_launch(Fused ? "step_fused" : "godunov");
}

template <class T, class I, class PS, class CS>
int run_godunov(
        hipStream_t s,
        PS const* prim,
        CS const* cons,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        T gamma,
        T dx0,
        T dx1,
        T dx2,
        T dt,
        T const* dt_dev,
        T cfl,
        bool reference_order)
{
    if (n0 < 3 || n1 < 3 || n2 < 3) {
        return EB_OK; // no interior cell (godunov.hpp:40-43 gives an empty range)
    }
    T ds0;
    T ds1;
    T ds2;
    T dv;
    Mesh::faces(dx0, dx1, dx2, ds0, ds1, ds2, dv);
    eb::DtArg<T> const dta {dt / dv, dt_dev, cfl, dv};
    auto const p = prim_ptrs<T>(prim);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    auto const cu = cons_ptrs<T>(cons);
    if (reference_order) {
        int64_t const blocks = ((n0 - 2 + 63) / 64) * ((n1 - 2 + 3) / 4) * (n2 - 2);
        if (blocks > std::numeric_limits<int32_t>::max()) {
            return fail(EB_ERR_INDEX_OVERFLOW, "grid too large for the reference-order kernel");
        }
        hipLaunchKernelGGL((eb::k_godunov_naive<T, I>), dim3(static_cast<unsigned int>(blocks)), dim3(64, 4), 0, s,
                           pc, cu, static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), gamma,
                           static_cast<T>(gamma - 1), ds0, ds1, ds2, dta);
        return check_launch("godunov_reference_order");
    }
    eb::DtSink<T> const none {nullptr, nullptr, nullptr, nullptr, nullptr, T(0)};
    return dispatch_tiled<T, I, false>(s, pc, cu, eb::PrimPtrs<T> {}, n0, n1, n2, gamma, ds0, ds1, ds2, dta, T(0),
                                       T(0), T(0), none);
}

template <class T, class I, class CS, class PS>
int run_ghost_refresh(
        eb_context* ctx,
        hipStream_t s,
        CS const* cons,
        PS const* prim_out,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        T gamma,
        T dx0,
        T dx1,
        T dx2,
        T* ghost_min)
{
    int64_t const n = n0 * n1 * n2;
    int64_t const inner = (n0 > 2 ? n0 - 2 : 0) * (n1 > 2 ? n1 - 2 : 0) * (n2 > 2 ? n2 - 2 : 0);
    auto const c = cons_ptrs<T>(cons);
    eb::ConsPtrs<T const> const cc {c.d, c.e, c.mx0, c.mx1, c.mx2};
    if (ghost_min == nullptr) {
        // ghost primitives only: no commit, so one thread per ghost cell
        int64_t const blocks = std::min<int64_t>(2048, (n - inner + 255) / 256);
        hipLaunchKernelGGL((eb::k_ghost_refresh<T, I, false>), dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, cc,
                           prim_ptrs<T>(prim_out), static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2),
                           gamma, static_cast<T>(gamma - 1), T(1) / dx0, T(1) / dx1, T(1) / dx2,
                           eb::DtSink<T> {nullptr, nullptr, nullptr, nullptr, nullptr, T(0)});
        return check_launch("ghost_refresh");
    }
    hipLaunchKernelGGL((eb::k_ghost_refresh<T, I>), dim3(commit_grid(n - inner)), dim3(256), 0, s, cc,
                       prim_ptrs<T>(prim_out), static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), gamma,
                       static_cast<T>(gamma - 1), T(1) / dx0, T(1) / dx1, T(1) / dx2,
                       plain_sink<T>(ctx, ghost_min));
    return check_launch("ghost_refresh");
}

template <class T, class I, class PS, class CS>
int run_step_fused(
        eb_context* ctx,
        hipStream_t s,
        PS const* prim_in,
        CS const* cons,
        PS const* prim_out,
        int64_t n0,
        int64_t n1,
        int64_t n2,
        T gamma,
        T dx0,
        T dx1,
        T dx2,
        T cfl,
        T const* dt_in,
        T const* ghost_min,
        T* dt_out,
        double* time_acc,
        int64_t const* region)
{
    T ds0;
    T ds1;
    T ds2;
    T dv;
    Mesh::faces(dx0, dx1, dx2, ds0, ds1, ds2, dv);
    eb::DtArg<T> const dta {T(0), dt_in, cfl, dv};
    auto const p = prim_ptrs<T>(prim_in);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    eb::DtSink<T> const sink {ctx->scratch, dt_out, ghost_min, time_acc, dt_in, cfl};
    if (region != nullptr) {
        int64_t lo[3] = {1, 1, 1};
        int64_t hi[3] = {n0 - 2, n1 - 2, n2 - 2};
        for (int d = 0; d < 3; ++d) {
            lo[d] = region[2 * d] > lo[d] ? region[2 * d] : lo[d];
            hi[d] = region[2 * d + 1] < hi[d] ? region[2 * d + 1] : hi[d];
        }
        if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) {
            return EB_OK; // nothing to advance
        }
        if (static_cast<__int128>(n0) * n1 * static_cast<int64_t>(sizeof(T)) >= (static_cast<__int128>(1) << 32)) {
            return fail(EB_ERR_UNSUPPORTED, "a k-plane must hold fewer than 2^32 bytes per field");
        }
        return launch_fused_region<T, I>(s, pc, cons_ptrs<T>(cons), prim_ptrs<T>(prim_out), n0, n1, n2, lo, hi,
                                         gamma, ds0, ds1, ds2, dta, T(1) / dx0, T(1) / dx1, T(1) / dx2, sink);
    }
    return dispatch_tiled<T, I, true>(s, pc, cons_ptrs<T>(cons), prim_ptrs<T>(prim_out), n0, n1, n2, gamma, ds0,
                                      ds1, ds2, dta, T(1) / dx0, T(1) / dx1, T(1) / dx2, sink);
}

template <class T, class I, class CS>
int run_conserved_totals(eb_context* ctx, hipStream_t s, CS const* cons, int64_t n0, int64_t n1, int64_t n2,
                         double* totals)
{
    int blocks = 0;
    if (n0 >= 3 && n1 >= 3 && n2 >= 3) {
        int64_t const rows = (n1 - 2) * (n2 - 2);
        blocks = static_cast<int>(rows < kDiagBlocks ? rows : kDiagBlocks);
        auto const c = cons_ptrs<T>(cons);
        eb::ConsPtrs<T const> const cc {c.d, c.e, c.mx0, c.mx1, c.mx2};
        hipLaunchKernelGGL((eb::k_conserved_partials<T, I>), dim3(blocks), dim3(eb::kDiagThreads), 0, s, cc,
                           static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), ctx->partials);
    }
    hipLaunchKernelGGL(eb::k_compensated_final, dim3(1), dim3(64), 0, s, ctx->partials, blocks, totals, 0);
    return check_launch("conserved_totals");
}

template <class T, class I, class PS>
int run_boundary_flux(eb_context* ctx, hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, T gamma,
                      T dx0, T dx1, T dx2, T const* dt_dev, T cfl, double* acc)
{
    if (n0 < 3 || n1 < 3 || n2 < 3) {
        return EB_OK; // no interior: nothing flows
    }
    T ds0;
    T ds1;
    T ds2;
    T dv;
    Mesh::faces(dx0, dx1, dx2, ds0, ds1, ds2, dv);
    int64_t const faces = 2 * ((n1 - 2) * (n2 - 2) + (n0 - 2) * (n2 - 2) + (n0 - 2) * (n1 - 2));
    int64_t const want = (faces + eb::kDiagThreads - 1) / eb::kDiagThreads;
    int const blocks = static_cast<int>(want < kDiagBlocks ? want : kDiagBlocks);
    auto const p = prim_ptrs<T>(prim);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    eb::DtArg<T> const dta {T(0), dt_dev, cfl, dv};
    hipLaunchKernelGGL((eb::k_boundary_partials<T, I>), dim3(blocks), dim3(eb::kDiagThreads), 0, s, pc,
                       static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), gamma,
                       static_cast<T>(gamma - 1), ds0, ds1, ds2, dta, ctx->partials);
    hipLaunchKernelGGL(eb::k_compensated_final, dim3(1), dim3(64), 0, s, ctx->partials, blocks, acc, 1);
    return check_launch("boundary_flux");
}

// Device address of `scratch` on the current device, looked up once per device
// (hipGetSymbolAddress allocates nothing) and kept for later calls.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
eb::MinScratch* g_noalloc_scratch[kMaxDevices] = {};

int noalloc_scratch(eb::MinScratch** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        return fail(EB_ERR_HIP, "no current HIP device (or device index >= 64)");
    }
    std::lock_guard<std::mutex> const lock(g_noalloc_mutex);
    if (g_noalloc_scratch[dev] == nullptr) {
        void* p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            return fail(EB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_noalloc_scratch[dev] = static_cast<eb::MinScratch*>(p);
    }
    *out = g_noalloc_scratch[dev];
    return EB_OK;
}

} // namespace

// ======================================================================= ABI

extern "C" {

int eb_context_create(eb_context** ctx)
{
    if (ctx == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "ctx is null");
    }
    *ctx = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        return fail(EB_ERR_HIP, "hipGetDevice failed");
    }
    eb::MinScratch* scratch = nullptr;
    double* dt_slot = nullptr;
    double* partials = nullptr;
    if (hipMalloc(&scratch, sizeof(eb::MinScratch)) != hipSuccess) {
        return fail(EB_ERR_OUT_OF_MEMORY, "hipMalloc of reduction scratch failed");
    }
    if (hipMalloc(&dt_slot, sizeof(double)) != hipSuccess
        || hipMalloc(&partials, sizeof(double) * 10 * kDiagBlocks) != hipSuccess) {
        (void)hipFree(scratch);
        (void)hipFree(dt_slot);
        return fail(EB_ERR_OUT_OF_MEMORY, "hipMalloc of context scratch failed");
    }
    hipLaunchKernelGGL(eb::min_scratch_init, dim3(1), dim3(64), 0, nullptr, scratch, 1);
    if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(scratch);
        (void)hipFree(dt_slot);
        (void)hipFree(partials);
        return fail(EB_ERR_HIP, "reduction scratch initialisation failed");
    }
    *ctx = new eb_context {dev, scratch, dt_slot, partials};
    return EB_OK;
}

int eb_context_destroy(eb_context* ctx)
{
    if (ctx == nullptr) {
        return EB_OK;
    }
    // every buffer is released even if one release fails
    bool ok = hipFree(ctx->scratch) == hipSuccess;
    ok = (hipFree(ctx->dt_slot) == hipSuccess) && ok;
    ok = (hipFree(ctx->partials) == hipSuccess) && ok;
    int const rc = ok ? EB_OK : fail(EB_ERR_HIP, "hipFree failed");
    delete ctx;
    return rc;
}

int eb_time_step_noalloc_f64(hipStream_t stream, const eb_prim_f64* prim, int64_t n0, int64_t n1, int64_t n2,
                             int index_bits, double gamma, double dx0, double dx1, double dx2, double* dt_dev)
{
    int rc = check_extents(n0, n1, n2, index_bits);
    if (rc != EB_OK) {
        return rc;
    }
    if (dt_dev == nullptr || (n0 * n1 * n2 > 0 && prim_null(prim))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    eb::MinScratch* words = nullptr;
    if ((rc = noalloc_scratch(&words)) != EB_OK) {
        return rc;
    }
    // The record starts from its initial state on every call, set by a fill kernel
    // (not a memset, so that the call can be captured in a graph).
    hipLaunchKernelGGL(eb::min_scratch_init, dim3(1), dim3(64), 0, stream, words, 1);
    if ((rc = check_launch("time_step_noalloc: scratch init")) != EB_OK) {
        return rc;
    }
    return index_bits == 32
                   ? run_time_step<double, int32_t>(words, stream, prim, n0, n1, n2, gamma, dx0, dx1, dx2, dt_dev)
                   : run_time_step<double, int64_t>(words, stream, prim, n0, n1, n2, gamma, dx0, dx1, dx2, dt_dev);
}

const char* eb_status_string(int status)
{
    switch (status) {
    case EB_OK: return "ok";
    case EB_ERR_INVALID_ARGUMENT: return "invalid argument";
    case EB_ERR_INDEX_OVERFLOW: return "index overflow";
    case EB_ERR_HIP: return "HIP error";
    case EB_ERR_OUT_OF_MEMORY: return "out of device memory";
    case EB_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown status";
    }
}

const char* eb_last_error(void)
{
    return g_last_error.c_str();
}

int eb_set_godunov_config(int variant, int by, int kchunk)
{
    if (kchunk < 0) {
        return fail(EB_ERR_INVALID_ARGUMENT, "kchunk must be >= 0 (0: automatic)");
    }
    if (!config_supported(variant, by)) {
        return fail(EB_ERR_UNSUPPORTED,
                    "unsupported (variant, tile height): (0-4, 0 = automatic), (0, 2|4), (1, 4|8|16), (2, 4|8), "
                    "(3, 4|8), (4, 8|16)");
    }
    g_config.store(pack_config(variant, by, kchunk));
    return EB_OK;
}

int eb_get_godunov_tile_height(int real_bytes, int fused, int* by)
{
    if (by == nullptr || (real_bytes != 4 && real_bytes != 8)) {
        return fail(EB_ERR_INVALID_ARGUMENT, "real_bytes must be 4 or 8 and by non-null");
    }
    GodunovConfig const cfg = load_config();
    *by = cfg.by > 0 ? cfg.by : auto_by(cfg.variant, static_cast<std::size_t>(real_bytes), fused != 0);
    return EB_OK;
}

int eb_get_godunov_config(int* variant, int* by, int* kchunk)
{
    if (variant == nullptr || by == nullptr || kchunk == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null output");
    }
    GodunovConfig const cfg = load_config();
    *variant = cfg.variant;
    *by = cfg.by;
    *kchunk = cfg.kchunk;
    return EB_OK;
}

} // extern "C"

namespace {

#define EB_CHECK(expr)          \
    do {                        \
        int const rc_ = (expr); \
        if (rc_ != EB_OK) {     \
            return rc_;         \
        }                       \
    } while (0)

// Entry-point bodies, shared by the f64 and f32 symbols. Extents are checked
// first; array pointers may be null only when there is no cell to touch.

template <class T, class CS, class PS>
int abi_cons_to_prim(hipStream_t s, CS const* cons, PS const* prim, int64_t n0, int64_t n1, int64_t n2, int bits,
                     T gamma)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (n0 * n1 * n2 > 0 && (cons_null(cons) || prim_null(prim))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    return bits == 32 ? run_cons_to_prim<T, int32_t>(s, cons, prim, n0, n1, n2, gamma)
                      : run_cons_to_prim<T, int64_t>(s, cons, prim, n0, n1, n2, gamma);
}

template <class T, class PS, class CS>
int abi_prim_to_cons(hipStream_t s, PS const* prim, CS const* cons, int64_t n0, int64_t n1, int64_t n2, int bits,
                     T gamma)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (n0 * n1 * n2 > 0 && (cons_null(cons) || prim_null(prim))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    return bits == 32 ? run_prim_to_cons<T, int32_t>(s, prim, cons, n0, n1, n2, gamma)
                      : run_prim_to_cons<T, int64_t>(s, prim, cons, n0, n1, n2, gamma);
}

template <class T, class PS>
int abi_time_step(eb_context* ctx, hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, int bits,
                  T gamma, T dx0, T dx1, T dx2, T* dt_dev)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (ctx == nullptr || dt_dev == nullptr || (n0 * n1 * n2 > 0 && prim_null(prim))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    return bits == 32 ? run_time_step<T, int32_t>(ctx->scratch, s, prim, n0, n1, n2, gamma, dx0, dx1, dx2, dt_dev)
                      : run_time_step<T, int64_t>(ctx->scratch, s, prim, n0, n1, n2, gamma, dx0, dx1, dx2, dt_dev);
}

template <class T, class PS>
int abi_time_step_sync(eb_context* ctx, hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, int bits,
                       T gamma, T dx0, T dx1, T dx2, T* dt_host)
{
    if (ctx == nullptr || dt_host == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    T* const dt_dev = reinterpret_cast<T*>(ctx->dt_slot);
    EB_CHECK(abi_time_step<T>(ctx, s, prim, n0, n1, n2, bits, gamma, dx0, dx1, dx2, dt_dev));
    if (hipMemcpyAsync(dt_host, dt_dev, sizeof(T), hipMemcpyDeviceToHost, s) != hipSuccess
        || hipStreamSynchronize(s) != hipSuccess) {
        return fail(EB_ERR_HIP, "dt copy to host failed");
    }
    return EB_OK;
}

template <class T, class PS, class CS>
int abi_godunov(hipStream_t s, PS const* prim, CS const* cons, int64_t n0, int64_t n1, int64_t n2, int bits,
                T gamma, T dx0, T dx1, T dx2, T dt, T const* dt_dev, T cfl, bool reference_order)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (n0 * n1 * n2 > 0 && (cons_null(cons) || prim_null(prim))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    return bits == 32
                   ? run_godunov<T, int32_t>(s, prim, cons, n0, n1, n2, gamma, dx0, dx1, dx2, dt, dt_dev, cfl,
                                             reference_order)
                   : run_godunov<T, int64_t>(s, prim, cons, n0, n1, n2, gamma, dx0, dx1, dx2, dt, dt_dev, cfl,
                                             reference_order);
}

template <class T, class CS, class PS>
int abi_ghost_refresh(eb_context* ctx, hipStream_t s, CS const* cons, PS const* prim_out, int64_t n0, int64_t n1,
                      int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T* ghost_min)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (ctx == nullptr || (n0 * n1 * n2 > 0 && (cons_null(cons) || prim_null(prim_out)))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    return bits == 32 ? run_ghost_refresh<T, int32_t>(ctx, s, cons, prim_out, n0, n1, n2, gamma, dx0, dx1, dx2,
                                                      ghost_min)
                      : run_ghost_refresh<T, int64_t>(ctx, s, cons, prim_out, n0, n1, n2, gamma, dx0, dx1, dx2,
                                                      ghost_min);
}

template <class T, class PS, class CS>
int abi_step_fused(eb_context* ctx, hipStream_t s, PS const* prim_in, CS const* cons, PS const* prim_out, int64_t n0,
                   int64_t n1, int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T cfl, T const* dt_in,
                   T const* ghost_min, T* dt_out, double* time_acc, int64_t const* region = nullptr)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (n0 < 3 || n1 < 3 || n2 < 3) {
        return fail(EB_ERR_INVALID_ARGUMENT, "the fused step needs at least one interior cell");
    }
    if (ctx == nullptr || prim_null(prim_in) || cons_null(cons) || prim_null(prim_out) || dt_in == nullptr
        || (dt_out == nullptr && region == nullptr)) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    if (dt_out == nullptr && time_acc != nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "time_acc needs dt_out (the time is accumulated by the dt reduction)");
    }
    if (prim_in->d == prim_out->d || dt_in == dt_out) {
        return fail(EB_ERR_INVALID_ARGUMENT, "prim_out / dt_out must not alias the inputs");
    }
    return bits == 32 ? run_step_fused<T, int32_t>(ctx, s, prim_in, cons, prim_out, n0, n1, n2, gamma, dx0, dx1,
                                                   dx2, cfl, dt_in, ghost_min, dt_out, time_acc, region)
                      : run_step_fused<T, int64_t>(ctx, s, prim_in, cons, prim_out, n0, n1, n2, gamma, dx0, dx1,
                                                   dx2, cfl, dt_in, ghost_min, dt_out, time_acc, region);
}

template <class T>
int abi_fill(hipStream_t s, T* a, int64_t n, T v)
{
    if (n < 0) {
        return fail(EB_ERR_INVALID_ARGUMENT, "negative element count");
    }
    if (n == 0) {
        return EB_OK;
    }
    if (a == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    hipLaunchKernelGGL((eb::k_fill<T>), dim3(pointwise_grid(n)), dim3(256), 0, s, a, n, v);
    return check_launch("fill");
}

template <class T, class PS, class CS>
int abi_stream_probe(hipStream_t s, PS const* a, CS const* b, int64_t n)
{
    if (n < 0) {
        return fail(EB_ERR_INVALID_ARGUMENT, "negative cell count");
    }
    if (n == 0) {
        return EB_OK;
    }
    if (prim_null(a) || cons_null(b)) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    auto const p = prim_ptrs<T>(a);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    hipLaunchKernelGGL((eb::k_stream_probe<T>), dim3(pointwise_grid(n)), dim3(256), 0, s, pc, cons_ptrs<T>(b), n);
    return check_launch("stream_probe");
}

template <class T, class I, class PS>
void launch_implode(hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, T dx0, T dx1, T dx2)
{
    hipLaunchKernelGGL((eb::k_init_implode<T, I>), dim3(pointwise_grid(n0 * n1 * n2)), dim3(256), 0, s,
                       prim_ptrs<T>(prim), static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), dx0, dx1,
                       dx2);
}

template <class T, class PS>
int abi_init_implode(hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, int bits, T dx0, T dx1,
                     T dx2)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (n0 * n1 * n2 == 0) {
        return EB_OK;
    }
    if (prim_null(prim)) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    if (bits == 32) {
        launch_implode<T, int32_t>(s, prim, n0, n1, n2, dx0, dx1, dx2);
    } else {
        launch_implode<T, int64_t>(s, prim, n0, n1, n2, dx0, dx1, dx2);
    }
    return check_launch("init_implode");
}

template <class T, class I, class PS>
void launch_diagonal(hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, T const* a)
{
    hipLaunchKernelGGL((eb::k_init_diagonal<T, I>), dim3(pointwise_grid(n0 * n1 * n2)), dim3(256), 0, s,
                       prim_ptrs<T>(prim), static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), a[0], a[1],
                       a[2], a[3], a[4], a[5], a[6], a[7]);
}

template <class T, class PS>
int abi_init_diagonal(hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, int bits, T dx0, T dx1,
                      T dx2, T offset, T d_hi, T p_hi, T d_lo, T p_lo)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (n0 * n1 * n2 == 0) {
        return EB_OK;
    }
    if (prim_null(prim)) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    T const a[8] = {dx0, dx1, dx2, offset, d_hi, p_hi, d_lo, p_lo};
    if (bits == 32) {
        launch_diagonal<T, int32_t>(s, prim, n0, n1, n2, a);
    } else {
        launch_diagonal<T, int64_t>(s, prim, n0, n1, n2, a);
    }
    return check_launch("init_diagonal");
}

template <class T, class PS>
int abi_init_random(hipStream_t s, PS const* prim, int64_t n, int bits, uint64_t seed, double d_lo, double d_hi,
                    double p_lo, double p_hi, double u_sigma)
{
    EB_CHECK(check_extents(n, 1, 1, bits));
    if (n == 0) {
        return EB_OK;
    }
    if (prim_null(prim)) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    auto const p = prim_ptrs<T>(prim);
    auto const sd = static_cast<unsigned long long>(seed);
    if (bits == 32) {
        hipLaunchKernelGGL((eb::k_init_random<T, int32_t>), dim3(pointwise_grid(n)), dim3(256), 0, s, p,
                           static_cast<int32_t>(n), sd, d_lo, d_hi, p_lo, p_hi, u_sigma);
    } else {
        hipLaunchKernelGGL((eb::k_init_random<T, int64_t>), dim3(pointwise_grid(n)), dim3(256), 0, s, p, n, sd,
                           d_lo, d_hi, p_lo, p_hi, u_sigma);
    }
    return check_launch("init_random");
}

// ---- adjoint --------------------------------------------------------------

template <class T, class PS, class CS>
int abi_godunov_vjp(hipStream_t s, PS const* prim, CS const* mu_in, CS const* mu_out, PS const* gprim, int64_t n0,
                    int64_t n1, int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T const* dt_dev, T cfl,
                    int64_t const* region)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    bool const first = gprim != nullptr;
    if (prim_null(prim) || cons_null(mu_in) || dt_dev == nullptr || (first ? prim_null(gprim) : cons_null(mu_out))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    if (!first && mu_in->d == mu_out->d) {
        return fail(EB_ERR_INVALID_ARGUMENT, "mu_out must not alias mu_in");
    }
    int64_t const n = n0 * n1 * n2;
    if (n == 0) {
        return EB_OK;
    }
    T ds0;
    T ds1;
    T ds2;
    T dv;
    Mesh::faces(dx0, dx1, dx2, ds0, ds1, ds2, dv);
    eb::DtArg<T> const dta {T(0), dt_dev, cfl, dv};
    auto const p = prim_ptrs<T>(prim);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    auto const mi = cons_ptrs<T>(mu_in);
    eb::ConsPtrs<T const> const mic {mi.d, mi.e, mi.mx0, mi.mx1, mi.mx2};
    eb::ConsPtrs<T> const mo = first ? eb::ConsPtrs<T> {} : cons_ptrs<T>(mu_out);
    eb::PrimPtrs<T> const gp = first ? prim_ptrs<T>(gprim) : eb::PrimPtrs<T> {};
    T const gm1 = gamma - 1;
    // Interior cells: the tiled face-once kernel; ghost cells: one face each.
#ifndef EB_VJP_BY
#define EB_VJP_BY 4
#endif
    constexpr int kBy = EB_VJP_BY;
    bool const interior = n0 >= 3 && n1 >= 3 && n2 >= 3;
    // Interior cells [lo, hi] per axis to process (all of them by default): cells
    // outside keep their mu_out / gprim untouched, which is exact when mu_in and
    // the result vanish there (a localised adjoint).
    int64_t lo[3] = {1, 1, 1};
    int64_t hi[3] = {n0 - 2, n1 - 2, n2 - 2};
    if (region != nullptr) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = region[2 * d] > lo[d] ? region[2 * d] : lo[d];
            hi[d] = region[2 * d + 1] < hi[d] ? region[2 * d + 1] : hi[d];
        }
    }
    bool const tiles = interior && lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2];
    int tiles_x = 0;
    int tiles_y = 0;
    int tiles_z = 0;
    int tile0_x = 0;
    int tile0_y = 0;
    int tile0_z = 0;
    int kchunk = 1;
    if (interior) {
        if (static_cast<__int128>(n0) * n1 * static_cast<int64_t>(sizeof(T)) >= (static_cast<__int128>(1) << 32)) {
            return fail(EB_ERR_UNSUPPORTED, "a k-plane must hold fewer than 2^32 bytes per field");
        }
    }
    if (tiles) {
        int64_t const x0 = (lo[0] - 1) / 64;
        int64_t const x1 = (hi[0] - 1) / 64 + 1;
        int64_t const y0 = (lo[1] - 1) / kBy;
        int64_t const y1 = (hi[1] - 1) / kBy + 1;
        int64_t const columns = (x1 - x0) * (y1 - y0);
        int64_t const nz = hi[2] - lo[2] + 1;
        // two blocks per CU (512 at a time), chunks of at most 8 planes (1024 blocks at
        // 128^3); small boxes get one round of short chunks
        int64_t const kc = pair_kchunk(columns, nz, 8, 512);
        kchunk = static_cast<int>(kc < n2 - 2 ? kc : n2 - 2);
        int64_t const z0 = (lo[2] - 1) / kchunk;
        int64_t const z1 = (hi[2] - 1) / kchunk + 1;
        if (columns * (z1 - z0) > std::numeric_limits<int32_t>::max() / 2) {
            return fail(EB_ERR_INDEX_OVERFLOW, "tile count exceeds the grid limit");
        }
        tile0_x = static_cast<int>(x0);
        tile0_y = static_cast<int>(y0);
        tile0_z = static_cast<int>(z0);
        tiles_x = static_cast<int>(x1 - x0);
        tiles_y = static_cast<int>(y1 - y0);
        tiles_z = static_cast<int>(z1 - z0);
    }
    int64_t const inner = interior ? (n0 - 2) * (n1 - 2) * (n2 - 2) : 0;
    // Ghost cells ride in the first blocks of the tiled launch (a multiple of 8,
    // so the tiles keep their XCD-contiguous map); without interior, alone.
    int64_t gb = (n - inner + 64 * kBy - 1) / (64 * kBy);
    gb = gb > 1024 ? 1024 : gb;
    int const ghost_blocks = static_cast<int>((gb + 7) / 8 * 8);
    dim3 const tgrid(static_cast<unsigned int>(ghost_blocks + tiles_x * tiles_y * tiles_z));
    dim3 const ggrid(reduction_grid(n - inner));
#define EB_VJP(I, M)                                                                                                 \
    do {                                                                                                             \
        if (interior) {                                                                                              \
            hipLaunchKernelGGL((eb::k_godunov_vjp_tiled<T, I, kBy, M>), tgrid, dim3(64, kBy), 0, s, pc, mic, mo, gp, \
                               static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), kchunk, tile0_x, tile0_y, \
                               tile0_z, tiles_x, tiles_y, tiles_z, ghost_blocks, gamma, gm1, ds0, ds1, ds2, dta);    \
        } else {                                                                                                     \
            hipLaunchKernelGGL((eb::k_godunov_vjp_ghost<T, I, M>), ggrid, dim3(256), 0, s, pc, mic, mo, gp,           \
                               static_cast<I>(n0), static_cast<I>(n1), static_cast<I>(n2), gamma, gm1, ds0, ds1, ds2, \
                               dta);                                                                                 \
        }                                                                                                            \
    } while (0)
    if (bits == 32) {
        if (first) {
            EB_VJP(int32_t, 1);
        } else {
            EB_VJP(int32_t, 0);
        }
    } else {
        if (first) {
            EB_VJP(int64_t, 1);
        } else {
            EB_VJP(int64_t, 0);
        }
    }
#undef EB_VJP
    return check_launch("godunov_vjp");
}

template <class T, class PS, class CS>
int abi_c2p_vjp(hipStream_t s, PS const* prim, PS const* gprim, CS const* gcons, int64_t n0, int64_t n1, int64_t n2,
                int bits, T gamma)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    int64_t const n = n0 * n1 * n2;
    if (n > 0 && (prim_null(prim) || prim_null(gprim) || cons_null(gcons))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    if (n == 0) {
        return EB_OK;
    }
    auto const p = prim_ptrs<T>(prim);
    auto const g = prim_ptrs<T>(gprim);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    eb::PrimPtrs<T const> const gc {g.d, g.p, g.ux0, g.ux1, g.ux2};
    if (bits == 32) {
        hipLaunchKernelGGL((eb::k_c2p_vjp<T, int32_t>), dim3(pointwise_grid(n)), dim3(256), 0, s, pc, gc,
                           cons_ptrs<T>(gcons), static_cast<int32_t>(n), static_cast<T>(gamma - 1));
    } else {
        hipLaunchKernelGGL((eb::k_c2p_vjp<T, int64_t>), dim3(pointwise_grid(n)), dim3(256), 0, s, pc, gc,
                           cons_ptrs<T>(gcons), n, static_cast<T>(gamma - 1));
    }
    return check_launch("cons_to_prim_vjp");
}

template <class T, class PS, class CS>
int abi_p2c_vjp(hipStream_t s, PS const* prim, CS const* gcons, PS const* gprim, int64_t n0, int64_t n1, int64_t n2,
                int bits, T gamma)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    int64_t const n = n0 * n1 * n2;
    if (n > 0 && (prim_null(prim) || prim_null(gprim) || cons_null(gcons))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    if (n == 0) {
        return EB_OK;
    }
    auto const p = prim_ptrs<T>(prim);
    auto const g = cons_ptrs<T>(gcons);
    eb::PrimPtrs<T const> const pc {p.d, p.p, p.ux0, p.ux1, p.ux2};
    eb::ConsPtrs<T const> const gc {g.d, g.e, g.mx0, g.mx1, g.mx2};
    if (bits == 32) {
        hipLaunchKernelGGL((eb::k_p2c_vjp<T, int32_t>), dim3(pointwise_grid(n)), dim3(256), 0, s, pc, gc,
                           prim_ptrs<T>(gprim), static_cast<int32_t>(n), static_cast<T>(gamma - 1));
    } else {
        hipLaunchKernelGGL((eb::k_p2c_vjp<T, int64_t>), dim3(pointwise_grid(n)), dim3(256), 0, s, pc, gc,
                           prim_ptrs<T>(gprim), n, static_cast<T>(gamma - 1));
    }
    return check_launch("prim_to_cons_vjp");
}

template <class T, class PS, class CS>
int abi_hllc(hipStream_t s, int dir, PS const* qL, PS const* qR, CS const* g, CS const* flux, PS const* gL,
             PS const* gR, int64_t nfaces, T gamma)
{
    if (dir < 0 || dir > 2 || nfaces < 0) {
        return fail(EB_ERR_INVALID_ARGUMENT, "dir must be 0, 1 or 2 and nfaces >= 0");
    }
    if (nfaces == 0) {
        return EB_OK;
    }
    if (nfaces > (int64_t(1) << 40)) {
        return fail(EB_ERR_INDEX_OVERFLOW, "too many faces");
    }
    bool const vjp = flux == nullptr;
    if (prim_null(qL) || prim_null(qR) || (vjp ? (cons_null(g) || prim_null(gL) || prim_null(gR)) : cons_null(flux))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null array");
    }
    auto cp = [](PS const* q) {
        auto const a = prim_ptrs<T>(q);
        return eb::PrimPtrs<T const> {a.d, a.p, a.ux0, a.ux1, a.ux2};
    };
    dim3 const grid(static_cast<unsigned int>((nfaces + 255) / 256));
    if (vjp) {
        auto const w = cons_ptrs<T>(g);
        eb::ConsPtrs<T const> const wc {w.d, w.e, w.mx0, w.mx1, w.mx2};
#define EB_HV(D)                                                                                                   \
    hipLaunchKernelGGL((eb::k_hllc_vjp<T, D>), grid, dim3(256), 0, s, cp(qL), cp(qR), wc, prim_ptrs<T>(gL),      \
                       prim_ptrs<T>(gR), nfaces, gamma)
        if (dir == 0) {
            EB_HV(0);
        } else if (dir == 1) {
            EB_HV(1);
        } else {
            EB_HV(2);
        }
#undef EB_HV
        return check_launch("hllc_vjp");
    }
#define EB_HF(D)                                                                                                   \
    hipLaunchKernelGGL((eb::k_hllc_flux<T, D>), grid, dim3(256), 0, s, cp(qL), cp(qR), cons_ptrs<T>(flux), nfaces, \
                       gamma)
    if (dir == 0) {
        EB_HF(0);
    } else if (dir == 1) {
        EB_HF(1);
    } else {
        EB_HF(2);
    }
#undef EB_HF
    return check_launch("hllc_flux");
}

template <class T, class CS>
int abi_conserved_totals(eb_context* ctx, hipStream_t s, CS const* cons, int64_t n0, int64_t n1, int64_t n2,
                         int bits, double* totals)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (ctx == nullptr || totals == nullptr || (n0 >= 3 && n1 >= 3 && n2 >= 3 && cons_null(cons))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    return bits == 32 ? run_conserved_totals<T, int32_t>(ctx, s, cons, n0, n1, n2, totals)
                      : run_conserved_totals<T, int64_t>(ctx, s, cons, n0, n1, n2, totals);
}

template <class T, class PS>
int abi_boundary_flux(eb_context* ctx, hipStream_t s, PS const* prim, int64_t n0, int64_t n1, int64_t n2, int bits,
                      T gamma, T dx0, T dx1, T dx2, T const* dt_dev, T cfl, double* acc)
{
    EB_CHECK(check_extents(n0, n1, n2, bits));
    if (ctx == nullptr || acc == nullptr || dt_dev == nullptr
        || (n0 >= 3 && n1 >= 3 && n2 >= 3 && prim_null(prim))) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    return bits == 32 ? run_boundary_flux<T, int32_t>(ctx, s, prim, n0, n1, n2, gamma, dx0, dx1, dx2, dt_dev, cfl, acc)
                      : run_boundary_flux<T, int64_t>(ctx, s, prim, n0, n1, n2, gamma, dx0, dx1, dx2, dt_dev, cfl,
                                                      acc);
}

#undef EB_CHECK

} // namespace

extern "C" {

// ---- f64 / f32 symbols: thin forwarding wrappers --------------------------

#define EB_ABI(SUF, T)                                                                                               \
    int eb_cons_to_prim_##SUF(hipStream_t s, const eb_cons_##SUF* c, const eb_prim_##SUF* p, int64_t n0,            \
                              int64_t n1, int64_t n2, int bits, T gamma)                                            \
    {                                                                                                                \
        return abi_cons_to_prim<T>(s, c, p, n0, n1, n2, bits, gamma);                                               \
    }                                                                                                                \
    int eb_prim_to_cons_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* c, int64_t n0,            \
                              int64_t n1, int64_t n2, int bits, T gamma)                                            \
    {                                                                                                                \
        return abi_prim_to_cons<T>(s, p, c, n0, n1, n2, bits, gamma);                                               \
    }                                                                                                                \
    int eb_time_step_##SUF(eb_context* ctx, hipStream_t s, const eb_prim_##SUF* p, int64_t n0, int64_t n1,          \
                           int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T* dt_dev)                           \
    {                                                                                                                \
        return abi_time_step<T>(ctx, s, p, n0, n1, n2, bits, gamma, dx0, dx1, dx2, dt_dev);                         \
    }                                                                                                                \
    int eb_time_step_sync_##SUF(eb_context* ctx, hipStream_t s, const eb_prim_##SUF* p, int64_t n0, int64_t n1,     \
                                int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T* dt_host)                     \
    {                                                                                                                \
        return abi_time_step_sync<T>(ctx, s, p, n0, n1, n2, bits, gamma, dx0, dx1, dx2, dt_host);                   \
    }                                                                                                                \
    int eb_godunov_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* c, int64_t n0, int64_t n1,     \
                         int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T dt)                                  \
    {                                                                                                                \
        return abi_godunov<T>(s, p, c, n0, n1, n2, bits, gamma, dx0, dx1, dx2, dt, nullptr, T(0), false);           \
    }                                                                                                                \
    int eb_godunov_dev_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* c, int64_t n0, int64_t n1, \
                             int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, const T* dt_dev, T cfl)            \
    {                                                                                                                \
        if (dt_dev == nullptr) {                                                                                     \
            return fail(EB_ERR_INVALID_ARGUMENT, "null dt_dev");                                                     \
        }                                                                                                            \
        return abi_godunov<T>(s, p, c, n0, n1, n2, bits, gamma, dx0, dx1, dx2, T(0), dt_dev, cfl, false);           \
    }                                                                                                                \
    int eb_godunov_reference_order_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* c, int64_t n0, \
                                         int64_t n1, int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, T dt)      \
    {                                                                                                                \
        return abi_godunov<T>(s, p, c, n0, n1, n2, bits, gamma, dx0, dx1, dx2, dt, nullptr, T(0), true);            \
    }                                                                                                                \
    int eb_ghost_refresh_##SUF(eb_context* ctx, hipStream_t s, const eb_cons_##SUF* c, const eb_prim_##SUF* p,      \
                               int64_t n0, int64_t n1, int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2,          \
                               T* ghost_min)                                                                         \
    {                                                                                                                \
        return abi_ghost_refresh<T>(ctx, s, c, p, n0, n1, n2, bits, gamma, dx0, dx1, dx2, ghost_min);               \
    }                                                                                                                \
    int eb_step_fused_##SUF(eb_context* ctx, hipStream_t s, const eb_prim_##SUF* pin, const eb_cons_##SUF* c,       \
                            const eb_prim_##SUF* pout, int64_t n0, int64_t n1, int64_t n2, int bits, T gamma,       \
                            T dx0, T dx1, T dx2, T cfl, const T* dt_in, const T* ghost_min, T* dt_out,              \
                            double* time_acc)                                                                        \
    {                                                                                                                \
        return abi_step_fused<T>(ctx, s, pin, c, pout, n0, n1, n2, bits, gamma, dx0, dx1, dx2, cfl, dt_in,          \
                                 ghost_min, dt_out, time_acc);                                                       \
    }                                                                                                                \
    int eb_step_fused_region_##SUF(eb_context* ctx, hipStream_t s, const eb_prim_##SUF* pin, const eb_cons_##SUF* c,\
                                   const eb_prim_##SUF* pout, int64_t n0, int64_t n1, int64_t n2, int bits, T gamma,\
                                   T dx0, T dx1, T dx2, T cfl, const T* dt_in, const T* ghost_min, T* dt_out,      \
                                   double* time_acc, const int64_t* region)                                          \
    {                                                                                                                \
        return abi_step_fused<T>(ctx, s, pin, c, pout, n0, n1, n2, bits, gamma, dx0, dx1, dx2, cfl, dt_in,          \
                                 ghost_min, dt_out, time_acc, region);                                               \
    }                                                                                                                \
    int eb_init_implode_##SUF(hipStream_t s, const eb_prim_##SUF* p, int64_t n0, int64_t n1, int64_t n2, int bits,  \
                              T dx0, T dx1, T dx2)                                                                   \
    {                                                                                                                \
        return abi_init_implode<T>(s, p, n0, n1, n2, bits, dx0, dx1, dx2);                                          \
    }                                                                                                                \
    int eb_init_diagonal_##SUF(hipStream_t s, const eb_prim_##SUF* p, int64_t n0, int64_t n1, int64_t n2, int bits, \
                               T dx0, T dx1, T dx2, T offset, T d_hi, T p_hi, T d_lo, T p_lo)                        \
    {                                                                                                                \
        return abi_init_diagonal<T>(s, p, n0, n1, n2, bits, dx0, dx1, dx2, offset, d_hi, p_hi, d_lo, p_lo);         \
    }                                                                                                                \
    int eb_stream_probe_##SUF(hipStream_t s, const eb_prim_##SUF* a, const eb_cons_##SUF* b, int64_t n)             \
    {                                                                                                                \
        return abi_stream_probe<T>(s, a, b, n);                                                                      \
    }                                                                                                                \
    int eb_fill_##SUF(hipStream_t s, T* a, int64_t n, T v)                                                          \
    {                                                                                                                \
        return abi_fill<T>(s, a, n, v);                                                                              \
    }                                                                                                                \
    int eb_init_random_##SUF(hipStream_t s, const eb_prim_##SUF* p, int64_t n, int bits, uint64_t seed,             \
                             double d_lo, double d_hi, double p_lo, double p_hi, double u_sigma)                     \
    {                                                                                                                \
        return abi_init_random<T>(s, p, n, bits, seed, d_lo, d_hi, p_lo, p_hi, u_sigma);                            \
    }                                                                                                                \
    int eb_godunov_vjp_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* mu_in,                    \
                             const eb_cons_##SUF* mu_out, int64_t n0, int64_t n1, int64_t n2, int bits, T gamma,     \
                             T dx0, T dx1, T dx2, const T* dt_dev, T cfl, const int64_t* region)                     \
    {                                                                                                                \
        return abi_godunov_vjp<T, eb_prim_##SUF, eb_cons_##SUF>(s, p, mu_in, mu_out, nullptr, n0, n1, n2, bits,     \
                                                                gamma, dx0, dx1, dx2, dt_dev, cfl, region);           \
    }                                                                                                                \
    int eb_godunov_vjp_first_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* mu_in,              \
                                   const eb_prim_##SUF* gprim, int64_t n0, int64_t n1, int64_t n2, int bits,        \
                                   T gamma, T dx0, T dx1, T dx2, const T* dt_dev, T cfl, const int64_t* region)      \
    {                                                                                                                \
        return abi_godunov_vjp<T, eb_prim_##SUF, eb_cons_##SUF>(s, p, mu_in, nullptr, gprim, n0, n1, n2, bits,      \
                                                                gamma, dx0, dx1, dx2, dt_dev, cfl, region);           \
    }                                                                                                                \
    int eb_cons_to_prim_vjp_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_prim_##SUF* gp,                 \
                                  const eb_cons_##SUF* gc, int64_t n0, int64_t n1, int64_t n2, int bits, T gamma)  \
    {                                                                                                                \
        return abi_c2p_vjp<T>(s, p, gp, gc, n0, n1, n2, bits, gamma);                                               \
    }                                                                                                                \
    int eb_prim_to_cons_vjp_##SUF(hipStream_t s, const eb_prim_##SUF* p, const eb_cons_##SUF* gc,                  \
                                  const eb_prim_##SUF* gp, int64_t n0, int64_t n1, int64_t n2, int bits, T gamma)  \
    {                                                                                                                \
        return abi_p2c_vjp<T>(s, p, gc, gp, n0, n1, n2, bits, gamma);                                               \
    }                                                                                                                \
    int eb_hllc_flux_##SUF(hipStream_t s, int dir, const eb_prim_##SUF* qL, const eb_prim_##SUF* qR,               \
                           const eb_cons_##SUF* flux, int64_t nfaces, T gamma)                                       \
    {                                                                                                                \
        return abi_hllc<T, eb_prim_##SUF, eb_cons_##SUF>(s, dir, qL, qR, nullptr, flux, nullptr, nullptr, nfaces,   \
                                                         gamma);                                                     \
    }                                                                                                                \
    int eb_hllc_vjp_##SUF(hipStream_t s, int dir, const eb_prim_##SUF* qL, const eb_prim_##SUF* qR,                \
                          const eb_cons_##SUF* g, const eb_prim_##SUF* gL, const eb_prim_##SUF* gR, int64_t nfaces, \
                          T gamma)                                                                                   \
    {                                                                                                                \
        return abi_hllc<T, eb_prim_##SUF, eb_cons_##SUF>(s, dir, qL, qR, g, nullptr, gL, gR, nfaces, gamma);        \
    }                                                                                                                \
    int eb_conserved_totals_##SUF(eb_context* ctx, hipStream_t s, const eb_cons_##SUF* c, int64_t n0, int64_t n1,   \
                                  int64_t n2, int bits, double* totals)                                              \
    {                                                                                                                \
        return abi_conserved_totals<T>(ctx, s, c, n0, n1, n2, bits, totals);                                        \
    }                                                                                                                \
    int eb_boundary_flux_##SUF(eb_context* ctx, hipStream_t s, const eb_prim_##SUF* p, int64_t n0, int64_t n1,      \
                               int64_t n2, int bits, T gamma, T dx0, T dx1, T dx2, const T* dt_dev, T cfl,          \
                               double* acc)                                                                          \
    {                                                                                                                \
        return abi_boundary_flux<T>(ctx, s, p, n0, n1, n2, bits, gamma, dx0, dx1, dx2, dt_dev, cfl, acc);           \
    }

EB_ABI(f64, double)
EB_ABI(f32, float)

#undef EB_ABI

} // extern "C"

#include "batch_abi.hpp"
