// Host side of the batched ensemble (C ABI in euler_capi.h). Included by capi.hip
// after its helpers (fail, check_extents, run_time_step, run_ghost_refresh,
// launch_tiled, Mesh), inside the same translation unit.
#pragma once

#include <algorithm>
#include <array>
#include <type_traits>
#include <vector>

#include "batch.hpp"

struct eb_batch
{
    int device = 0;
    int real_bits = 64; // 64: double, 32: float
    int index_bits = 64;
    int ninst = 0;
    int layout = 0; // 0: 64 x by tiles, 1: linear runs of 64*by cells, 2: 64 x by tiles, two rows per lane
    int by = 8;
    int kchunk = 32;
    int ntiles = 0;
    int nq = 1;
    int grid = 0;
    std::vector<unsigned char> host_insts; // BatchInstance<T> x ninst
    void* insts = nullptr;                 // device copy
    eb::BatchTile* tiles = nullptr;
    int* ranges = nullptr;
    eb::BatchQueue* queues = nullptr;
    void* dt = nullptr;       // 3 x ninst reals
    double* time = nullptr;   // ninst
    eb::MinScratch* scratch = nullptr;
    void* tune_dt = nullptr;  // ninst zeros
    void* tune_out = nullptr; // ninst
    int tiles_cap = 0;
};

namespace {

template <class T>
using prim_c = std::conditional_t<std::is_same_v<T, double>, eb_prim_f64, eb_prim_f32>;
template <class T>
using cons_c = std::conditional_t<std::is_same_v<T, double>, eb_cons_f64, eb_cons_f32>;
template <class T>
using inst_c = std::conditional_t<std::is_same_v<T, double>, eb_batch_instance_f64, eb_batch_instance_f32>;

constexpr int kBatchQueues = 8; // one per XCD

// Linear tiles of 1024 cells stage 147 KB of LDS in FP32; FP64 would not fit.
bool batch_by_supported(int layout, int by, int real_bits)
{
    return layout == 0   ? (by == 4 || by == 8 || by == 16)
           : layout == 1 ? (by == 4 || by == 8 || (by == 16 && real_bits == 32))
                         : (layout == 2 && (by == 8 || by == 16));
}

template <class T>
eb::BatchInstance<T>* batch_host(eb_batch* b)
{
    return reinterpret_cast<eb::BatchInstance<T>*>(b->host_insts.data());
}

void batch_free(eb_batch* b)
{
    for (void* p : {b->insts, static_cast<void*>(b->tiles), static_cast<void*>(b->ranges),
                    static_cast<void*>(b->queues), b->dt, static_cast<void*>(b->time),
                    static_cast<void*>(b->scratch), b->tune_dt, b->tune_out}) {
        if (p != nullptr) {
            (void)hipFree(p);
        }
    }
}

template <class T, class I, int BY, int L>
int batch_resident_blocks(int* out)
{
    int dev = 0;
    int cus = 0;
    int per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess
        || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess
        || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, eb::k_batch_step<T, I, BY, L>,
                                                        L == 2 ? 32 * BY : 64 * BY, 0)
                   != hipSuccess) {
        return fail(EB_ERR_HIP, "occupancy query failed");
    }
    *out = std::max(1, per_cu) * std::max(1, cus);
    return EB_OK;
}

template <class T, class I>
int batch_resident(int layout, int by, int* out)
{
    if (layout == 1) {
        if constexpr (sizeof(T) == 4) {
            if (by == 16) {
                return batch_resident_blocks<T, I, 16, 1>(out);
            }
        }
        return by == 4 ? batch_resident_blocks<T, I, 4, 1>(out) : batch_resident_blocks<T, I, 8, 1>(out);
    }
    if (layout == 2) {
        return by == 8 ? batch_resident_blocks<T, I, 8, 2>(out) : batch_resident_blocks<T, I, 16, 2>(out);
    }
    if (by == 4) {
        return batch_resident_blocks<T, I, 4, 0>(out);
    }
    if (by == 8) {
        return batch_resident_blocks<T, I, 8, 0>(out);
    }
    return batch_resident_blocks<T, I, 16, 0>(out);
}

// Tiles of every instance for tile height `by` and `kchunk` planes per tile (0: a
// whole column), split into queues of about equal cell count; uploads descriptors,
// tiles and ranges (blocking: waits for all work on the device first;
// configuration time only).
template <class T>
int batch_layout(eb_batch* b, int layout, int by, int kchunk)
{
    eb::BatchInstance<T>* hi = batch_host<T>(b);
    std::vector<eb::BatchTile> tiles;
    std::vector<double> cost;
    for (int i = 0; i < b->ninst; ++i) {
        eb::BatchInstance<T>& in = hi[i];
        int64_t const nx = in.n0 - 2;
        int64_t const ny = in.n1 - 2;
        int64_t const nz = in.n2 - 2;
        int64_t const kc = kchunk > 0 ? std::min<int64_t>(kchunk, nz) : nz;
        in.kchunk = static_cast<int>(kc);
        if (layout == 1) {
            if (in.n0 > 64 * by) {
                return fail(EB_ERR_UNSUPPORTED, "linear tiles need rows of at most 64*by cells");
            }
            in.tiles_x = static_cast<int>((in.n0 * ny + 64 * by - 1) / (64 * by));
            in.tiles_y = 1;
        } else {
            in.tiles_x = static_cast<int>((nx + 63) / 64);
            in.tiles_y = static_cast<int>((ny + by - 1) / by);
        }
        in.tiles_z = static_cast<int>((nz + kc - 1) / kc);
        int64_t const nt = int64_t(in.tiles_x) * in.tiles_y * in.tiles_z;
        if (nt + static_cast<int64_t>(tiles.size()) > std::numeric_limits<int32_t>::max() / 2) {
            return fail(EB_ERR_INDEX_OVERFLOW, "too many tiles in the ensemble");
        }
        in.ntiles = static_cast<int>(nt);
        for (int t = 0; t < in.ntiles; ++t) {
            int const bz = t / (in.tiles_x * in.tiles_y);
            double const planes = static_cast<double>(std::min<int64_t>(kc, nz - kc * bz));
            double w;
            if (layout == 1) {
                int const tp = t % in.tiles_x;
                w = static_cast<double>(std::min<int64_t>(64 * by, in.n0 * ny - int64_t(64 * by) * tp)) * planes;
            } else {
                int const bx = t % in.tiles_x;
                int const byy = (t / in.tiles_x) % in.tiles_y;
                w = static_cast<double>(std::min<int64_t>(64, nx - 64 * int64_t(bx)))
                    * static_cast<double>(std::min<int64_t>(by, ny - int64_t(by) * byy)) * planes;
            }
            tiles.push_back({i, t});
            cost.push_back(w);
        }
    }
    int const ntiles = static_cast<int>(tiles.size());
    int resident = 0;
    int const st = b->index_bits == 32 ? batch_resident<T, int32_t>(layout, by, &resident)
                                       : batch_resident<T, int64_t>(layout, by, &resident);
    if (st != EB_OK) {
        return st;
    }
    int grid = std::min(ntiles, resident);
    int nq = 1;
    if (grid >= 8 * kBatchQueues) {
        nq = kBatchQueues;
        grid -= grid % nq;
    }
    std::vector<int> ranges(nq + 1, ntiles);
    ranges[0] = 0;
    double total = 0;
    for (double w : cost) {
        total += w;
    }
    double acc = 0;
    int q = 1;
    for (int t = 0; t < ntiles && q < nq; ++t) {
        acc += cost[t];
        while (q < nq && acc >= total * q / nq) {
            ranges[q++] = t + 1;
        }
    }
    // Largest tiles first within each queue (tiles pulled late are the small ones,
    // which shortens the tail of a step); equal costs keep their instance order.
    {
        std::vector<int> order(ntiles);
        for (int t = 0; t < ntiles; ++t) {
            order[t] = t;
        }
        for (int qq = 0; qq < nq; ++qq) {
            std::stable_sort(order.begin() + ranges[qq], order.begin() + ranges[qq + 1],
                             [&](int x, int y) { return cost[x] > cost[y]; });
        }
        std::vector<eb::BatchTile> sorted(ntiles);
        for (int t = 0; t < ntiles; ++t) {
            sorted[t] = tiles[order[t]];
        }
        tiles.swap(sorted);
    }
    // Launches already enqueued, on any stream, read the descriptors replaced (and
    // possibly reallocated) below: they finish first.
    if (hipDeviceSynchronize() != hipSuccess) {
        return fail(EB_ERR_HIP, "device synchronisation before the layout upload failed");
    }
    if (ntiles > b->tiles_cap) {
        if (b->tiles != nullptr) {
            (void)hipFree(b->tiles);
            b->tiles = nullptr;
        }
        if (hipMalloc(reinterpret_cast<void**>(&b->tiles), sizeof(eb::BatchTile) * std::max(ntiles, 1))
            != hipSuccess) {
            b->tiles_cap = 0;
            return fail(EB_ERR_OUT_OF_MEMORY, "tile list allocation failed");
        }
        b->tiles_cap = ntiles;
    }
    if (hipMemcpy(b->tiles, tiles.data(), sizeof(eb::BatchTile) * ntiles, hipMemcpyHostToDevice) != hipSuccess
        || hipMemcpy(b->ranges, ranges.data(), sizeof(int) * (nq + 1), hipMemcpyHostToDevice) != hipSuccess
        || hipMemcpy(b->insts, hi, sizeof(eb::BatchInstance<T>) * b->ninst, hipMemcpyHostToDevice)
                   != hipSuccess) {
        return fail(EB_ERR_HIP, "batch layout upload failed");
    }
    b->layout = layout;
    b->by = by;
    b->kchunk = kchunk;
    b->ntiles = ntiles;
    b->nq = nq;
    b->grid = grid;
    return EB_OK;
}

template <class T>
int batch_create(eb_batch** out, inst_c<T> const* inst, int n, int index_bits)
{
    if (out == nullptr || (n > 0 && inst == nullptr) || n < 1) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null output, null instances or empty ensemble");
    }
    *out = nullptr;
    if (index_bits != 32 && index_bits != 64) {
        return fail(EB_ERR_INVALID_ARGUMENT, "index_bits must be 32 or 64");
    }
    auto* b = new eb_batch;
    b->real_bits = std::is_same_v<T, double> ? 64 : 32;
    b->index_bits = index_bits;
    b->ninst = n;
    b->host_insts.resize(sizeof(eb::BatchInstance<T>) * n);
    eb::BatchInstance<T>* hi = batch_host<T>(b);
    for (int i = 0; i < n; ++i) {
        inst_c<T> const& s = inst[i];
        int const st = check_extents(s.n0, s.n1, s.n2, index_bits);
        if (st != EB_OK) {
            delete b;
            return st;
        }
        if (s.n0 < 3 || s.n1 < 3 || s.n2 < 3) {
            delete b;
            return fail(EB_ERR_INVALID_ARGUMENT, "every instance needs at least one interior cell");
        }
        if (static_cast<__int128>(s.n0) * s.n1 * static_cast<int64_t>(sizeof(T)) >= (static_cast<__int128>(1) << 32)) {
            delete b;
            return fail(EB_ERR_UNSUPPORTED, "a k-plane must hold fewer than 2^32 bytes per field");
        }
        if (prim_null(&s.prim_a) || prim_null(&s.prim_b) || cons_null(&s.cons) || s.prim_a.d == s.prim_b.d) {
            delete b;
            return fail(EB_ERR_INVALID_ARGUMENT, "null or aliased instance arrays");
        }
        eb::BatchInstance<T>& d = hi[i];
        d = eb::BatchInstance<T> {};
        d.buf[0] = prim_ptrs<T>(&s.prim_a);
        d.buf[1] = prim_ptrs<T>(&s.prim_b);
        d.cons = cons_ptrs<T>(&s.cons);
        d.n0 = s.n0;
        d.n1 = s.n1;
        d.n2 = s.n2;
        d.gamma = s.gamma;
        d.gm1 = static_cast<T>(s.gamma - 1);
        d.dx0 = s.dx0;
        d.dx1 = s.dx1;
        d.dx2 = s.dx2;
        Mesh::faces(s.dx0, s.dx1, s.dx2, d.ds0, d.ds1, d.ds2, d.dv);
        d.invdx0 = T(1) / s.dx0;
        d.invdx1 = T(1) / s.dx1;
        d.invdx2 = T(1) / s.dx2;
        d.cfl = s.cfl;
    }
    bool ok = hipGetDevice(&b->device) == hipSuccess;
    ok = ok && hipMalloc(&b->insts, sizeof(eb::BatchInstance<T>) * n) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&b->ranges), sizeof(int) * (kBatchQueues + 1)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&b->queues), sizeof(eb::BatchQueue) * kBatchQueues) == hipSuccess;
    ok = ok && hipMalloc(&b->dt, sizeof(T) * 3 * n) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&b->time), sizeof(double) * n) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&b->scratch), sizeof(eb::MinScratch) * n) == hipSuccess;
    ok = ok && hipMalloc(&b->tune_dt, sizeof(T) * n) == hipSuccess;
    ok = ok && hipMalloc(&b->tune_out, sizeof(T) * n) == hipSuccess;
    if (!ok) {
        batch_free(b);
        delete b;
        return fail(EB_ERR_OUT_OF_MEMORY, "batch allocation failed");
    }
    ok = hipMemset(b->queues, 0, sizeof(eb::BatchQueue) * kBatchQueues) == hipSuccess
         && hipMemset(b->dt, 0, sizeof(T) * 3 * n) == hipSuccess
         && hipMemset(b->time, 0, sizeof(double) * n) == hipSuccess
         && hipMemset(b->tune_dt, 0, sizeof(T) * n) == hipSuccess;
    hipLaunchKernelGGL(eb::min_scratch_init, dim3((n * (eb::kMinSlots + 1) + 255) / 256), dim3(256), 0, 0, b->scratch, n);
    ok = ok && hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        batch_free(b);
        delete b;
        return fail(EB_ERR_HIP, "batch initialisation failed");
    }
    T* dt = static_cast<T*>(b->dt);
    for (int i = 0; i < n; ++i) {
        hi[i].dt = dt + 3 * i;
        hi[i].time = b->time + i;
        hi[i].scratch = b->scratch + i;
    }
    int const st = batch_layout<T>(b, 0, 8, 32);
    if (st != EB_OK) {
        batch_free(b);
        delete b;
        return st;
    }
    *out = b;
    return EB_OK;
}

inline eb_context instance_context(eb_batch* b, int i)
{
    return eb_context {b->device, b->scratch + i, nullptr, nullptr};
}

template <class T, class I>
int batch_start(eb_batch* b, hipStream_t s)
{
    eb::BatchInstance<T>* hi = batch_host<T>(b);
    for (int i = 0; i < b->ninst; ++i) {
        eb::BatchInstance<T> const& in = hi[i];
        eb_context ctx = instance_context(b, i);
        prim_c<T> const pa {in.buf[0].d, in.buf[0].p, in.buf[0].ux0, in.buf[0].ux1, in.buf[0].ux2};
        prim_c<T> const pb {in.buf[1].d, in.buf[1].p, in.buf[1].ux0, in.buf[1].ux1, in.buf[1].ux2};
        cons_c<T> const c {in.cons.d, in.cons.e, in.cons.mx0, in.cons.mx1, in.cons.mx2};
        int st = run_time_step<T, I>(ctx.scratch, s, &pa, in.n0, in.n1, in.n2, in.gamma, in.dx0, in.dx1, in.dx2, in.dt);
        if (st == EB_OK) {
            st = run_ghost_refresh<T, I>(&ctx, s, &c, &pb, in.n0, in.n1, in.n2, in.gamma, in.dx0, in.dx1, in.dx2,
                                         in.dt + 2);
        }
        if (st != EB_OK) {
            return st;
        }
    }
    return EB_OK;
}

template <class T, class I>
int batch_refresh(eb_batch* b, hipStream_t s, int buffer)
{
    eb::BatchInstance<T>* hi = batch_host<T>(b);
    for (int i = 0; i < b->ninst; ++i) {
        eb::BatchInstance<T> const& in = hi[i];
        eb_context ctx = instance_context(b, i);
        eb::PrimPtrs<T> const& d = in.buf[buffer];
        prim_c<T> const p {d.d, d.p, d.ux0, d.ux1, d.ux2};
        cons_c<T> const c {in.cons.d, in.cons.e, in.cons.mx0, in.cons.mx1, in.cons.mx2};
        int const st = run_ghost_refresh<T, I>(&ctx, s, &c, &p, in.n0, in.n1, in.n2, in.gamma, in.dx0, in.dx1,
                                               in.dx2, in.dt + 2);
        if (st != EB_OK) {
            return st;
        }
    }
    return EB_OK;
}

template <class T, class I>
int batch_launch(eb_batch* b, hipStream_t s, int parity, bool tune)
{
    auto const* insts = static_cast<eb::BatchInstance<T> const*>(b->insts);
    T const* tdt = tune ? static_cast<T const*>(b->tune_dt) : nullptr;
    T* tout = tune ? static_cast<T*>(b->tune_out) : nullptr;
    dim3 const grid(static_cast<unsigned int>(std::max(b->grid, 1)));
#define EB_BATCH_LAUNCH(BY, L)                                                                                    \
    hipLaunchKernelGGL((eb::k_batch_step<T, I, BY, L>), grid, dim3(64, BY), 0, s, insts, b->tiles, b->ranges, \
                       b->nq, b->queues, parity, tdt, tout)
#define EB_BATCH_LAUNCH_PAIR(BY)                                                                                  \
    hipLaunchKernelGGL((eb::k_batch_step<T, I, BY, 2>), grid, dim3(64, BY / 2), 0, s, insts, b->tiles, b->ranges, \
                       b->nq, b->queues, parity, tdt, tout)
    if (b->layout == 2) {
        if (b->by == 8) {
            EB_BATCH_LAUNCH_PAIR(8);
        } else {
            EB_BATCH_LAUNCH_PAIR(16);
        }
    } else if (b->layout == 1) {
        if constexpr (sizeof(T) == 4) {
            if (b->by == 16) {
                EB_BATCH_LAUNCH(16, 1);
                return check_launch("batch_step");
            }
        }
        if (b->by == 4) {
            EB_BATCH_LAUNCH(4, 1);
        } else {
            EB_BATCH_LAUNCH(8, 1);
        }
    } else if (b->by == 4) {
        EB_BATCH_LAUNCH(4, 0);
    } else if (b->by == 8) {
        EB_BATCH_LAUNCH(8, 0);
    } else {
        EB_BATCH_LAUNCH(16, 0);
    }
#undef EB_BATCH_LAUNCH
#undef EB_BATCH_LAUNCH_PAIR
    return check_launch("batch_step");
}

// The same step as one launch per instance: exactly eb_step_fused_* on each
// instance with the process-wide single-grid configuration, for comparison with
// the single persistent launch.
template <class T, class I>
int batch_per_instance(eb_batch* b, hipStream_t s, int parity)
{
    eb::BatchInstance<T> const* hi = batch_host<T>(b);
    for (int i = 0; i < b->ninst; ++i) {
        eb::BatchInstance<T> const& in = hi[i];
        eb::PrimPtrs<T> const& src = in.buf[parity];
        eb::PrimPtrs<T const> const pin {src.d, src.p, src.ux0, src.ux1, src.ux2};
        eb::DtArg<T> const dta {T(0), in.dt + parity, in.cfl, in.dv};
        eb::DtSink<T> const sink {in.scratch, in.dt + (1 - parity), in.dt + 2, in.time, in.dt + parity, in.cfl};
        int const st = dispatch_tiled<T, I, true>(s, pin, in.cons, in.buf[1 - parity], in.n0, in.n1, in.n2,
                                                  in.gamma, in.ds0, in.ds1, in.ds2, dta, in.invdx0, in.invdx1,
                                                  in.invdx2, sink);
        if (st != EB_OK) {
            return st;
        }
    }
    return EB_OK;
}

// Times every (tile height, planes per tile) candidate on the live ensemble with a
// zero time step (values unchanged, same work), keeps the fastest layout.
template <class T, class I>
int batch_autotune(eb_batch* b, hipStream_t s, int parity, int reps, float* best_ms)
{
    static constexpr int kShapes[][2] = {{0, 4}, {0, 8}, {0, 16}, {1, 4}, {1, 8}, {1, 16}, {2, 8}, {2, 16}}; // (layout, by)
    static constexpr int kKc[] = {0, 64, 32, 16, 8};
    int64_t max_n0 = 0;
    for (int i = 0; i < b->ninst; ++i) {
        max_n0 = std::max<int64_t>(max_n0, batch_host<T>(b)[i].n0);
    }
    hipEvent_t e0 = nullptr;
    hipEvent_t e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) {
        return fail(EB_ERR_HIP, "event creation failed");
    }
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(EB_ERR_HIP, "event creation failed");
    }
    int best_layout = b->layout;
    int best_by = b->by;
    int best_kc = b->kchunk;
    float best = -1.0f;
    int st = EB_OK;
    for (auto const& shape : kShapes) {
        int const layout = shape[0];
        int const by = shape[1];
        if ((layout == 1 && max_n0 > 64 * by) || !batch_by_supported(layout, by, 8 * int(sizeof(T)))) {
            continue;
        }
        for (int kc : kKc) {
            st = batch_layout<T>(b, layout, by, kc);
            if (st != EB_OK) {
                break;
            }
            st = batch_launch<T, I>(b, s, parity, true); // warm-up
            if (st != EB_OK) {
                break;
            }
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps && st == EB_OK; ++r) {
                st = batch_launch<T, I>(b, s, parity, true);
            }
            (void)hipEventRecord(e1, s);
            if (st != EB_OK || hipEventSynchronize(e1) != hipSuccess) {
                st = st != EB_OK ? st : fail(EB_ERR_HIP, "autotune timing failed");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= static_cast<float>(reps);
            if (best < 0.0f || ms < best) {
                best = ms;
                best_layout = layout;
                best_by = by;
                best_kc = kc;
            }
        }
        if (st != EB_OK) {
            break;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int const st2 = batch_layout<T>(b, best_layout, best_by, best_kc);
    if (st != EB_OK) {
        return st;
    }
    if (best_ms != nullptr) {
        *best_ms = best;
    }
    return st2;
}

template <class T>
int batch_dispatch(eb_batch* b, int op, hipStream_t s, int arg, int reps, float* out_ms)
{
    bool const i32 = b->index_bits == 32;
    switch (op) {
    case 0:
        return i32 ? batch_start<T, int32_t>(b, s) : batch_start<T, int64_t>(b, s);
    case 1:
        return i32 ? batch_refresh<T, int32_t>(b, s, arg) : batch_refresh<T, int64_t>(b, s, arg);
    case 2:
        return i32 ? batch_launch<T, int32_t>(b, s, arg, false) : batch_launch<T, int64_t>(b, s, arg, false);
    case 3:
        return i32 ? batch_per_instance<T, int32_t>(b, s, arg) : batch_per_instance<T, int64_t>(b, s, arg);
    default:
        return i32 ? batch_autotune<T, int32_t>(b, s, arg, reps, out_ms)
                   : batch_autotune<T, int64_t>(b, s, arg, reps, out_ms);
    }
}

int batch_op(eb_batch* b, int op, hipStream_t s, int arg, int reps, float* out_ms)
{
    if (b == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null batch");
    }
    if ((op == 1 || op == 2 || op == 3 || op == 4) && arg != 0 && arg != 1) {
        return fail(EB_ERR_INVALID_ARGUMENT, "buffer / parity must be 0 or 1");
    }
    if (op == 4 && reps < 1) {
        return fail(EB_ERR_INVALID_ARGUMENT, "reps must be >= 1");
    }
    return b->real_bits == 64 ? batch_dispatch<double>(b, op, s, arg, reps, out_ms)
                              : batch_dispatch<float>(b, op, s, arg, reps, out_ms);
}

} // namespace

extern "C" {

int eb_batch_create_f64(eb_batch** batch, const eb_batch_instance_f64* instances, int ninstances, int index_bits)
{
    return batch_create<double>(batch, instances, ninstances, index_bits);
}

int eb_batch_create_f32(eb_batch** batch, const eb_batch_instance_f32* instances, int ninstances, int index_bits)
{
    return batch_create<float>(batch, instances, ninstances, index_bits);
}

int eb_batch_destroy(eb_batch* batch)
{
    if (batch == nullptr) {
        return EB_OK;
    }
    batch_free(batch);
    delete batch;
    return EB_OK;
}

int eb_batch_set_tiles(eb_batch* batch, int layout, int by, int kchunk)
{
    if (batch == nullptr || kchunk < 0) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null batch or negative kchunk");
    }
    if (!batch_by_supported(layout, by, batch->real_bits)) {
        return fail(EB_ERR_UNSUPPORTED, "tiles: layout 0 with by 4, 8 or 16; layout 1 with by 4 or 8 (or 16 in "
                                        "FP32); layout 2 with by 8 or 16");
    }
    int const keep_layout = batch->layout;
    int const keep_by = batch->by;
    int const keep_kc = batch->kchunk;
    int const st = batch->real_bits == 64 ? batch_layout<double>(batch, layout, by, kchunk)
                                          : batch_layout<float>(batch, layout, by, kchunk);
    if (st != EB_OK) {
        // leave the previous, valid layout in place
        (void)(batch->real_bits == 64 ? batch_layout<double>(batch, keep_layout, keep_by, keep_kc)
                                      : batch_layout<float>(batch, keep_layout, keep_by, keep_kc));
    }
    return st;
}

int eb_batch_get_tiles(const eb_batch* batch, int* layout, int* by, int* kchunk, int* ntiles, int* grid)
{
    if (batch == nullptr || layout == nullptr || by == nullptr || kchunk == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    *layout = batch->layout;
    *by = batch->by;
    *kchunk = batch->kchunk;
    if (ntiles != nullptr) {
        *ntiles = batch->ntiles;
    }
    if (grid != nullptr) {
        *grid = batch->grid;
    }
    return EB_OK;
}

int eb_batch_device_state(const eb_batch* batch, void** dt, double** time)
{
    if (batch == nullptr || dt == nullptr || time == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null argument");
    }
    *dt = batch->dt;
    *time = batch->time;
    return EB_OK;
}

int eb_batch_copy_state(const eb_batch* batch, hipStream_t stream, void* dt_dst, double* time_dst)
{
    if (batch == nullptr) {
        return fail(EB_ERR_INVALID_ARGUMENT, "null batch");
    }
    std::size_t const rs = batch->real_bits == 64 ? sizeof(double) : sizeof(float);
    if (dt_dst != nullptr
        && hipMemcpyAsync(dt_dst, batch->dt, rs * 3 * batch->ninst, hipMemcpyDefault, stream) != hipSuccess) {
        return fail(EB_ERR_HIP, "dt copy failed");
    }
    if (time_dst != nullptr
        && hipMemcpyAsync(time_dst, batch->time, sizeof(double) * batch->ninst, hipMemcpyDefault, stream)
                   != hipSuccess) {
        return fail(EB_ERR_HIP, "time copy failed");
    }
    return EB_OK;
}

int eb_batch_start(eb_batch* batch, hipStream_t stream)
{
    return batch_op(batch, 0, stream, 0, 1, nullptr);
}

int eb_batch_refresh_ghosts(eb_batch* batch, hipStream_t stream, int buffer)
{
    return batch_op(batch, 1, stream, buffer, 1, nullptr);
}

int eb_batch_step(eb_batch* batch, hipStream_t stream, int parity)
{
    return batch_op(batch, 2, stream, parity, 1, nullptr);
}

int eb_batch_step_per_instance(eb_batch* batch, hipStream_t stream, int parity)
{
    return batch_op(batch, 3, stream, parity, 1, nullptr);
}

int eb_batch_autotune(eb_batch* batch, hipStream_t stream, int parity, int reps, float* best_ms)
{
    return batch_op(batch, 4, stream, parity, reps, best_ms);
}

} // extern "C"
