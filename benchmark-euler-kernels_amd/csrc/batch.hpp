// Batched ensemble of independent Godunov problems: one persistent launch per step
// over a flattened queue of tiles drawn from every instance. Each instance has its
// own extents, gamma, cell widths, CFL factor, time step and simulated time; the
// tile work is the single-grid fused step (godunov + cons_to_prim + next time step,
// godunov_tile in kernels.hpp), so every instance's results are bitwise those of a
// launch of its own.
//
// Tiles are 64 x BY columns of cells (layout 0), runs of 64*BY consecutive cells
// of a plane (layout 1, godunov_tile_lin), which lose no lanes to ragged x
// extents, or 64 x BY columns with two rows per lane (layout 2,
// godunov_tile_pair).
//
// Scheduling: the tile list is split into one contiguous range per XCD (ranges of
// about equal cell count, instances kept together so halo re-reads stay in that
// XCD's L2), each range ordered largest tile first so that the tiles pulled last are
// small; block b serves queue b % nq and pulls tiles with an atomic counter. The
// last block of a queue to finish re-arms the counter, so consecutive launches (and
// graph replays) need no memset. Each instance's next dt is completed by the last
// of its tiles through the instance's own self-resetting min scratch.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace eb {

template <class T>
struct BatchInstance
{
    PrimPtrs<T> buf[2]; // ping-pong primitives: step parity p reads buf[p], writes buf[1-p]
    ConsPtrs<T> cons;
    int64_t n0;
    int64_t n1;
    int64_t n2;
    T gamma;
    T gm1;
    T dx0; // cell widths as given (host-side launches derive 1/dx from them)
    T dx1;
    T dx2;
    T ds0;
    T ds1;
    T ds2;
    T dv;
    T invdx0;
    T invdx1;
    T invdx2;
    T cfl;
    T* dt;               // [0], [1]: ping-pong raw dt; [2]: ghost-cell dt minimum
    double* time;        // FP64 simulated time
    MinScratch* scratch; // this instance's reduction words
    int kchunk;
    int tiles_x;
    int tiles_y;
    int tiles_z;
    int ntiles;
    int pad;
};

struct BatchTile
{
    int inst;  // instance index
    int local; // tile index inside the instance: bx + tiles_x*(by + tiles_y*bz)
};

struct alignas(128) BatchQueue
{
    unsigned int next; // tiles handed out (plus one overshoot per block)
    unsigned int done; // blocks that have finished
};

// Marks a block-uniform value as wave-uniform so that the descriptor loads that
// depend on it are scalar loads and the plane bases stay in scalar registers.
__device__ __forceinline__ int wave_uniform(int v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

// One step of every instance. tune_dt != nullptr times the launch on live state:
// dt is read from tune_dt (zeros), the conservatives are stored back unchanged,
// the next dt lands in tune_out and no simulated time is accumulated.
template <class T, class I, int BY, int Layout>
__global__ __launch_bounds__(Layout == 2 ? 32 * BY : 64 * BY)
__attribute__((amdgpu_waves_per_eu(Layout == 2 ? 2 : 1))) void k_batch_step(
        BatchInstance<T> const* __restrict__ insts,
        BatchTile const* __restrict__ tiles,
        int const* __restrict__ ranges,
        int nq,
        BatchQueue* queues,
        int parity,
        T const* tune_dt,
        T* tune_out)
{
    __shared__ int s_tile;
    int const q = nq > 1 ? static_cast<int>(blockIdx.x) % nq : 0;
    int const beg = ranges[q];
    int const end = ranges[q + 1];
    int const lid = threadIdx.x + 64 * threadIdx.y;
    unsigned int const blocks_in_q
            = gridDim.x / nq + ((static_cast<int>(gridDim.x % nq) > q) ? 1u : 0u);
    bool const tune = tune_dt != nullptr;

    // Invariant: no single-lane region may span the loop-head barrier (a region
    // merged across the back edge would put that barrier on two divergent paths of
    // one wave, i.e. make the wave reach it twice, and hang the block). The time-step
    // commit at the end of an iteration is therefore closed by a barrier of its own,
    // and, as a second guard, the tile is fetched by another lane (the first lane of
    // wave 1) than the one that commits (the first lane of wave 0).
    int const fetch_lane = 64;
    for (;;) {
        if (lid == fetch_lane) {
            s_tile = beg + static_cast<int>(atomicAdd(&queues[q].next, 1u));
        }
        __syncthreads();
        int const t = wave_uniform(s_tile);
        __syncthreads();
        if (t >= end) {
            break;
        }
        BatchTile const tl = tiles[t];
        BatchInstance<T> const& in = insts[tl.inst];
        PrimPtrs<T> const src = in.buf[parity];
        PrimPtrs<T const> const pin {src.d, src.p, src.ux0, src.ux1, src.ux2};
        T const* dt_in = tune ? tune_dt + tl.inst : in.dt + parity;
        // dt/dv exactly as for a single grid (euler_simulation.cpp:65, godunov.hpp:32)
        T const dt = in.cfl * dt_in[0];
        T const dtodv = dt / in.dv;
        T lm;
        if constexpr (Layout == 2) {
            // paired rows: 64 x BY tiles, BY/2 waves of two rows each
            int const bx = tl.local % in.tiles_x;
            int const by = (tl.local / in.tiles_x) % in.tiles_y;
            int const bz = tl.local / (in.tiles_x * in.tiles_y);
            lm = godunov_tile_pair<T, I, BY, true>(pin, in.cons, in.buf[1 - parity], static_cast<I>(in.n0),
                                                   static_cast<I>(in.n1), static_cast<I>(in.n2), in.kchunk, bx, by,
                                                   bz, in.gamma, in.gm1, in.ds0, in.ds1, in.ds2, dtodv, in.invdx0,
                                                   in.invdx1, in.invdx2, !tune);
        } else if constexpr (Layout == 1) {
            // linear tiles: tiles_x tiles per plane, tiles_y = 1
            lm = godunov_tile_lin<T, I, BY, true>(pin, in.cons, in.buf[1 - parity], static_cast<I>(in.n0),
                                                  static_cast<I>(in.n1), static_cast<I>(in.n2), in.kchunk,
                                                  tl.local % in.tiles_x, tl.local / in.tiles_x, in.gamma, in.gm1,
                                                  in.ds0, in.ds1, in.ds2, dtodv, in.invdx0, in.invdx1, in.invdx2,
                                                  !tune);
        } else {
            int const bx = tl.local % in.tiles_x;
            int const by = (tl.local / in.tiles_x) % in.tiles_y;
            int const bz = tl.local / (in.tiles_x * in.tiles_y);
            lm = godunov_tile<T, I, BY, true>(pin, in.cons, in.buf[1 - parity], static_cast<I>(in.n0),
                                              static_cast<I>(in.n1), static_cast<I>(in.n2), in.kchunk, bx, by, bz,
                                              in.gamma, in.gm1, in.ds0, in.ds1, in.ds2, dtodv, in.invdx0, in.invdx1,
                                              in.invdx2, !tune);
        }
        T const bm = block_min<T, BY>(lm);
        if (lid == 0) {
            DtSink<T> const sink {in.scratch, tune ? tune_out + tl.inst : in.dt + (1 - parity), in.dt + 2,
                                  tune ? nullptr : in.time, dt_in, in.cfl};
            // an instance's tiles come off the queue spread out in time
            grid_min_commit(sink, bm, static_cast<unsigned int>(tl.local), static_cast<unsigned int>(in.ntiles), true);
        }
        __syncthreads(); // closes the commit region inside this iteration
    }
    if (lid == 0) {
        unsigned int const d = atomicAdd(&queues[q].done, 1u);
        if (d == blocks_in_q - 1) {
            atomicExch(&queues[q].next, 0u);
            atomicExch(&queues[q].done, 0u);
        }
    }
}

} // namespace eb
