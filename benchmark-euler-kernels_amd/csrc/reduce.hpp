// Min-reduction machinery for the CFL time step: wavefront shuffles, an LDS
// combine across the waves of a block, and a single-pass grid completion in which
// the last block to arrive publishes the result to device memory and re-arms the
// scratch words for the next launch. No host synchronisation is involved and no
// memset is needed between launches, so the sequence is safe inside a hipGraph.
//
// Semantics follow the reference's Kokkos::Min<T> reduction (time_step.hpp:37-39):
// identity = largest finite value, candidate kept only when it compares strictly
// smaller, so NaN candidates never enter. The minimum of non-NaN values is exact,
// hence independent of the combine order.
#pragma once

#include <cfloat>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace eb {

// One accumulator of the grid reduction, alone on a 128-byte L2 line.
struct alignas(128) MinSlot
{
    unsigned long long acc; // ordered encoding of the running minimum
    unsigned int count;     // contributions so far
    unsigned int pad;
};

// Blocks of one launch are spread over this many slots, so that the device-scope
// atomics of a launch's blocks queue on 32 lines rather than one.
constexpr unsigned int kMinSlots = 32;

// Scratch words owned by an eb_context (one per concurrent reduction stream):
// the slots and the top accumulator that the slots feed.
struct MinScratch
{
    MinSlot top;
    MinSlot slot[kMinSlots];
};

template <class T>
struct Limits;

template <>
struct Limits<double>
{
    static constexpr double max() { return DBL_MAX; }
};

template <>
struct Limits<float>
{
    static constexpr float max() { return FLT_MAX; }
};

// Order-preserving map from IEEE bits to unsigned integers (non-NaN inputs).
__device__ __forceinline__ unsigned long long ord_encode(double x)
{
    unsigned long long const b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ unsigned long long ord_encode(float x)
{
    unsigned int const b = __float_as_uint(x);
    return static_cast<unsigned long long>((b >> 31) ? ~b : (b | 0x80000000u));
}

template <class T>
__device__ __forceinline__ T ord_decode(unsigned long long e);

template <>
__device__ __forceinline__ double ord_decode<double>(unsigned long long e)
{
    unsigned long long const b = (e >> 63) ? (e & 0x7fffffffffffffffull) : ~e;
    return __longlong_as_double(static_cast<long long>(b));
}

template <>
__device__ __forceinline__ float ord_decode<float>(unsigned long long e)
{
    unsigned int const u = static_cast<unsigned int>(e);
    unsigned int const b = (u >> 31) ? (u & 0x7fffffffu) : ~u;
    return __uint_as_float(b);
}

// Kokkos::Min join: keep the incumbent unless the candidate is strictly smaller.
template <class T>
__device__ __forceinline__ T min_join(T incumbent, T candidate)
{
    return (candidate < incumbent) ? candidate : incumbent;
}

// Time-step candidates of one thread's cells, folded. A cell's candidate is 1/x,
// x = sum_d (c + |u_d|) / dx_d, and the step is the Kokkos::Min of the candidates
// from max(), a NaN candidate never winning (time_step.hpp:28-38). With every
// 1/dx_d >= 0, x is +0, positive, +inf or NaN; a correctly rounded 1/x never
// increases as x grows, so the smallest candidate is 1/(largest x), and one
// division per thread gives the bits of one per cell (x = +0, candidate +inf, and
// NaN never win either way). With a negative or NaN 1/dx_d (`fold` false, uniform
// over the launch) each candidate is divided out as before.
template <class T>
struct DtFold
{
    bool fold;
    T xmax = T(0);
    T dtmin = Limits<T>::max();

    __device__ __forceinline__ explicit DtFold(bool f) : fold(f) {}

    __device__ __forceinline__ void add(T x)
    {
        if (fold) {
            xmax = (xmax < x) ? x : xmax;
        } else {
            dtmin = min_join(dtmin, T(1) / x);
        }
    }

    __device__ __forceinline__ T result() const { return fold ? min_join(dtmin, T(1) / xmax) : dtmin; }
};

template <class T>
__device__ __forceinline__ bool dt_foldable(T invdx0, T invdx1, T invdx2)
{
    return invdx0 >= T(0) && invdx1 >= T(0) && invdx2 >= T(0);
}

template <class T>
__device__ __forceinline__ T wave_min(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v = min_join(v, __shfl_xor(v, off, 64));
    }
    return v;
}

// Block-wide minimum; every thread must call it. Result valid in thread 0.
template <class T, int MaxWaves>
__device__ __forceinline__ T block_min(T v)
{
    __shared__ T partial[MaxWaves];
    int const lid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    int const nthreads = blockDim.x * blockDim.y * blockDim.z;
    int const wave = lid >> 6;
    int const lane = lid & 63;
    v = wave_min(v);
    if (lane == 0) {
        partial[wave] = v;
    }
    __syncthreads();
    T r = Limits<T>::max();
    if (lid == 0) {
        int const nw = (nthreads + 63) >> 6;
        for (int w = 0; w < nw; ++w) {
            r = min_join(r, partial[w]);
        }
    }
    return r;
}

// Arguments for the grid completion performed by the last block.
template <class T>
struct DtSink
{
    MinScratch* scratch;     // self-resetting accumulator words
    T* dt_out;               // receives the raw minimum (no CFL factor)
    T const* extra_min;      // optional constant candidate merged at the end (nullable)
    double* time_acc;        // optional FP64 simulated-time accumulator (nullable)
    T const* dt_in;          // dt of the step just taken (for time_acc), raw
    T cfl;                   // time_acc += double(cfl * dt_in[0])
};

// Adds `enc` to one accumulator of `expected` contributions; true in the caller
// that contributes last, which then holds every contribution's minimum in `enc`
// and re-arms the accumulator for the next launch. Release before the count,
// acquire only in the last contributor: each contributor's fence synchronises
// with the last one's through the chain of read-modify-writes on `count`. A
// device-scope acquire invalidates the XCD's L2, so the others do not pay it.
__device__ __forceinline__ bool slot_commit(MinSlot* s, unsigned long long& enc, unsigned int expected)
{
    atomicMin(&s->acc, enc);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (atomicAdd(&s->count, 1u) != expected - 1) {
        return false;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    // All-ones sits above every encoding of either precision, so the words can be
    // shared by FP32 and FP64 launches.
    enc = atomicExch(&s->acc, 0xffffffffffffffffull);
    atomicExch(&s->count, 0u);
    return true;
}

// FP32 form of slot_commit with no fence: the running minimum's 32-bit ordered
// encoding (high half) and the complement of the contribution count (low half)
// share the accumulator word, updated by compare-and-swap. Every contribution is
// then a read-modify-write of that one word, whose modifications are totally
// ordered, so the contributor that brings the count to `expected` holds every
// contribution's minimum with no release or acquire (a device-scope release
// writes back the XCD's L2). All-ones is the empty state of both layouts (minimum
// above every encoding, count 0), so FP32 and FP64 launches can share the words.
// A failed compare-and-swap costs a round trip, so it suits contributions spread
// out in time: with the 256 blocks of a one-round launch arriving together, the
// 32 slot completers queue on the top word (FP32 fused step 64^3 15 -> 24 us),
// while spread ones gain (batch step 5.26 -> 4.79 ms, 512^3 FP32 fused step
// 2.73 -> 2.68 ms, time step 0.439 -> 0.419 ms).
__device__ __forceinline__ bool packed_commit(MinSlot* s, unsigned int& enc, unsigned int expected)
{
    unsigned long long seen = 0xffffffffffffffffull; // first guess: empty
    for (;;) {
        unsigned int const m = static_cast<unsigned int>(seen >> 32);
        unsigned int const mn = enc < m ? enc : m;
        unsigned int const left = static_cast<unsigned int>(seen) - 1u; // ~(count + 1)
        unsigned long long const next = (static_cast<unsigned long long>(mn) << 32) | left;
        unsigned long long const was = atomicCAS(&s->acc, seen, next);
        if (was == seen) {
            if (~left != expected) {
                return false;
            }
            enc = mn;
            atomicExch(&s->acc, 0xffffffffffffffffull);
            return true;
        }
        seen = was;
    }
}

// Called by one thread of every block with that block's minimum; `rank` numbers
// the blocks of the reduction 0..nblocks-1. Block r feeds slot r % kMinSlots; the
// block that completes a slot feeds the top accumulator, and the one that
// completes the top publishes. Same-address atomics serialise at the coherence
// point (~10-20 ns each): with one word for all blocks, the 256 blocks of a 64^3
// fused step, which arrive together, spent ~6 us there (21.0 -> 18.8 us FP64,
// 17.4 -> 14.7 us FP32 with the two levels of at most ~nblocks/32 and 32).
// FP32 takes the fence-free packed form when the blocks arrive spread out: when
// the caller says so (`spread`, uniform over the launch), or when the launch has
// more blocks than one per CU (several rounds).
template <class T>
__device__ __forceinline__ void grid_min_commit(DtSink<T> const& sink, T block_value, unsigned int rank,
                                                unsigned int nblocks, bool spread = false)
{
    unsigned int const slot = rank % kMinSlots;
    unsigned int const in_slot = (nblocks - 1 - slot) / kMinSlots + 1;
    unsigned int const slots = nblocks < kMinSlots ? nblocks : kMinSlots;
    T m;
    if (sizeof(T) == 4 && (spread || nblocks > 256)) {
        unsigned int enc = static_cast<unsigned int>(ord_encode(block_value));
        if (!packed_commit(&sink.scratch->slot[slot], enc, in_slot)) {
            return;
        }
        if (!packed_commit(&sink.scratch->top, enc, slots)) {
            return;
        }
        m = ord_decode<T>(enc);
    } else {
        unsigned long long enc = ord_encode(block_value);
        if (!slot_commit(&sink.scratch->slot[slot], enc, in_slot)) {
            return;
        }
        if (!slot_commit(&sink.scratch->top, enc, slots)) {
            return;
        }
        m = ord_decode<T>(enc);
    }
    if (sink.extra_min != nullptr) {
        m = min_join(m, *sink.extra_min);
    }
    sink.dt_out[0] = m;
    if (sink.time_acc != nullptr && sink.dt_in != nullptr) {
        T const dt_step = sink.cfl * sink.dt_in[0];
        sink.time_acc[0] += static_cast<double>(dt_step);
    }
}

// Initial state of n records (every accumulator empty).
__global__ void min_scratch_init(MinScratch* s, int n)
{
    int const t = blockIdx.x * blockDim.x + threadIdx.x;
    int const per = static_cast<int>(kMinSlots) + 1;
    if (t < n * per) {
        MinSlot& w = (t % per == 0) ? s[t / per].top : s[t / per].slot[t % per - 1];
        w.acc = 0xffffffffffffffffull; // replaced by the first commit's atomicMin
        w.count = 0u;
        w.pad = 0u;
    }
}

} // namespace eb
