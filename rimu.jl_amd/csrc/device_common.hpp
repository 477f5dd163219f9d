// device_common.hpp -- block-level primitives shared by the step kernels: wavefront and
// workgroup scans/reductions, and the decoupled look-back that orders the output of
// compacting kernels (each tile publishes its count; a tile's output offset is the sum
// over all earlier tiles).  Tiles take ids from an atomic counter, so every earlier tile
// is already resident when a tile waits on it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rimu {

constexpr int WAVE = 64;

// Inclusive scan across the 64 lanes of a wavefront.
template <class T>
__device__ inline T wave_inclusive_scan(T v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        T n = __shfl_up(v, off, WAVE);
        if (lane >= off) v += n;
    }
    return v;
}

// Exclusive scan across a workgroup of NT threads (NT a multiple of 64, at most 1024).
// Returns the exclusive prefix of this thread; *total receives the workgroup sum.
template <class T, int NT>
__device__ inline T block_exclusive_scan(T v, T* total) {
    __shared__ T s_wave[NT / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    T inc = wave_inclusive_scan(v);
    if (lane == WAVE - 1) s_wave[wid] = inc;
    __syncthreads();
    T wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / WAVE; ++w) {
        const T x = s_wave[w];
        if (w < wid) wave_off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return wave_off + inc - v;
}

// Deterministic sum over a workgroup (fixed tree order).  Result valid in thread 0.
template <int NT>
__device__ inline double block_sum(double v) {
    __shared__ double s[NT];
    s[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int off = NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// N deterministic workgroup sums at once in scratch (N * NT doubles of LDS that the caller
// no longer needs), each in the tree order of block_sum (the same bits).  Results valid in
// thread 0.
template <int NT, int N>
__device__ inline void block_sum_n(double (&v)[N], double* scratch) {
#pragma unroll
    for (int j = 0; j < N; ++j) scratch[j * NT + threadIdx.x] = v[j];
    __syncthreads();
#pragma unroll
    for (int off = NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int j = 0; j < N; ++j) scratch[j * NT + threadIdx.x] += scratch[j * NT + threadIdx.x + off];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = scratch[j * NT];
    __syncthreads();
}

// Three deterministic workgroup sums at once.  Results valid in thread 0.
template <int NT>
__device__ inline void block_sum3(double& a, double& b, double& c) {
    __shared__ double s[3][NT];
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    __syncthreads();
#pragma unroll
    for (int off = NT / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s[0][threadIdx.x] += s[0][threadIdx.x + off];
            s[1][threadIdx.x] += s[1][threadIdx.x + off];
            s[2][threadIdx.x] += s[2][threadIdx.x + off];
        }
        __syncthreads();
    }
    a = s[0][0];
    b = s[1][0];
    c = s[2][0];
    __syncthreads();
}

// Status words of the look-back: top two bits flag (0 empty, 1 tile aggregate, 2 inclusive
// prefix), low 62 bits the value.
constexpr uint64_t LB_FLAG_AGG = 1ull << 62;
constexpr uint64_t LB_FLAG_PFX = 2ull << 62;
constexpr uint64_t LB_MASK = (1ull << 62) - 1;
// Default bound of every look-back wait, in rounds of one s_sleep: far beyond any wait a
// resident predecessor causes; past it the waiting tile gives up and raises an error flag.
constexpr uint32_t LB_SPIN_LIMIT = 1u << 25;

__device__ inline uint64_t lb_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lb_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wavefront-parallel look-back, called by all 64 lanes of one wavefront of tile `tile`.
// Each probe reads the status words of 64 predecessors at once (lane l reads tile
// window_end - l); the nearest inclusive prefix ends the walk, aggregates in front of it
// are added.  Returns the exclusive prefix in every lane.  Every wait is bounded by
// spin_limit rounds: if a predecessor never publishes (which the tile-id ordering rules out)
// the error flag is raised and the loop ends.
__device__ inline uint64_t lookback_exclusive_wave(uint64_t* status, uint32_t tile, uint64_t agg,
                                                   int* err_flag, uint32_t spin_limit) {
    const int lane = threadIdx.x & (WAVE - 1);
    if (tile == 0) {
        if (lane == 0) lb_store(&status[0], LB_FLAG_PFX | agg);
        return 0;
    }
    if (lane == 0) lb_store(&status[tile], LB_FLAG_AGG | agg);
    uint64_t prefix = 0;
    int64_t window_end = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (true) {
        const int64_t j = window_end - lane;
        const uint64_t w = (j >= 0) ? lb_load(&status[j]) : LB_FLAG_PFX;  // before tile 0: prefix 0
        const uint32_t f = (uint32_t)(w >> 62);
        const uint64_t empty = __ballot(f == 0);
        const uint64_t pfx = __ballot(f == 2);
        const int first = pfx ? (__ffsll((long long)pfx) - 1) : WAVE;  // nearest prefix lane
        const uint64_t needed = first >= WAVE - 1 ? ~0ull : ((2ull << first) - 1ull);
        if (empty & needed) {
            if (++spins > spin_limit) {
                if (lane == 0) atomicOr(err_flag, 2);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t v = (lane <= first) ? (w & LB_MASK) : 0ull;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) v += __shfl_xor(v, off, WAVE);
        prefix += v;
        if (first < WAVE) break;
        window_end -= WAVE;
    }
    if (lane == 0) lb_store(&status[tile], LB_FLAG_PFX | (prefix + agg));
    return prefix;
}

// Dynamic tile id shared through LDS.
__device__ inline uint32_t next_tile_id(uint32_t* counter) {
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    __syncthreads();
    return t;
}

}  // namespace rimu
