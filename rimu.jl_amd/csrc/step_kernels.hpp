// step_kernels.hpp -- the kernels of one FCIQMC step on a sorted device vector.
//
//   plan     : diagonal step per entry (death/clone with stochastic rounding) and the
//              number of spawn work items of each entry          (spawning.jl:73-93)
//   spawn    : all spawn attempts, Philox-seeded, compacted in (entry, attempt) order
//              with a workgroup scan and a decoupled look-back    (spawning.jl:174-243,
//                                                                  :364-378)
//   radix    : LSD radix sort of the spawn records on the address key, stable: one
//              histogram pass for every digit, then per pass a scatter whose tiles rank in
//              LDS and find their offsets by look-back (or, selectable, a count and scan
//              per pass)                                           (deposit!, dvec.jl:129-136)
//   merge    : merge-path union of the sorted spawns with the diagonal survivors, a
//              sequential left fold per address (diagonal value first, then spawns in list
//              order), ThresholdCompression, norm and length      (compression.jl:18-33)
//   project  : vproj / hproj of ProjectedEnergy by binary search (poststepstrategy.jl:110-121)
#pragma once

#include "bosefs_device.hpp"
#include "device_common.hpp"

namespace rimu {

constexpr int NT = 256;          // workgroup size of every step kernel
constexpr int MERGE_IPT = 8;     // merged positions per thread (one-chunk keys)
// Merged positions per thread for key type K: two-chunk keys take 12 (3072-position tiles,
// 75 KiB of LDS).  Measured at 1e8 walkers: the 128-bit merge takes 3.16-3.22 ms with 12
// against 3.67 ms with 8, 3.20 with 10, 3.17 with 14 and 5.36 ms with 16; the 64-bit merge
// is fastest with 8 (1.09 ms against 1.82 ms with 12).
template <class K> struct merge_ipt { static constexpr int value = sizeof(K) == 8 ? MERGE_IPT : 12; };
constexpr int SPAWN_TILE_SIZE = 1024;  // attempts per spawn tile (NT * SPAWN_ITEMS)
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SORT_IPT = 16;     // keys per thread per radix tile
constexpr int SORT_TILE = NT * SORT_IPT;
// Keys per thread of a radix tile for key type K.  Two-chunk keys take 24 (6144-record
// tiles, 144 KiB of LDS, one workgroup per CU): with their top bits mixed, every digit is
// spread evenly, and longer digit runs per tile measured faster than smaller tiles at
// higher occupancy (per scatter pass of 1e8 records: 1.36 ms at 6144 records, 1.41 at
// 5120, 1.43 at 4096, 1.82 at 3072, 2.16 at 2048).  One-chunk keys are fastest at 4096
// (0.56 ms per pass of 7e7 records against 0.74 at 6144 and 0.71 at 8192).
template <class K> struct sort_ipt { static constexpr int value = sizeof(K) == 8 ? SORT_IPT : 24; };
template <class K> struct sort_tile { static constexpr int value = NT * sort_ipt<K>::value; };
static_assert(RADIX == NT, "radix kernels use one thread per digit");

enum { FLAG_SPAWN_OVERFLOW = 1, FLAG_LOOKBACK = 2, FLAG_VECTOR_OVERFLOW = 4, FLAG_FIXUP_FALLBACK = 8,
       FLAG_SPEC = 32 };

// A count a kernel works on: known on the host (dev == nullptr), or written to device memory
// by an earlier kernel of the same step, with the grid sized for at most cap (a larger count
// raises FLAG_SPEC: the kernels after it do nothing and the host repeats the step with grids
// sized from the counts it has read back).
struct Count {
    int64_t n;
    const int64_t* dev;
    int64_t cap;
};
__device__ inline int64_t count_value(const Count& c) { return c.dev ? *c.dev : c.n; }

struct StepArgs {
    ModelDev h;
    uint64_t seed;
    int64_t step;
    double shift, dtau, boost;
    const double* shift_dev;  // the shift in device memory (device shift update), or null: shift
    int style;        // 1 integer, 2 deterministic, 3 dynamic semistochastic
    int compression;  // 0 none, 1 threshold
    double comp_threshold, proj_threshold, rel_threshold, abs_threshold;
    int init_rule;    // 0 NonInitiator, 1 Initiator, 2 SimpleInitiator, 3 CoherentInitiator
    double init_threshold;
    KeyMix mix;       // stored keys are mix(address)
    uint32_t spin_limit;  // bound of the plan kernel's look-back waits (LB_SPIN_LIMIT)
};

// Bounded waits of the radix look-back: spin_limit rounds per wait; withhold_first makes
// tile 0 of every pass keep its status words to itself, so that every later tile's wait
// runs into the bound (a host-forced test of the give-up path).
struct LookbackCtl {
    uint32_t spin_limit;
    int withhold_first;
};

// Initiator class of a record, carried in key bit 63 (addresses then have B <= 63 bits).
// The bit marks a deposit whose parent p_val has |p_val| <= threshold, i.e. is not an
// initiator (to_initiator_value, initiators.jl:143-157): on a spawn record it selects the
// `unsafe` component (else `safe`); on the diagonal deposit of an entry, staged in the
// merge, it selects `safe` (else `initiator`).  Radix passes and merge comparisons ignore it.
constexpr uint64_t CLASS_BIT = 1ull << 63;
template <class K> __host__ __device__ constexpr K class_bit() { return (K)1 << (8 * sizeof(K) - 1); }

// Wavefront shuffles of keys (two 64-bit shuffles for two-chunk keys).
__device__ inline uint64_t shfl_xor_key(uint64_t x, int off) { return __shfl_xor(x, off, WAVE); }
__device__ inline u128 shfl_xor_key(u128 x, int off) {
    const uint64_t lo = __shfl_xor((uint64_t)x, off, WAVE);
    const uint64_t hi = __shfl_xor((uint64_t)(x >> 64), off, WAVE);
    return ((u128)hi << 64) | lo;
}

__device__ inline bool below_initiator_threshold(const StepArgs& a, double v) {
    return !(fabs(v) > a.init_threshold);
}

template <class V> struct is_int_v { static constexpr bool value = false; };
template <> struct is_int_v<int64_t> { static constexpr bool value = true; };

__device__ inline double sgn(double x) { return (double)((x > 0) - (x < 0)); }

// projected_deposit!, spawning.jl:13-45: integers are rounded stochastically, floats are
// projected to +-threshold when smaller than a nonzero threshold.
template <class V>
__device__ inline V project(double x, double thresh, double U) {
    if constexpr (is_int_v<V>::value) {
        return (V)sgn(x) * (V)floor(fabs(x) + U);
    } else {
        if (thresh > 0 && fabs(x) < thresh) return (U < fabs(x) / thresh) ? sgn(x) * thresh : 0.0;
        return x;
    }
}

// Spawn-side decisions of one entry: exact (all n_od neighbours) or stochastic
// (n_att draws with replacement), spawning.jl:232-243 and :364-378.
struct ColumnPlan {
    int64_t n_od;
    int64_t n_items;  // neighbours (exact) or attempts (stochastic)
    bool exact;
};

template <class K, class V, bool RS>
__device__ inline ColumnPlan plan_column(const StepArgs& a, K k, V v) {
    ColumnPlan c;
    c.n_od = model_num_offdiagonals<RS>(a.h, k);
    const double av = fabs((double)v);
    if (a.style == 2) {
        c.exact = true;
    } else if (a.style == 3) {
        const double thresh = fmin(a.abs_threshold, (double)c.n_od);
        c.exact = (a.boost * av * a.rel_threshold) >= thresh;
    } else {
        c.exact = false;
    }
    if (c.exact) c.n_items = c.n_od;
    else c.n_items = max((int64_t)floor(av * a.boost), (int64_t)1);
    if (c.n_od == 0) c.n_items = 0;
    return c;
}

// ---------------------------------------------------------------------------------------
// plan: the diagonal step and the number of spawn attempts of every entry, the exclusive
// scan of the attempt counts (aoff, L + 1 values; aoff[L] = W, also written to *W_out), and
// the first entry of every spawn tile (tile_first[t] = entry holding attempt t*SPAWN_TILE,
// for t < tf_cap).  Thread x of tile t handles the PLAN_IPT consecutive entries
// t*PLAN_TILE + x*PLAN_IPT + j; tiles are placed by the decoupled look-back (tile ids in
// start order, so every tile waited on is running).
// partials[j * gridDim.x + tile], j = 0..5: deaths, clones, zombies, exact, inexact, attempts.
// ---------------------------------------------------------------------------------------
constexpr int PLAN_IPT = 8;
constexpr int PLAN_TILE = NT * PLAN_IPT;

template <class K, class V, bool RS>
__global__ __launch_bounds__(NT) void plan_kernel(StepArgs a, const K* __restrict__ keys,
                                                  const V* __restrict__ vals, int64_t L,
                                                  V* __restrict__ dvals, int64_t* __restrict__ aoff,
                                                  int64_t* __restrict__ tile_first, int64_t tf_cap,
                                                  uint64_t* status, uint32_t* tile_counter, int* flags,
                                                  double* __restrict__ partials, int64_t* W_out) {
    __shared__ int64_t s_cnt[PLAN_TILE];  // counts, then exclusive offsets, in entry order
    __shared__ int64_t s_prefix;
    const uint32_t tile = next_tile_id(tile_counter);
    const int64_t t0 = (int64_t)tile * PLAN_TILE;
    const double shift = a.shift_dev ? a.shift_dev[0] : a.shift;
    double st[6] = {0, 0, 0, 0, 0, 0};
    // striped: entry t0 + j*NT + x (coalesced loads and stores)
    K kk[PLAN_IPT];
    V vv[PLAN_IPT];
#pragma unroll
    for (int j = 0; j < PLAN_IPT; ++j) {  // all loads in flight before any is used
        const int64_t i = min(t0 + j * NT + threadIdx.x, max(L - 1, (int64_t)0));
        kk[j] = L > 0 ? keys[i] : 0;
        vv[j] = L > 0 ? vals[i] : (V)0;
    }
#pragma unroll
    for (int j = 0; j < PLAN_IPT; ++j) {
        const int64_t i = t0 + j * NT + threadIdx.x;
        int64_t n_items = 0;
        if (i < L) {
            const K k = unmix_key(a.mix, kk[j]);  // the address
            const V v = vv[j];
            // diagonal_step!, spawning.jl:73-77 with diagonal_element(T) fciqmc.jl:93-96
            const double hd = model_diagonal_element<RS>(a.h, k);
            const double d = 1 - a.dtau * (hd - shift);
            const double x = d * (double)v;
            double U = 0;
            if (is_int_v<V>::value || a.proj_threshold > 0) {
                const U4 r = draw(a.seed, k, 0, a.step, TAG_DIAG);
                U = u53(r.x, r.y);
            }
            const V nv = project<V>(x, a.proj_threshold, U);
            dvals[i] = nv;
            // clones_deaths_zombies, spawning.jl:79-93
            const double res = (double)nv, val = (double)v;
            if (res > val) st[1] += fabs(res - val);
            else if (sgn(res) != sgn(val)) { st[0] += fabs(val); st[2] += fabs(res); }
            else st[0] += fabs(res - val);
            const ColumnPlan c = plan_column<K, V, RS>(a, k, v);
            if (c.n_od > 0) {
                if (a.style != 1) { if (c.exact) st[3] += 1; else st[4] += 1; }
                st[5] += (double)c.n_items;
            }
            n_items = c.n_items;
        }
        s_cnt[j * NT + threadIdx.x] = n_items;
    }
    __syncthreads();
    // blocked: thread x scans the PLAN_IPT consecutive entries x*PLAN_IPT + j
    const int b0 = (int)threadIdx.x * PLAN_IPT;
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < PLAN_IPT; ++j) sum += s_cnt[b0 + j];
    int64_t total;
    const int64_t ex = block_exclusive_scan<int64_t, NT>(sum, &total);  // ends with a barrier
    if (threadIdx.x < WAVE) {
        const uint64_t b = lookback_exclusive_wave(status, tile, (uint64_t)total, flags, a.spin_limit);
        if (threadIdx.x == 0) s_prefix = (int64_t)b;
    }
    __syncthreads();
    int64_t run = s_prefix + ex;
#pragma unroll
    for (int j = 0; j < PLAN_IPT; ++j) {
        const int64_t i = t0 + b0 + j;
        const int64_t c = s_cnt[b0 + j];
        s_cnt[b0 + j] = run;
        if (i < L) {
            const int64_t end = run + c;
            for (int64_t t = (run + SPAWN_TILE_SIZE - 1) / SPAWN_TILE_SIZE; t * SPAWN_TILE_SIZE < end && t < tf_cap; ++t)
                tile_first[t] = i;
        } else if (i == L) {
            aoff[L] = run;
            *W_out = run;
        }
        run += c;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PLAN_IPT; ++j) {
        const int64_t i = t0 + j * NT + threadIdx.x;
        if (i < L) aoff[i] = s_cnt[j * NT + threadIdx.x];
    }
    // the six tile sums at once, in s_cnt once every offset has been read
    static_assert(sizeof(s_cnt) >= 6 * NT * sizeof(double), "s_cnt holds the six sums");
    __syncthreads();
    block_sum_n<NT, 6>(st, reinterpret_cast<double*>(s_cnt));
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j) partials[(int64_t)j * gridDim.x + tile] = st[j];
    }
}

// ---------------------------------------------------------------------------------------
// Exclusive scan of int64 counts (three kernels; deterministic).
// ---------------------------------------------------------------------------------------
constexpr int SCAN_IPT = 8;
constexpr int SCAN_TILE = NT * SCAN_IPT;

template <class T>
__global__ __launch_bounds__(NT) void scan_tiles_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                        int64_t n, T* __restrict__ tile_sums) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
    T loc[SCAN_IPT];
    T sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_IPT; ++j) {
        const int64_t idx = base + j;
        loc[j] = idx < n ? in[idx] : (T)0;
        sum += loc[j];
    }
    T total;
    T run = block_exclusive_scan<T, NT>(sum, &total);
#pragma unroll
    for (int j = 0; j < SCAN_IPT; ++j) {
        const int64_t idx = base + j;
        if (idx < n) out[idx] = run;
        run += loc[j];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

template <class T>
__global__ __launch_bounds__(1024) void scan_sums_kernel(T* __restrict__ sums, int64_t n,
                                                         T* __restrict__ grand_total) {
    T carry = 0;
    for (int64_t base = 0; base < n; base += 1024) {
        const int64_t idx = base + threadIdx.x;
        const T v = idx < n ? sums[idx] : (T)0;
        T tot;
        const T ex = block_exclusive_scan<T, 1024>(v, &tot);
        if (idx < n) sums[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && grand_total) *grand_total = carry;
}

template <class T>
__global__ __launch_bounds__(NT) void scan_add_kernel(T* __restrict__ out, int64_t n,
                                                      const T* __restrict__ tile_sums) {
    const T add = tile_sums[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    T x[SCAN_IPT];
#pragma unroll
    for (int j = 0; j < SCAN_IPT; ++j) {
        const int64_t idx = base + (int64_t)j * NT + threadIdx.x;
        x[j] = out[min(idx, n - 1)];  // all loads issued before the first store
    }
#pragma unroll
    for (int j = 0; j < SCAN_IPT; ++j) {
        const int64_t idx = base + (int64_t)j * NT + threadIdx.x;
        if (idx < n) out[idx] = x[j] + add;
    }
}

// ---------------------------------------------------------------------------------------
// spawn: one thread per spawn attempt.  Attempts are numbered entry by entry (attempt j of
// entry i is aoff[i] + j), so consecutive lanes hold consecutive attempts and the lane
// order is the (entry, attempt) order.  Each thread handles SPAWN_ITEMS attempts, strided
// by NT within the tile.  Nonzero records are compacted with a wavefront ballot and
// popcount prefix; the (item, wavefront) counts are scanned across the workgroup and the
// tile's records go to its own slot range with their count in tile_cnt (compact_kernel
// places them).  When every attempt yields a record (DENSE: floating values without
// projection) the record of attempt a is simply written at position a.
// Also writes sum |spawn| and the key OR/AND (which radix digits vary) of the tile.
// ---------------------------------------------------------------------------------------
constexpr int SPAWN_ITEMS = 4;
constexpr int SPAWN_TILE = NT * SPAWN_ITEMS;
static_assert(SPAWN_TILE % NT == 0, "whole slots per thread in the entry scan");
static_assert(SPAWN_TILE == SPAWN_TILE_SIZE, "plan and spawn agree on the spawn tile");

template <class K, class V, bool DENSE, bool RS>
__global__ __launch_bounds__(NT) void spawn_kernel(StepArgs a, const K* __restrict__ keys,
                                                   const V* __restrict__ vals, int64_t L,
                                                   const int64_t* __restrict__ aoff, Count cA,
                                                   int64_t out_cap, int* flags,
                                                   K* __restrict__ out_keys,
                                                   V* __restrict__ out_vals,
                                                   int64_t* __restrict__ tile_cnt,
                                                   double* __restrict__ partials,
                                                   const int64_t* __restrict__ tile_first,
                                                   K* __restrict__ tile_or,
                                                   K* __restrict__ tile_and) {
    constexpr int NW = NT / WAVE;
    constexpr int WIN = SPAWN_TILE + 1;
    __shared__ int64_t s_aoff[WIN];
    __shared__ int s_head[SPAWN_TILE];  // per attempt slot: window index of its entry
    __shared__ int s_wmax[NW];
    __shared__ int s_cnt[SPAWN_ITEMS * NW];
    __shared__ int s_total;
    __shared__ K s_kor[NW], s_kand[NW];

    const uint32_t tile = blockIdx.x;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int64_t a0 = (int64_t)tile * SPAWN_TILE;
    const int64_t A = count_value(cA);
    // attempts beyond the grid (FLAG_SPEC) or, dense, beyond the spawn buffer: nothing is
    // written; tiles past the attempts leave neutral partials (uniform branches)
    const bool spec_miss = A > cA.cap;
    const bool no_room = DENSE && A > out_cap;
    if (spec_miss || no_room || a0 >= A) {
        if (threadIdx.x == 0) {
            if (tile == 0 && spec_miss) atomicOr(flags, FLAG_SPEC);
            if (tile == 0 && no_room) atomicOr(flags, FLAG_SPAWN_OVERFLOW);
            partials[tile] = 0;
            tile_or[tile] = 0;
            tile_and[tile] = (K)~(K)0;
            if (!DENSE) tile_cnt[tile] = 0;
        }
        return;
    }
    const int64_t e0 = tile_first[tile];
    {
        constexpr int WR = (WIN + NT - 1) / NT;
        int64_t w[WR];
#pragma unroll
        for (int r = 0; r < WR; ++r) {
            const int t = (int)threadIdx.x + r * NT;
            if (t < WIN) w[r] = aoff[min(e0 + t, L)];
        }
#pragma unroll
        for (int r = 0; r < WR; ++r) {
            const int t = (int)threadIdx.x + r * NT;
            if (t < WIN) s_aoff[t] = w[r];
        }
#pragma unroll
        for (int r = 0; r < SPAWN_ITEMS; ++r) s_head[(int)threadIdx.x + r * NT] = 0;
        __syncthreads();
        // entry t of the window (0 < t < WIN-1) starts at attempt s_aoff[t]: mark that slot.
        // Entries without attempts share the start of the next one; the larger index wins,
        // which is the entry that owns the attempt.
#pragma unroll
        for (int r = 0; r < WR; ++r) {
            const int t = (int)threadIdx.x + r * NT;
            if (t >= 1 && t < WIN - 1 && e0 + t < L) {
                const int64_t p = w[r] - a0;
                if (p >= 0 && p < SPAWN_TILE) atomicMax(&s_head[p], t);
            }
        }
        __syncthreads();
        // inclusive max-scan over the slots: slot q then holds the entry of attempt a0 + q
        {
            constexpr int Q = SPAWN_TILE / NT;
            const int q0 = (int)threadIdx.x * Q;
            int m[Q];
            int run = 0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                run = max(run, s_head[q0 + q]);
                m[q] = run;
            }
            int incl = m[Q - 1];
#pragma unroll
            for (int off = 1; off < WAVE; off <<= 1) {
                const int o = __shfl_up(incl, off, WAVE);
                if (lane >= off) incl = max(incl, o);
            }
            int before = __shfl_up(incl, 1, WAVE);
            if (lane == 0) before = 0;
            if (lane == WAVE - 1) s_wmax[wid] = incl;
            __syncthreads();
#pragma unroll
            for (int ww = 0; ww < NW; ++ww)
                if (ww < wid) before = max(before, s_wmax[ww]);
#pragma unroll
            for (int q = 0; q < Q; ++q) s_head[q0 + q] = max(m[q], before);
        }
    }
    __syncthreads();

    K rk[SPAWN_ITEMS];
    V rv[SPAWN_ITEMS];
    int rpos[SPAWN_ITEMS];
    bool rok[SPAWN_ITEMS];
    double spawned = 0;
    K kor = 0, kand = (K)~(K)0;
    // entry of each attempt: last i with aoff[i] <= att (every entry has at least one
    // attempt unless it has no off-diagonals; outside the window: global search)
    int64_t ent[SPAWN_ITEMS], aidx[SPAWN_ITEMS];
#pragma unroll
    for (int it = 0; it < SPAWN_ITEMS; ++it) {
        const int64_t att = a0 + (int64_t)it * NT + threadIdx.x;
        int64_t i = e0, ai = att;
        if (att < A) {
            if (att < s_aoff[WIN - 1]) {
                const int w = s_head[it * NT + threadIdx.x];
                i = e0 + w;
                ai = s_aoff[w];
            } else {
                int64_t lo = e0, hi = L;
                while (hi - lo > 1) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (aoff[mid] <= att) lo = mid; else hi = mid;
                }
                i = lo;
                ai = aoff[lo];
            }
        }
        ent[it] = i;
        aidx[it] = ai;
    }
    // the entries' keys and values, all loads in flight at once (e0 < L whenever A > 0)
    K ekey[SPAWN_ITEMS];
    V eval[SPAWN_ITEMS];
#pragma unroll
    for (int it = 0; it < SPAWN_ITEMS; ++it) {
        const int64_t i = min(ent[it], L - 1);
        ekey[it] = keys[i];
        eval[it] = vals[i];
    }
#pragma unroll
    for (int it = 0; it < SPAWN_ITEMS; ++it) ekey[it] = unmix_key(a.mix, ekey[it]);  // addresses
#pragma unroll
    for (int it = 0; it < SPAWN_ITEMS; ++it) {
        const int64_t att = a0 + (int64_t)it * NT + threadIdx.x;
        bool ok = false;
        K nb = 0;
        V nv = (V)0;
        if (att < A) {
            const int64_t j = att - aidx[it];
            const K k = ekey[it];
            const V v = eval[it];
            const ColumnPlan c = plan_column<K, V, RS>(a, k, v);
            const bool need_u = is_int_v<V>::value || a.proj_threshold > 0;
            U4 r = {0, 0, 0, 0};
            if (!c.exact || need_u) r = draw(a.seed, k, (uint32_t)j, a.step, TAG_SPAWN);
            const int64_t chosen = c.exact ? j + 1 : pick1(r.x, c.n_od);
            double hv;
            nb = mix_key(a.mix, model_offdiagonal<RS>(a.h, k, chosen, &hv));
            const double mel = -hv * a.dtau;  // FirstOrderOffdiagonals, fciqmc.jl:109-112
            double x;
            if (c.exact) {
                x = (double)v * mel;
            } else {
                const double magnitude = (double)v / (double)c.n_items;
                const double prob = 1.0 / (double)c.n_od;
                x = mel * magnitude / prob;
            }
            nv = project<V>(x, a.proj_threshold, need_u ? u53(r.y, r.z) : 0.0);
            spawned += fabs((double)nv);
            ok = DENSE || nv != (V)0;
            if (ok) { kor |= nb; kand &= nb; }
        }
        rk[it] = nb;
        if (a.init_rule != 0 && att < A && below_initiator_threshold(a, (double)eval[it])) rk[it] |= class_bit<K>();
        rv[it] = nv;
        rok[it] = ok;
        if (!DENSE) {
            const uint64_t m = __ballot(ok);
            rpos[it] = __popcll(m & lt_mask);
            if (lane == 0) s_cnt[it * NW + wid] = __popcll(m);
        }
    }

    // DENSE: the record of attempt a goes to position a.  Otherwise the tile's records go,
    // in attempt order, to its own slot range [a0, a0 + total); compact_kernel places them.
    int total = 0;
    if (!DENSE) {
        __syncthreads();
        // exclusive scan of the (item, wavefront) counts: attempt order within the tile
        if (threadIdx.x == 0) {
            int run = 0;
            for (int q = 0; q < SPAWN_ITEMS * NW; ++q) { const int c = s_cnt[q]; s_cnt[q] = run; run += c; }
            s_total = run;
        }
        __syncthreads();
        total = s_total;
    }
#pragma unroll
    for (int it = 0; it < SPAWN_ITEMS; ++it) {
        if (rok[it]) {
            const int64_t pos = DENSE ? a0 + (int64_t)it * NT + threadIdx.x
                                      : a0 + s_cnt[it * NW + wid] + rpos[it];
            out_keys[pos] = rk[it];
            out_vals[pos] = rv[it];
        }
    }
    const double s = block_sum<NT>(spawned);
    // key OR/AND of the tile (which radix digits vary; reduced over the tiles afterwards)
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        kor |= shfl_xor_key(kor, off);
        kand &= shfl_xor_key(kand, off);
    }
    if (lane == 0) {
        s_kor[wid] = kor;
        s_kand[wid] = kand;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        K bo = 0, ba = (K)~(K)0;
        for (int q = 0; q < NW; ++q) { bo |= s_kor[q]; ba &= s_kand[q]; }
        partials[tile] = s;
        tile_or[tile] = bo;
        tile_and[tile] = ba;
        if (!DENSE) tile_cnt[tile] = total;
    }
}

// ---------------------------------------------------------------------------------------
// Ordering after a radix sort on the key bits [lo, B) only (mixed keys, see KeyMix): runs of
// records that agree on those bits but not below them are put in key order by a stable
// insertion sort, one thread per run.  Equal keys keep their list order.  Runs longer than
// FIXUP_MAX_RUN, or more runs than the list holds, raise FLAG_FIXUP_FALLBACK and the host
// sorts on all bits instead.
// ---------------------------------------------------------------------------------------
constexpr int FIXUP_MAX_RUN = 512;
constexpr int64_t FIXUP_LIST = 1 << 20;

// Adjacent records i, i+1 that share bits [lo, B) but differ below: the start of the run
// holding them is listed once (by the first such pair of the run).
template <class K>
__global__ __launch_bounds__(NT) void fixup_find_kernel(const K* __restrict__ keys, Count cn, int lo,
                                                        K kmask, int64_t* __restrict__ runs,
                                                        uint32_t* count, int* flags) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (*flags & FLAG_SPEC) return;  // raised by earlier kernels only
    const int64_t n = count_value(cn);
    if (i + 1 >= n) return;
    const K a = keys[i] & kmask, b = keys[i + 1] & kmask;
    if (a == b || ((a ^ b) >> lo) != 0) return;
    // run start: the first record with these top bits (bounded walk: runs are short)
    int64_t s = i;
    while (s > 0 && i - s <= FIXUP_MAX_RUN && (((keys[s - 1] & kmask) ^ a) >> lo) == 0) --s;
    if (i - s > FIXUP_MAX_RUN) { atomicOr(flags, FLAG_FIXUP_FALLBACK); return; }
    for (int64_t j = s; j < i; ++j)   // an earlier unequal pair in the run lists it instead
        if ((keys[j] & kmask) != (keys[j + 1] & kmask)) return;
    const uint32_t slot = atomicAdd(count, 1u);
    if (slot < FIXUP_LIST) runs[slot] = s;
    else atomicOr(flags, FLAG_FIXUP_FALLBACK);
}

// One thread per listed run: stable insertion sort of [s, end of run) on the masked key.
template <class K, class V>
__global__ __launch_bounds__(NT) void fixup_sort_kernel(K* __restrict__ keys, V* __restrict__ vals, Count cn,
                                                        int lo, K kmask, const int64_t* __restrict__ runs,
                                                        const uint32_t* count, int* flags) {
    if (*flags & FLAG_SPEC) return;  // raised by earlier kernels only
    const int64_t n = count_value(cn);
    const uint32_t nr = min(*count, (uint32_t)FIXUP_LIST);
    for (uint32_t r = blockIdx.x * NT + threadIdx.x; r < nr; r += gridDim.x * NT) {
        const int64_t s = runs[r];
        const K top = (keys[s] & kmask) >> lo;
        int64_t e = s + 1;
        while (e < n && e - s <= FIXUP_MAX_RUN && ((keys[e] & kmask) >> lo) == top) ++e;
        if (e - s > FIXUP_MAX_RUN) { atomicOr(flags, FLAG_FIXUP_FALLBACK); continue; }
        for (int64_t j = s + 1; j < e; ++j) {
            const K kj = keys[j];
            const V vj = vals[j];
            int64_t q = j;
            while (q > s && (keys[q - 1] & kmask) > (kj & kmask)) {
                keys[q] = keys[q - 1];
                vals[q] = vals[q - 1];
                --q;
            }
            keys[q] = kj;
            vals[q] = vj;
        }
    }
}

// First entry of every spawn tile: entry i owns attempts [aoff[i], aoff[i+1]); the tile
// starting at attempt t*SPAWN_TILE belongs first to the entry whose range contains it.
__global__ __launch_bounds__(NT) void tile_first_kernel(const int64_t* __restrict__ aoff, int64_t L,
                                                        int64_t* __restrict__ tile_first) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= L) return;
    const int64_t c0 = aoff[i], c1 = aoff[i + 1];
    for (int64_t t = (c0 + SPAWN_TILE - 1) / SPAWN_TILE; t * SPAWN_TILE < c1; ++t) tile_first[t] = i;
}

// Merge-path split of every merge tile boundary d = t*TILE (clamped to S+L): number of A
// elements among the first d merged elements (B before A on equal keys).
// kmask clears the initiator class bit of the spawn keys (all ones without initiators).
template <class K>
__global__ __launch_bounds__(NT) void merge_split_kernel(const K* __restrict__ ak, Count cS,
                                                         const K* __restrict__ bk, int64_t L,
                                                         int64_t tile_size, int64_t ntiles,
                                                         K kmask,
                                                         int64_t* __restrict__ splits,
                                                         K* __restrict__ prevk) {
    const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (t > ntiles) return;
    const int64_t S = min(count_value(cS), cS.cap);  // more records raise FLAG_SPEC earlier
    const int64_t d = min(t * tile_size, S + L);
    int64_t lo = max((int64_t)0, d - L), hi = min(d, S);
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((ak[mid] & kmask) < bk[d - mid - 1]) lo = mid + 1; else hi = mid;
    }
    splits[t] = lo;
    // key of the merged element just before the tile (read only when the tile starts past 0)
    const int64_t bi = d - lo;
    const K pa = lo > 0 ? (K)(ak[lo - 1] & kmask) : (K)0;
    const K pb = bi > 0 ? bk[bi - 1] : (K)0;
    prevk[t] = (lo > 0 && bi > 0) ? max(pa, pb) : (lo > 0 ? pa : pb);
}

// OR and AND of the per-tile key words (one workgroup).
template <class K>
__global__ __launch_bounds__(1024) void key_bits_kernel(const K* __restrict__ tile_or,
                                                        const K* __restrict__ tile_and,
                                                        int64_t ntiles, K* key_or, K* key_and) {
    __shared__ K s_o[1024 / WAVE], s_a[1024 / WAVE];
    K o = 0, an = (K)~(K)0;
    for (int64_t base = 0; base < ntiles; base += 4 * 1024) {
        K xo[4], xa[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t t = base + r * 1024 + threadIdx.x;
            xo[r] = t < ntiles ? tile_or[t] : (K)0;
            xa[r] = t < ntiles ? tile_and[t] : (K)~(K)0;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) { o |= xo[r]; an &= xa[r]; }
    }
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        o |= shfl_xor_key(o, off);
        an &= shfl_xor_key(an, off);
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        s_o[threadIdx.x / WAVE] = o;
        s_a[threadIdx.x / WAVE] = an;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 1024 / WAVE; ++q) { o |= s_o[q]; an &= s_a[q]; }
        *key_or = o;
        *key_and = an;
    }
}

// ---------------------------------------------------------------------------------------
// Radix sort pass: count (per-tile digit histogram, digit-major) and scatter (stable
// ranks within the tile from wavefront ballots, staged in LDS in digit order, then
// written to the global digit offsets).
// ---------------------------------------------------------------------------------------
// dmask selects the digit bits of the pass: RADIX - 1, or fewer bits for a top digit that
// would reach the initiator class bit.
// Digit counts go to HIST_COPIES copies of the histogram, lane l adding to copy
// l % HIST_COPIES: lanes of a wavefront with the same digit (the top digits of the keys take
// few values) then hit different words, and the odd stride puts the copies of one digit in
// different LDS banks.
constexpr int HIST_COPIES = 8;
constexpr int HIST_STRIDE = RADIX + 1;
template <class K>
__global__ __launch_bounds__(NT) void radix_count_kernel(const K* __restrict__ keys, int64_t n,
                                                         int shift, uint32_t dmask,
                                                         uint32_t* __restrict__ counts,
                                                         int64_t num_tiles) {
    constexpr int IPT = sort_ipt<K>::value;
    __shared__ uint32_t hist[HIST_COPIES * HIST_STRIDE];
    for (int j = threadIdx.x; j < HIST_COPIES * HIST_STRIDE; j += NT) hist[j] = 0;
    __syncthreads();
    uint32_t* mine = hist + (threadIdx.x % HIST_COPIES) * HIST_STRIDE;
    const int64_t base = (int64_t)blockIdx.x * sort_tile<K>::value;
    K k[IPT];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = base + (int64_t)r * NT + threadIdx.x;
        k[r] = keys[min(idx, n - 1)];  // all loads issued before the first atomic
    }
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = base + (int64_t)r * NT + threadIdx.x;
        if (idx < n) atomicAdd(&mine[(uint32_t)(k[r] >> shift) & dmask], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < HIST_COPIES; ++q) c += hist[q * HIST_STRIDE + threadIdx.x];
    counts[(int64_t)blockIdx.x * RADIX + threadIdx.x] = c;  // tile-major row
}

// Offsets of every (tile, digit) for the stable scatter, from tile-major counts:
//   offs[t][d] = sum over d' < d of all counts of d'  +  sum over t' < t of counts[t'][d].
// Tiles are grouped in chunks of RADIX_CHUNK.  radix_chunk_sums_kernel writes the chunk
// sums digit-major (sums[d * nchunks + c]); an exclusive scan of that small array gives
// each (digit, chunk) its base; radix_chunk_offsets_kernel walks the tiles of a chunk and
// turns the counts into offsets in place.  Every pass reads and writes whole rows.
constexpr int RADIX_CHUNK = 64;

__global__ __launch_bounds__(RADIX) void radix_chunk_sums_kernel(const uint32_t* __restrict__ counts,
                                                                 int64_t num_tiles, int64_t nchunks,
                                                                 uint32_t* __restrict__ sums) {
    const int64_t c = blockIdx.x;
    const int d = threadIdx.x;
    const int64_t t0 = c * RADIX_CHUNK, t1 = min(t0 + RADIX_CHUNK, num_tiles);
    uint32_t x[RADIX_CHUNK];
#pragma unroll
    for (int i = 0; i < RADIX_CHUNK; ++i) x[i] = t0 + i < t1 ? counts[(t0 + i) * RADIX + d] : 0u;
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < RADIX_CHUNK; ++i) acc += x[i];
    sums[(int64_t)d * nchunks + c] = acc;
}

__global__ __launch_bounds__(RADIX) void radix_chunk_offsets_kernel(uint32_t* __restrict__ counts,
                                                                    int64_t num_tiles, int64_t nchunks,
                                                                    const uint32_t* __restrict__ base) {
    const int64_t c = blockIdx.x;
    const int d = threadIdx.x;
    const int64_t t0 = c * RADIX_CHUNK, t1 = min(t0 + RADIX_CHUNK, num_tiles);
    uint32_t x[RADIX_CHUNK];
#pragma unroll
    for (int i = 0; i < RADIX_CHUNK; ++i) x[i] = t0 + i < t1 ? counts[(t0 + i) * RADIX + d] : 0u;
    uint32_t run = base[(int64_t)d * nchunks + c];
#pragma unroll
    for (int i = 0; i < RADIX_CHUNK; ++i) {
        if (t0 + i < t1) counts[(t0 + i) * RADIX + d] = run;
        run += x[i];
    }
}

template <class K, class V>
__global__ __launch_bounds__(NT) void radix_scatter_kernel(const K* __restrict__ keys_in,
                                                           const V* __restrict__ vals_in, int64_t n,
                                                           int shift, uint32_t dmask,
                                                           const uint32_t* __restrict__ offs,
                                                           int64_t num_tiles,
                                                           K* __restrict__ keys_out,
                                                           V* __restrict__ vals_out) {
    // Wavefront w ranks the contiguous sub-tile [w*IPT*64, (w+1)*IPT*64) of the
    // tile with private digit counters: round r handles keys w*IPT*64 + r*64 + lane,
    // so (round, lane) is the element order and ranks are stable without barriers.
    constexpr int NW = NT / WAVE;
    constexpr int IPT = sort_ipt<K>::value;
    constexpr int TILE = sort_tile<K>::value;
    __shared__ K s_k[TILE];
    __shared__ V s_v[TILE];
    __shared__ uint32_t s_wcnt[NW][RADIX];  // per-wavefront digit counts, then prefixes
    __shared__ uint32_t s_start[RADIX];
    __shared__ uint32_t s_goff[RADIX];

    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int64_t base = (int64_t)blockIdx.x * TILE;
    const int64_t wbase = base + (int64_t)wid * IPT * WAVE;

    for (int j = threadIdx.x; j < NW * RADIX; j += NT) (&s_wcnt[0][0])[j] = 0;
    s_goff[threadIdx.x] = offs[(int64_t)blockIdx.x * RADIX + threadIdx.x];
    __syncthreads();

    K k[IPT];
    V v[IPT];
    uint32_t rank[IPT];
    uint32_t* cnt = s_wcnt[wid];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = wbase + (int64_t)r * WAVE + lane;
        const bool valid = idx < n;
        k[r] = valid ? keys_in[idx] : (K)0;
        v[r] = valid ? vals_in[idx] : (V)0;
    }
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = wbase + (int64_t)r * WAVE + lane;
        const bool valid = idx < n;
        const uint32_t digit = (uint32_t)(k[r] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RADIX_BITS; ++b) {
            const uint64_t bb = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t before = valid ? cnt[digit] : 0u;
        const uint32_t in_wave = (uint32_t)__popcll(peers & lt_mask);
        rank[r] = before + in_wave;
        // the lowest lane of each peer group advances the counter after every lane read it
        __builtin_amdgcn_wave_barrier();
        if (valid && in_wave == 0) cnt[digit] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: prefix over wavefronts (element order) and the tile's digit total
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = s_wcnt[w][d];
            s_wcnt[w][d] = run;
            run += c;
        }
        uint32_t tot;
        s_start[d] = block_exclusive_scan<uint32_t, NT>(run, &tot);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = wbase + (int64_t)r * WAVE + lane;
        if (idx < n) {
            const uint32_t digit = (uint32_t)(k[r] >> shift) & dmask;
            const uint32_t pos = s_start[digit] + s_wcnt[wid][digit] + rank[r];
            s_k[pos] = k[r];
            s_v[pos] = v[r];
        }
    }
    __syncthreads();
    const int64_t cnt_tile = min((int64_t)TILE, n - base);
    for (int i = threadIdx.x; i < cnt_tile; i += NT) {
        const K kk = s_k[i];
        const uint32_t digit = (uint32_t)(kk >> shift) & dmask;
        const int64_t g = (int64_t)s_goff[digit] + (i - (int64_t)s_start[digit]);
        keys_out[g] = kk;
        vals_out[g] = s_v[i];
    }
}

// ---------------------------------------------------------------------------------------
// Radix sort without per-pass count kernels.  One pass over the keys histograms every
// digit the sort will use (radix_hist_all_kernel, radix_hist_base_kernel: the global
// exclusive prefix over digits of each pass).  Each scatter tile then publishes its digit
// counts and looks back over its predecessors' status words, one digit per thread, for the
// count of each digit in the tiles before it (decoupled look-back).  A tile waits only for
// tiles of lower id, which have started before it.  Every wait is bounded (LookbackCtl), past
// which FLAG_LOOKBACK is raised;
// a tile that gave up publishes a prefix no larger than the true one, so every write of the
// pass stays inside the output, and the host repeats the step with the per-pass sort.  Ranks
// within the tile and the scatter are those of radix_scatter_kernel, so the output is the
// same.
// ---------------------------------------------------------------------------------------
constexpr int MAX_PASSES = 16;
constexpr int LB_WIN = 8;   // predecessor status words read per look-back round
// Status words of the radix look-back: flag (2 bits, as LB_FLAG_*), the epoch of the sort
// call that wrote them (30 bits; words of other calls read as empty, so the status array is
// cleared only when it grows or the epochs wrap), and a count (32 bits).
constexpr uint32_t RLB_EPOCH_MAX = (1u << 30) - 1;
constexpr uint64_t RLB_EPOCH_FIELD = (uint64_t)RLB_EPOCH_MAX << 32;
constexpr uint64_t RLB_VALUE = 0xffffffffull;
struct PassSet {
    int n;
    int shift[MAX_PASSES];
    uint32_t dmask[MAX_PASSES];
};

// part[(b % HIST_ROWS * ps.n + p) * RADIX + d] += count of digit d of pass p over the tiles
// of workgroup b (grid-stride over tiles).  The workgroups add into HIST_ROWS rows, not
// one: 2048 workgroups adding into the same 256 words per pass serialize on them.
// Dynamic LDS: ps.n * copies * HIST_STRIDE words (copies as in radix_count_kernel; the
// host takes hist_all_copies(ps.n), which keeps the array within 64 KiB).
__host__ __device__ inline int hist_all_copies(int npasses) { return npasses <= 8 ? 4 : 2; }
constexpr int HIST_ALL_BLOCKS = 2048;
constexpr int HIST_ROWS = 64;
template <class K>
__global__ __launch_bounds__(NT) void radix_hist_all_kernel(const K* __restrict__ keys, Count cn, int* flags,
                                                            PassSet ps, int copies, uint32_t* __restrict__ part) {
    constexpr int IPT = sort_ipt<K>::value;
    constexpr int TILE = sort_tile<K>::value;
    extern __shared__ uint32_t h_all[];
    // every thread computes the same n: the early return is uniform
    const int64_t n = count_value(cn);
    if (n > cn.cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flags, FLAG_SPEC);
        return;
    }
    const int words = ps.n * copies * HIST_STRIDE;
    for (int j = threadIdx.x; j < words; j += NT) h_all[j] = 0;
    __syncthreads();
    const int copy = threadIdx.x % copies;
    const int64_t tiles = (n + TILE - 1) / TILE;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base = t * TILE;
        K k[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) k[r] = keys[min(base + (int64_t)r * NT + threadIdx.x, n - 1)];
        for (int p = 0; p < ps.n; ++p) {
            const int shift = ps.shift[p];
            const uint32_t dmask = ps.dmask[p];
            uint32_t* mine = h_all + (p * copies + copy) * HIST_STRIDE;
#pragma unroll
            for (int r = 0; r < IPT; ++r)
                if (base + (int64_t)r * NT + threadIdx.x < n)
                    atomicAdd(&mine[(uint32_t)(k[r] >> shift) & dmask], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < ps.n; ++p) {
        uint32_t c = 0;
        for (int q = 0; q < copies; ++q) c += h_all[(p * copies + q) * HIST_STRIDE + threadIdx.x];
        if (c) atomicAdd(&part[((int64_t)(blockIdx.x % HIST_ROWS) * ps.n + p) * RADIX + threadIdx.x], c);
    }
}

// hist[p][d] = sum over d' < d of the digit counts of pass p, the counts summed over the
// HIST_ROWS rows of part (one workgroup per pass, one thread per digit).
__global__ __launch_bounds__(RADIX) void radix_hist_base_kernel(const uint32_t* __restrict__ part,
                                                                int npasses, uint32_t* __restrict__ hist) {
    const int p = blockIdx.x;
    uint32_t c = 0;
#pragma unroll 16
    for (int b = 0; b < HIST_ROWS; ++b) c += part[((int64_t)b * npasses + p) * RADIX + threadIdx.x];
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<uint32_t, RADIX>(c, &tot);
    hist[p * RADIX + threadIdx.x] = ex;
}

template <class K, class V>
__global__ __launch_bounds__(NT) void radix_scatter_lb_kernel(const K* __restrict__ keys_in,
                                                              const V* __restrict__ vals_in, Count cn,
                                                              int shift, uint32_t dmask,
                                                              const uint32_t* __restrict__ gbase,
                                                              uint64_t* status, uint32_t epoch,
                                                              uint32_t* tile_counter, int* flags,
                                                              LookbackCtl ctl,
                                                              K* __restrict__ keys_out,
                                                              V* __restrict__ vals_out) {
    constexpr int NW = NT / WAVE;
    constexpr int IPT = sort_ipt<K>::value;
    constexpr int TILE = sort_tile<K>::value;
    __shared__ K s_k[TILE];
    __shared__ V s_v[TILE];
    __shared__ uint32_t s_wcnt[NW][RADIX];
    __shared__ uint32_t s_start[RADIX];
    __shared__ uint32_t s_run[RADIX];
    __shared__ uint32_t s_goff[RADIX];

    // tile ids in start order from a counter: every tile a tile waits for has started,
    // whatever order the workgroups are dispatched in (also measured faster than tile =
    // workgroup index: 0.70 against 0.755 ms per pass of 7e7 one-word records, same GPU)
    // FLAG_SPEC is raised only by earlier kernels, so every wave reads the same bit here
    if (*flags & FLAG_SPEC) return;
    const int64_t n = count_value(cn);
    const uint32_t tile = next_tile_id(tile_counter);
    if ((int64_t)tile * TILE >= n) return;  // past the records (ids of later tiles are larger)
    const uint64_t tag = (uint64_t)epoch << 32;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int64_t base = (int64_t)tile * TILE;
    const int64_t wbase = base + (int64_t)wid * IPT * WAVE;

    for (int j = threadIdx.x; j < NW * RADIX; j += NT) (&s_wcnt[0][0])[j] = 0;
    __syncthreads();

    K k[IPT];
    V v[IPT];
    uint32_t rank[IPT];
    uint32_t* cnt = s_wcnt[wid];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = wbase + (int64_t)r * WAVE + lane;
        const bool valid = idx < n;
        k[r] = valid ? keys_in[idx] : (K)0;
        v[r] = valid ? vals_in[idx] : (V)0;
    }
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = wbase + (int64_t)r * WAVE + lane;
        const bool valid = idx < n;
        const uint32_t digit = (uint32_t)(k[r] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RADIX_BITS; ++b) {
            const uint64_t bb = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t before = valid ? cnt[digit] : 0u;
        const uint32_t in_wave = (uint32_t)__popcll(peers & lt_mask);
        rank[r] = before + in_wave;
        __builtin_amdgcn_wave_barrier();
        if (valid && in_wave == 0) cnt[digit] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = s_wcnt[w][d];
            s_wcnt[w][d] = run;
            run += c;
        }
        uint32_t tot;
        s_start[d] = block_exclusive_scan<uint32_t, NT>(run, &tot);
        s_run[d] = run;
        if (tile != 0 || !ctl.withhold_first)
            lb_store(&status[(int64_t)tile * RADIX + d], (tile == 0 ? LB_FLAG_PFX : LB_FLAG_AGG) | tag | run);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int64_t idx = wbase + (int64_t)r * WAVE + lane;
        if (idx < n) {
            const uint32_t digit = (uint32_t)(k[r] >> shift) & dmask;
            const uint32_t pos = s_start[digit] + s_wcnt[wid][digit] + rank[r];
            s_k[pos] = k[r];
            s_v[pos] = v[r];
        }
    }
    // look-back for digit d: the count of d in all tiles before this one.  Each round reads
    // LB_WIN predecessors at once (the nearest one seldom holds its inclusive prefix yet) and
    // adds them up to the nearest inclusive prefix.  A word of another epoch or an empty one
    // ends the round and is read again.
    {
        const int d = threadIdx.x;
        uint64_t pre = 0;
        if (tile > 0) {
            int64_t j = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (true) {
                uint64_t w[LB_WIN];
#pragma unroll
                for (int q = 0; q < LB_WIN; ++q) {
                    w[q] = j - q >= 0 ? lb_load(&status[(j - q) * RADIX + d]) : (LB_FLAG_PFX | tag);
                }
                int used = 0;
                bool done = false, empty = false;
#pragma unroll
                for (int q = 0; q < LB_WIN; ++q) {
                    if (done || empty) continue;
                    const uint64_t f = w[q] >> 62;
                    if (f == 0 || (w[q] & RLB_EPOCH_FIELD) != tag) { empty = true; continue; }
                    pre += w[q] & RLB_VALUE;
                    ++used;
                    if (f == 2) done = true;
                }
                if (done) break;
                j -= used;
                if (empty) {
                    if (++spins > ctl.spin_limit) { atomicOr(flags, FLAG_LOOKBACK); break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            lb_store(&status[(int64_t)tile * RADIX + d], LB_FLAG_PFX | tag | (pre + s_run[d]));
        }
        s_goff[d] = gbase[d] + (uint32_t)pre;
    }
    __syncthreads();
    const int64_t cnt_tile = min((int64_t)TILE, n - base);
    for (int i = threadIdx.x; i < cnt_tile; i += NT) {
        const K kk = s_k[i];
        const uint32_t digit = (uint32_t)(kk >> shift) & dmask;
        const int64_t g = (int64_t)s_goff[digit] + (i - (int64_t)s_start[digit]);
        keys_out[g] = kk;
        vals_out[g] = s_v[i];
    }
}

// ---------------------------------------------------------------------------------------
// merge: union of A (sorted spawns, duplicates) and B (vector keys with diagonal values,
// unique).  Merged order: by key, B before A on equal keys.  Each address is one segment;
// the thread owning the segment's first position folds it sequentially:
//   acc = 0 (+ diag) + spawn_1 + spawn_2 + ...
// then applies ThresholdCompression and keeps the survivor.  A tile writes its survivors,
// in merged order, to its own slot range [tile * TILE, tile * TILE + count) of the slot
// buffer and its count to tile_cnt[tile]; compact_kernel then moves them to their final
// positions, so no tile waits on another.  partials[j * gridDim.x + tile], j = 0..2:
// len_before, len, norm.
// ---------------------------------------------------------------------------------------
constexpr int MERGE_EXTRA = 64;  // spawn records staged past the tile for runs that cross it
static_assert(MERGE_IPT <= 32 && merge_ipt<unsigned __int128>::value <= 32, "merge candidate masks are 32-bit");
static_assert(MERGE_EXTRA <= NT, "one extra record per thread at most");

// from_initiator_value of the rule for the components (safe, unsafe, initiator):
// Initiator initiators.jl:138-140, SimpleInitiator :180-182, CoherentInitiator :208-214.
template <class V>
__device__ inline V initiator_value(const StepArgs& a, V s, V u, V i) {
    if (a.init_rule == 2) return s + i;
    if (a.init_rule == 3) return (i != (V)0 || fabs((double)u) > a.init_threshold) ? (i + s) + u : i + s;
    return (s + i) + (i != (V)0 ? u : (V)0);
}

// INIT: the initiator approximation.  Spawn keys carry CLASS_BIT; the vector keys are
// staged with CLASS_BIT set where the old value bo[i] is not above the threshold.  Each
// address then folds three components (safe, unsafe, initiator), each a left fold of its
// deposits in list order; an address exists when a component is nonzero (deposit! deletes
// all-zero entries, pdworkingmemory.jl:21-31) and takes the rule's value, which
// NoCompression keeps even when it is zero (stochasticstyles.jl:116-121).
template <class K, class V, bool INIT>
__global__ __launch_bounds__(NT) void merge_kernel(StepArgs a, const K* __restrict__ ak,
                                                   const V* __restrict__ av, Count cS,
                                                   const K* __restrict__ bk,
                                                   const V* __restrict__ bv, int64_t L,
                                                   K* __restrict__ slot_k, V* __restrict__ slot_v,
                                                   int64_t* __restrict__ tile_cnt,
                                                   double* __restrict__ partials,
                                                   const int64_t* __restrict__ splits,
                                                   const K* __restrict__ prevk,
                                                   const V* __restrict__ bo, const int* flags) {
    constexpr int IPT = merge_ipt<K>::value;
    constexpr int TILE = NT * IPT;
    constexpr K KM = INIT ? (K)~class_bit<K>() : (K)~(K)0;
    // the tile's A range, its B range, then up to MERGE_EXTRA records of A past the tile
    __shared__ K s_key[TILE + MERGE_EXTRA];
    __shared__ V s_val[TILE + MERGE_EXTRA];

    const int64_t tile = blockIdx.x;
    // the flags are raised by earlier kernels only: every wave reads the same bits
    if (*flags & (FLAG_LOOKBACK | FLAG_SPEC | FLAG_FIXUP_FALLBACK)) {  // no sorted list to merge
        if (threadIdx.x == 0) {
            tile_cnt[tile] = 0;
            for (int j = 0; j < 3; ++j) partials[(int64_t)j * gridDim.x + tile] = 0;
        }
        return;
    }
    const int64_t S = count_value(cS);
    const int64_t total = S + L;
    const int64_t d0 = tile * TILE;
    const int64_t d1 = min(d0 + TILE, total);
    const int64_t a0 = splits[tile], a1 = splits[tile + 1];
    const int64_t b0 = d0 - a0, b1 = d1 - a1;
    const int na = (int)(a1 - a0), nb = (int)(b1 - b0);
    const int ne = (int)min((int64_t)MERGE_EXTRA, S - a1);
    const K tile_prev = prevk[tile];
    // stage A, B and the extra records: all loads of a thread are issued before the first
    // LDS store, so a wave waits for one memory round trip instead of one per element
    {
        K rk[IPT];
        V rv[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int i = (int)threadIdx.x + r * NT;
            if (i < na + nb) {
                const bool in_a = i < na;
                const K* kp = in_a ? ak + (a0 + i) : bk + (b0 + (i - na));
                const V* vp = in_a ? av + (a0 + i) : bv + (b0 + (i - na));
                rk[r] = *kp;
                rv[r] = *vp;
                if constexpr (INIT) {
                    if (!in_a && below_initiator_threshold(a, (double)bo[b0 + (i - na)])) rk[r] |= class_bit<K>();
                }
            }
        }
        K ek = 0;
        V ev = (V)0;
        const int x = (int)threadIdx.x;
        if (x < ne) { ek = ak[a1 + x]; ev = av[a1 + x]; }
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int i = (int)threadIdx.x + r * NT;
            if (i < na + nb) { s_key[i] = rk[r]; s_val[i] = rv[r]; }
        }
        if (x < ne) { s_key[na + nb + x] = ek; s_val[na + nb + x] = ev; }
    }
    __syncthreads();
    const K* s_a = s_key;
    const K* s_b = s_key + na;

    // this thread's positions within the tile
    const int dd = (int)threadIdx.x * IPT;
    const int n_here = (int)max((int64_t)0, min((int64_t)IPT, (d1 - d0) - dd));
    int ia = 0;
    {
        int lo = max(0, dd - nb), hi = min(dd, na);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((s_a[mid] & KM) < (s_b[dd - mid - 1] & KM)) lo = mid + 1; else hi = mid;
        }
        ia = lo;
    }
    int ib = dd - ia;
    // key of the merged element just before this thread's first position
    bool have_prev = false;
    K prev = 0;
    if (dd > 0) {
        have_prev = true;
        const K pa = ia > 0 ? (K)(s_a[ia - 1] & KM) : (K)0;
        const K pb = ib > 0 ? (K)(s_b[ib - 1] & KM) : (K)0;
        prev = (ia > 0 && ib > 0) ? max(pa, pb) : (ia > 0 ? pa : pb);
    } else if (d0 > 0) {
        have_prev = true;
        prev = tile_prev;
    }

    // candidates stay in the slot of their position j (static register indices); bit j of
    // cvalid marks a segment head with a nonzero folded value
    K ck[IPT];
    V cv[IPT];
    uint32_t cvalid = 0;
    double lb = 0, ln = 0, nrm = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        ck[j] = 0;
        cv[j] = (V)0;
        if (j < n_here) {
            const bool take_b = ib < nb && (ia >= na || (s_b[ib] & KM) <= (s_a[ia] & KM));
            const K key = (take_b ? s_b[ib] : s_a[ia]) & KM;
            if ((!have_prev || key != prev) && !INIT) {
                // segment head: fold diag (if any) then the run of spawns in A
                V acc = (V)0;
                if (take_b) acc = acc + s_val[na + ib];
                int la = ia;
                while (la < na && s_a[la] == key) { acc = acc + s_val[la]; ++la; }
                if (la == na) {  // the run may continue past the tile: staged records first
                    int x = 0;
                    while (x < ne && s_key[na + nb + x] == key) { acc = acc + s_val[na + nb + x]; ++x; }
                    if (x == ne) {
                        int64_t ga = a1 + ne;
                        while (ga < S && ak[ga] == key) { acc = acc + av[ga]; ++ga; }
                    }
                }
                if (acc != (V)0) {
                    lb += 1;
                    ck[j] = key;
                    cv[j] = acc;
                    cvalid |= 1u << j;
                }
            } else if ((!have_prev || key != prev) && INIT) {
                // the same walk with three components
                V cs = (V)0, cu = (V)0, ci = (V)0;
                auto dep = [&](K rawk, V x) {
                    if (rawk & class_bit<K>()) cu = cu + x; else cs = cs + x;
                };
                if (take_b) {
                    const V d = s_val[na + ib];
                    if (s_b[ib] & class_bit<K>()) cs = cs + d; else ci = ci + d;
                }
                int la = ia;
                while (la < na && (s_a[la] & KM) == key) { dep(s_a[la], s_val[la]); ++la; }
                if (la == na) {
                    int x = 0;
                    while (x < ne && (s_key[na + nb + x] & KM) == key) { dep(s_key[na + nb + x], s_val[na + nb + x]); ++x; }
                    if (x == ne) {
                        for (int64_t ga = a1 + ne; ga < S; ++ga) {
                            const K rk = ak[ga];
                            if ((rk & KM) != key) break;
                            dep(rk, av[ga]);
                        }
                    }
                }
                if (cs != (V)0 || cu != (V)0 || ci != (V)0) {
                    lb += 1;
                    const V val = initiator_value<V>(a, cs, cu, ci);
                    const bool keep_zero = is_int_v<V>::value || a.compression != 1;
                    if (val != (V)0 || keep_zero) {
                        ck[j] = key;
                        cv[j] = val;
                        cvalid |= 1u << j;
                    }
                }
            }
            prev = key;
            have_prev = true;
            if (take_b) ++ib; else ++ia;
        }
    }
    if constexpr (!is_int_v<V>::value) {
        if (a.compression == 1) {
            // ThresholdCompression: the values below the threshold of a wavefront are listed
            // in LDS and drawn with every lane busy (a draw depends only on the address)
            const double th = a.comp_threshold;
            uint32_t need = 0;
#pragma unroll
            for (int j = 0; j < IPT; ++j)
                if (((cvalid >> j) & 1u) && fabs((double)cv[j]) / th < 1) need |= 1u << j;
            const int lane = threadIdx.x & (WAVE - 1);
            const int wid = threadIdx.x / WAVE;
            const int nneed = __popc(need);
            const int incl = wave_inclusive_scan(nneed);
            const int wtot = __shfl(incl, WAVE - 1, WAVE);
            const int mybase = incl - nneed;
            K* wk = s_key + wid * (WAVE * IPT);
            V* wv = s_val + wid * (WAVE * IPT);
            __syncthreads();  // the tile's inputs are no longer read
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if ((need >> j) & 1u) {
                    const int q = mybase + __popc(need & ((1u << j) - 1u));
                    wk[q] = ck[j];
                    wv[q] = cv[j];
                }
            }
            __syncthreads();
            for (int q = lane; q < wtot; q += WAVE) {
                const K key = wk[q];
                const double x = (double)wv[q];
                const U4 r = draw(a.seed, unmix_key(a.mix, key), 0, a.step, TAG_COMPRESS);
                wv[q] = (fabs(x) / th > u53(r.x, r.y)) ? th * sgn(x) : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if ((need >> j) & 1u) {
                    const int q = mybase + __popc(need & ((1u << j) - 1u));
                    cv[j] = wv[q];
                    if (cv[j] == (V)0) cvalid &= ~(1u << j);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if ((cvalid >> j) & 1u) {
            ln += 1;
            nrm += fabs((double)cv[j]);
        }
    }
    const int count = __popc(cvalid);
    int btotal;
    const int ex = block_exclusive_scan<int, NT>(count, &btotal);  // ends with a barrier
    // stage the survivors in LDS (the input keys are no longer read) for coalesced writes
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if ((cvalid >> j) & 1u) {
            const int q = ex + __popc(cvalid & ((1u << j) - 1u));
            s_key[q] = ck[j];
            s_val[q] = cv[j];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < btotal; i += NT) {
        slot_k[d0 + i] = s_key[i];
        slot_v[d0 + i] = s_val[i];
    }
    block_sum3<NT>(lb, ln, nrm);
    if (threadIdx.x == 0) {
        tile_cnt[tile] = btotal;
        partials[(int64_t)0 * gridDim.x + tile] = lb;
        partials[(int64_t)1 * gridDim.x + tile] = ln;
        partials[(int64_t)2 * gridDim.x + tile] = nrm;
    }
}

// compact: tile t's records (merge survivors or spawns; NT * IPT slots per tile) move from
// its slot range to [offs[t], offs[t+1]) (offs is the exclusive scan of the tile counts,
// offs[ntiles] the new length).  If the new length exceeds cap nothing is written and
// FLAG is raised.
template <class K, class V, int IPT, int FLAG>
__global__ __launch_bounds__(NT) void compact_kernel(const K* __restrict__ slot_k,
                                                     const V* __restrict__ slot_v,
                                                     const int64_t* __restrict__ offs, int64_t ntiles,
                                                     K* __restrict__ ok, V* __restrict__ ov,
                                                     int64_t cap, int64_t* total_out, int* flags) {
    constexpr int TILE = NT * IPT;
    const int64_t tile = blockIdx.x;
    const int64_t total = offs[ntiles];
    if (total > cap) {
        if (tile == 0 && threadIdx.x == 0) {
            atomicOr(flags, FLAG);
            *total_out = total;
        }
        return;
    }
    if (tile == 0 && threadIdx.x == 0) *total_out = total;
    const int64_t off = offs[tile];
    const int n = (int)(offs[tile + 1] - off);
    const int64_t src = tile * TILE;
    K k[IPT];
    V v[IPT];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int i = (int)threadIdx.x + r * NT;
        if (i < n) { k[r] = slot_k[src + i]; v[r] = slot_v[src + i]; }
    }
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int i = (int)threadIdx.x + r * NT;
        if (i < n) { ok[off + i] = k[r]; ov[off + i] = v[r]; }
    }
}

// ---------------------------------------------------------------------------------------
// project: terms[i] = coef_i * v[key_i] (0 if absent) by binary search in the new vector.
// ---------------------------------------------------------------------------------------
template <class K, class V>
__global__ __launch_bounds__(NT) void project_kernel(const K* __restrict__ keys,
                                                     const V* __restrict__ vals, const int64_t* len_p,
                                                     const K* __restrict__ pk,
                                                     const double* __restrict__ pc, int64_t np,
                                                     double* __restrict__ terms) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= np) return;
    const int64_t L = *len_p;
    const K key = pk[i];
    int64_t lo = 0, hi = L;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    const double x = (lo < L && keys[lo] == key) ? (double)vals[lo] : 0.0;
    terms[i] = x * pc[i];
}

// ---------------------------------------------------------------------------------------
// Statistics: per-tile partial sums are reduced in two deterministic levels.  Level one
// (grid PARTIAL_BLOCKS x 12) deals each partial array out in rows of NT: row r goes to
// workgroup r mod PARTIAL_BLOCKS, element r*NT+i to its thread i; level two (one workgroup)
// adds the chunk sums in order.  Where each element lands does not depend on the array's
// length, so trailing zero partials (tiles of a grid sized for more work than came) add
// exact zeros and leave every sum bit-identical to the one over the exact tile count.
// out: [0] attempts [1] exact [2] inexact [3] spawns [4] deaths [5] clones [6] zombies
//      [7] len_before [8] len [9] norm [10] vproj [11] hproj
// ---------------------------------------------------------------------------------------
constexpr int NSTAT = 12;
constexpr int PARTIAL_BLOCKS = 64;

struct PartialSet {
    const double* p[NSTAT];
    int64_t n[NSTAT];
};

__global__ __launch_bounds__(NT) void partial_sums_kernel(PartialSet ps, double* __restrict__ chunk_sums) {
    const int j = blockIdx.y;
    const int64_t n = ps.n[j];
    const int64_t stride = (int64_t)gridDim.x * NT;
    double acc = 0;
    for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < n; t += stride) acc += ps.p[j][t];
    const double s = block_sum<NT>(acc);
    if (threadIdx.x == 0) chunk_sums[(int64_t)j * gridDim.x + blockIdx.x] = s;
}

// DoubleLogUpdate (shiftstrategy.jl:174-181) for an engine that keeps its shift in device
// memory: from the step's norm N, S' = S - (xi/dtau ln(N/N_target) + zeta/dtau ln(N/N_prev))
// and N_prev' = N.  in holds (S, N_prev) of this step, out receives (S', N_prev'): two halves
// of one buffer, so that a step repeated from the same vector updates from the same state.
// ln of a quotient <= 0 is -inf, as on the host.
struct ShiftUpdate {
    const double* in;  // null: the host keeps the shift
    double* out;
    double dtau, target, zeta, xi;
};

__device__ inline double log_or_minf(double x) { return x > 0 ? log(x) : -INFINITY; }

__device__ inline double double_log_update(const ShiftUpdate& su, double norm) {
#pragma clang fp contract(off)  // the host's operation order and roundings
    return su.in[0] - (su.xi / su.dtau * log_or_minf(norm / su.target) +
                       su.zeta / su.dtau * log_or_minf(norm / su.in[1]));
}

// out[0..NSTAT-1]: the statistics; with a device shift update also out[NSTAT] = S' and
// out[NSTAT + 1] = N_prev' (the norm is statistic 9).
__global__ __launch_bounds__(64) void finalize_kernel(const double* __restrict__ chunk_sums, int nchunks,
                                                      double* __restrict__ out, ShiftUpdate su) {
    const int j = threadIdx.x;
    if (j >= NSTAT) return;
    double acc = 0;
    for (int b = 0; b < nchunks; ++b) acc += chunk_sums[j * nchunks + b];
    out[j] = acc;
    if (j == 9 && su.in) {
        const double s = double_log_update(su, acc);
        su.out[0] = s;
        su.out[1] = acc;
        out[NSTAT] = s;
        out[NSTAT + 1] = acc;
    }
}

}  // namespace rimu
