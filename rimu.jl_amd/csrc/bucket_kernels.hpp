// bucket_kernels.hpp -- annihilation of a step's spawn records by buckets of the stored key
// (one-word mixed keys).
//
// The stored keys are a bijective mix of the addresses, so their top bits are spread evenly.
// The spawn records are radix-sorted (stable, look-back scatters of step_kernels.hpp) on the
// top T bits only, which puts every address's records, in list order, into bucket
// key >> (B - T) -- and the vector, kept in stored-key order, holds each bucket's entries as
// one contiguous segment.  A persistent grid of workgroups takes the buckets in a fixed
// stride; for each bucket a workgroup
//   1. holds the bucket's vector entries and records in registers (loaded while the previous
//      bucket was being finished), numbers them entries first, then records in list order,
//      stores the values in LDS by that item index and packs every item as (key bits below
//      the bucket bits, item index) -- so that sorting the packed words puts each address's
//      diagonal survivor first and its spawns after it in list order, the order of deposit!
//      over the list (Interfaces/dictvectors.jl:49-51, dvec.jl:129-136),
//   2. sorts the packed words in LDS: LDS atomics drop every word into one of 2048
//      sub-buckets (the 11 key bits below the bucket bits; arrival order inside a sub-bucket
//      is arbitrary), and
//      each word's place inside its sub-bucket is the number of smaller words there -- exact
//      whatever the arrival order was, since the item indices make all words distinct,
//   3. folds each address left to right in one pass over a window of the sorted words,
//      applies ThresholdCompression (compression.jl:18-33) and counts len_before, len and
//      the norm,
//   4. writes its survivors to its own slot range with their count; bucket_compact_kernel
//      places them after a scan of the counts.
// The result is that of the full sort and merge: same addresses, same fold order, same
// draws; with an initiator rule each address folds three components instead of one, as on
// the merge path.  A bucket larger than its LDS arrays raises FLAG_BUCKET_OVERFLOW and writes
// nothing; the host then sorts on every bit and merges along the merge path.
#pragma once

#include "step_kernels.hpp"

namespace rimu {

constexpr int BK_NT = 512;                    // threads of a bucket workgroup
constexpr int BK_CAP = 2560;                  // records + vector entries per bucket in LDS
constexpr int BK_IPT = BK_CAP / BK_NT;        // items per thread
constexpr int BK_IDX_BITS = 12;               // item index below the key bits of a packed word
constexpr int BK_SUB_BITS = 11;               // sub-buckets of the LDS sort
constexpr int BK_NSUB = 1 << BK_SUB_BITS;
constexpr int BK_WG_PER_CU = 2;               // resident bucket workgroups per CU (LDS-bound)
constexpr int BK_SORT_WIN = 2;                // sub-bucket words read at once by the ranking
constexpr int BK_FOLD_EXTRA = 4;              // positions read past a thread's own for its runs
static_assert(BK_CAP <= (1 << BK_IDX_BITS), "item indices fit their bits");
static_assert(BK_CAP % BK_NT == 0 && BK_NSUB % BK_NT == 0, "whole rounds per thread");
static_assert(BK_IPT < 8, "survivor masks are 32-bit; per-thread counts fit three bits");
constexpr int FLAG_BUCKET_OVERFLOW = 16;

// First record and first vector entry of every bucket: rstart[b] = lower bound of b among
// the records' top T bits (records sorted on them), vstart[b] likewise in the vector;
// b = 0 .. 2^T (the last is S, resp. L).
__global__ __launch_bounds__(NT) void bucket_bounds_kernel(const uint64_t* __restrict__ ak, Count cS,
                                                           const uint64_t* __restrict__ bk, int64_t L,
                                                           int shift, int64_t nbuckets,
                                                           int64_t* __restrict__ rstart,
                                                           int64_t* __restrict__ vstart) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i > 2 * nbuckets + 1) return;
    const int64_t S = min(count_value(cS), cS.cap);  // more records raise FLAG_SPEC earlier
    const bool rec = i <= nbuckets;
    const int64_t b = rec ? i : i - nbuckets - 1;
    const uint64_t* k = rec ? ak : bk;
    int64_t lo = 0, hi = rec ? S : L;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)((k[mid] & ~CLASS_BIT) >> shift) < b) lo = mid + 1; else hi = mid;
    }
    (rec ? rstart : vstart)[b] = lo;
}

// LDS stores of this wavefront's lanes are seen by all its lanes after this point (the LDS
// performs one wavefront's operations in order; the fences keep the compiler from moving
// loads or stores across it).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Scans of the bucket workgroup, written with the lane and (uniform) wavefront index the
// caller derives from its per-bucket thread index, so that no lane masks are held in
// registers across buckets.
template <class T>
__device__ inline T bk_wave_inclusive_scan(T v, int lane) {
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const T n = __shfl_up(v, off, WAVE);
        if (lane >= off) v += n;
    }
    return v;
}

// Exclusive prefix of this thread over the workgroup; *total receives the sum.
template <class T>
__device__ inline T bk_block_exclusive_scan(T v, int lane, int wid, T* s_wave, T* total) {
    const T inc = bk_wave_inclusive_scan(v, lane);
    if (lane == WAVE - 1) s_wave[wid] = inc;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BK_NT / WAVE; ++w) {
        const T x = s_wave[w];
        if (w < wid) off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// LDS of one bucket workgroup.  `vin` holds the items' values by item index and `place`
// their packed words sub-bucket by sub-bucket; `sk` receives the words in order.  After the
// fold `place` and `vin` stage the values that ThresholdCompression draws for.
template <class V, bool INIT>
struct BucketLds {
    uint64_t place[BK_CAP];
    uint64_t sk[BK_CAP];
    V vin[BK_CAP];
    uint8_t cls[INIT ? BK_CAP : 1];  // initiator runs: the class of every item (see below)
    uint32_t sub[BK_NSUB + 1];      // sub-bucket counts, then their starts
    double red[2][BK_NT / WAVE];
    uint32_t scan_u[BK_NT / WAVE];
    int scan_i[BK_NT / WAVE];
};

// Bucket b = blockIdx.x, blockIdx.x + gridDim.x, ...; nbk buckets of which rstart/vstart give
// the record and vector ranges.  Outputs per bucket: survivors at slot[rstart[b] + vstart[b]
// + i], their count in bucket_cnt[b], and partials[b], [nbk + b], [2 nbk + b] = len_before,
// len, norm.
// INIT (initiator rules): a record's key carries CLASS_BIT when its parent was not above the
// initiator threshold, a vector entry has the class when its old value bo is not above it;
// each address folds three components in list order -- safe (entries with the class, records
// without), unsafe (records with the class), initiator (entries without) -- exists when one
// is nonzero and takes the rule's value (initiator_value), as merge_kernel<K, V, true> does.
template <class V, bool INIT>
__global__ __launch_bounds__(BK_NT, 2 * BK_WG_PER_CU) void bucket_merge_kernel(
    StepArgs a, const uint64_t* __restrict__ ak, const V* __restrict__ av, const uint64_t* __restrict__ bk,
    const V* __restrict__ bv, const V* __restrict__ bo, const int64_t* __restrict__ rstart, const int64_t* __restrict__ vstart, int64_t nbk,
    int lowbits, uint64_t* __restrict__ slot_k, V* __restrict__ slot_v, int64_t* __restrict__ bucket_cnt,
    double* __restrict__ partials, int* flags) {
    constexpr int NW = BK_NT / WAVE;
    constexpr uint64_t IDX_MASK = (1ull << BK_IDX_BITS) - 1ull;
    __shared__ BucketLds<V, INIT> sm;
    // no sorted list to merge (raised by earlier kernels only: every wave reads the same bits)
    if (*flags & (FLAG_LOOKBACK | FLAG_SPEC)) {
        for (int64_t b = blockIdx.x; b < nbk; b += gridDim.x) {
            if (threadIdx.x == 0) {
                bucket_cnt[b] = 0;
                partials[b] = 0;
                partials[nbk + b] = 0;
                partials[2 * nbk + b] = 0;
            }
        }
        return;
    }
    const uint64_t lowmask = lowbits >= 64 ? ~0ull : (1ull << lowbits) - 1ull;
    // sub-bucket digit: the top BK_SUB_BITS key bits below the bucket bits (all of them when
    // fewer remain)
    const int dsh = BK_IDX_BITS + (lowbits > BK_SUB_BITS ? lowbits - BK_SUB_BITS : 0);
    const uint32_t dmask = lowbits >= BK_SUB_BITS ? (uint32_t)(BK_NSUB - 1) : (1u << lowbits) - 1u;
    for (int j = threadIdx.x; j < BK_NSUB + 1; j += BK_NT) sm.sub[j] = 0;

    // items of the first bucket; later ones are loaded during the bucket before them.  Item
    // i < nb is vector entry v0 + i, item nb + r is record r0 + r: the item index orders each
    // address's items as the fold takes them (diagonal value first, spawns in list order).
    int64_t b = blockIdx.x;
    uint64_t ck[BK_IPT];
    V cv[BK_IPT];
    uint32_t ccls = 0;  // INIT: bit j = class of item slot j
    int64_t r0 = 0, v0 = 0;
    int na = 0, nb = 0;
    auto load = [&](int t, int64_t bb, uint64_t* qk, V* qv, uint32_t& q_cls, int64_t& q_r0, int64_t& q_v0,
                    int& q_na, int& q_nb) {
        q_r0 = rstart[bb];
        q_v0 = vstart[bb];
        q_na = (int)(rstart[bb + 1] - q_r0);
        q_nb = (int)(vstart[bb + 1] - q_v0);
        const int q_n = q_na + q_nb;
        if (q_n > BK_CAP) return;  // overflow: nothing is read
        const uint64_t* ek_b = bk + q_v0;  // uniform bases, 32-bit offsets
        const V* ev_b = bv + q_v0;
        const uint64_t* rk_b = ak + q_r0;
        const V* rv_b = av + q_r0;
        q_cls = 0;
#pragma unroll
        for (int j = 0; j < BK_IPT; ++j) {
            const uint32_t i = (uint32_t)t + (uint32_t)(j * BK_NT);
            if (i < (uint32_t)q_nb) {
                qk[j] = ek_b[i];
                qv[j] = ev_b[i];
                if constexpr (INIT) {
                    if (below_initiator_threshold(a, (double)bo[q_v0 + i])) q_cls |= 1u << j;
                }
            }
            if (i >= (uint32_t)q_nb && i < (uint32_t)q_n) {
                qk[j] = rk_b[i - (uint32_t)q_nb];
                qv[j] = rv_b[i - (uint32_t)q_nb];
                if constexpr (INIT) {
                    if (qk[j] & CLASS_BIT) q_cls |= 1u << j;
                }
            }
        }
    };
    if (b < nbk) load((int)threadIdx.x, b, ck, cv, ccls, r0, v0, na, nb);
    __syncthreads();  // sub-bucket counters cleared

    for (; b < nbk; b += gridDim.x) {
        // the thread index, re-read every bucket: values derived from it stay inside the loop
        // instead of being kept in registers across buckets
        int tid;
        asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
        const int lane = tid & (WAVE - 1);
        const int wid = __builtin_amdgcn_readfirstlane(tid / WAVE);
        uint64_t nk[BK_IPT];
        V nv[BK_IPT];
        uint32_t ncls = 0;
        int64_t n_r0 = 0, n_v0 = 0;
        int n_na = 0, n_nb = 0;
        const int64_t bn = b + gridDim.x;

        const int n = na + nb;
        const bool fits = n <= BK_CAP;  // uniform over the workgroup
        uint64_t okey[BK_IPT];
        V oval[BK_IPT];
        uint32_t cvalid = 0;
        int n_before = 0;  // len_before: folded values that are not zero
        if (fits) {
            const uint64_t bucket_key = lowbits >= 64 ? 0ull : (uint64_t)b << lowbits;
            // ---- 1. pack (key bits below the bucket bits, item index), values by item
            // index, drop every word into its sub-bucket (arrival order is arbitrary) ---------
            uint64_t w[BK_IPT];
            uint32_t dg[BK_IPT], slot[BK_IPT];
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j) {
                const int i = tid + j * BK_NT;
                if (i < n) {
                    w[j] = ((ck[j] & lowmask) << BK_IDX_BITS) | (uint64_t)i;
                    sm.vin[i] = cv[j];
                    if constexpr (INIT) sm.cls[i] = (uint8_t)((ccls >> j) & 1u);
                    dg[j] = (uint32_t)(w[j] >> dsh) & dmask;
                    slot[j] = atomicAdd(&sm.sub[dg[j]], 1u);
                }
            }
            __syncthreads();
            // ---- 2. sub-bucket starts (exclusive scan of the counts; sub[BK_NSUB] = n) and
            // placement; every item keeps the bounds of its sub-bucket -----------------------
            {
                constexpr int SPT = BK_NSUB / BK_NT;
                uint32_t x[SPT];
                uint32_t sum = 0;
#pragma unroll
                for (int q = 0; q < SPT; ++q) { x[q] = sm.sub[tid * SPT + q]; sum += x[q]; }
                uint32_t tot;
                uint32_t run = bk_block_exclusive_scan<uint32_t>(sum, lane, wid, sm.scan_u, &tot);
#pragma unroll
                for (int q = 0; q < SPT; ++q) { sm.sub[tid * SPT + q] = run; run += x[q]; }
                if (tid == 0) sm.sub[BK_NSUB] = tot;
            }
            __syncthreads();
            int s0[BK_IPT], s1[BK_IPT];
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j) {
                const int i = tid + j * BK_NT;
                if (i < n) {
                    s0[j] = (int)sm.sub[dg[j]];
                    s1[j] = (int)sm.sub[dg[j] + 1];
                    sm.place[s0[j] + slot[j]] = w[j];
                }
            }
            __syncthreads();
            // counters of the next bucket start at zero (their last reader was the placement)
            for (int j = tid; j < BK_NSUB + 1; j += BK_NT) sm.sub[j] = 0;
            // ---- 3. an item's place in its sub-bucket is the number of smaller words there
            // (the words are distinct: item indices), read BK_SORT_WIN words at a time (wider
            // windows measured no faster) --------------------------------------------------
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j) {
                const int i = tid + j * BK_NT;
                if (i < n) {
                    int less = 0;
                    for (int c = s0[j]; c < s1[j]; c += BK_SORT_WIN) {
                        uint64_t q[BK_SORT_WIN];
#pragma unroll
                        for (int k = 0; k < BK_SORT_WIN; ++k) q[k] = c + k < s1[j] ? sm.place[c + k] : ~0ull;
#pragma unroll
                        for (int k = 0; k < BK_SORT_WIN; ++k) less += q[k] < w[j] ? 1 : 0;
                    }
                    sm.sk[s0[j] + less] = w[j];
                }
                __builtin_amdgcn_sched_barrier(0);  // one item's window in registers at a time
            }
            __syncthreads();

            // ---- 4. fold: thread t owns the runs of equal keys whose first position lies in
            // [t * BK_IPT, (t + 1) * BK_IPT).  It reads those positions and BK_FOLD_EXTRA more
            // at once and adds each run left to right -- (0 + the diagonal value, if any) +
            // the spawns in list order -- in one pass over the window; a run's result is taken
            // at its last position k, into slot min(k, BK_IPT - 1) (only one owned run can end
            // at or after BK_IPT - 1), so the slots keep the runs in key order.  A run going
            // past the window reads on from LDS.
            const int p0 = tid * BK_IPT;
            constexpr int NWIN = BK_IPT + BK_FOLD_EXTRA;
            constexpr uint64_t NOKEY = ~0ull;  // no position: differs from every key (< 2^52)
            uint64_t kk[NWIN + 1];
            V vv[NWIN];
            uint32_t emask = 0, cmask = 0;  // INIT: window positions holding a vector entry / the class
#pragma unroll
            for (int k = 0; k <= NWIN; ++k) {
                const int p = p0 + k;
                const uint64_t wd = p < n ? sm.sk[p] : NOKEY;
                kk[k] = p < n ? wd >> BK_IDX_BITS : NOKEY;
                if (k < NWIN) {
                    const int idx = (int)(wd & IDX_MASK);
                    vv[k] = p < n ? sm.vin[idx] : (V)0;
                    if constexpr (INIT) {
                        if (p < n) {
                            emask |= (uint32_t)(idx < nb) << k;
                            cmask |= (uint32_t)sm.cls[idx] << k;
                        }
                    }
                }
            }
            const uint64_t kprev = p0 > 0 && p0 < n ? sm.sk[p0 - 1] >> BK_IDX_BITS : NOKEY;
            if constexpr (!INIT) {
                bool own = false;
                V acc = (V)0;
#pragma unroll
                for (int k = 0; k < NWIN; ++k) {
                    const bool valid = p0 + k < n;
                    const bool start = valid && kk[k] != (k > 0 ? kk[k - 1] : kprev);
                    if (start) own = k < BK_IPT;
                    acc = (start ? (V)0 : acc) + vv[k];
                    const bool end = valid && kk[k + 1] != kk[k];
                    const int slot_k = k < BK_IPT - 1 ? k : BK_IPT - 1;
                    if (own && end) {
                        okey[slot_k] = bucket_key | kk[k];
                        oval[slot_k] = acc;
                        if (acc != (V)0) cvalid |= 1u << slot_k;
                        own = false;
                    }
                }
                if (own) {  // the owned run goes on past the window: it ends in slot BK_IPT - 1
                    const uint64_t key = kk[NWIN - 1];
                    for (int q = p0 + NWIN; q < n; ++q) {
                        const uint64_t wd = sm.sk[q];
                        if ((wd >> BK_IDX_BITS) != key) break;
                        acc = acc + sm.vin[(int)(wd & IDX_MASK)];
                    }
                    okey[BK_IPT - 1] = bucket_key | key;
                    oval[BK_IPT - 1] = acc;
                    if (acc != (V)0) cvalid |= 1u << (BK_IPT - 1);
                }
                n_before = __popc(cvalid);
            } else {
                // three components per address (safe, unsafe, initiator), each its own left fold
                V cs = (V)0, cu = (V)0, ci = (V)0;
                bool own = false;
                uint32_t cpres = 0;  // addresses with a nonzero component (len_before)
                const bool keep_zero = is_int_v<V>::value || a.compression != 1;
                // entries with the class and records without go to safe, records with it to
                // unsafe, entries without it to initiator (selects: no divergent branches)
                auto dep = [&](bool entry, bool c, V x) {
                    const bool to_s = entry == c, to_u = !entry && c, to_i = entry && !c;
                    cs = to_s ? cs + x : cs;
                    cu = to_u ? cu + x : cu;
                    ci = to_i ? ci + x : ci;
                };
                auto close_run = [&](int slot_k, uint64_t key) {
                    if (cs != (V)0 || cu != (V)0 || ci != (V)0) {
                        cpres |= 1u << slot_k;
                        const V val = initiator_value<V>(a, cs, cu, ci);
                        if (val != (V)0 || keep_zero) {
                            okey[slot_k] = bucket_key | key;
                            oval[slot_k] = val;
                            cvalid |= 1u << slot_k;
                        }
                    }
                };
#pragma unroll
                for (int k = 0; k < NWIN; ++k) {
                    const int p = p0 + k;
                    const bool valid = p < n;
                    const bool start = valid && kk[k] != (k > 0 ? kk[k - 1] : kprev);
                    if (start) {
                        own = k < BK_IPT;
                        cs = cu = ci = (V)0;
                    }
                    if (valid && own) dep((emask >> k) & 1u, (cmask >> k) & 1u, vv[k]);
                    const bool end = valid && kk[k + 1] != kk[k];
                    if (own && end) {
                        close_run(k < BK_IPT - 1 ? k : BK_IPT - 1, kk[k]);
                        own = false;
                    }
                }
                if (own) {  // the owned run goes on past the window: it ends in slot BK_IPT - 1
                    const uint64_t key = kk[NWIN - 1];
                    for (int q = p0 + NWIN; q < n; ++q) {
                        const uint64_t wd = sm.sk[q];
                        if ((wd >> BK_IDX_BITS) != key) break;
                        const int idx = (int)(wd & IDX_MASK);
                        dep(idx < nb, sm.cls[idx] != 0, sm.vin[idx]);
                    }
                    close_run(BK_IPT - 1, key);
                }
                n_before = __popc(cpres);
            }
            __syncthreads();  // the sorted words and the values are no longer read
        }
        // the next bucket's items travel while this one is compressed and written out
        if (bn < nbk) load(tid, bn, nk, nv, ncls, n_r0, n_v0, n_na, n_nb);
        if (!fits) {
            if (tid == 0) {
                atomicOr(flags, FLAG_BUCKET_OVERFLOW);
                bucket_cnt[b] = 0;
                partials[b] = 0;
                partials[nbk + b] = 0;
                partials[2 * nbk + b] = 0;
            }
        } else {
            if constexpr (!is_int_v<V>::value) {
                if (a.compression == 1) {
                    // ThresholdCompression: each wavefront lists its values below the threshold
                    // and draws them with every lane busy (a draw depends only on the address)
                    const double th = a.comp_threshold;
                    uint32_t need = 0;
#pragma unroll
                    for (int j = 0; j < BK_IPT; ++j)
                        if (((cvalid >> j) & 1u) && fabs((double)oval[j]) / th < 1) need |= 1u << j;
                    const int nneed = __popc(need);
                    const int incl = bk_wave_inclusive_scan(nneed, lane);
                    const int wtot = __shfl(incl, WAVE - 1, WAVE);
                    const int mybase = incl - nneed;
                    uint64_t* wk = sm.place + wid * (WAVE * BK_IPT);
                    V* wv = sm.vin + wid * (WAVE * BK_IPT);
#pragma unroll
                    for (int j = 0; j < BK_IPT; ++j) {
                        if ((need >> j) & 1u) {
                            const int q = mybase + __popc(need & ((1u << j) - 1u));
                            wk[q] = okey[j];
                            wv[q] = oval[j];
                        }
                    }
                    wave_lds_sync();  // the staging area is this wavefront's own
                    // the seed re-read here: its round keys are not held in registers across buckets
                    uint64_t seed;
                    asm volatile("s_mov_b64 %0, %1" : "=s"(seed) : "s"(a.seed));
                    for (int q = lane; q < wtot; q += WAVE) {
                        const uint64_t key = wk[q];
                        const double x = (double)wv[q];
                        const U4 r = draw(seed, unmix_key(a.mix, key), 0, a.step, TAG_COMPRESS);
                        wv[q] = (fabs(x) / th > u53(r.x, r.y)) ? th * sgn(x) : 0.0;
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int j = 0; j < BK_IPT; ++j) {
                        if ((need >> j) & 1u) {
                            const int q = mybase + __popc(need & ((1u << j) - 1u));
                            oval[j] = wv[q];
                            if (oval[j] == (V)0) cvalid &= ~(1u << j);
                        }
                    }
                }
            }
            double nrm = 0;
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j)
                if ((cvalid >> j) & 1u) nrm += fabs((double)oval[j]);
            // ---- 5. survivors in order: each thread writes its own after those of the threads
            // before it (a wavefront's writes cover one contiguous range) ---------------------
            const int count = __popc(cvalid);
            int btotal;
            const int ex = bk_block_exclusive_scan<int>(count, lane, wid, sm.scan_i, &btotal);
            const int64_t base = r0 + v0 + ex;  // the bucket's slot range: its merged positions
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j) {
                if ((cvalid >> j) & 1u) {
                    const int q = __popc(cvalid & ((1u << j) - 1u));
                    slot_k[base + q] = okey[j];
                    slot_v[base + q] = oval[j];
                }
            }
            // per-bucket sums: counts from ballots (n_before <= BK_IPT < 8: three bits), the
            // norm in a fixed order (wavefront trees, then the wavefronts in order)
            const int nb_w = __popcll(__ballot(n_before & 1)) + 2 * __popcll(__ballot(n_before & 2)) +
                             4 * __popcll(__ballot(n_before & 4));
#pragma unroll
            for (int off = WAVE / 2; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off, WAVE);
            if (lane == 0) {
                sm.red[0][wid] = (double)nb_w;
                sm.red[1][wid] = nrm;
            }
            __syncthreads();
            if (tid == 0) {
                double s0 = 0, s1 = 0;
                for (int q = 0; q < NW; ++q) { s0 += sm.red[0][q]; s1 += sm.red[1][q]; }
                bucket_cnt[b] = btotal;
                partials[b] = s0;
                partials[nbk + b] = (double)btotal;
                partials[2 * nbk + b] = s1;
            }
            __syncthreads();  // sums read before the next bucket writes LDS
        }
#pragma unroll
        for (int j = 0; j < BK_IPT; ++j) { ck[j] = nk[j]; cv[j] = nv[j]; }
        ccls = ncls;
        r0 = n_r0;
        v0 = n_v0;
        na = n_na;
        nb = n_nb;
    }
}

// Survivors of bucket b move from its slot range [rstart[b] + vstart[b], ...) to
// [offs[b], offs[b + 1]) (offs: exclusive scan of the bucket counts; offs[nbuckets] is the new
// length).  A new length above cap writes nothing and raises FLAG_VECTOR_OVERFLOW.
template <class V>
__global__ __launch_bounds__(NT) void bucket_compact_kernel(const uint64_t* __restrict__ slot_k,
                                                            const V* __restrict__ slot_v,
                                                            const int64_t* __restrict__ rstart,
                                                            const int64_t* __restrict__ vstart,
                                                            const int64_t* __restrict__ offs, int64_t nbuckets,
                                                            uint64_t* __restrict__ ok, V* __restrict__ ov,
                                                            int64_t cap, int64_t* total_out, int* flags) {
    const int64_t b = blockIdx.x;
    const int64_t total = offs[nbuckets];
    if (total > cap) {
        if (b == 0 && threadIdx.x == 0) {
            atomicOr(flags, FLAG_VECTOR_OVERFLOW);
            *total_out = total;
        }
        return;
    }
    if (b == 0 && threadIdx.x == 0) *total_out = total;
    const int64_t off = offs[b];
    const int n = (int)(offs[b + 1] - off);
    const int64_t src = rstart[b] + vstart[b];
    constexpr int CPT = (BK_CAP + NT - 1) / NT;
    uint64_t k[CPT];
    V v[CPT];
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
        const int i = (int)threadIdx.x + r * NT;
        if (i < n) { k[r] = slot_k[src + i]; v[r] = slot_v[src + i]; }
    }
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
        const int i = (int)threadIdx.x + r * NT;
        if (i < n) { ok[off + i] = k[r]; ov[off + i] = v[r]; }
    }
}

}  // namespace rimu
