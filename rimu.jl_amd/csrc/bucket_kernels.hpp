// bucket_kernels.hpp -- annihilation of a step's spawn records by buckets of the stored key
// (one-word mixed keys).
//
// The stored keys are a bijective mix of the addresses, so their top bits are spread evenly.
// The spawn records are radix-sorted (stable, look-back scatters of step_kernels.hpp) on the
// top T bits only, which puts every address's records, in list order, into bucket
// key >> (B - T) -- and the vector, kept in stored-key order, holds each bucket's entries as
// one contiguous segment.  One workgroup per bucket then
//   1. loads the bucket's records and the vector segment (keys, diagonal survivors),
//   2. orders the records by key in LDS, equal keys in list order: one stable radix pass on the
//      8 key bits below the bucket bits (ranks from wavefront ballots, as the global scatter
//      passes do) makes 256 sub-buckets of a few records each, and every record then counts
//      the smaller (key, list index) pairs of its sub-bucket,
//   3. merges them with the segment along the merge path and folds every address left to
//      right -- diagonal value first, then its spawns in list order, the order of deposit!
//      over the list (Interfaces/dictvectors.jl:49-51, dvec.jl:129-136) -- applies
//      ThresholdCompression (compression.jl:18-33) and counts len_before, len and the norm,
//   4. writes its survivors to its own slot range with their count; bucket_compact_kernel
//      places them after a scan of the counts.
// The result is that of the full sort and merge: same addresses, same fold order, same
// draws.  A bucket larger than its LDS arrays raises FLAG_BUCKET_OVERFLOW and writes
// nothing; the host then sorts on every bit and merges along the merge path.
#pragma once

#include "step_kernels.hpp"

namespace rimu {

constexpr int BK_NT = 512;                    // threads of a bucket workgroup
constexpr int BK_CAP_A = 1536;                // spawn records per bucket in LDS
constexpr int BK_CAP_B = 1024;                // vector entries per bucket in LDS
constexpr int BK_RPT = BK_CAP_A / BK_NT;      // records per thread in the LDS sort
constexpr int BK_IPT = (BK_CAP_A + BK_CAP_B) / BK_NT;  // merged positions per thread
constexpr int BK_IDX_BITS = 11;               // list index packed below the key bits
static_assert(BK_CAP_A <= (1 << BK_IDX_BITS), "list indices fit their bits");
static_assert(BK_CAP_A % BK_NT == 0 && BK_CAP_B % BK_NT == 0 && (BK_CAP_A + BK_CAP_B) % BK_NT == 0,
              "whole rounds per thread");
static_assert(BK_IPT <= 32, "survivor masks are 32-bit");
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

// LDS of one bucket workgroup.  Records are sorted as packed words
// (key bits below the bucket bits) << BK_IDX_BITS | list index, so that equal keys stay in
// list order; their values stay in list order in rv.  After the merge walk, rk stages keys
// and rv..kb stage values (contiguous: 2 * BK_CAP_A keys, BK_CAP_A + BK_CAP_B values).
template <class V>
struct BucketLds {
    uint64_t rk[2][BK_CAP_A];
    V rv[BK_CAP_A];
    V vb[BK_CAP_B];                 // diagonal survivors of the bucket's vector entries
    uint64_t kb[BK_CAP_B];          // the entries' keys (ascending)
    uint32_t wcnt[BK_NT / WAVE][RADIX];
    uint32_t start[RADIX + 1];
    double red[3][BK_NT / WAVE];
};
static_assert(2 * BK_CAP_A >= BK_CAP_A + BK_CAP_B, "staged keys fit the sort arrays");

template <class V>
__global__ __launch_bounds__(BK_NT) void bucket_merge_kernel(StepArgs a, const uint64_t* __restrict__ ak,
                                                             const V* __restrict__ av,
                                                             const uint64_t* __restrict__ bk,
                                                             const V* __restrict__ bv,
                                                             const int64_t* __restrict__ rstart,
                                                             const int64_t* __restrict__ vstart,
                                                             int lowbits,
                                                             uint64_t* __restrict__ slot_k,
                                                             V* __restrict__ slot_v,
                                                             int64_t* __restrict__ bucket_cnt,
                                                             double* __restrict__ partials,
                                                             int* flags) {
    constexpr int NW = BK_NT / WAVE;
    __shared__ BucketLds<V> sm;
    const int64_t b = blockIdx.x;
    const int64_t nbk = gridDim.x;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = threadIdx.x / WAVE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int64_t r0 = rstart[b], v0 = vstart[b];
    const int na = (int)(rstart[b + 1] - r0), nb = (int)(vstart[b + 1] - v0);
    // no sorted list to merge (raised by earlier kernels only: every wave reads the same bits)
    const bool failed_sort = (*flags & (FLAG_LOOKBACK | FLAG_SPEC)) != 0;
    if (failed_sort || na > BK_CAP_A || nb > BK_CAP_B) {     // uniform over the workgroup
        if (threadIdx.x == 0) {
            if (!failed_sort) atomicOr(flags, FLAG_BUCKET_OVERFLOW);
            bucket_cnt[b] = 0;
            partials[b] = 0;
            partials[nbk + b] = 0;
            partials[2 * nbk + b] = 0;
        }
        return;
    }
    const uint64_t lowmask = lowbits >= 64 ? ~0ull : (1ull << lowbits) - 1ull;
    const uint64_t bucket_key = lowbits >= 64 ? 0ull : (uint64_t)b << lowbits;
    // ---- load: records (packed) and values, list order; the vector segment --------------
    {
        uint64_t rk[BK_RPT];
        V rv[BK_RPT];
        uint64_t qk[BK_CAP_B / BK_NT];
        V qv[BK_CAP_B / BK_NT];
#pragma unroll
        for (int r = 0; r < BK_RPT; ++r) {
            const int i = (int)threadIdx.x + r * BK_NT;
            if (i < na) { rk[r] = ak[r0 + i]; rv[r] = av[r0 + i]; }
        }
#pragma unroll
        for (int r = 0; r < BK_CAP_B / BK_NT; ++r) {
            const int i = (int)threadIdx.x + r * BK_NT;
            if (i < nb) { qk[r] = bk[v0 + i]; qv[r] = bv[v0 + i]; }
        }
#pragma unroll
        for (int r = 0; r < BK_RPT; ++r) {
            const int i = (int)threadIdx.x + r * BK_NT;
            if (i < na) {
                sm.rk[0][i] = ((rk[r] & lowmask) << BK_IDX_BITS) | (uint64_t)i;
                sm.rv[i] = rv[r];
            }
        }
#pragma unroll
        for (int r = 0; r < BK_CAP_B / BK_NT; ++r) {
            const int i = (int)threadIdx.x + r * BK_NT;
            if (i < nb) { sm.kb[i] = qk[r]; sm.vb[i] = qv[r]; }
        }
    }
    __syncthreads();

    // ---- order the packed records by key bits, in LDS: one stable radix pass on the top 8
    // key bits below the bucket bits (wavefront w ranks the contiguous records
    // [w * BK_RPT * 64, (w + 1) * BK_RPT * 64) with its own digit counters, round r taking
    // record w * BK_RPT * 64 + r * 64 + lane), which splits the bucket into 256 sub-buckets
    // of a few records each; then every record finds its place inside its sub-bucket by
    // counting the smaller packed words there (equal keys: smaller list index first).
    const int wbase = wid * BK_RPT * WAVE;
    const int dshift = lowbits > RADIX_BITS ? lowbits - RADIX_BITS : 0;
    const int sh = dshift + BK_IDX_BITS;
    const uint32_t dmask = lowbits >= RADIX_BITS ? (uint32_t)(RADIX - 1) : (1u << lowbits) - 1u;
    uint32_t dig[BK_RPT];
    uint64_t w[BK_RPT];
    {
        for (int j = threadIdx.x; j < NW * RADIX; j += BK_NT) (&sm.wcnt[0][0])[j] = 0;
        __syncthreads();
        uint32_t* cnt = sm.wcnt[wid];
        uint32_t rank[BK_RPT];
#pragma unroll
        for (int r = 0; r < BK_RPT; ++r) {
            const int idx = wbase + r * WAVE + lane;
            w[r] = idx < na ? sm.rk[0][idx] : 0ull;
        }
#pragma unroll
        for (int r = 0; r < BK_RPT; ++r) {
            const int idx = wbase + r * WAVE + lane;
            const bool valid = idx < na;
            const uint32_t digit = valid ? (uint32_t)(w[r] >> sh) & dmask : 0u;
            dig[r] = digit;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < RADIX_BITS; ++bit) {
                const uint64_t bb = __ballot((digit >> bit) & 1u);
                peers &= ((digit >> bit) & 1u) ? bb : ~bb;
            }
            const uint32_t before = valid ? cnt[digit] : 0u;
            const uint32_t in_wave = (uint32_t)__popcll(peers & lt_mask);
            rank[r] = before + in_wave;
            __builtin_amdgcn_wave_barrier();
            if (valid && in_wave == 0) cnt[digit] = before + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        if (threadIdx.x < RADIX) {
            const int d = threadIdx.x;
            uint32_t run = 0;
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const uint32_t c = sm.wcnt[q][d];
                sm.wcnt[q][d] = run;
                run += c;
            }
            sm.start[d] = run;
        }
        __syncthreads();
        if (threadIdx.x < WAVE) {  // exclusive scan of the 256 digit totals by one wavefront
            uint32_t x[RADIX / WAVE];
            uint32_t sum = 0;
#pragma unroll
            for (int q = 0; q < RADIX / WAVE; ++q) { x[q] = sm.start[lane * (RADIX / WAVE) + q]; sum += x[q]; }
            const uint32_t incl = wave_inclusive_scan(sum);
            uint32_t run = incl - sum;
#pragma unroll
            for (int q = 0; q < RADIX / WAVE; ++q) { sm.start[lane * (RADIX / WAVE) + q] = run; run += x[q]; }
            if (lane == WAVE - 1) sm.start[RADIX] = run;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < BK_RPT; ++r) {
            const int idx = wbase + r * WAVE + lane;
            if (idx < na) sm.rk[1][sm.start[dig[r]] + sm.wcnt[wid][dig[r]] + rank[r]] = w[r];
        }
        __syncthreads();
    }
    if (dshift > 0) {  // bits below the sub-bucket digit: rank inside the sub-bucket
#pragma unroll
        for (int r = 0; r < BK_RPT; ++r) {
            const int idx = wbase + r * WAVE + lane;
            if (idx < na) {
                const int s0 = (int)sm.start[dig[r]], s1 = (int)sm.start[dig[r] + 1];
                int less = 0;
                for (int q = s0; q < s1; ++q) less += sm.rk[1][q] < w[r];
                sm.rk[0][s0 + less] = w[r];
            }
        }
        __syncthreads();
    }
    const int cur = dshift > 0 ? 0 : 1;
    const uint64_t* s_a = sm.rk[cur];
    auto akey = [&](int i) { return bucket_key | (s_a[i] >> BK_IDX_BITS); };
    auto aval = [&](int i) { return sm.rv[(int)(s_a[i] & ((1u << BK_IDX_BITS) - 1u))]; };

    // ---- merge path inside the bucket, a left fold per address ---------------------------
    const int total = na + nb;
    const int dd = (int)threadIdx.x * BK_IPT;
    const int n_here = max(0, min(BK_IPT, total - dd));
    int ia = 0;
    {
        int lo = max(0, dd - nb), hi = min(dd, na);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (akey(mid) < sm.kb[dd - mid - 1]) lo = mid + 1; else hi = mid;
        }
        ia = lo;
    }
    int ib = dd - ia;
    bool have_prev = false;
    uint64_t prev = 0;
    if (dd > 0 && dd < total) {  // keys of other buckets differ from this bucket's
        have_prev = true;
        const uint64_t pa = ia > 0 ? akey(ia - 1) : 0;
        const uint64_t pb = ib > 0 ? sm.kb[ib - 1] : 0;
        prev = (ia > 0 && ib > 0) ? max(pa, pb) : (ia > 0 ? pa : pb);
    }
    uint64_t ck[BK_IPT];
    V cv[BK_IPT];
    uint32_t cvalid = 0;
    double lb = 0, ln = 0, nrm = 0;
#pragma unroll
    for (int j = 0; j < BK_IPT; ++j) {
        ck[j] = 0;
        cv[j] = (V)0;
        if (j < n_here) {
            const uint64_t ka_ = ia < na ? akey(ia) : ~0ull;
            const bool take_b = ib < nb && (ia >= na || sm.kb[ib] <= ka_);
            const uint64_t key = take_b ? sm.kb[ib] : ka_;
            if (!have_prev || key != prev) {
                // segment head: the diagonal value (if any), then the spawns in list order
                V acc = (V)0;
                if (take_b) acc = acc + sm.vb[ib];
                for (int la = ia; la < na && akey(la) == key; ++la) acc = acc + aval(la);
                if (acc != (V)0) {
                    lb += 1;
                    ck[j] = key;
                    cv[j] = acc;
                    cvalid |= 1u << j;
                }
            }
            prev = key;
            have_prev = true;
            if (take_b) ++ib; else ++ia;
        }
    }
    __syncthreads();  // the sorted records and the segment are no longer read
    uint64_t* stage_k = &sm.rk[0][0];                 // 2 * BK_CAP_A keys
    V* stage_v = &sm.rv[0];                          // rv, vb (and kb): BK_CAP_A + BK_CAP_B values
    if constexpr (!is_int_v<V>::value) {
        if (a.compression == 1) {
            // ThresholdCompression: each wavefront lists its values below the threshold and
            // draws them with every lane busy (a draw depends only on the address)
            const double th = a.comp_threshold;
            uint32_t need = 0;
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j)
                if (((cvalid >> j) & 1u) && fabs((double)cv[j]) / th < 1) need |= 1u << j;
            const int nneed = __popc(need);
            const int incl = wave_inclusive_scan(nneed);
            const int wtot = __shfl(incl, WAVE - 1, WAVE);
            const int mybase = incl - nneed;
            uint64_t* wk = stage_k + wid * (WAVE * BK_IPT);
            V* wv = stage_v + wid * (WAVE * BK_IPT);
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j) {
                if ((need >> j) & 1u) {
                    const int q = mybase + __popc(need & ((1u << j) - 1u));
                    wk[q] = ck[j];
                    wv[q] = cv[j];
                }
            }
            __syncthreads();
            for (int q = lane; q < wtot; q += WAVE) {
                const uint64_t key = wk[q];
                const double x = (double)wv[q];
                const U4 r = draw(a.seed, unmix_key(a.mix, key), 0, a.step, TAG_COMPRESS);
                wv[q] = (fabs(x) / th > u53(r.x, r.y)) ? th * sgn(x) : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < BK_IPT; ++j) {
                if ((need >> j) & 1u) {
                    const int q = mybase + __popc(need & ((1u << j) - 1u));
                    cv[j] = wv[q];
                    if (cv[j] == (V)0) cvalid &= ~(1u << j);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < BK_IPT; ++j) {
        if ((cvalid >> j) & 1u) {
            ln += 1;
            nrm += fabs((double)cv[j]);
        }
    }
    // survivors in merged order, staged for coalesced writes
    const int count = __popc(cvalid);
    int btotal;
    const int ex = block_exclusive_scan<int, BK_NT>(count, &btotal);  // ends with a barrier
#pragma unroll
    for (int j = 0; j < BK_IPT; ++j) {
        if ((cvalid >> j) & 1u) {
            const int q = ex + __popc(cvalid & ((1u << j) - 1u));
            stage_k[q] = ck[j];
            stage_v[q] = cv[j];
        }
    }
    // per-bucket sums (fixed order: wavefront trees, then the wavefronts in order)
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        lb += __shfl_xor(lb, off, WAVE);
        ln += __shfl_xor(ln, off, WAVE);
        nrm += __shfl_xor(nrm, off, WAVE);
    }
    if (lane == 0) {
        sm.red[0][wid] = lb;
        sm.red[1][wid] = ln;
        sm.red[2][wid] = nrm;
    }
    __syncthreads();
    const int64_t base = r0 + v0;  // the bucket's slot range: its merged positions
    for (int i = threadIdx.x; i < btotal; i += BK_NT) {
        slot_k[base + i] = stage_k[i];
        slot_v[base + i] = stage_v[i];
    }
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int q = 0; q < NW; ++q) { s0 += sm.red[0][q]; s1 += sm.red[1][q]; s2 += sm.red[2][q]; }
        bucket_cnt[b] = btotal;
        partials[b] = s0;
        partials[nbk + b] = s1;
        partials[2 * nbk + b] = s2;
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
    constexpr int CPT = (BK_CAP_A + BK_CAP_B + NT - 1) / NT;
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
