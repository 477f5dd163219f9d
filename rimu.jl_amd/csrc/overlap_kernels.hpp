// overlap_kernels.hpp -- overlaps of two device vectors: dot(t, u) and the three-way
// dot(t, op, u) = dot_from_right(t, op, u) (DictVectors/abstractdvec.jl:313-324), a sum over
// the entries (k, v) of u of  t[k] * diagonal_element(op, k) * v  plus, for an operator with
// off-diagonals,  t[k'] * h' * v  over every off-diagonal (k', h') of k.  These are the
// reductions behind AllOverlaps (replicastrategy.jl:112-183).
//
// One workgroup covers OV_TILE consecutive entries of u.  Lookups of t[key] are binary
// searches in t's sorted keys: t[k] for u's own keys within the range of t that the tile's
// first and last key bound, staged in LDS when it holds at most OV_LDS keys; t[k'] for
// neighbours over all of t, the first levels taken from OV_SPLIT evenly spaced keys of t
// staged in LDS -- or, when the keys are stored mixed (spread evenly over their B bits), from
// an index of t by the top P key bits, which leaves a range of a few dozen keys to search.
// Per-workgroup partial sums are reduced in a fixed order (deterministic).
#pragma once

#include "step_kernels.hpp"

namespace rimu {

constexpr int OV_IPT = 8;
constexpr int OV_TILE = NT * OV_IPT;
constexpr int OV_SPLIT = 1024;
constexpr int OV_MAX_OPS = 8;
constexpr int OV_LDS = 3072;   // t's keys of a tile's range staged in LDS when they fit
constexpr int OV_INDEX_MAX_BITS = 22;   // top-bit index of t for mixed keys: at most 2^22 + 1 entries
constexpr int64_t OV_INDEX_MIN = 1 << 16;  // smaller t: the LDS splitters suffice
template <class K> struct ov_lds { static constexpr int value = sizeof(K) == 8 ? OV_LDS : 2048; };

enum { OP_HAMILTONIAN = 1, OP_G2_REAL = 2, OP_DENSITY = 3 };

struct OpDev {
    int kind;
    int param;  // G2RealCorrelator: d; DensityMatrixDiagonal: mode (1-based)
    ModelDev h; // OP_HAMILTONIAN
    KeyMix mix;  // stored keys are mix(address) (the same for both vectors)
};

static_assert(sizeof(OpDev) % 8 == 0, "operators are staged in LDS as 8-byte words");

// G2RealCorrelator(d): (1/M) sum_i n_i (n_{i+d} - delta_{0d}), periodic
// (correlation_functions.jl:82-92); the integer sum is exact, then one division.
template <class A>
__device__ inline double g2_real(A bits, int M, int d) {
    const int dd = ((d % M) + M) % M;
    int64_t acc = 0;
    A x = bits;
    int mode = 1;
    while (x) {
        const int z = ctzA(x);
        mode += z;
        x = shrA(x, z);
        const int n = ctzA((A)~x);
        if (dd == 0) {
            acc += (int64_t)n * (n - 1);
        } else {
            int j = mode + dd;
            if (j > M) j -= M;
            acc += (int64_t)n * occupation(bits, j);
        }
        x = shrA(x, n);
    }
    return (double)acc / (double)M;
}

// Position of key in keys[a, b), or -1.
template <class K>
__device__ inline int64_t find_in(const K* __restrict__ keys, int64_t a, int64_t b, K key) {
    const int64_t end = b;
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (keys[mid] < key) a = mid + 1; else b = mid;
    }
    return (a < end && keys[a] == key) ? a : -1;
}

// Position of key in keys[0, L), or -1; s_split[j] = keys[j * L / OV_SPLIT].
template <class K>
__device__ inline int64_t find_split(const K* __restrict__ keys, int64_t L, const K* s_split, K key) {
    int lo = 0, hi = OV_SPLIT;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_split[mid] <= key) lo = mid + 1; else hi = mid;
    }
    if (lo == 0) return -1;
    const int64_t a = (int64_t)(lo - 1) * L / OV_SPLIT;
    const int64_t b = lo < OV_SPLIT ? (int64_t)lo * L / OV_SPLIT : L;
    return find_in(keys, a, b, key);
}

// idx[b] = the first position in keys[0, L) whose top P of the B key bits are at least b,
// b = 0 .. 2^P (idx[2^P] = L).  Position i writes idx[b] for b in (top(i - 1), top(i)]
// (i = 0 from b = 0; i = L up to 2^P), one coalesced pass over the keys.
template <class K>
__global__ __launch_bounds__(NT) void overlap_index_kernel(const K* __restrict__ keys, int64_t L, int B, int P,
                                                           int64_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i > L) return;
    const int sh = B - P;
    const int64_t nb = (int64_t)1 << P;
    const int64_t lo = i == 0 ? 0 : (int64_t)(keys[i - 1] >> sh) + 1;
    const int64_t hi = i == L ? nb : (int64_t)(keys[i] >> sh);
    for (int64_t b = lo; b <= hi; ++b) idx[b] = i;
}

// t[key] * diag * v plus the off-diagonal terms of a Hamiltonian (RS: HubbardRealSpace).
// With an index (tidx != nullptr) a neighbour is searched in its top-bit range only.
template <bool RS, class K>
__device__ inline double ham_term(const ModelDev& h, const KeyMix& mix, K key, double tval, double v,
                                  const K* __restrict__ tk, const double* tvd,
                                  const int64_t* tvi, int64_t Lt, const K* s_split,
                                  const int64_t* __restrict__ tidx, int tsh) {
    double res = tval * model_diagonal_element<RS>(h, key) * v;
    const int64_t n = model_num_offdiagonals<RS>(h, key);
    for (int64_t c = 1; c <= n; ++c) {
        double hv;
        const K nb = mix_key(mix, model_offdiagonal<RS>(h, key, c, &hv));
        if (hv == 0.0) continue;
        int64_t q;
        if (tidx) {
            const int64_t b = (int64_t)(nb >> tsh);
            q = find_in(tk, tidx[b], tidx[b + 1], nb);
        } else {
            q = find_split(tk, Lt, s_split, nb);
        }
        if (q >= 0) res += (tvd ? tvd[q] : (double)tvi[q]) * hv * v;
    }
    return res;
}

// One entry (stored key, v) of u against operator op: t[key] * diag * v plus the
// off-diagonal terms.
template <class K>
__device__ __noinline__ double op_term(const OpDev& op, int M, K skey, double tval, double v,
                                       const K* __restrict__ tk, const double* tvd,
                                       const int64_t* tvi, int64_t Lt, const K* s_split,
                                       const int64_t* __restrict__ tidx, int tsh) {
    const K key = unmix_key(op.mix, skey);  // the address
    if (op.kind == OP_G2_REAL) return tval * g2_real(key, M, op.param) * v;
    if (op.kind == OP_DENSITY) return tval * (double)occupation(key, op.param) * v;
    if (op.h.kind == 3) return ham_term<true>(op.h, op.mix, key, tval, v, tk, tvd, tvi, Lt, s_split, tidx, tsh);
    return ham_term<false>(op.h, op.mix, key, tval, v, tk, tvd, tvi, Lt, s_split, tidx, tsh);
}

// ranges[t] = lower bound in t of u's key at t * OV_TILE (t < ntiles), ranges[ntiles] = Lt.
// u's keys are unique, so tile t's keys lie in t[ranges[t], ranges[t + 1]).
template <class K>
__global__ __launch_bounds__(NT) void overlap_ranges_kernel(const K* __restrict__ tk, int64_t Lt,
                                                            const K* __restrict__ uk,
                                                            int64_t ntiles, int64_t* __restrict__ ranges) {
    const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) { ranges[t] = Lt; return; }
    const K key = uk[t * OV_TILE];
    int64_t a = 0, b = Lt;
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        if (tk[mid] < key) a = mid + 1; else b = mid;
    }
    ranges[t] = a;
}

// partials[o * gridDim.x + block], o = 0 (dot) .. nops (operators).  Lt >= 1, Lu >= 1.
// need_split: an operator has off-diagonals (the splitters are only read then, and not at
// all when the top-bit index tidx is given: neighbour b's range is tidx[b >> tsh ...]).
template <class K, class VT, class VU>
__global__ __launch_bounds__(NT) void overlap_kernel(const K* __restrict__ tk,
                                                     const VT* __restrict__ tv, int64_t Lt,
                                                     const K* __restrict__ uk,
                                                     const VU* __restrict__ uv, int64_t Lu,
                                                     const OpDev* __restrict__ ops, int nops, int M,
                                                     const int64_t* __restrict__ ranges, int need_split,
                                                     const int64_t* __restrict__ tidx, int tsh,
                                                     double* __restrict__ partials) {
    constexpr int LDSN = ov_lds<K>::value;
    __shared__ K s_split[OV_SPLIT];
    __shared__ OpDev s_ops[OV_MAX_OPS];
    __shared__ K s_lk[LDSN];
    const int64_t t0 = (int64_t)blockIdx.x * OV_TILE;
    if (need_split && !tidx)
        for (int j = threadIdx.x; j < OV_SPLIT; j += NT) s_split[j] = tk[(int64_t)j * Lt / OV_SPLIT];
    for (int j = threadIdx.x; j < nops * (int)(sizeof(OpDev) / 8); j += NT)
        ((uint64_t*)s_ops)[j] = ((const uint64_t*)ops)[j];
    const int64_t lo = ranges[blockIdx.x], hi = ranges[blockIdx.x + 1];
    const bool staged = hi - lo <= LDSN;
    if (staged) {
        const int n = (int)(hi - lo);
        for (int j = threadIdx.x; j < n; j += NT) s_lk[j] = tk[lo + j];
    }
    __syncthreads();

    K k[OV_IPT];
    double v[OV_IPT];
#pragma unroll
    for (int j = 0; j < OV_IPT; ++j) {
        const int64_t i = t0 + (int64_t)j * NT + threadIdx.x;
        const int64_t ic = min(i, Lu - 1);
        k[j] = uk[ic];
        v[j] = i < Lu ? (double)uv[ic] : 0.0;
    }
    const double* tvd = is_int_v<VT>::value ? nullptr : (const double*)tv;
    const int64_t* tvi = is_int_v<VT>::value ? (const int64_t*)tv : nullptr;
    double acc[1 + OV_MAX_OPS];
#pragma unroll
    for (int o = 0; o <= OV_MAX_OPS; ++o) acc[o] = 0.0;
#pragma unroll
    for (int j = 0; j < OV_IPT; ++j) {
        const int64_t i = t0 + (int64_t)j * NT + threadIdx.x;
        if (i >= Lu) continue;
        int64_t p;
        if (staged) {
            int a = 0, b = (int)(hi - lo);
            const int n = b;
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (s_lk[mid] < k[j]) a = mid + 1; else b = mid;
            }
            p = (a < n && s_lk[a] == k[j]) ? lo + a : -1;
        } else {
            p = find_in(tk, lo, hi, k[j]);
        }
        const double tval = p >= 0 ? (double)tv[p] : 0.0;
        acc[0] += tval * v[j];
#pragma unroll
        for (int o = 0; o < OV_MAX_OPS; ++o)
            if (o < nops)
                acc[1 + o] += op_term<K>(s_ops[o], M, k[j], tval, v[j], tk, tvd, tvi, Lt, s_split, tidx, tsh);
    }
#pragma unroll
    for (int o = 0; o <= OV_MAX_OPS; ++o) {
        if (o <= nops) {   // nops is uniform: every thread takes the same branches
            const double s = block_sum<NT>(acc[o]);
            if (threadIdx.x == 0) partials[(int64_t)o * gridDim.x + blockIdx.x] = s;
        }
    }
}

// out[o] = sum of partials[o * nparts + b] over b in a fixed order; one workgroup per o.
__global__ __launch_bounds__(NT) void overlap_final_kernel(const double* __restrict__ partials,
                                                           int64_t nparts, double* __restrict__ out) {
    const int o = blockIdx.x;
    double acc = 0.0;
    for (int64_t b = threadIdx.x; b < nparts; b += NT) acc += partials[(int64_t)o * nparts + b];
    const double s = block_sum<NT>(acc);
    if (threadIdx.x == 0) out[o] = s;
}

}  // namespace rimu
