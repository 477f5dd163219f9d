// dense_hv.hip -- deterministic H.v on the ranked basis of HubbardReal1D, its adjoint, and
// the checkpointed adjoint of a K-step projection (the C ABI rimu_dense_* of rimu_hip.h).
//
// Row r of a vector is the address of colex rank r: the N one-bits of the address at
// positions p_0 < p_1 < ... give rank = sum_j C(p_j, j+1), and colex order equals
// ascending numeric order of the bit strings.  A hop moves one boson by one bit, so the
// rank of a neighbour follows from the row's own rank with one binomial: a one moving
// from bit q to q+1 (the j-th one) adds C(q, j); from p to p-1 it subtracts C(p-1, j).
// Hops across the periodic boundary rotate the string and are ranked from scratch.
//
// Vectors are row-major [dim][R] (R right-hand sides, one per Hamiltonian): a neighbour's
// R amplitudes are one contiguous 8R-byte load.  Every reduction is a fixed-order tree,
// so results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "bosefs_device.hpp"
#include "capi_util.hpp"
#include "device_common.hpp"

using namespace rimu;

namespace {

constexpr int DT = 256;          // workgroup size
constexpr int MAXR = 8;          // right-hand sides per launch
constexpr int BINOM_N = 65;      // C(n, k) for 0 <= n, k <= 64

struct DenseParams {
    double a[MAXR], b[MAXR], u[MAXR], t[MAXR];
    int M, B;
};

// Colex rank from scratch.
__device__ inline int64_t colex_rank(uint64_t bits, const int64_t* __restrict__ binom) {
    int64_t r = 0;
    int j = 1;
    while (bits) {
        const int p = __builtin_ctzll(bits);
        r += binom[p * BINOM_N + j];
        ++j;
        bits &= bits - 1;
    }
    return r;
}

// The hops of row r (address `bits`, rank `rank`) in the order of hopnextneighbour's
// `chosen` (bosefs.jl:297-345): for the s-th occupied mode, right hop then left hop.
// f(neighbour_rank, sqrt_prod) is called for each; the matrix element is -t*sqrt_prod.
template <class F>
__device__ inline void for_each_hop(uint64_t bits, int64_t rank, int M, int B,
                                    const int64_t* __restrict__ binom, F&& f) {
    const uint64_t maskB = low_mask(B);
    uint64_t st = run_starts(bits);
    while (st) {
        const int p = __builtin_ctzll(st);
        st &= st - 1;
        const int n = ones_from(bits, p);
        const int m = p - popc64(bits & low_mask(p)) + 1;
        // right hop
        if (m == M) {
            const uint64_t nb = ((bits << 1) | 1ull) & maskB;
            f(colex_rank(nb, binom), sqrt((double)((int64_t)n * (ones_from(bits, 0) + 1))));
        } else {
            const int q = p + n - 1;  // the last one of the run moves to q+1
            const int j = popc64(bits & low_mask(q));
            f(rank + binom[q * BINOM_N + j], sqrt((double)((int64_t)n * (ones_from(bits, q + 2) + 1))));
        }
        // left hop
        if (m == 1) {
            const uint64_t nb = (bits >> 1) | (1ull << (B - 1));
            const int hz = 63 - clz64(~bits & maskB);
            f(colex_rank(nb, binom), sqrt((double)((int64_t)n * ((B - 1 - hz) + 1))));
        } else {
            const int j = popc64(bits & low_mask(p));  // the first one of the run moves to p-1
            int nprev = 0;
            if (p >= 2) {
                const uint64_t z = ~bits & low_mask(p - 1);
                nprev = z ? (p - 2 - (63 - clz64(z))) : (p - 1);
            }
            f(rank - binom[(p - 1) * BINOM_N + j], sqrt((double)((int64_t)n * (nprev + 1))));
        }
    }
}

// w = (a + b H) v, one thread per row.
template <int R>
__global__ __launch_bounds__(DT) void dense_apply_kernel(const uint64_t* __restrict__ basis, int64_t D,
                                                         DenseParams p, const int64_t* __restrict__ binom,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ w) {
    const int64_t r = (int64_t)blockIdx.x * DT + threadIdx.x;
    if (r >= D) return;
    const uint64_t bits = basis[r];
    const double dr = (double)bose_hubbard_interaction(bits) / 2;
    double acc[R];
#pragma unroll
    for (int i = 0; i < R; ++i) acc[i] = 0.0;
    for_each_hop(bits, r, p.M, p.B, binom, [&](int64_t k, double sp) {
        const double* vk = v + k * R;
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i] += -sp * vk[i];
    });
#pragma unroll
    for (int i = 0; i < R; ++i)
        w[r * R + i] = (p.a[i] + p.b[i] * (p.u[i] * dr)) * v[r * R + i] + p.b[i] * (p.t[i] * acc[i]);
}

// Pullback: vbar = (a + b H) ybar; per-block partial sums of b <ybar|D|v> and b <ybar|K|v>
// written to partials[blockIdx][2R].  K is real symmetric, so <ybar|K|v> = sum_r v_r (K ybar)_r
// and only ybar is gathered at the neighbours.
template <int R>
__global__ __launch_bounds__(DT) void dense_pullback_kernel(const uint64_t* __restrict__ basis, int64_t D,
                                                            DenseParams p, const int64_t* __restrict__ binom,
                                                            const double* __restrict__ v,
                                                            const double* __restrict__ ybar,
                                                            double* __restrict__ vbar,
                                                            double* __restrict__ partials) {
    const int64_t r = (int64_t)blockIdx.x * DT + threadIdx.x;
    double pu[R], pt[R];
#pragma unroll
    for (int i = 0; i < R; ++i) { pu[i] = 0.0; pt[i] = 0.0; }
    if (r < D) {
        const uint64_t bits = basis[r];
        const double dr = (double)bose_hubbard_interaction(bits) / 2;
        double ky[R];
#pragma unroll
        for (int i = 0; i < R; ++i) ky[i] = 0.0;
        for_each_hop(bits, r, p.M, p.B, binom, [&](int64_t k, double sp) {
            const double* yk = ybar + k * R;
#pragma unroll
            for (int i = 0; i < R; ++i) ky[i] += -sp * yk[i];
        });
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const double yr = ybar[r * R + i];
            const double vr = v[r * R + i];
            vbar[r * R + i] = (p.a[i] + p.b[i] * (p.u[i] * dr)) * yr + p.b[i] * (p.t[i] * ky[i]);
            pu[i] = p.b[i] * (yr * (dr * vr));
            pt[i] = p.b[i] * (vr * ky[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const double su = block_sum<DT>(pu[i]);
        const double stt = block_sum<DT>(pt[i]);
        if (threadIdx.x == 0) {
            partials[(int64_t)blockIdx.x * 2 * R + i] = su;
            partials[(int64_t)blockIdx.x * 2 * R + R + i] = stt;
        }
    }
}

// out[i] = sum_r x[r][i] * y[r][i] as per-block partials[blockIdx][R].
template <int R>
__global__ __launch_bounds__(DT) void dense_dot_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                       int64_t D, double* __restrict__ partials) {
    const int64_t r = (int64_t)blockIdx.x * DT + threadIdx.x;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const double s = block_sum<DT>(r < D ? x[r * R + i] * y[r * R + i] : 0.0);
        if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * R + i] = s;
    }
}

// Sum `rows` rows of width `width` in order: out[c] (+)= sum_b partials[b][c].
__global__ __launch_bounds__(DT) void reduce_rows_kernel(const double* __restrict__ partials, int64_t rows,
                                                         int width, double* __restrict__ out, int accumulate) {
    const int c = blockIdx.x;
    double acc = 0;
    for (int64_t b = threadIdx.x; b < rows; b += DT) acc += partials[b * width + c];
    const double s = block_sum<DT>(acc);
    if (threadIdx.x == 0) out[c] = accumulate ? out[c] + s : s;
}

// lambda = 2 (Hv - E v) / <v|v>, per right-hand side (gradient of the Rayleigh quotient).
template <int R>
__global__ __launch_bounds__(DT) void rq_gradient_kernel(const double* __restrict__ v, const double* __restrict__ hv,
                                                         const double* __restrict__ dots /* [vv R][vHv R] */,
                                                         int64_t D, double* __restrict__ lam) {
    const int64_t r = (int64_t)blockIdx.x * DT + threadIdx.x;
    if (r >= D) return;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const double vv = dots[i], e = dots[R + i] / vv;
        lam[r * R + i] = 2.0 * (hv[r * R + i] - e * v[r * R + i]) / vv;
    }
}

// One three-term Lanczos update per right-hand side: w -= alpha_j v_j + beta_j v_{j-1},
// with alpha_j and beta_j^2 read from device memory, and the partial sums of |w|^2 of each
// block written to partials[blockIdx][R] (reduced in order afterwards).
template <int R>
__global__ __launch_bounds__(DT) void lanczos_orth_kernel(double* __restrict__ w, const double* __restrict__ v,
                                                          const double* __restrict__ vprev,
                                                          const double* __restrict__ alpha,
                                                          const double* __restrict__ beta2, int64_t D,
                                                          double* __restrict__ partials) {
    const int64_t r = (int64_t)blockIdx.x * DT + threadIdx.x;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        double x = 0;
        if (r < D) {
            x = w[r * R + i] - alpha[i] * v[r * R + i] - sqrt(beta2[i]) * vprev[r * R + i];
            w[r * R + i] = x;
        }
        const double s = block_sum<DT>(x * x);
        if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * R + i] = s;
    }
}

// out = w / sqrt(norm2) per right-hand side; a right-hand side whose norm fell to
// `floor2` or below (its Krylov space is exhausted) gets the zero vector.
template <int R>
__global__ __launch_bounds__(DT) void lanczos_scale_kernel(const double* __restrict__ w,
                                                           const double* __restrict__ norm2, double floor2,
                                                           int64_t D, double* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * DT + threadIdx.x;
    if (r >= D) return;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const double n2 = norm2[i];
        out[r * R + i] = n2 > floor2 ? w[r * R + i] / sqrt(n2) : 0.0;
    }
}

}  // namespace

struct rimu_dense {
    int R = 0;
    int M = 0, N = 0, B = 0;
    int64_t D = 0;
    rimu_model models[MAXR];
    hipStream_t stream = nullptr;
    DevBuf<uint64_t> basis;
    DevBuf<int64_t> binom;
    DevBuf<double> x, y, z, partials, scal;
    double* h_scal = nullptr;  // pinned, 8 * MAXR
};

namespace {

DenseParams params(const rimu_dense* d, const double* a, const double* b) {
    DenseParams p{};
    for (int i = 0; i < d->R; ++i) {
        p.a[i] = a[i];
        p.b[i] = b[i];
        p.u[i] = d->models[i].u;
        p.t[i] = d->models[i].t;
    }
    p.M = d->M;
    p.B = d->B;
    return p;
}

template <int R>
int launch_apply(rimu_dense* d, const DenseParams& p, const double* v, double* w) {
    dense_apply_kernel<R><<<ceil_div(d->D, DT), DT, 0, d->stream>>>(d->basis.p, d->D, p, d->binom.p, v, w);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

int apply(rimu_dense* d, const DenseParams& p, const double* v, double* w) {
    switch (d->R) {
        case 1: return launch_apply<1>(d, p, v, w);
        case 2: return launch_apply<2>(d, p, v, w);
        case 3: return launch_apply<3>(d, p, v, w);
        case 4: return launch_apply<4>(d, p, v, w);
        case 5: return launch_apply<5>(d, p, v, w);
        case 6: return launch_apply<6>(d, p, v, w);
        case 7: return launch_apply<7>(d, p, v, w);
        default: return launch_apply<8>(d, p, v, w);
    }
}

// vbar and the (u, t) cotangents; cot receives [ubar R][tbar R] (added to it if accumulate)
template <int R>
int launch_pullback(rimu_dense* d, const DenseParams& p, const double* v, const double* ybar,
                    double* vbar, double* cot, int accumulate) {
    const int64_t blocks = ceil_div(d->D, DT);
    HIP_TRY(d->partials.alloc((size_t)blocks * 2 * R));
    dense_pullback_kernel<R><<<blocks, DT, 0, d->stream>>>(d->basis.p, d->D, p, d->binom.p, v, ybar,
                                                            vbar, d->partials.p);
    reduce_rows_kernel<<<2 * R, DT, 0, d->stream>>>(d->partials.p, blocks, 2 * R, cot, accumulate);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

int pullback(rimu_dense* d, const DenseParams& p, const double* v, const double* ybar, double* vbar,
             double* cot, int accumulate) {
    switch (d->R) {
        case 1: return launch_pullback<1>(d, p, v, ybar, vbar, cot, accumulate);
        case 2: return launch_pullback<2>(d, p, v, ybar, vbar, cot, accumulate);
        case 3: return launch_pullback<3>(d, p, v, ybar, vbar, cot, accumulate);
        case 4: return launch_pullback<4>(d, p, v, ybar, vbar, cot, accumulate);
        case 5: return launch_pullback<5>(d, p, v, ybar, vbar, cot, accumulate);
        case 6: return launch_pullback<6>(d, p, v, ybar, vbar, cot, accumulate);
        case 7: return launch_pullback<7>(d, p, v, ybar, vbar, cot, accumulate);
        default: return launch_pullback<8>(d, p, v, ybar, vbar, cot, accumulate);
    }
}

template <int R>
int launch_dot(rimu_dense* d, const double* x, const double* y, double* out) {
    const int64_t blocks = ceil_div(d->D, DT);
    HIP_TRY(d->partials.alloc((size_t)blocks * R));
    dense_dot_kernel<R><<<blocks, DT, 0, d->stream>>>(x, y, d->D, d->partials.p);
    reduce_rows_kernel<<<R, DT, 0, d->stream>>>(d->partials.p, blocks, R, out, 0);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

int dot(rimu_dense* d, const double* x, const double* y, double* out) {
    switch (d->R) {
        case 1: return launch_dot<1>(d, x, y, out);
        case 2: return launch_dot<2>(d, x, y, out);
        case 3: return launch_dot<3>(d, x, y, out);
        case 4: return launch_dot<4>(d, x, y, out);
        case 5: return launch_dot<5>(d, x, y, out);
        case 6: return launch_dot<6>(d, x, y, out);
        case 7: return launch_dot<7>(d, x, y, out);
        default: return launch_dot<8>(d, x, y, out);
    }
}

int rq_gradient(rimu_dense* d, const double* v, const double* hv, const double* dots, double* lam) {
    const int64_t blocks = ceil_div(d->D, DT);
    switch (d->R) {
        case 1: rq_gradient_kernel<1><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        case 2: rq_gradient_kernel<2><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        case 3: rq_gradient_kernel<3><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        case 4: rq_gradient_kernel<4><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        case 5: rq_gradient_kernel<5><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        case 6: rq_gradient_kernel<6><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        case 7: rq_gradient_kernel<7><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
        default: rq_gradient_kernel<8><<<blocks, DT, 0, d->stream>>>(v, hv, dots, d->D, lam); break;
    }
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

template <int R>
int launch_lanczos(rimu_dense* d, double* w, const double* v, const double* vprev, const double* alpha,
                   const double* beta2, double* norm2_out, double* vnext, double floor2) {
    const int64_t blocks = ceil_div(d->D, DT);
    HIP_TRY(d->partials.alloc((size_t)blocks * R));
    lanczos_orth_kernel<R><<<blocks, DT, 0, d->stream>>>(w, v, vprev, alpha, beta2, d->D, d->partials.p);
    reduce_rows_kernel<<<R, DT, 0, d->stream>>>(d->partials.p, blocks, R, norm2_out, 0);
    lanczos_scale_kernel<R><<<blocks, DT, 0, d->stream>>>(w, norm2_out, floor2, d->D, vnext);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

int lanczos_update(rimu_dense* d, double* w, const double* v, const double* vprev, const double* alpha,
                   const double* beta2, double* norm2_out, double* vnext, double floor2) {
    switch (d->R) {
        case 1: return launch_lanczos<1>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        case 2: return launch_lanczos<2>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        case 3: return launch_lanczos<3>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        case 4: return launch_lanczos<4>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        case 5: return launch_lanczos<5>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        case 6: return launch_lanczos<6>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        case 7: return launch_lanczos<7>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
        default: return launch_lanczos<8>(d, w, v, vprev, alpha, beta2, norm2_out, vnext, floor2);
    }
}

// Lowest eigenvalue of the symmetric tridiagonal matrix with diagonal a[0..m) and squared
// off-diagonal b2[1..m) (b2[k] couples k-1 and k), by bisection on the Sturm count: the
// number of negative pivots of T - x I is the number of eigenvalues below x.
double tridiag_lowest(const double* a, const double* b2, int64_t m) {
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t k = 0; k < m; ++k) {
        const double off = (k > 0 ? sqrt(b2[k]) : 0.0) + (k + 1 < m ? sqrt(b2[k + 1]) : 0.0);
        lo = std::min(lo, a[k] - off);
        hi = std::max(hi, a[k] + off);
    }
    auto below = [&](double x) {
        int64_t c = 0;
        double piv = 1.0;
        for (int64_t k = 0; k < m; ++k) {
            piv = a[k] - x - (k > 0 ? b2[k] / piv : 0.0);
            if (piv == 0.0) piv = -1e-300;
            if (piv < 0) ++c;
        }
        return c;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (below(mid) >= 1) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

bool binomial64(int n, int k, int64_t* out) {
    if (k < 0 || k > n) { *out = 0; return true; }
    unsigned __int128 c = 1;
    for (int i = 1; i <= k; ++i) {
        c = c * (unsigned)(n - k + i) / (unsigned)i;
        if (c > (unsigned __int128)INT64_MAX) return false;
    }
    *out = (int64_t)c;
    return true;
}

}  // namespace

extern "C" {

int rimu_basis_size(int32_t M, int32_t N, int64_t* dim) {
    if (!dim) return fail(RIMU_ERR_ARG, "dim is null");
    if (M < 1 || N < 0 || M + N - 1 > 64) return fail(RIMU_ERR_ARG, "need N + M - 1 <= 64");
    if (!binomial64(N + M - 1, N, dim)) return fail(RIMU_ERR_ARG, "dimension overflows int64");
    return RIMU_OK;
}

int rimu_dense_create(const rimu_model* models, int32_t nrhs, rimu_dense** out) {
    if (!out || !models) return fail(RIMU_ERR_ARG, "null argument");
    *out = nullptr;
    if (nrhs < 1 || nrhs > MAXR) return fail(RIMU_ERR_ARG, "nrhs must be in 1..8");
    const rimu_model& m0 = models[0];
    for (int i = 0; i < nrhs; ++i) {
        if (models[i].kind != RIMU_MODEL_HUBBARD_REAL_1D)
            return fail(RIMU_ERR_ARG, "ranked-basis products support HubbardReal1D");
        if (models[i].M != m0.M || models[i].N != m0.N)
            return fail(RIMU_ERR_ARG, "all right-hand sides must share M and N");
    }
    if (m0.M < 2 || m0.N < 1 || m0.M + m0.N - 1 > 64) return fail(RIMU_ERR_ARG, "need M >= 2, N >= 1, N + M - 1 <= 64");
    int64_t D;
    if (int rc = rimu_basis_size(m0.M, m0.N, &D)) return rc;
    if (D > (int64_t)1 << 31) return fail(RIMU_ERR_ARG, "basis larger than 2^31 addresses");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RIMU_ERR_NODEVICE, "no HIP device");
    rimu_dense* d = new rimu_dense();
    d->R = nrhs;
    d->M = m0.M;
    d->N = m0.N;
    d->B = m0.M + m0.N - 1;
    d->D = D;
    for (int i = 0; i < nrhs; ++i) d->models[i] = models[i];
    auto bad = [&](const char* what, hipError_t e) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        rimu_dense_destroy(d);
        return fail(RIMU_ERR_HIP, msg);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) return bad("stream", e);
    // binomial table and the sorted basis (next bit pattern with the same popcount)
    std::vector<int64_t> binom(BINOM_N * BINOM_N, 0);
    for (int n = 0; n < BINOM_N; ++n)
        for (int k = 0; k < BINOM_N; ++k) {
            int64_t c;
            binom[n * BINOM_N + k] = binomial64(n, k, &c) ? c : INT64_MAX;
        }
    std::vector<uint64_t> keys((size_t)D);
    uint64_t x = (d->N >= 64) ? ~0ull : ((1ull << d->N) - 1);
    for (int64_t r = 0; r < D; ++r) {
        keys[r] = x;
        const uint64_t c = x & (~x + 1), rr = x + c;  // Gosper's hack
        x = c ? (((rr ^ x) >> 2) / c) | rr : 0;
    }
    if ((e = d->basis.alloc(D)) != hipSuccess) return bad("basis", e);
    if ((e = d->binom.alloc(binom.size())) != hipSuccess) return bad("binomials", e);
    if ((e = d->scal.alloc(8 * MAXR)) != hipSuccess) return bad("scalars", e);
    if ((e = hipHostMalloc((void**)&d->h_scal, 8 * MAXR * sizeof(double), hipHostMallocDefault)) != hipSuccess)
        return bad("pinned scalars", e);
    if ((e = hipMemcpy(d->basis.p, keys.data(), D * 8, hipMemcpyHostToDevice)) != hipSuccess) return bad("basis upload", e);
    if ((e = hipMemcpy(d->binom.p, binom.data(), binom.size() * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return bad("binomial upload", e);
    *out = d;
    return RIMU_OK;
}

int rimu_dense_destroy(rimu_dense* d) {
    if (!d) return RIMU_OK;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    d->basis.release(); d->binom.release(); d->x.release(); d->y.release(); d->z.release();
    d->partials.release(); d->scal.release();
    if (d->h_scal) (void)hipHostFree(d->h_scal);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
    return RIMU_OK;
}

int rimu_dense_dim(rimu_dense* d, int64_t* dim) {
    if (!d || !dim) return fail(RIMU_ERR_ARG, "null argument");
    *dim = d->D;
    return RIMU_OK;
}

int rimu_dense_basis(rimu_dense* d, uint64_t* keys) {
    if (!d || !keys) return fail(RIMU_ERR_ARG, "null argument");
    HIP_TRY(hipMemcpy(keys, d->basis.p, d->D * 8, hipMemcpyDeviceToHost));
    return RIMU_OK;
}

int rimu_dense_apply(rimu_dense* d, const double* a, const double* b, const double* v, double* w) {
    if (!d || !a || !b || !v || !w) return fail(RIMU_ERR_ARG, "null argument");
    const size_t bytes = (size_t)d->D * d->R * 8;
    HIP_TRY(d->x.alloc((size_t)d->D * d->R));
    HIP_TRY(d->y.alloc((size_t)d->D * d->R));
    HIP_TRY(hipMemcpyAsync(d->x.p, v, bytes, hipMemcpyHostToDevice, d->stream));
    if (int rc = apply(d, params(d, a, b), d->x.p, d->y.p)) return rc;
    HIP_TRY(hipMemcpyAsync(w, d->y.p, bytes, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return RIMU_OK;
}

int rimu_dense_pullback(rimu_dense* d, const double* a, const double* b, const double* v,
                        const double* ybar, double* vbar, double* ubar, double* tbar) {
    if (!d || !a || !b || !v || !ybar || !vbar) return fail(RIMU_ERR_ARG, "null argument");
    const size_t n = (size_t)d->D * d->R, bytes = n * 8;
    HIP_TRY(d->x.alloc(n));
    HIP_TRY(d->y.alloc(n));
    HIP_TRY(d->z.alloc(n));
    HIP_TRY(hipMemcpyAsync(d->x.p, v, bytes, hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->y.p, ybar, bytes, hipMemcpyHostToDevice, d->stream));
    if (int rc = pullback(d, params(d, a, b), d->x.p, d->y.p, d->z.p, d->scal.p, 0)) return rc;
    HIP_TRY(hipMemcpyAsync(vbar, d->z.p, bytes, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipMemcpyAsync(d->h_scal, d->scal.p, 2 * d->R * 8, hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (int i = 0; i < d->R; ++i) {
        if (ubar) ubar[i] = d->h_scal[i];
        if (tbar) tbar[i] = d->h_scal[d->R + i];
    }
    return RIMU_OK;
}

int rimu_dense_projector(rimu_dense* d, int64_t K, double dtau, const double* shift, int64_t ckpt,
                         const double* v0, double* energy, double* grad_u, double* grad_t,
                         double* hf_u, double* hf_t, double* v_final) {
    if (!d || !shift || !v0) return fail(RIMU_ERR_ARG, "null argument");
    if (K < 0 || ckpt < 1 || !(dtau > 0)) return fail(RIMU_ERR_ARG, "need K >= 0, ckpt >= 1, dtau > 0");
    const int R = d->R;
    const size_t n = (size_t)d->D * R;
    const int64_t nseg = K == 0 ? 0 : ceil_div(K, ckpt);
    double a_t[MAXR], b_t[MAXR], a_h[MAXR], b_h[MAXR];
    for (int i = 0; i < R; ++i) {
        a_t[i] = 1 + dtau * shift[i];  // T = 1 + dt (S - H)
        b_t[i] = -dtau;
        a_h[i] = 0;                    // H itself
        b_h[i] = 1;
    }
    const DenseParams pT = params(d, a_t, b_t), pH = params(d, a_h, b_h);
    DevBuf<double> ck, seg, cur, nxt, hv, lam, lam2;
    auto cleanup = [&]() { ck.release(); seg.release(); cur.release(); nxt.release(); hv.release(); lam.release(); lam2.release(); };
    auto run = [&]() -> int {
        HIP_TRY(ck.alloc(std::max<int64_t>(nseg, 1) * n));
        HIP_TRY(seg.alloc((size_t)std::min<int64_t>(ckpt, std::max<int64_t>(K, 1)) * n));
        HIP_TRY(cur.alloc(n));
        HIP_TRY(nxt.alloc(n));
        HIP_TRY(hv.alloc(n));
        HIP_TRY(lam.alloc(n));
        HIP_TRY(lam2.alloc(n));
        HIP_TRY(hipMemcpyAsync(cur.p, v0, n * 8, hipMemcpyHostToDevice, d->stream));
        // forward sweep, checkpoint every ckpt steps
        for (int64_t k = 0; k < K; ++k) {
            if (k % ckpt == 0)
                HIP_TRY(hipMemcpyAsync(ck.p + (k / ckpt) * n, cur.p, n * 8, hipMemcpyDeviceToDevice, d->stream));
            if (int rc = apply(d, pT, cur.p, nxt.p)) return rc;
            std::swap(cur.p, nxt.p);
        }
        // energy: scal[0:R] = <v|v>, scal[R:2R] = <v|Hv>; HF: scal[2R:4R] = <v|D|v>, <v|K|v>
        if (int rc = apply(d, pH, cur.p, hv.p)) return rc;
        if (int rc = dot(d, cur.p, cur.p, d->scal.p)) return rc;
        if (int rc = dot(d, cur.p, hv.p, d->scal.p + R)) return rc;
        {
            double one[MAXR], unit_b[MAXR];
            for (int i = 0; i < R; ++i) { one[i] = 0; unit_b[i] = 1; }
            // pullback with ybar = v, b = 1 gives <v|D|v> and <v|K|v>; vbar is discarded
            if (int rc = pullback(d, params(d, one, unit_b), cur.p, cur.p, lam2.p, d->scal.p + 2 * R, 0)) return rc;
        }
        // cotangent of the Rayleigh quotient, and the (u, t) accumulators start at zero
        if (int rc = rq_gradient(d, cur.p, hv.p, d->scal.p, lam.p)) return rc;
        HIP_TRY(hipMemsetAsync(d->scal.p + 4 * R, 0, 2 * R * 8, d->stream));
        // reverse sweep: recompute each segment from its checkpoint, then pull back
        for (int64_t s = nseg - 1; s >= 0; --s) {
            const int64_t k0 = s * ckpt, k1 = std::min(K, k0 + ckpt);
            HIP_TRY(hipMemcpyAsync(seg.p, ck.p + s * n, n * 8, hipMemcpyDeviceToDevice, d->stream));
            for (int64_t k = k0; k < k1 - 1; ++k)
                if (int rc = apply(d, pT, seg.p + (k - k0) * n, seg.p + (k - k0 + 1) * n)) return rc;
            for (int64_t k = k1 - 1; k >= k0; --k) {
                if (int rc = pullback(d, pT, seg.p + (k - k0) * n, lam.p, lam2.p, d->scal.p + 4 * R, 1)) return rc;
                std::swap(lam.p, lam2.p);
            }
        }
        HIP_TRY(hipMemcpyAsync(d->h_scal, d->scal.p, 6 * R * 8, hipMemcpyDeviceToHost, d->stream));
        if (v_final) HIP_TRY(hipMemcpyAsync(v_final, cur.p, n * 8, hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        const double* h = d->h_scal;
        for (int i = 0; i < R; ++i) {
            const double vv = h[i], e = h[R + i] / vv;
            const double du = h[2 * R + i] / vv, dt = h[3 * R + i] / vv;
            if (energy) energy[i] = e;
            if (hf_u) hf_u[i] = du;
            if (hf_t) hf_t[i] = dt;
            // direct dependence of the final Rayleigh quotient plus the projection steps
            if (grad_u) grad_u[i] = du + h[4 * R + i];
            if (grad_t) grad_t[i] = dt + h[5 * R + i];
        }
        return RIMU_OK;
    };
    const int rc = run();
    cleanup();
    return rc;
}

int rimu_dense_ground_state(rimu_dense* d, const double* v0, int64_t max_iters, double tol,
                            double* energy, int64_t* iters) {
    if (!d || !v0 || !energy) return fail(RIMU_ERR_ARG, "null argument");
    if (max_iters < 1 || !(tol > 0)) return fail(RIMU_ERR_ARG, "need max_iters >= 1 and tol > 0");
    const int R = d->R;
    const size_t n = (size_t)d->D * R;
    const int64_t cap = std::min<int64_t>(max_iters, d->D);
    constexpr int64_t CHECK = 16;
    // beta^2 at or below this ends a right-hand side's Krylov space (rounding leaves
    // beta ~ 1e-14 |H| there; a true beta this small would move the Ritz value by ~beta^2)
    constexpr double FLOOR2 = 1e-20;
    double a_h[MAXR], b_h[MAXR];
    for (int i = 0; i < R; ++i) { a_h[i] = 0; b_h[i] = 1; }
    const DenseParams pH = params(d, a_h, b_h);
    DevBuf<double> v, vprev, w, hist;
    // hist: alpha [cap][R], then beta^2 [cap + 1][R] (beta^2[0] = 0), then |v0|^2 [R]
    auto cleanup = [&]() { v.release(); vprev.release(); w.release(); hist.release(); };
    auto run = [&]() -> int {
        HIP_TRY(v.alloc(n));
        HIP_TRY(vprev.alloc(n));
        HIP_TRY(w.alloc(n));
        HIP_TRY(hist.alloc((size_t)(2 * cap + 2) * R));
        double* alpha = hist.p;
        double* beta2 = hist.p + cap * R;
        double* n0 = hist.p + (2 * cap + 1) * R;
        HIP_TRY(hipMemsetAsync(hist.p, 0, (size_t)(2 * cap + 2) * R * 8, d->stream));
        HIP_TRY(hipMemsetAsync(vprev.p, 0, n * 8, d->stream));
        HIP_TRY(hipMemcpyAsync(w.p, v0, n * 8, hipMemcpyHostToDevice, d->stream));
        // v_0 = v0 / |v0| (lanczos_update with alpha = beta = 0 only normalises)
        if (int rc = lanczos_update(d, w.p, vprev.p, vprev.p, beta2, beta2, n0, v.p, 0.0)) return rc;
        std::vector<double> h_a, h_b2, h_n0(R), last(R, INFINITY), cur(R);
        std::vector<int64_t> m_eff(R, -1);
        int64_t j = 0;
        for (; j < cap; ++j) {
            if (int rc = apply(d, pH, v.p, w.p)) return rc;
            if (int rc = dot(d, v.p, w.p, alpha + j * R)) return rc;
            if (int rc = lanczos_update(d, w.p, v.p, vprev.p, alpha + j * R, beta2 + j * R,
                                        beta2 + (j + 1) * R, vprev.p, FLOOR2)) return rc;
            std::swap(v.p, vprev.p);
            const bool last_iter = j + 1 == cap;
            if ((j + 1) % CHECK != 0 && !last_iter) continue;
            const int64_t m = j + 1;
            h_a.resize((size_t)m * R);
            h_b2.resize((size_t)(m + 1) * R);
            HIP_TRY(hipMemcpyAsync(h_a.data(), alpha, (size_t)m * R * 8, hipMemcpyDeviceToHost, d->stream));
            HIP_TRY(hipMemcpyAsync(h_b2.data(), beta2, (size_t)(m + 1) * R * 8, hipMemcpyDeviceToHost, d->stream));
            HIP_TRY(hipMemcpyAsync(h_n0.data(), n0, R * 8, hipMemcpyDeviceToHost, d->stream));
            HIP_TRY(hipStreamSynchronize(d->stream));
            bool done = true;
            std::vector<double> a(m), b2(m + 1);
            for (int i = 0; i < R; ++i) {
                if (!(h_n0[i] > 0)) return fail(RIMU_ERR_ARG, "start vector is zero");
                int64_t mi = m;
                for (int64_t k = 1; k <= m; ++k) {
                    if (!(h_b2[k * R + i] > FLOOR2)) { mi = k; break; }  // exhausted after k vectors
                }
                for (int64_t k = 0; k < mi; ++k) {
                    a[k] = h_a[k * R + i];
                    b2[k] = h_b2[k * R + i];
                }
                cur[i] = tridiag_lowest(a.data(), b2.data(), mi);
                const bool exhausted = !(h_b2[mi * R + i] > FLOOR2);
                const bool conv = exhausted || std::fabs(cur[i] - last[i]) <= tol * std::max(1.0, std::fabs(cur[i]));
                if (!conv) done = false;
                last[i] = cur[i];
            }
            if (done) { ++j; break; }
        }
        for (int i = 0; i < R; ++i) energy[i] = cur[i];
        if (iters) *iters = j;
        return RIMU_OK;
    };
    const int rc = run();
    cleanup();
    return rc;
}

}  // extern "C"
