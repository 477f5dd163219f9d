// batched.hip -- many independent IsStochasticInteger FCIQMC runs of HubbardReal1D advanced
// together (C ABI rimu_batch_*).
//
// A step of every run has five phases, the plain engine's algorithm at the size of a run:
// spawn slices (a run's spawn attempts split evenly over K workgroups, one per 2048 target
// walkers: diagonal step and spawn attempts, ballot compaction into the slice's spawn
// buffer), the run's sort (one workgroup: gather of the slices in order, stable LSD radix
// sort with LDS histograms and ballot ranks, merge tile boundaries), merge slices (the
// run's merge tiles over K workgroups), copy slices (the tiles' survivors into the new
// vector and the next step's slice bounds, over K workgroups) and the run's final phase
// (one workgroup: norm, ProjectedEnergy values, history row and the DoubleLogUpdate shift
// update, shiftstrategy.jl:174-181), each phase one launch covering all runs, or all five
// in one persistent launch.  The step number, shift and lengths live in device memory,
// so a sequence of steps is captured once in a HIP graph and replayed.
//
// Random draws are addressed by (seed, address, attempt, step) and integer walker sums are
// exact, so the result of a step does not depend on the workgroup size or radix width.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "bosefs_device.hpp"
#include "capi_util.hpp"
#include "device_common.hpp"

using namespace rimu;

namespace {

constexpr int MIPT = 2;          // merged positions per thread per merge tile
constexpr int SPAWN_CH = 1024;   // entries per spawn chunk (SPAWN_CH / NT per thread)
// attempts per thread per spawn group (eight measured slower: more registers, one
// wavefront less per SIMD)
constexpr int AIT = 4;
constexpr double WALKERS_PER_SLICE = 2048;  // target walkers per spawn workgroup of a run
constexpr int64_t MAX_SLICES = 64;
constexpr int SPAWN_NT = 256;                // spawn workgroup size (several per CU)
constexpr int64_t MIN_TILE = 256 * MIPT;     // merge tile of the smallest workgroup size
constexpr int64_t MAX_TILE = 1024 * MIPT;    // and of the largest
constexpr int MAXPROJ = 160;     // projector entries per run (start address and neighbours)
constexpr int NHIST = 6;         // shift, norm, len, vproj, hproj, attempts

enum { RUN_OK = 0, RUN_DEAD = 1, RUN_VEC_OVERFLOW = 2, RUN_SPAWN_OVERFLOW = 3 };

struct RunState {
    int M, N, B, status;
    double u, t, dtau, shift, pnorm, target, zeta, xi;
    uint64_t seed;
    int64_t len, cur, step, nrec;     // nrec: history rows written
    int64_t capV, capS, offV, offS, offP, nproj;
    int64_t K, slice0, offX;          // spawn slices, first slice's global index, slice pool offset
    int64_t offB;                     // the slices' bounds (K + 1 per run)
    int64_t offT, offM;               // merge tile arrays, merge slot pool
    // handed from the sort kernel to the merge and final kernels of the same step
    int64_t S_cur, ntl, sbuf, flag;
    double att;
};

struct Pools {
    uint64_t* vk[2];
    int64_t* vv[2];
    int64_t* dv;
    uint64_t* sk[2];
    int64_t* sv[2];
    const uint64_t* pk;
    const double* pc;    // coefficients of hproj (vproj uses the start address, pc index 0 row)
    double* hist;        // [NHIST][hist_cap][nruns]
    int64_t hist_cap;
    uint64_t* xk;        // spawn records of slices 1..K-1 (capS each; slice 0 writes to sk[0])
    int64_t* xv;
    int64_t* xcnt;       // records of every slice
    double* xatt;        // spawn attempts of every slice
    const int* wg_run;   // run of every spawn / merge workgroup
    // Spawn slice s of a run takes the attempts [A s / K, A (s + 1) / K) of its vector (A =
    // walkers, attempt numbers run entry by entry): sbe[s] is the entry holding attempt
    // A s / K and sbp[s] the walkers before that entry; sbe[K] = length, sbp[K] = A.  Two
    // copies of totB each, like the vectors: the bounds of vector cur at cur * totB, so a
    // step that is not completed leaves the current vector's bounds as they were.
    int64_t* sbe;
    int64_t* sbp;
    int64_t totB;
    int64_t* msplit;     // per merge tile: spawns before its start, then survivor offsets
    uint64_t* mprev;     // per merge tile: merged key just before it
    int64_t* mcnt;       // per merge tile: survivors
    double* mnrm;        // per merge tile: norm of the survivors
    uint64_t* mk;        // per merge tile: survivor slots (TILE each)
    int64_t* mv;
    int64_t* mw;         // per merge slot: walkers before the survivor in its tile
};

__device__ inline int64_t isgn(double x) { return (int64_t)((x > 0) - (x < 0)); }

// Diagonal step and spawning of one slice of a run: its attempts [A s / K, A (s + 1) / K)
// (A walkers; attempts numbered entry by entry) and the entries holding them.  Records go,
// in (entry, attempt) order, to the run's spawn buffer (slice 0) or to the slice's own
// region; the run's sort gathers the slices in order.
template <int NT>
struct SpawnLds {
    int cnt[2][AIT * (NT / WAVE)];  // double-buffered by attempt group
    int tot[2];
    uint64_t k[SPAWN_CH];        // the chunk's keys
    int64_t v[SPAWN_CH];         // and values
    int64_t aoff[SPAWN_CH + 1];  // attempt offsets within the chunk (counts first)
};

template <int NT>
__device__ void spawn_phase(RunState* __restrict__ runs, const Pools& P, int64_t gslice, SpawnLds<NT>& sp,
                            int tid) {
    constexpr int NW = NT / WAVE;
    constexpr int EPT = SPAWN_CH / NT;    // entries per thread per chunk
    constexpr int CH = SPAWN_CH;
    int64_t S_run = 0;  // records of the slice so far (the same in every thread)
    int grp = 0;        // attempt groups so far: their parity picks the count buffers

    const int run = P.wg_run[gslice];
    const RunState& st = runs[run];
    if (st.status != RUN_OK) return;  // uniform across the workgroup
    const int64_t slice = gslice - st.slice0;
    const int lane = tid & (WAVE - 1);
    const int wid = tid / WAVE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;

    const int M = st.M, B = st.B;
    const double u = st.u, t = st.t, dtau = st.dtau, shift = st.shift;
    const uint64_t seed = st.seed;
    const int64_t Lrun = st.len, capS = st.capS;
    const int cur = (int)st.cur;
    const int64_t step = st.step + 1;
    // the slice's attempts [q_lo, q_hi) and the entries holding them: e_lo (whose first d0
    // attempts belong to the slice before) up to the entry holding attempt q_hi - 1
    const int64_t K = st.K;
    const int64_t* sbe = P.sbe + cur * P.totB + st.offB;
    const int64_t* sbp = P.sbp + cur * P.totB + st.offB;
    const int64_t A = sbp[K];
    const int64_t q_lo = A * slice / K, q_hi = A * (slice + 1) / K;
    const int64_t e_lo = sbe[slice], d0 = q_lo - sbp[slice];
    const int64_t e_nx = sbe[slice + 1];
    const int64_t e_hi = (e_nx < Lrun && q_hi > sbp[slice + 1]) ? e_nx + 1 : e_nx;
    int64_t left = q_hi - q_lo;  // attempts of the slice not yet taken (uniform)
    const uint64_t* vk = P.vk[cur] + st.offV + e_lo;
    const int64_t* vv = P.vv[cur] + st.offV + e_lo;
    int64_t* dv = P.dv + st.offV + e_lo;
    const int64_t L = e_hi - e_lo;
    uint64_t* sk0 = slice == 0 ? P.sk[0] + st.offS : P.xk + st.offX + (slice - 1) * capS;
    int64_t* sv0 = slice == 0 ? P.sv[0] + st.offS : P.xv + st.offX + (slice - 1) * capS;


    // ---- diagonal step and spawning, CH entries at a time -------------------------------
    // Entries are loaded striped (coalesced) and kept in LDS; attempts are numbered entry by
    // entry, so the records come out in (entry, attempt) order.
    double st_spawn = 0, st_death = 0, st_clone = 0, st_zombie = 0;
    for (int64_t c0 = 0; c0 < L; c0 += CH) {
        const int n_ch = (int)min((int64_t)CH, L - c0);
        uint64_t kq[EPT];
        int64_t vq[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {  // all loads in flight before any is used
            const int64_t e = min(c0 + j * NT + (int64_t)tid, L - 1);
            kq[j] = vk[e];
            vq[j] = vv[e];
        }
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int loc = j * NT + tid;
            if (loc < n_ch) {
                const uint64_t k = kq[j];
                const int64_t v = vq[j];
                const bool first = c0 + loc == 0;  // the slice's first entry
                if (!first || d0 == 0) {  // the entry's diagonal step: the slice of its first attempt
                    const double hd = u * (double)bose_hubbard_interaction(k) / 2;
                    const double x = (1 - dtau * (hd - shift)) * (double)v;
                    const U4 r = draw(seed, k, 0, step, TAG_DIAG);
                    const int64_t res = isgn(x) * (int64_t)floor(fabs(x) + u53(r.x, r.y));
                    dv[c0 + loc] = res;
                    if (res > v) st_clone += (double)(res - v);
                    else if (isgn((double)res) != isgn((double)v)) { st_death += fabs((double)v); st_zombie += fabs((double)res); }
                    else st_death += fabs((double)(res - v));
                }
                sp.k[loc] = k;
                sp.v[loc] = v;
                // attempts: |v| (every entry has one at least), less those of the slice before
                sp.aoff[loc] = (v < 0 ? -v : v) - (first ? d0 : 0);
            }
        }
        __syncthreads();
        // blocked scan: thread x owns chunk entries x*EPT .. x*EPT+EPT-1
        int64_t c[EPT];
        int64_t sum = 0;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int loc = tid * EPT + j;
            c[j] = loc < n_ch ? sp.aoff[loc] : 0;
            sum += c[j];
        }
        int64_t T;
        int64_t run_ = block_exclusive_scan<int64_t, NT>(sum, &T);  // ends with a barrier
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            sp.aoff[tid * EPT + j] = run_;
            run_ += c[j];
        }
        if (tid == 0) sp.aoff[CH] = T;
        // the last entry's attempts past q_hi belong to the next slice
        T = min(T, left);
        left -= T;
        __syncthreads();
        // attempts in groups of NT * AIT: one ballot per attempt, then a scan of the
        // (attempt item, wavefront) counts gives every record its place in attempt order
        for (int64_t a0 = 0; a0 < T; a0 += NT * AIT, ++grp) {
            // counts of this group; the other buffer is still read by the group before
            int* s_cnt = sp.cnt[grp & 1];
            bool ok[AIT];
            uint64_t nbk[AIT];
            int64_t nbv[AIT];
            int rpos[AIT];
#pragma unroll
            for (int it = 0; it < AIT; ++it) {
                const int64_t att = a0 + it * NT + tid;
                ok[it] = false;
                nbk[it] = 0;
                nbv[it] = 0;
                if (att < T) {
                    // the entry of the attempt: the last lo with aoff[lo] <= att (offsets
                    // rise strictly over the chunk's entries and equal T past them); a fixed
                    // number of halvings, so the AIT searches of a thread overlap
                    static_assert((CH & (CH - 1)) == 0, "power-of-two chunk");
                    int lo = 0;
#pragma unroll
                    for (int s = CH / 2; s >= 1; s >>= 1)
                        if (sp.aoff[lo + s] <= att) lo += s;
                    const uint64_t k = sp.k[lo];
                    const int64_t v = sp.v[lo];
                    const int64_t j = att - sp.aoff[lo] + (c0 + lo == 0 ? d0 : 0);  // attempt of the entry
                    const int64_t n_od = 2 * (int64_t)num_occupied(k);
                    const int64_t na = v < 0 ? -v : v;
                    const double magnitude = (double)v / (double)na;
                    const double prob = 1.0 / (double)n_od;
                    const U4 r = draw(seed, k, (uint32_t)j, step, TAG_SPAWN);
                    int64_t prod;
                    nbk[it] = real1d_hop(k, pick1(r.x, n_od), M, B, &prod);
                    const double mel = -(-t * sqrt((double)prod)) * dtau;
                    const double x = mel * magnitude / prob;
                    nbv[it] = isgn(x) * (int64_t)floor(fabs(x) + u53(r.y, r.z));
                    st_spawn += fabs((double)nbv[it]);
                    ok[it] = nbv[it] != 0;
                }
                const uint64_t bal = __ballot(ok[it]);
                rpos[it] = __popcll(bal & lt_mask);
                if (lane == 0) s_cnt[it * NW + wid] = __popcll(bal);
            }
            __syncthreads();
            // exclusive prefix of the (attempt item, wavefront) counts in one wavefront
            static_assert(AIT * NW <= WAVE, "one wavefront scans the counts");
            if (tid < WAVE) {
                const int c = tid < AIT * NW ? s_cnt[tid] : 0;
                const int inc = wave_inclusive_scan(c);
                if (tid < AIT * NW) s_cnt[tid] = inc - c;
                if (tid == AIT * NW - 1) sp.tot[grp & 1] = inc;
            }
            __syncthreads();
            const int64_t base = S_run;
#pragma unroll
            for (int it = 0; it < AIT; ++it) {
                if (ok[it]) {
                    const int64_t pos = base + s_cnt[it * NW + wid] + rpos[it];
                    if (pos < capS) { sk0[pos] = nbk[it]; sv0[pos] = nbv[it]; }
                }
            }
            S_run += sp.tot[grp & 1];
        }
        __syncthreads();  // the chunk's entries in LDS are read until here
    }
    // record count (it may exceed capS: the gather then reports the overflow) and attempts
    if (tid == 0) {
        P.xcnt[st.slice0 + slice] = S_run;
        P.xatt[st.slice0 + slice] = (double)(q_hi - q_lo);
    }
    (void)st_spawn; (void)st_death; (void)st_clone; (void)st_zombie; (void)u;
}

template <int NT>
__global__ __launch_bounds__(NT) void batch_spawn_kernel(RunState* __restrict__ runs, Pools P) {
    __shared__ SpawnLds<NT> sp;
    spawn_phase<NT>(runs, P, blockIdx.x, sp, threadIdx.x);
}

// Gather of the slices' spawn records, their radix sort, and the boundaries of every merge
// tile (one workgroup per run).  Hands S, the sorted buffer and the tile count to the next
// kernels through the run state.
constexpr int SORT_LDS_CAP = 4096;  // records a run's sort keeps in LDS for all its passes

template <int NT, bool LS>
struct SortLds {
    uint32_t hist[2][256];              // digit offsets, double-buffered by chunk
    uint16_t wcnt[2][NT / WAVE][256];   // (wavefront, digit) counts, double-buffered by chunk
    int64_t xoff[MAX_SLICES + 1];  // first record of each slice in the gathered list
    double att;
    int flag;
    uint64_t lk[LS ? 2 : 1][LS ? SORT_LDS_CAP : 1];  // LS: both buffers of the passes
    int64_t lv[LS ? 2 : 1][LS ? SORT_LDS_CAP : 1];
};

// LS: a run of at most SORT_LDS_CAP records is gathered into LDS, sorted there and written
// once to its spawn buffer; larger runs pass through their global buffers.
template <int NT, bool LS>
__device__ void sort_phase(RunState* __restrict__ runs, const Pools& P, int run, int radix_bits, SortLds<NT, LS>& so,
                           int tid) {
    constexpr int NW = NT / WAVE;
    constexpr int TILE = NT * MIPT;
    uint32_t* s_hist = so.hist[0];
    int& s_flag = so.flag;

    RunState& st = runs[run];
    if (st.status != RUN_OK) return;  // uniform across the workgroup
    const int lane = tid & (WAVE - 1);
    const int wid = tid / WAVE;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int B = st.B;
    const int64_t L = st.len, capS = st.capS;
    const int cur = (int)st.cur;
    const uint64_t* vk = P.vk[cur] + st.offV;
    uint64_t* sk0 = P.sk[0] + st.offS;
    int64_t* sv0 = P.sv[0] + st.offS;
    uint64_t* sk1 = P.sk[1] + st.offS;
    int64_t* sv1 = P.sv[1] + st.offS;

    if (tid == 0) s_flag = 0;
    __syncthreads();

    // ---- the slices' spawn records, gathered after slice 0's in slice order -------------
    // slice counts and attempts: lane q of the first wavefront reads slice q (K <= 64)
    const int K = (int)st.K;
    if (tid < WAVE) {
        const int64_t c = tid < K ? P.xcnt[st.slice0 + tid] : 0;
        double a = tid < K ? P.xatt[st.slice0 + tid] : 0.0;
        const int64_t inc = wave_inclusive_scan(c);
        so.xoff[tid + 1] = inc;
        if (tid == 0) so.xoff[0] = 0;
#pragma unroll
        for (int off = WAVE / 2; off > 0; off >>= 1) a += __shfl_xor(a, off, WAVE);  // integers: exact
        if (tid == 0) so.att = a;
    }
    __syncthreads();
    int64_t S = so.xoff[K];
    const double att_sum = so.att;
    // LS: a run of at most SORT_LDS_CAP records is gathered into LDS and sorted there
    const bool in_lds = LS && S <= SORT_LDS_CAP && S <= capS;  // uniform
    uint64_t* gk = in_lds ? so.lk[0] : sk0;
    int64_t* gv = in_lds ? so.lv[0] : sv0;
    if (in_lds) {  // slice 0's records, written in place by its spawn workgroup
        for (int64_t i = tid; i < min(so.xoff[1], S); i += NT) {
            gk[i] = sk0[i];
            gv[i] = sv0[i];
        }
    }
    if (S > capS) {
        if (tid == 0) s_flag = RUN_SPAWN_OVERFLOW;
        S = capS;
    } else {
        // record j >= (slice 0's count) comes from slice q, xoff[q] <= j < xoff[q + 1]; four
        // records per thread at a time, all loads in flight before the stores
        constexpr int CU = 4;
        for (int64_t j0 = so.xoff[1]; j0 < S; j0 += CU * NT) {
            uint64_t kq[CU];
            int64_t vq[CU];
#pragma unroll
            for (int c = 0; c < CU; ++c) {
                const int64_t j = j0 + c * NT + tid;
                if (j < S) {
                    // the last slice lo with xoff[lo] <= j (xoff[q] = S for q >= K): six
                    // halvings, so the CU searches of a thread overlap
                    static_assert(MAX_SLICES == 64, "six halvings");
                    int lo = 0;
#pragma unroll
                    for (int s = 32; s >= 1; s >>= 1)
                        if (so.xoff[lo + s] <= j) lo += s;
                    const int64_t src = st.offX + (int64_t)(lo - 1) * capS + (j - so.xoff[lo]);
                    kq[c] = P.xk[src];
                    vq[c] = P.xv[src];
                }
            }
#pragma unroll
            for (int c = 0; c < CU; ++c) {
                const int64_t j = j0 + c * NT + tid;
                if (j < S) {
                    gk[j] = kq[c];
                    gv[j] = vq[c];
                }
            }
        }
    }
    __syncthreads();

    // ---- stable LSD radix sort of the spawn records on bits [0, B) ------------------------
    const int nbins = 1 << radix_bits;
    uint64_t* srck = gk;
    int64_t* srcv = gv;
    uint64_t* dstk = in_lds ? so.lk[LS ? 1 : 0] : sk1;
    int64_t* dstv = in_lds ? so.lv[LS ? 1 : 0] : sv1;
    for (int sh = 0; sh < B; sh += radix_bits) {
        const uint64_t dmask = (uint64_t)(nbins - 1);
        for (int d = tid; d < nbins; d += NT) s_hist[d] = 0;
        for (int j = tid; j < NW * nbins; j += NT) so.wcnt[0][j / nbins][j % nbins] = 0;
        __syncthreads();
        for (int64_t i = tid; i < S; i += NT) atomicAdd(&s_hist[(srck[i] >> sh) & dmask], 1u);
        __syncthreads();
        {  // exclusive digit offsets (at most 256 bins <= NT threads)
            const uint32_t c = tid < nbins ? s_hist[tid] : 0u;
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan<uint32_t, NT>(c, &tot);  // ends with a barrier
            if (tid < nbins) s_hist[tid] = ex;
        }
        __syncthreads();
        // chunk c reads offsets hist[c & 1] and counts wcnt[c & 1], writes the next chunk's
        // offsets into hist[1 - (c & 1)] and clears wcnt[1 - (c & 1)] (last read by chunk
        // c - 1 before this chunk's first barrier): two barriers per chunk
        int p = 0;
        for (int64_t t0 = 0; t0 < S; t0 += NT, p ^= 1) {
            const int64_t i = t0 + tid;
            const bool valid = i < S;
            const uint64_t key = valid ? srck[i] : 0;
            const int64_t val = valid ? srcv[i] : 0;
            const uint32_t digit = (uint32_t)((key >> sh) & dmask);
            uint64_t peers = __ballot(valid);
            for (int b = 0; b < radix_bits; ++b) {
                const uint64_t bb = __ballot((digit >> b) & 1u);
                peers &= ((digit >> b) & 1u) ? bb : ~bb;
            }
            const uint32_t in_wave = (uint32_t)__popcll(peers & lt_mask);
            if (valid && in_wave == 0) so.wcnt[p][wid][digit] = (uint16_t)__popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = so.hist[p][digit] + in_wave;
                for (int w = 0; w < wid; ++w) pos += so.wcnt[p][w][digit];
                dstk[pos] = key;
                dstv[pos] = val;
            }
            for (int d = tid; d < nbins; d += NT) {
                uint32_t add = 0;
                for (int w = 0; w < NW; ++w) {
                    add += so.wcnt[p][w][d];
                    so.wcnt[p ^ 1][w][d] = 0;
                }
                so.hist[p ^ 1][d] = so.hist[p][d] + add;
            }
            __syncthreads();
        }
        uint64_t* tk = srck; srck = dstk; dstk = tk;
        int64_t* tv = srcv; srcv = dstv; dstv = tv;
    }
    if (in_lds) {  // the sorted records to the run's spawn buffer, where the merge reads them
        for (int64_t i = tid; i < S; i += NT) {
            sk0[i] = srck[i];
            sv0[i] = srcv[i];
        }
    }

    // ---- merge tile boundaries: spawns before each tile start and the key before it -------
    const uint64_t* ak = srck;
    const int64_t ntl = (S + L + TILE - 1) / TILE;
    for (int64_t q = tid; q <= ntl; q += NT) {
        const int64_t d = min(q * TILE, S + L);
        int64_t lo = max((int64_t)0, d - L), hi = min(d, S);
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (ak[mid] < vk[d - mid - 1]) lo = mid + 1; else hi = mid;
        }
        P.msplit[st.offT + q] = lo;
        const int64_t bi = d - lo;
        const uint64_t pa = lo > 0 ? ak[lo - 1] : 0, pb = bi > 0 ? vk[bi - 1] : 0;
        P.mprev[st.offT + q] = (lo > 0 && bi > 0) ? max(pa, pb) : (lo > 0 ? pa : pb);
    }
    if (tid == 0) {
        st.S_cur = S;
        st.ntl = ntl;
        st.sbuf = (in_lds || srck == sk0) ? 0 : 1;
        st.flag = s_flag;
        st.att = att_sum;
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void batch_sort_kernel(RunState* __restrict__ runs, Pools P, int radix_bits) {
    __shared__ SortLds<NT, true> so;
    sort_phase<NT, true>(runs, P, blockIdx.x, radix_bits, so, threadIdx.x);
}

// Merge of the sorted spawns (A) with the diagonal survivors (B): workgroup w of a run takes
// the tiles [ntl*w/K, ntl*(w+1)/K).  Merged order: by key, B before A on equal keys; each
// address folds diag + spawns.  Survivors go to the tile's slots with their count and norm.
template <int NT>
struct MergeLds {
    uint64_t a[NT * MIPT], b[NT * MIPT];   // tile keys: spawns, vector entries
    int64_t av[NT * MIPT], bv[NT * MIPT];  // their values (diagonal survivors for b)
};

template <int NT>
__device__ void merge_phase(RunState* __restrict__ runs, const Pools& P, int64_t gslice, MergeLds<NT>& me, int tid) {
    constexpr int TILE = NT * MIPT;
    uint64_t* s_a = me.a;
    uint64_t* s_b = me.b;
    int64_t* s_av = me.av;
    int64_t* s_bv = me.bv;

    const int run = P.wg_run[gslice];
    const RunState& st = runs[run];
    if (st.status != RUN_OK) return;  // uniform across the workgroup
    const int64_t w = gslice - st.slice0;
    const int64_t ntl = st.ntl, S = st.S_cur, L = st.len;
    const int64_t t_lo = ntl * w / st.K, t_hi = ntl * (w + 1) / st.K;
    const int cur = (int)st.cur;
    const uint64_t* vk = P.vk[cur] + st.offV;
    const int64_t* dv = P.dv + st.offV;
    const uint64_t* ak = P.sk[st.sbuf] + st.offS;
    const int64_t* av = P.sv[st.sbuf] + st.offS;
    uint64_t* mk = P.mk + st.offM;
    int64_t* mv = P.mv + st.offM;
    int64_t* mw = P.mw + st.offM;

    for (int64_t tl = t_lo; tl < t_hi; ++tl) {
        const int64_t d0 = tl * TILE;
        const int64_t d1 = min(d0 + TILE, S + L);
        const int64_t a0 = P.msplit[st.offT + tl], a1 = P.msplit[st.offT + tl + 1];
        const uint64_t tile_prev = P.mprev[st.offT + tl];
        const int64_t b0 = d0 - a0, b1 = d1 - a1;
        const int na = (int)(a1 - a0), nb = (int)(b1 - b0);
        {
            uint64_t rk[MIPT];
            int64_t rv[MIPT];
#pragma unroll
            for (int r = 0; r < MIPT; ++r) {  // all loads in flight before any is used
                const int i = tid + r * NT;
                if (i < na + nb) {
                    const bool in_a = i < na;
                    const uint64_t* kp = in_a ? ak + (a0 + i) : vk + (b0 + (i - na));
                    const int64_t* vp = in_a ? av + (a0 + i) : dv + (b0 + (i - na));
                    rk[r] = *kp;
                    rv[r] = *vp;
                }
            }
#pragma unroll
            for (int r = 0; r < MIPT; ++r) {
                const int i = tid + r * NT;
                if (i < na) { s_a[i] = rk[r]; s_av[i] = rv[r]; }
                else if (i < na + nb) { s_b[i - na] = rk[r]; s_bv[i - na] = rv[r]; }
            }
        }
        __syncthreads();
        const int dd = tid * MIPT;
        const int n_here = (int)max((int64_t)0, min((int64_t)MIPT, (d1 - d0) - dd));
        int ia;
        {
            int lo = max(0, dd - nb), hi = min(dd, na);
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_a[mid] < s_b[dd - mid - 1]) lo = mid + 1; else hi = mid;
            }
            ia = lo;
        }
        int ib = dd - ia;
        bool have_prev = false;
        uint64_t prev = 0;
        if (dd > 0 && n_here > 0) {
            have_prev = true;
            const uint64_t pa = ia > 0 ? s_a[ia - 1] : 0, pb = ib > 0 ? s_b[ib - 1] : 0;
            prev = (ia > 0 && ib > 0) ? max(pa, pb) : (ia > 0 ? pa : pb);
        } else if (d0 > 0 && n_here > 0) {
            have_prev = true;
            prev = tile_prev;
        }
        uint64_t ok_k[MIPT];
        int64_t ok_v[MIPT];
        int cnt = 0;
        int64_t wk = 0;  // walkers of the thread's survivors
#pragma unroll
        for (int j = 0; j < MIPT; ++j) {
            if (j < n_here) {
                const bool take_b = ib < nb && (ia >= na || s_b[ib] <= s_a[ia]);
                const uint64_t key = take_b ? s_b[ib] : s_a[ia];
                if (!have_prev || key != prev) {
                    int64_t acc = 0;
                    if (take_b) acc += s_bv[ib];
                    int la = ia;
                    while (la < na && s_a[la] == key) { acc += s_av[la]; ++la; }
                    if (la == na) {  // the run may continue past the tile
                        int64_t ga = a0 + la;
                        while (ga < S && ak[ga] == key) { acc += av[ga]; ++ga; }
                    }
                    if (acc != 0) {
                        ok_k[cnt] = key;
                        ok_v[cnt] = acc;
                        ++cnt;
                        wk += acc < 0 ? -acc : acc;
                    }
                }
                prev = key;
                have_prev = true;
                if (take_b) ++ib; else ++ia;
            }
        }
        // one scan of (walkers << 16 | survivors): places, the walkers before every survivor
        // in the tile, and the tile's count and norm (at most 2 NT < 2^16 survivors, so the
        // count never carries into the walkers)
        int64_t ptot;
        const int64_t pex = block_exclusive_scan<int64_t, NT>(wk << 16 | cnt, &ptot);  // ends with a barrier
        const int ex = (int)(pex & 0xffff);
        int64_t wx = pex >> 16;
        for (int j = 0; j < cnt; ++j) {
            mk[d0 + ex + j] = ok_k[j];
            mv[d0 + ex + j] = ok_v[j];
            mw[d0 + ex + j] = wx;
            wx += ok_v[j] < 0 ? -ok_v[j] : ok_v[j];
        }
        if (tid == 0) {
            P.mcnt[st.offT + tl] = ptot & 0xffff;
            P.mnrm[st.offT + tl] = (double)(ptot >> 16);
        }
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void batch_merge_kernel(RunState* __restrict__ runs, Pools P) {
    __shared__ MergeLds<NT> me;
    merge_phase<NT>(runs, P, blockIdx.x, me, threadIdx.x);
}

// The tiles' survivors into the new vector: copy slice w of a run takes the tiles
// [ntl*w/K, ntl*(w+1)/K), their offset the survivors of the tiles before.  With the walkers
// before every entry (those of the tiles before, and before it in its tile from the merge)
// it writes the next step's spawn slice bounds that fall in its entries.
template <int NT>
__device__ void copy_phase(RunState* __restrict__ runs, const Pools& P, int64_t gslice, int tid) {
    constexpr int TILE = NT * MIPT;
    constexpr int CU = 4;  // elements per thread at a time, all loads in flight before the stores
    const int run = P.wg_run[gslice];
    const RunState& st = runs[run];
    if (st.status != RUN_OK) return;  // uniform across the workgroup
    const int64_t w = gslice - st.slice0;
    const int64_t ntl = st.ntl, capB = st.capV + st.capS;  // the vector buffer holds L + S
    const int64_t t_lo = ntl * w / st.K, t_hi = ntl * (w + 1) / st.K;
    if (t_lo >= t_hi) return;
    const int cur = (int)st.cur;
    uint64_t* nk = P.vk[1 - cur] + st.offV;
    int64_t* nv = P.vv[1 - cur] + st.offV;
    const uint64_t* mk = P.mk + st.offM;
    const int64_t* mv = P.mv + st.offM;
    const int64_t* mw = P.mw + st.offM;
    // survivors and walkers before the range, and all walkers (tile norms: exact integers),
    // summed by the first wavefront
    __shared__ int64_t s_pre[3];
    if (tid < WAVE) {
        int64_t part = 0, wpart = 0, wall = 0;
        for (int64_t t = tid; t < ntl; t += WAVE) {
            const int64_t m = (int64_t)P.mnrm[st.offT + t];
            if (t < t_lo) {
                part += P.mcnt[st.offT + t];
                wpart += m;
            }
            wall += m;
        }
        part = wave_inclusive_scan(part);
        wpart = wave_inclusive_scan(wpart);
        wall = wave_inclusive_scan(wall);
        if (tid == WAVE - 1) {
            s_pre[0] = part;
            s_pre[1] = wpart;
            s_pre[2] = wall;
        }
    }
    __syncthreads();
    int64_t o = s_pre[0], W = s_pre[1];
    const int64_t A = s_pre[2];
    const int64_t K = st.K;
    int64_t* sbe = P.sbe + (1 - cur) * P.totB + st.offB;  // the new vector's bounds
    int64_t* sbp = P.sbp + (1 - cur) * P.totB + st.offB;
    for (int64_t t = t_lo; t < t_hi; ++t) {
        const int64_t n = P.mcnt[st.offT + t];
        if (o + n > capB) break;  // not reached: L <= capV and S <= capS
        for (int64_t i0 = 0; i0 < n; i0 += CU * NT) {
            uint64_t kq[CU];
            int64_t vq[CU], wq[CU];
#pragma unroll
            for (int c = 0; c < CU; ++c) {
                const int64_t i = i0 + c * NT + tid;
                if (i < n) {
                    kq[c] = mk[t * TILE + i];
                    vq[c] = mv[t * TILE + i];
                    wq[c] = mw[t * TILE + i];
                }
            }
#pragma unroll
            for (int c = 0; c < CU; ++c) {
                const int64_t i = i0 + c * NT + tid;
                if (i < n) {
                    nk[o + i] = kq[c];
                    nv[o + i] = vq[c];
                    const int64_t Pg = W + wq[c];  // walkers before the entry
                    const int64_t a = vq[c] < 0 ? -vq[c] : vq[c];
                    // slices q whose first attempt floor(A q / K) lies in [Pg, Pg + a), i.e.
                    // Pg K <= A q < (Pg + a) K.  The smallest is ceil(Pg K / A), exact in
                    // doubles while A K < 2^53: a quotient that is not an integer lies at
                    // least 1 / A from one, far beyond the rounding of the division.
                    int64_t q = max((int64_t)1, (int64_t)ceil((double)(Pg * K) / (double)A));
                    for (; q < K && A * q < (Pg + a) * K; ++q) {
                        sbe[q] = o + i;
                        sbp[q] = Pg;
                    }
                }
            }
        }
        W += (int64_t)P.mnrm[st.offT + t];
        o += n;
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void batch_copy_kernel(RunState* __restrict__ runs, Pools P) {
    copy_phase<NT>(runs, P, blockIdx.x, threadIdx.x);
}

// ProjectedEnergy, statistics, the history row and the DoubleLogUpdate shift (one
// workgroup per run), from the new vector that copy_phase wrote.
template <int NT>
__device__ void final_phase(RunState* __restrict__ runs, int nruns, const Pools& P, int run, int tid) {
    RunState& st = runs[run];
    if (st.status != RUN_OK) return;  // uniform across the workgroup
    const double dtau = st.dtau, shift = st.shift;
    const int64_t capV = st.capV, ntl = st.ntl;
    if (st.flag == RUN_SPAWN_OVERFLOW) {  // uniform: the step is void, no history row
        if (tid == 0) st.status = RUN_SPAWN_OVERFLOW;
        return;
    }
    const int cur = (int)st.cur;
    const int64_t step = st.step + 1;
    const uint64_t* nk = P.vk[1 - cur] + st.offV;  // the new vector (copy_phase)
    const int64_t* nv = P.vv[1 - cur] + st.offV;

    // survivors of all tiles
    int64_t out = 0;
    for (int64_t t0 = 0; t0 < ntl; t0 += NT) {
        const int64_t t = t0 + tid;
        const int64_t c = t < ntl ? P.mcnt[st.offT + t] : 0;
        int64_t tot;
        (void)block_exclusive_scan<int64_t, NT>(c, &tot);
        out += tot;
    }
    double nrm = 0;  // integer walkers: the sum of the tile norms is exact in any order
    for (int64_t t = tid; t < ntl; t += NT) nrm += P.mnrm[st.offT + t];
    // the whole new vector: a step that outgrows max_length (capV) is completed, reported
    // and kept, and then the run stops
    const int64_t Lnew = out;
    const double att_sum = st.att;

    // ---- ProjectedEnergy, statistics and the shift update -----------------------------
    double vp = 0, hp = 0;
    {
        // each projector key's first position >= key, found by a group of G lanes: every
        // round samples G positions of [lo, hi), so a round of dependent loads divides it by
        // G + 1 (about four rounds instead of a binary search's 15 on the sweep's largest runs)
        constexpr int G = 16;
        __shared__ double s_x[MAXPROJ];
        const int64_t np = st.nproj;
        const int g = tid / G, l = tid % G;
        const int gshift = (tid % WAVE) / G * G;  // the group's bits in a wavefront ballot
        for (int64_t q0 = 0; q0 < np; q0 += NT / G) {
            const int64_t q = q0 + g;
            const bool act = q < np;  // uniform over the group
            const uint64_t key = act ? P.pk[st.offP + q] : 0;
            int64_t lo = 0, hi = act ? Lnew : 0;  // nk[lo - 1] < key <= nk[hi] (where they exist)
            while (__ballot(hi > lo)) {
                const int64_t span = hi - lo;
                const int64_t sp_l = lo + span * (l + 1) / (G + 1);  // in [lo, hi) when span > 0
                const bool ge = span > 0 && nk[sp_l] >= key;
                const uint32_t m = (uint32_t)(__ballot(ge) >> gshift) & ((1u << G) - 1u);
                if (span > 0) {
                    if (m == 0) {
                        lo = lo + span * G / (G + 1) + 1;  // past the last sample
                    } else {
                        const int f = __builtin_ctz(m);  // the first sample >= key
                        hi = lo + span * (f + 1) / (G + 1);
                        if (f > 0) lo = lo + span * f / (G + 1) + 1;
                    }
                }
            }
            if (act && l == 0) s_x[q] = (lo < Lnew && nk[lo] == key) ? (double)nv[lo] : 0.0;
        }
        __syncthreads();
        for (int64_t q = tid; q < np; q += NT) {
            const double x = s_x[q];
            hp += P.pc[st.offP + q] * x;
            if (q == 0) vp = x;   // entry 0 is the start address itself
        }
    }
    block_sum3<NT>(nrm, vp, hp);
    if (tid == 0) {
        double sh = shift;
        if (Lnew > 0) sh = shift - (st.xi / dtau * log(nrm / st.target) + st.zeta / dtau * log(nrm / st.pnorm));
        const int64_t row = st.nrec;
        if (row < P.hist_cap) {
            const int64_t base = row * nruns + run;
            const int64_t plane = P.hist_cap * nruns;
            P.hist[0 * plane + base] = sh;
            P.hist[1 * plane + base] = nrm;
            P.hist[2 * plane + base] = (double)Lnew;
            P.hist[3 * plane + base] = vp;
            P.hist[4 * plane + base] = hp;
            P.hist[5 * plane + base] = att_sum;
            st.nrec = row + 1;
        }
        if (Lnew == 0) {
            st.status = RUN_DEAD;
        } else {
            st.shift = sh;
            st.pnorm = nrm;
            st.len = Lnew;
            st.cur = 1 - cur;
            // the end bound of the new vector, whose other bounds the copy phase wrote
            P.sbe[(1 - cur) * P.totB + st.offB + st.K] = Lnew;
            P.sbp[(1 - cur) * P.totB + st.offB + st.K] = (int64_t)nrm;
            if (Lnew > capV) st.status = RUN_VEC_OVERFLOW;
        }
        st.step = step;
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void batch_final_kernel(RunState* __restrict__ runs, int nruns, Pools P) {
    final_phase<NT>(runs, nruns, P, blockIdx.x, threadIdx.x);
}

// ---------------------------------------------------------------------------------------
// The whole step in one persistent launch, each run on one XCD.  The runs are dealt to the
// XCDs (largest first, to the least loaded); a workgroup reads the id of the XCD it runs on
// and takes work items from that XCD's queue, in the order: the spawn slices of its runs,
// their sorts, merge slices, copy slices and final phases.  An item waits until the items
// of its run's phase before have finished (a sort for the K spawn slices, a merge for the
// sort, a copy for the K merge slices, the final phase for the K copy slices).  Those items
// hold lower tickets of the same queue, so they were taken by running workgroups of the
// same XCD that wait only on lower tickets themselves: every wait ends, whatever the grid
// size.  All hand-overs of a run stay inside one XCD, whose L2 every CU of it shares: a
// producing workgroup waits for its stores to reach the L2 before it signals, a consuming
// one invalidates its CU's L1 before it reads.  Every wait is bounded (spin_limit rounds of
// s_sleep); a wait that gives up raises the abort word, after which every item is skipped
// and the host reports the error.  The last workgroup to run out of work clears the queue
// counters for the next launch; a run's final phase clears its phase counters.
// ---------------------------------------------------------------------------------------
constexpr int MAX_XCD = 16;
struct StepSync {
    uint32_t* words;         // [0, MAX_XCD) queue tickets, [MAX_XCD] workgroups done,
                             // [MAX_XCD + 1] abort; from [MAX_XCD + 4], 4 per run: spawn,
                             // sort, merge and copy items done
    const int* items;        // per queue, in ticket order: phase << 24 | slice or run
    const int* queue_start;  // queue q holds items[queue_start[q], queue_start[q + 1])
    int nq;                  // queues (XCDs)
    uint32_t spin_limit;
    int withhold;            // test of the give-up path: the first run of queue 0 never
                             // reports its spawn slices
};

template <int NT>
union StepLds {
    SpawnLds<NT> sp;
    SortLds<NT, false> so;
    MergeLds<NT> me;
};

__device__ inline uint32_t sync_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void sync_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t sync_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// id of the XCD this wavefront runs on
__device__ inline int xcd_id() {
    int x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & 15;
}
// The thread index, read afresh: values derived from it stay inside the phase that
// computes them instead of being hoisted out of the persistent loop and held in registers
// across all phases.
__device__ inline int thread_index() {
    int t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
    return t;
}

template <int NT>
__global__ __launch_bounds__(NT, 4) void batch_step_kernel(RunState* __restrict__ runs, int nruns, Pools P,
                                                           int radix_bits, StepSync C) {
    __shared__ StepLds<NT> sm;
    __shared__ uint32_t s_ticket;
    __shared__ int s_go;
    uint32_t* const done = C.words + MAX_XCD;
    uint32_t* const abort_word = C.words + MAX_XCD + 1;
    uint32_t* const rs = C.words + MAX_XCD + 4;
    const int q = __builtin_amdgcn_readfirstlane(xcd_id() % C.nq);
    const int q0 = C.queue_start[q], qn = C.queue_start[q + 1] - q0;
    while (true) {
        if (threadIdx.x == 0) s_ticket = sync_add(C.words + q, 1u);
        __syncthreads();
        const int t = (int)__builtin_amdgcn_readfirstlane(s_ticket);
        if (t >= qn) {
            if (threadIdx.x == 0 && sync_add(done, 1u) == gridDim.x - 1) {
                // every workgroup has drawn past its queue: the counters start the next launch at zero
                for (int j = 0; j < C.nq; ++j) sync_store(C.words + j, 0u);
                sync_store(done, 0u);
            }
            return;
        }
        const int code = C.items[q0 + t];
        const int phase = code >> 24, idx = code & 0xffffff;
        const int run = (phase == 1 || phase == 4) ? idx : P.wg_run[idx];
        uint32_t* const word = rs + 4 * run;
        if (threadIdx.x == 0) {
            // the run's status is that of the last launch until its final phase, which runs
            // after every other item of the run
            int go = runs[run].status == RUN_OK && sync_load(abort_word) == 0 ? 1 : 0;
            if (go && phase > 0) {
                const uint32_t need = phase == 2 ? 1u : (uint32_t)runs[run].K;
                const uint32_t* w = word + (phase - 1);
                uint32_t spins = 0;
                while (sync_load(w) < need) {
                    if (sync_load(abort_word) != 0) { go = 0; break; }
                    if (++spins > C.spin_limit) {
                        __hip_atomic_fetch_or(abort_word, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        go = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            if (go && phase > 0) {
                // the earlier phase's writes: invalidate this CU's L1 (and this XCD's L2) once,
                // and wait for the invalidation before the barrier releases the other waves
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                __builtin_amdgcn_s_waitcnt(0);
            }
            s_go = go;
        }
        __syncthreads();
        const bool go = __builtin_amdgcn_readfirstlane(s_go) != 0;
        __syncthreads();  // s_ticket and s_go are written again by the next round
        if (!go) continue;
        if (phase == 0) spawn_phase<NT>(runs, P, idx, sm.sp, thread_index());
        else if (phase == 1) sort_phase<NT, false>(runs, P, run, radix_bits, sm.so, thread_index());
        else if (phase == 2) merge_phase<NT>(runs, P, idx, sm.me, thread_index());
        else if (phase == 3) copy_phase<NT>(runs, P, idx, thread_index());
        else final_phase<NT>(runs, nruns, P, run, thread_index());
        __builtin_amdgcn_s_waitcnt(0);  // this wavefront's stores have reached the L2
        __syncthreads();
        if (threadIdx.x == 0) {
            if (phase == 0) {
                if (!(C.withhold && q == 0 && run == P.wg_run[C.items[C.queue_start[0]] & 0xffffff])) sync_add(word, 1u);
            } else if (phase == 1) {
                sync_store(word + 1, 1u);
            } else if (phase == 2) {
                sync_add(word + 2, 1u);
            } else if (phase == 3) {
                sync_add(word + 3, 1u);
            } else {  // every item of the run has finished: its counters start the next step at zero
                sync_store(word, 0u);
                sync_store(word + 1, 0u);
                sync_store(word + 2, 0u);
                sync_store(word + 3, 0u);
            }
        }
    }
}

// Largest XCD id a workgroup reports (the host takes that + 1 as the number of queues).
__global__ void xcd_count_kernel(int* out) {
    if (threadIdx.x == 0) atomicMax(out, xcd_id());
}

// Projector of every run: the start address (coefficient diag) and its neighbours.
__global__ void batch_projector_kernel(RunState* __restrict__ runs, int nruns, const uint64_t* __restrict__ start,
                                       uint64_t* __restrict__ pk, double* __restrict__ pc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    RunState& st = runs[r];
    const uint64_t a = start[r];
    uint64_t* k = pk + st.offP;
    double* c = pc + st.offP;
    k[0] = a;
    c[0] = st.u * (double)bose_hubbard_interaction(a) / 2;  // hproj = H' * projector
    const int64_t n_od = 2 * (int64_t)num_occupied(a);
    int64_t n = 1;
    for (int64_t ch = 1; ch <= n_od; ++ch) {
        int64_t prod;
        const uint64_t nb = real1d_hop(a, ch, st.M, st.B, &prod);
        const double val = -st.t * sqrt((double)prod);
        int64_t q = 1;
        while (q < n && k[q] != nb) ++q;  // repeated targets (M = 2) are summed
        if (q == n) { k[n] = nb; c[n] = val; ++n; }
        else c[q] += val;
    }
    st.nproj = n;
}

}  // namespace

struct rimu_batch {
    int nruns = 0;
    int block = 512, radix = 8;
    int persistent = 0;             // 1: the step is one launch of batch_step_kernel
    uint32_t spin_limit = LB_SPIN_LIMIT;
    int withhold = 0;               // test of the give-up path (rimu_batch_set_wait)
    int n_cu = 0, nq = 1;           // compute units, XCDs (work queues of the persistent step)
    bool persistent_ok = true;      // every XCD has its own queue
    int occupancy[3] = {1, 1, 1};   // resident batch_step_kernel workgroups per CU, by block size
    DevBuf<uint32_t> sync;          // StepSync words
    DevBuf<int> items, queue_start; // the persistent step's work queues
    hipStream_t stream = nullptr;
    std::vector<RunState> host_init;
    DevBuf<RunState> runs;
    DevBuf<uint64_t> vk[2], sk[2], pk, start;
    DevBuf<int64_t> vv[2], dv, sv[2];
    DevBuf<double> pc, hist;
    DevBuf<uint64_t> xk;
    DevBuf<int64_t> xv, xcnt;
    DevBuf<double> xatt;
    DevBuf<int> wg_run;
    DevBuf<int64_t> msplit, mcnt, mv, mw;
    DevBuf<uint64_t> mprev, mk;
    DevBuf<double> mnrm;
    DevBuf<int64_t> sbe, sbp;       // spawn slice bounds
    int64_t totV = 0, totS = 0, totP = 0, totX = 0, totK = 0, totT = 0, totM = 0, totB = 0, hist_cap = 0;
    std::map<std::pair<int, std::pair<int, int>>, hipGraphExec_t> graphs;  // (steps, (block, radix)); cleared
                                                                           // when the launch kind changes
    Pools pools() {
        Pools P{};
        for (int b = 0; b < 2; ++b) {
            P.vk[b] = vk[b].p;
            P.vv[b] = vv[b].p;
            P.sk[b] = sk[b].p;
            P.sv[b] = sv[b].p;
        }
        P.dv = dv.p;
        P.pk = pk.p;
        P.pc = pc.p;
        P.hist = hist.p;
        P.hist_cap = hist_cap;
        P.xk = xk.p;
        P.xv = xv.p;
        P.xcnt = xcnt.p;
        P.xatt = xatt.p;
        P.wg_run = wg_run.p;
        P.sbe = sbe.p;
        P.sbp = sbp.p;
        P.totB = totB;
        P.mw = mw.p;
        P.msplit = msplit.p;
        P.mprev = mprev.p;
        P.mcnt = mcnt.p;
        P.mnrm = mnrm.p;
        P.mk = mk.p;
        P.mv = mv.p;
        return P;
    }
};

namespace {

int sync_words(const rimu_batch* b) { return MAX_XCD + 4 + 4 * b->nruns; }

template <int NT>
void launch_persistent(rimu_batch* b, int occ) {
    const int64_t total = 3 * b->totK + 2 * (int64_t)b->nruns;
    const int64_t grid = std::max<int64_t>(b->nq, std::min<int64_t>(total, (int64_t)b->n_cu * occ));
    const StepSync C{b->sync.p, b->items.p, b->queue_start.p, b->nq, b->spin_limit, b->withhold};
    batch_step_kernel<NT><<<grid, NT, 0, b->stream>>>(b->runs.p, b->nruns, b->pools(), b->radix, C);
}

int launch_step(rimu_batch* b) {
    const Pools P = b->pools();
    const int K = (int)b->totK;
    if (b->persistent) {
        switch (b->block) {
            case 256: launch_persistent<256>(b, b->occupancy[0]); break;
            case 1024: launch_persistent<1024>(b, b->occupancy[2]); break;
            default: launch_persistent<512>(b, b->occupancy[1]); break;
        }
        HIP_TRY(hipGetLastError());
        return RIMU_OK;
    }
    batch_spawn_kernel<SPAWN_NT><<<K, SPAWN_NT, 0, b->stream>>>(b->runs.p, P);
    switch (b->block) {
        case 256:
            batch_sort_kernel<256><<<b->nruns, 256, 0, b->stream>>>(b->runs.p, P, b->radix);
            batch_merge_kernel<256><<<K, 256, 0, b->stream>>>(b->runs.p, P);
            batch_copy_kernel<256><<<K, 256, 0, b->stream>>>(b->runs.p, P);
            batch_final_kernel<256><<<b->nruns, 256, 0, b->stream>>>(b->runs.p, b->nruns, P);
            break;
        case 1024:
            batch_sort_kernel<1024><<<b->nruns, 1024, 0, b->stream>>>(b->runs.p, P, b->radix);
            batch_merge_kernel<1024><<<K, 1024, 0, b->stream>>>(b->runs.p, P);
            batch_copy_kernel<1024><<<K, 1024, 0, b->stream>>>(b->runs.p, P);
            batch_final_kernel<1024><<<b->nruns, 1024, 0, b->stream>>>(b->runs.p, b->nruns, P);
            break;
        default:
            batch_sort_kernel<512><<<b->nruns, 512, 0, b->stream>>>(b->runs.p, P, b->radix);
            batch_merge_kernel<512><<<K, 512, 0, b->stream>>>(b->runs.p, P);
            batch_copy_kernel<512><<<K, 512, 0, b->stream>>>(b->runs.p, P);
            batch_final_kernel<512><<<b->nruns, 512, 0, b->stream>>>(b->runs.p, b->nruns, P);
            break;
    }
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

void clear_graphs(rimu_batch* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

int graph_for(rimu_batch* b, int steps, hipGraphExec_t* out) {
    const auto key = std::make_pair(steps, std::make_pair(b->block, b->radix));
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) { *out = it->second; return RIMU_OK; }
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps; ++s) {
        if (int rc = launch_step(b)) {
            hipGraph_t dummy;
            (void)hipStreamEndCapture(b->stream, &dummy);
            return rc;
        }
    }
    HIP_TRY(hipStreamEndCapture(b->stream, &g));
    hipGraphExec_t exec;
    HIP_TRY(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    b->graphs[key] = exec;
    *out = exec;
    return RIMU_OK;
}

int run_steps(rimu_batch* b, int64_t nsteps, int spg) {
    if (spg <= 1) {
        for (int64_t s = 0; s < nsteps; ++s)
            if (int rc = launch_step(b)) return rc;
        return RIMU_OK;
    }
    hipGraphExec_t exec;
    if (int rc = graph_for(b, spg, &exec)) return rc;
    int64_t done = 0;
    for (; done + spg <= nsteps; done += spg) HIP_TRY(hipGraphLaunch(exec, b->stream));
    for (; done < nsteps; ++done)
        if (int rc = launch_step(b)) return rc;
    return RIMU_OK;
}

}  // namespace

extern "C" {

int rimu_batch_create(const rimu_batch_run* runs, int32_t nruns, int64_t history_steps, rimu_batch** out) {
    if (!runs || !out) return fail(RIMU_ERR_ARG, "null argument");
    *out = nullptr;
    if (nruns < 1 || history_steps < 0) return fail(RIMU_ERR_ARG, "need nruns >= 1 and history_steps >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RIMU_ERR_NODEVICE, "no HIP device");
    std::vector<RunState> st(nruns);
    std::vector<uint64_t> start(nruns);
    std::vector<int> wg_run;
    int64_t totV = 0, totS = 0, totP = 0, totX = 0, totT = 0, totM = 0, totB = 0;
    for (int r = 0; r < nruns; ++r) {
        const rimu_batch_run& q = runs[r];
        if (q.M < 2 || q.N < 1 || q.M + q.N - 1 > 64) return fail(RIMU_ERR_ARG, "run: need M >= 2, N >= 1, N + M - 1 <= 64");
        if (!(q.dtau > 0) || !(q.target_walkers > 0) || q.max_length < 1 || q.start_population == 0)
            return fail(RIMU_ERR_ARG, "run: need dtau > 0, target_walkers > 0, max_length >= 1, population != 0");
        const int B = q.M + q.N - 1;
        if ((q.start_address >> B) != 0 || __builtin_popcountll(q.start_address) != q.N)
            return fail(RIMU_ERR_ARG, "run: start_address is not an N-boson M-mode address");
        if (2 * q.M + 1 > MAXPROJ) return fail(RIMU_ERR_ARG, "run: too many modes for the projector");
        RunState& s = st[r];
        std::memset(&s, 0, sizeof(s));
        s.M = q.M; s.N = q.N; s.B = B; s.status = RUN_OK;
        s.u = q.u; s.t = q.t; s.dtau = q.dtau; s.shift = q.shift;
        s.pnorm = std::fabs((double)q.start_population);
        s.target = q.target_walkers; s.zeta = q.zeta; s.xi = q.xi; s.seed = q.seed;
        s.len = 1; s.cur = 0; s.step = 0; s.nrec = 0;
        s.capV = q.max_length;
        s.capS = (int64_t)std::min(4.0 * q.target_walkers + 4096.0, 4.0e9);
        s.offV = totV; s.offS = totS; s.offP = totP; s.nproj = 0;
        // spawning of a run is split over K workgroups, one per WALKERS_PER_SLICE target walkers
        s.K = std::min<int64_t>(MAX_SLICES, std::max<int64_t>(1, (int64_t)std::ceil(q.target_walkers / WALKERS_PER_SLICE)));
        s.offX = totX;
        s.offB = totB;
        totB += s.K + 1;
        // merge tiles: boundaries for the smallest tile, slots for the largest
        s.offT = totT;
        s.offM = totM;
        totT += (s.capS + s.capV) / MIN_TILE + 3;
        totM += s.capS + s.capV + MAX_TILE;
        totV += s.capV + s.capS; totS += s.capS; totP += 2 * q.M + 1; totX += (s.K - 1) * s.capS;
        start[r] = q.start_address;
    }
    // slices numbered run by run, runs with the most slices first (their workgroups are
    // dispatched first)
    std::vector<int> order(nruns);
    for (int r = 0; r < nruns; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return st[x].K > st[y].K; });
    for (int r : order) {
        st[r].slice0 = (int64_t)wg_run.size();
        for (int64_t k = 0; k < st[r].K; ++k) wg_run.push_back(r);
    }
    rimu_batch* b = new rimu_batch();
    b->nruns = nruns;
    b->host_init = st;
    b->totV = totV; b->totS = totS; b->totP = totP; b->hist_cap = history_steps;
    b->totX = totX; b->totK = (int64_t)wg_run.size(); b->totT = totT; b->totM = totM; b->totB = totB;
    auto bad = [&](const char* what, hipError_t e) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        rimu_batch_destroy(b);
        return fail(RIMU_ERR_HIP, m);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bad("stream", e);
    if ((e = b->runs.alloc(nruns)) != hipSuccess) return bad("runs", e);
    for (int k = 0; k < 2; ++k) {
        if ((e = b->vk[k].alloc(totV)) != hipSuccess) return bad("vector keys", e);
        if ((e = b->vv[k].alloc(totV)) != hipSuccess) return bad("vector values", e);
        if ((e = b->sk[k].alloc(totS)) != hipSuccess) return bad("spawn keys", e);
        if ((e = b->sv[k].alloc(totS)) != hipSuccess) return bad("spawn values", e);
    }
    if ((e = b->dv.alloc(totV)) != hipSuccess) return bad("diagonal values", e);
    if ((e = b->pk.alloc(totP)) != hipSuccess) return bad("projector keys", e);
    if ((e = b->pc.alloc(totP)) != hipSuccess) return bad("projector coefficients", e);
    if ((e = b->start.alloc(nruns)) != hipSuccess) return bad("start addresses", e);
    if ((e = b->hist.alloc(std::max<int64_t>(1, NHIST * history_steps * nruns))) != hipSuccess) return bad("history", e);
    // rows a run does not write (it stopped, or a step of it was cut off) read zero
    if ((e = hipMemsetAsync(b->hist.p, 0, std::max<int64_t>(1, NHIST * history_steps * nruns) * sizeof(double), b->stream)) != hipSuccess)
        return bad("history", e);
    if ((e = b->xk.alloc(std::max<int64_t>(1, totX))) != hipSuccess) return bad("slice spawn keys", e);
    if ((e = b->xv.alloc(std::max<int64_t>(1, totX))) != hipSuccess) return bad("slice spawn values", e);
    if ((e = b->xcnt.alloc(b->totK)) != hipSuccess) return bad("slice counts", e);
    if ((e = b->xatt.alloc(b->totK)) != hipSuccess) return bad("slice attempts", e);
    if ((e = b->wg_run.alloc(b->totK)) != hipSuccess) return bad("slice runs", e);
    if ((e = b->msplit.alloc(totT)) != hipSuccess) return bad("merge splits", e);
    if ((e = b->mprev.alloc(totT)) != hipSuccess) return bad("merge keys", e);
    if ((e = b->mcnt.alloc(totT)) != hipSuccess) return bad("merge counts", e);
    if ((e = b->mnrm.alloc(totT)) != hipSuccess) return bad("merge norms", e);
    if ((e = b->mk.alloc(totM)) != hipSuccess) return bad("merge slot keys", e);
    if ((e = b->mv.alloc(totM)) != hipSuccess) return bad("merge slot values", e);
    if ((e = b->mw.alloc(totM)) != hipSuccess) return bad("merge slot walker offsets", e);
    if ((e = b->sbe.alloc(2 * totB)) != hipSuccess) return bad("spawn slice bounds", e);
    if ((e = b->sbp.alloc(2 * totB)) != hipSuccess) return bad("spawn slice bounds", e);
    if ((e = hipMemcpyAsync(b->wg_run.p, wg_run.data(), b->totK * sizeof(int), hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
    // the persistent step: one work queue per XCD, the runs dealt to the least loaded queue
    // (by spawn slices), largest first
    std::vector<int> items, qstart;
    {
        int dev = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return bad("device", e);
        if ((e = hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return bad("device", e);
        DevBuf<int> xmax;
        if ((e = xmax.alloc(1)) != hipSuccess) return bad("xcd count", e);
        int zero = 0, xm = 0;
        if ((e = hipMemcpyAsync(xmax.p, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("xcd count", e);
        xcd_count_kernel<<<std::max(1, b->n_cu), 64, 0, b->stream>>>(xmax.p);
        if ((e = hipMemcpyAsync(&xm, xmax.p, sizeof(int), hipMemcpyDeviceToHost, b->stream)) != hipSuccess) return bad("xcd count", e);
        if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bad("xcd count", e);
        xmax.release();
        b->persistent_ok = xm + 1 <= MAX_XCD;  // one queue per XCD, or no persistent step
        b->nq = std::min(MAX_XCD, xm + 1);
        std::vector<std::vector<int>> qruns(b->nq);
        std::vector<int64_t> load(b->nq, 0);
        for (int r : order) {
            const int qq = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            qruns[qq].push_back(r);
            load[qq] += st[r].K;
        }
        for (int qq = 0; qq < b->nq; ++qq) {
            qstart.push_back((int)items.size());
            for (int ph = 0; ph < 5; ++ph)
                for (int r : qruns[qq]) {
                    if (ph == 1 || ph == 4) items.push_back(ph << 24 | r);
                    else for (int64_t k = 0; k < st[r].K; ++k) items.push_back(ph << 24 | (int)(st[r].slice0 + k));
                }
        }
        qstart.push_back((int)items.size());
        if ((e = b->items.alloc(items.size())) != hipSuccess) return bad("work queues", e);
        if ((e = b->queue_start.alloc(qstart.size())) != hipSuccess) return bad("work queues", e);
        if ((e = hipMemcpyAsync(b->items.p, items.data(), items.size() * sizeof(int), hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
        if ((e = hipMemcpyAsync(b->queue_start.p, qstart.data(), qstart.size() * sizeof(int), hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
        if ((e = b->sync.alloc(sync_words(b))) != hipSuccess) return bad("step sync", e);
        if ((e = hipMemsetAsync(b->sync.p, 0, sync_words(b) * sizeof(uint32_t), b->stream)) != hipSuccess) return bad("step sync", e);
        if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->occupancy[0], batch_step_kernel<256>, 256, 0)) != hipSuccess ||
            (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->occupancy[1], batch_step_kernel<512>, 512, 0)) != hipSuccess ||
            (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->occupancy[2], batch_step_kernel<1024>, 1024, 0)) != hipSuccess)
            return bad("occupancy", e);
        for (int& o : b->occupancy) o = std::max(o, 1);
    }
    if ((e = hipMemcpyAsync(b->runs.p, st.data(), nruns * sizeof(RunState), hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
    if ((e = hipMemcpyAsync(b->start.p, start.data(), nruns * 8, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
    std::vector<int64_t> pops(nruns);  // host buffers live until the stream is synchronized below
    // slice bounds of the start vector (one entry holding every attempt) in the first copy;
    // both copies start with sbe[0] = sbp[0] = 0, which no step writes
    std::vector<int64_t> sb_e(2 * totB, 0), sb_p(2 * totB, 0);
    for (int r = 0; r < nruns; ++r) {
        sb_e[st[r].offB + st[r].K] = 1;
        sb_p[st[r].offB + st[r].K] = std::llabs(runs[r].start_population);
    }
    if ((e = hipMemcpyAsync(b->sbe.p, sb_e.data(), 2 * totB * 8, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
    if ((e = hipMemcpyAsync(b->sbp.p, sb_p.data(), 2 * totB * 8, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
    for (int r = 0; r < nruns; ++r) {  // the start vector: start_address => start_population
        pops[r] = runs[r].start_population;
        if ((e = hipMemcpyAsync(b->vk[0].p + st[r].offV, &start[r], 8, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
        if ((e = hipMemcpyAsync(b->vv[0].p + st[r].offV, &pops[r], 8, hipMemcpyHostToDevice, b->stream)) != hipSuccess) return bad("upload", e);
    }
    batch_projector_kernel<<<ceil_div(nruns, 64), 64, 0, b->stream>>>(b->runs.p, nruns, b->start.p, b->pk.p, b->pc.p);
    if ((e = hipGetLastError()) != hipSuccess) return bad("projector", e);
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bad("projector", e);
    *out = b;
    return RIMU_OK;
}

int rimu_batch_destroy(rimu_batch* b) {
    if (!b) return RIMU_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    clear_graphs(b);
    b->runs.release();
    for (int k = 0; k < 2; ++k) { b->vk[k].release(); b->vv[k].release(); b->sk[k].release(); b->sv[k].release(); }
    b->dv.release(); b->pk.release(); b->pc.release(); b->start.release(); b->hist.release();
    b->xk.release(); b->xv.release(); b->xcnt.release(); b->xatt.release(); b->wg_run.release();
    b->msplit.release(); b->mprev.release(); b->mcnt.release(); b->mnrm.release(); b->mk.release(); b->mv.release();
    b->sbe.release(); b->sbp.release(); b->mw.release();
    b->sync.release(); b->items.release(); b->queue_start.release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return RIMU_OK;
}

int rimu_batch_set_config(rimu_batch* b, int32_t block_size, int32_t radix_bits) {
    if (!b) return fail(RIMU_ERR_ARG, "batch is null");
    if (block_size != 256 && block_size != 512 && block_size != 1024)
        return fail(RIMU_ERR_ARG, "block_size must be 256, 512 or 1024");
    if (radix_bits < 4 || radix_bits > 8) return fail(RIMU_ERR_ARG, "radix_bits must be in 4..8");
    b->block = block_size;
    b->radix = radix_bits;
    return RIMU_OK;
}

int rimu_batch_set_persistent(rimu_batch* b, int32_t persistent) {
    if (!b) return fail(RIMU_ERR_ARG, "batch is null");
    if (persistent != 0 && persistent != 1) return fail(RIMU_ERR_ARG, "persistent must be 0 or 1");
    if (persistent && !b->persistent_ok) return fail(RIMU_ERR_ARG, "more XCDs than the persistent step has queues");
    if (persistent != b->persistent) clear_graphs(b);
    b->persistent = persistent;
    return RIMU_OK;
}

int rimu_batch_set_wait(rimu_batch* b, uint32_t spin_limit, int32_t withhold_first) {
    if (!b) return fail(RIMU_ERR_ARG, "batch is null");
    if (spin_limit < 1) return fail(RIMU_ERR_ARG, "spin_limit must be positive");
    HIP_TRY(hipMemsetAsync(b->sync.p, 0, sync_words(b) * sizeof(uint32_t), b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    clear_graphs(b);  // the bound and the flag are arguments of the captured launches
    b->spin_limit = spin_limit;
    b->withhold = withhold_first != 0;
    return RIMU_OK;
}

int rimu_batch_step(rimu_batch* b, int64_t nsteps, int32_t steps_per_graph) {
    if (!b) return fail(RIMU_ERR_ARG, "batch is null");
    if (nsteps < 0 || steps_per_graph < 1) return fail(RIMU_ERR_ARG, "need nsteps >= 0 and steps_per_graph >= 1");
    if (int rc = run_steps(b, nsteps, steps_per_graph)) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    uint32_t abort_word = 0;
    HIP_TRY(hipMemcpy(&abort_word, b->sync.p + MAX_XCD + 1, sizeof(abort_word), hipMemcpyDeviceToHost));
    if (abort_word) return fail(RIMU_ERR_INTERNAL, "a wait of the persistent batched step gave up");
    return RIMU_OK;
}

int rimu_batch_history(rimu_batch* b, int64_t* nsteps, double* shift, double* norm, double* len,
                       double* vproj, double* hproj, double* attempts) {
    if (!b || !nsteps) return fail(RIMU_ERR_ARG, "null argument");
    std::vector<RunState> st(b->nruns);
    HIP_TRY(hipMemcpy(st.data(), b->runs.p, b->nruns * sizeof(RunState), hipMemcpyDeviceToHost));
    int64_t rows = 0;
    for (auto& s : st) rows = std::max(rows, s.nrec);
    *nsteps = rows;
    double* outs[NHIST] = {shift, norm, len, vproj, hproj, attempts};
    const int64_t plane = b->hist_cap * b->nruns;
    for (int f = 0; f < NHIST; ++f)
        if (outs[f] && rows > 0)
            HIP_TRY(hipMemcpy(outs[f], b->hist.p + f * plane, rows * b->nruns * sizeof(double), hipMemcpyDeviceToHost));
    return RIMU_OK;
}

int rimu_batch_status(rimu_batch* b, int32_t* status, int64_t* len) {
    if (!b) return fail(RIMU_ERR_ARG, "batch is null");
    std::vector<RunState> st(b->nruns);
    HIP_TRY(hipMemcpy(st.data(), b->runs.p, b->nruns * sizeof(RunState), hipMemcpyDeviceToHost));
    for (int r = 0; r < b->nruns; ++r) {
        if (status) status[r] = st[r].status;
        if (len) len[r] = st[r].len;
    }
    return RIMU_OK;
}

int rimu_batch_vector(rimu_batch* b, int32_t run, uint64_t* keys, int64_t* vals, int64_t cap, int64_t* n) {
    if (!b || !n) return fail(RIMU_ERR_ARG, "null argument");
    if (run < 0 || run >= b->nruns) return fail(RIMU_ERR_ARG, "run index out of range");
    RunState s;
    HIP_TRY(hipMemcpy(&s, b->runs.p + run, sizeof(RunState), hipMemcpyDeviceToHost));
    *n = s.len;
    if (s.len > cap) return fail(RIMU_ERR_CAPACITY, "capacity too small");
    HIP_TRY(hipMemcpy(keys, b->vk[s.cur].p + s.offV, s.len * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(vals, b->vv[s.cur].p + s.offV, s.len * 8, hipMemcpyDeviceToHost));
    return RIMU_OK;
}

int rimu_batch_autotune(rimu_batch* b, int64_t trial_steps, int32_t* best_block, int32_t* best_radix,
                        int32_t* best_persistent, double* ms) {
    if (!b) return fail(RIMU_ERR_ARG, "batch is null");
    if (trial_steps < 1) return fail(RIMU_ERR_ARG, "trial_steps must be >= 1");
    // snapshot: run states, both vector buffers (the step reads one, writes the other) and
    // the spawn slice bounds
    DevBuf<RunState> s_runs;
    DevBuf<uint64_t> s_vk[2];
    DevBuf<int64_t> s_vv[2], s_sbe, s_sbp;
    auto cleanup = [&]() {
        s_runs.release();
        s_sbe.release();
        s_sbp.release();
        for (int k = 0; k < 2; ++k) { s_vk[k].release(); s_vv[k].release(); }
    };
    auto body = [&]() -> int {
        HIP_TRY(hipStreamSynchronize(b->stream));
        HIP_TRY(s_runs.alloc(b->nruns));
        // copies on the batch's stream, ordered with its kernels
        HIP_TRY(hipMemcpyAsync(s_runs.p, b->runs.p, b->nruns * sizeof(RunState), hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(s_sbe.alloc(2 * b->totB));
        HIP_TRY(s_sbp.alloc(2 * b->totB));
        HIP_TRY(hipMemcpyAsync(s_sbe.p, b->sbe.p, 2 * b->totB * 8, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(s_sbp.p, b->sbp.p, 2 * b->totB * 8, hipMemcpyDeviceToDevice, b->stream));
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(s_vk[k].alloc(b->totV));
            HIP_TRY(s_vv[k].alloc(b->totV));
            HIP_TRY(hipMemcpyAsync(s_vk[k].p, b->vk[k].p, b->totV * 8, hipMemcpyDeviceToDevice, b->stream));
            HIP_TRY(hipMemcpyAsync(s_vv[k].p, b->vv[k].p, b->totV * 8, hipMemcpyDeviceToDevice, b->stream));
        }
        auto restore = [&]() -> int {
            HIP_TRY(hipMemcpyAsync(b->runs.p, s_runs.p, b->nruns * sizeof(RunState), hipMemcpyDeviceToDevice, b->stream));
            HIP_TRY(hipMemcpyAsync(b->sbe.p, s_sbe.p, 2 * b->totB * 8, hipMemcpyDeviceToDevice, b->stream));
            HIP_TRY(hipMemcpyAsync(b->sbp.p, s_sbp.p, 2 * b->totB * 8, hipMemcpyDeviceToDevice, b->stream));
            for (int k = 0; k < 2; ++k) {
                HIP_TRY(hipMemcpyAsync(b->vk[k].p, s_vk[k].p, b->totV * 8, hipMemcpyDeviceToDevice, b->stream));
                HIP_TRY(hipMemcpyAsync(b->vv[k].p, s_vv[k].p, b->totV * 8, hipMemcpyDeviceToDevice, b->stream));
            }
            return RIMU_OK;
        };
        const int blocks[3] = {256, 512, 1024};
        const int radixes[3] = {4, 6, 8};
        double best = 1e300;
        const int p0 = b->persistent;
        int bb = b->block, br = b->radix, bp = b->persistent;
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        int idx = 0;
        if (ms)
            for (int i = 0; i < 18; ++i) ms[i] = HUGE_VAL;  // candidates that are not timed
        for (int pi = 0; pi < (b->persistent_ok ? 2 : 1); ++pi)
        for (int bi = 0; bi < 3; ++bi)
            for (int ri = 0; ri < 3; ++ri, ++idx) {
                b->persistent = pi;
                b->block = blocks[bi];
                b->radix = radixes[ri];
                if (int rc = launch_step(b)) return rc;  // first launch outside the timing
                HIP_TRY(hipStreamSynchronize(b->stream));
                if (int rc = restore()) return rc;
                HIP_TRY(hipEventRecord(e0, b->stream));
                for (int64_t s = 0; s < trial_steps; ++s)
                    if (int rc = launch_step(b)) return rc;
                HIP_TRY(hipEventRecord(e1, b->stream));
                HIP_TRY(hipEventSynchronize(e1));
                float t = 0;
                HIP_TRY(hipEventElapsedTime(&t, e0, e1));
                double per = t / (double)trial_steps;
                if (int rc = restore()) return rc;
                if (pi) {  // a persistent trial whose wait gave up is not a candidate
                    uint32_t abort_word = 0;
                    HIP_TRY(hipStreamSynchronize(b->stream));
                    HIP_TRY(hipMemcpy(&abort_word, b->sync.p + MAX_XCD + 1, sizeof(abort_word), hipMemcpyDeviceToHost));
                    if (abort_word) {
                        per = HUGE_VAL;
                        HIP_TRY(hipMemsetAsync(b->sync.p, 0, sync_words(b) * sizeof(uint32_t), b->stream));
                    }
                }
                if (ms) ms[idx] = per;
                if (per < best) { best = per; bb = b->block; br = b->radix; bp = b->persistent; }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_TRY(hipStreamSynchronize(b->stream));  // the last restore
        if (bp != p0) clear_graphs(b);
        b->block = bb;
        b->radix = br;
        b->persistent = bp;
        if (best_block) *best_block = bb;
        if (best_radix) *best_radix = br;
        if (best_persistent) *best_persistent = bp;
        return RIMU_OK;
    };
    const int rc = body();
    cleanup();
    return rc;
}

}  // extern "C"
