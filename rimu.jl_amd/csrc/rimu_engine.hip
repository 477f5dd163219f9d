// rimu_engine.hip -- host side of the FCIQMC step engine and the C ABI of include/rimu_hip.h.
//
// One engine = one FCIQMC state: a sorted device vector (keys u64, values int64 or f64),
// double buffered, plus the working memory of a step (diagonal values, work offsets, the
// spawn list and its sort buffers, look-back status words, per-tile partial sums).
// All work of an engine is ordered on its own HIP stream.  A step synchronises with the
// host three times (work-item count, spawn count, statistics) so that every grid is sized
// to the data it covers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rimu_hip.h"
#include "capi_util.hpp"
#include "overlap_kernels.hpp"
#include "bucket_kernels.hpp"
#include "step_kernels.hpp"

using namespace rimu;

namespace rimu {
thread_local std::string g_last_error;
}

namespace {

struct KernelTimer {
    bool enabled = false;
    struct Rec { double ms = 0; int64_t launches = 0; };
    std::map<std::string, Rec> totals;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(const char* name, hipStream_t s, hipEvent_t* a) {
        if (!enabled) return;
        *a = get();
        (void)hipEventRecord(*a, s);
        pending.push_back({name, {*a, nullptr}});
    }
    void end(hipStream_t s) {
        if (!enabled) return;
        hipEvent_t b = get();
        (void)hipEventRecord(b, s);
        pending.back().second.second = b;
    }
    void collect() {  // after a stream synchronisation
        for (auto& p : pending) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
            auto& r = totals[p.first];
            r.ms += ms;
            r.launches += 1;
            pool.push_back(p.second.first);
            pool.push_back(p.second.second);
        }
        pending.clear();
    }
    ~KernelTimer() {
        for (auto& e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace

struct rimu_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    rimu_model model{};
    rimu_style style{};
    ModelDev hdev{};
    KeyMix mix{};                   // stored keys are mix(address) above MIX_MIN_BITS bits
    int sort_bits = MIX_SORT_BITS;  // top bits of mixed keys the radix passes cover
    // 2: sort on the top bits + bucket merge where eligible, else as 1; 1: one histogram pass
    // + look-back scatters over every bit, merge path; 0: counts per pass, merge path
    int radix_mode = 2;
    int n_cu = 0;                   // compute units of the device (queried on first use)
    uint64_t seed = 0;
    bool int_vals = false;
    size_t vsize = 8;
    int kw = 1;                     // 64-bit words per address: 1 (B <= 64) or 2 (B <= 128)
    int64_t max_entries = 0, max_spawns = 0;
    int32_t init_rule = 0;          // RIMU_INITIATOR_*
    double init_threshold = 0;

    // vector (double buffered); key buffers hold kw words per address
    DevBuf<uint64_t> vkeys[2];
    DevBuf<unsigned char> vvals[2];
    int cur = 0;
    int64_t len = 0;

    // step working memory
    DevBuf<unsigned char> dvals;
    DevBuf<int64_t> woff;
    DevBuf<int64_t> scan_tmp;
    DevBuf<uint32_t> rcounts, rchunks, rtmp;
    DevBuf<uint64_t> skeys[2];
    DevBuf<unsigned char> svals[2];
    DevBuf<uint64_t> plan_status, tile_or, tile_and;
    DevBuf<int64_t> tile_first, splits, tile_cnt;
    DevBuf<uint64_t> prevk, slot_k, slot_v;  // merge: per-tile survivor slots (values as raw 8 bytes)
    DevBuf<double> plan_p, spawn_p, merge_p, terms, stats_dev, chunk_sums;
    DevBuf<unsigned char> misc;  // counters and totals (see layout below)
    // projector
    DevBuf<uint64_t> pkeys;
    DevBuf<double> pcoef;
    // overlaps (this engine as the right-hand vector)
    DevBuf<unsigned char> ov_ops;
    DevBuf<double> ov_part, ov_out;
    DevBuf<int64_t> ov_ranges, ov_index;
    DevBuf<int64_t> fix_runs;       // runs to reorder after a sort on the top key bits
    DevBuf<int64_t> bk_rstart, bk_vstart;  // bucket merge: first record / vector entry per bucket
    DevBuf<uint64_t> lb_status;     // look-back radix: status words [pass][tile][digit]
    uint32_t lb_epoch = 0;          // sort calls since lb_status was last cleared
    int bucket_cooldown = 0;        // steps left on the merge path after a bucket overflow
    bool spec_ok = false;           // the grids of the next step may be sized from W_prev, S_prev
    int64_t W_prev = 0, S_prev = 0; // spawn attempts and records of the last step
    uint32_t lb_spin_limit = LB_SPIN_LIMIT;  // bound of every look-back wait
    int lb_withhold = 0;            // test of the give-up path (rimu_engine_set_lookback)
    DevBuf<uint32_t> lb_hist;       // look-back radix: digit bases [pass][digit], tile counters
    DevBuf<uint32_t> lb_part;       // look-back radix: digit counts [row][pass][digit]
    int64_t np = 0, nh = 0;
    // device shift update (rimu_engine_set_shift_update): shift_state holds (shift, pnorm)
    // twice; a step reads half shift_cur and writes the other, and a kept step flips shift_cur
    int shift_mode = 0;
    double dlu_target = 0, dlu_zeta = 0, dlu_xi = 0;
    DevBuf<double> shift_state;
    int shift_cur = 0;
    double h_shift = 0, h_pnorm = 0;  // the device state after the last kept step

    // host staging
    struct Misc {
        int32_t flags;
        uint32_t plan_counter;
        int64_t W;
        int64_t S;
        int64_t L_new;
        u128 key_or;     // as wide as the widest key; a one-chunk key fills the low word
        u128 key_and;
        uint32_t fix_count;  // runs listed by fixup_find_kernel
    };
    Misc* h_misc = nullptr;    // pinned
    double* h_stats = nullptr; // pinned: 12 statistics, then shift and pnorm

    KernelTimer timer;
    double last_alg_bytes = 0, last_records = 0;
};

namespace {

// Bits of one HubbardRealSpace component: M for a FermiFS, N + M - 1 for a BoseFS.
int rs_component_bits(const rimu_model& m, int c) {
    const int n = c == 0 ? m.N : m.N2;
    return ((m.fermi >> c) & 1) ? m.M : n + m.M - 1;
}

ModelDev to_dev(const rimu_model& m) {
    ModelDev h{};
    h.kind = m.kind;
    h.M = m.M;
    h.N = m.N;
    h.B = m.M + m.N - 1;
    h.u = m.u;
    h.t = m.t;
    for (int j = 0; j < 64; ++j) h.kes[j] = m.kes[j];
    if (m.kind == RIMU_MODEL_HUBBARD_REAL_SPACE) {
        h.ncomp = m.ncomp;
        h.N2 = m.N2;
        h.fermi = m.fermi;
        h.ndir = m.ndir;
        h.has_u = m.has_u;
        h.has_pot = m.has_pot;
        h.cbits2 = m.ncomp == 2 ? rs_component_bits(m, 1) : 0;
        h.B = rs_component_bits(m, 0) + h.cbits2;
        h.t2 = m.t2;
        h.u12 = m.u12;
        h.u22 = m.u22;
        for (int j = 0; j < 64; ++j) h.pot2[j] = m.pot2[j];
        std::memcpy(h.nbr, m.nbr, sizeof(h.nbr));
    }
    return h;
}

// Whether two models act on the same addresses (sites, particles per component, component
// types).
bool same_space(const ModelDev& a, const ModelDev& b) {
    const bool ra = a.kind == RIMU_MODEL_HUBBARD_REAL_SPACE, rb = b.kind == RIMU_MODEL_HUBBARD_REAL_SPACE;
    if (a.M != b.M || a.N != b.N || a.B != b.B || ra != rb) return false;
    return !ra || (a.ncomp == b.ncomp && a.N2 == b.N2 && a.fermi == b.fermi);
}

// Downloaded vectors: stored keys back to addresses, then ascending address order (the
// order the C ABI promises).
void unmix_and_sort(rimu_engine* e, uint64_t* keys, void* vals, int64_t n);

// Whether key is an address of the model: inside B bits, with every component's particle
// number (a BoseFS component of N particles has N ones among its N + M - 1 bits).
bool address_ok(const ModelDev& h, u128 key) {
    const u128 maskB = h.B >= 128 ? ~(u128)0 : (((u128)1 << h.B) - 1);
    if ((key & ~maskB) != 0) return false;
    auto ones = [](u128 x) { return __builtin_popcountll((uint64_t)x) + __builtin_popcountll((uint64_t)(x >> 64)); };
    if (h.kind != RIMU_MODEL_HUBBARD_REAL_SPACE || h.ncomp == 1) return ones(key) == h.N;
    const u128 low = ((u128)1 << h.cbits2) - 1;
    return ones(key >> h.cbits2) == h.N && ones(key & low) == h.N2;
}

void unmix_and_sort(rimu_engine* e, uint64_t* keys, void* vals, int64_t n) {
    const int kw = e->kw;
    std::vector<u128> addr((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        addr[i] = kw == 1 ? (u128)unmix_bits<uint64_t>(e->mix, keys[i])
                          : unmix_bits<u128>(e->mix, ((u128)keys[2 * i + 1] << 64) | keys[2 * i]);
    std::vector<int64_t> ord((size_t)n);
    for (int64_t i = 0; i < n; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return addr[a] < addr[b]; });
    std::vector<uint64_t> v((const uint64_t*)vals, (const uint64_t*)vals + n);  // 8-byte values
    for (int64_t j = 0; j < n; ++j) {
        const u128 x = addr[ord[j]];
        keys[(size_t)kw * j] = (uint64_t)x;
        if (kw == 2) keys[(size_t)kw * j + 1] = (uint64_t)(x >> 64);
        ((uint64_t*)vals)[j] = v[ord[j]];
    }
}

int check_real_space(const rimu_model* m) {
    if (m->ncomp != 1 && m->ncomp != 2) return fail(RIMU_ERR_ARG, "HubbardRealSpace needs 1 or 2 components");
    if (m->M < 2 || m->M > 64) return fail(RIMU_ERR_ARG, "HubbardRealSpace needs 2 <= M <= 64 sites");
    if (m->ndir != 2 && m->ndir != 4 && m->ndir != 6) return fail(RIMU_ERR_ARG, "HubbardRealSpace geometry needs 1 to 3 dimensions");
    if ((m->fermi & ~((1 << m->ncomp) - 1)) != 0) return fail(RIMU_ERR_ARG, "fermi flags a missing component");
    const int n2 = m->ncomp == 2 ? m->N2 : 0;
    if (m->N < 0 || n2 < 0 || m->N + n2 < 1) return fail(RIMU_ERR_ARG, "HubbardRealSpace needs particles");
    for (int c = 0; c < m->ncomp; ++c)
        if (((m->fermi >> c) & 1) && (c == 0 ? m->N : m->N2) > m->M)
            return fail(RIMU_ERR_ARG, "a FermiFS holds at most M particles");
    int bits = 0;
    for (int c = 0; c < m->ncomp; ++c) bits += rs_component_bits(*m, c);
    if (bits > 128) return fail(RIMU_ERR_ARG, "address must fit two 64-bit chunks (at most 128 bits)");
    for (int s = 0; s < m->M; ++s)
        for (int d = 0; d < m->ndir; ++d)
            if (m->nbr[s][d] != 255 && m->nbr[s][d] >= m->M)
                return fail(RIMU_ERR_ARG, "neighbour table names a site outside 1..M");
    return RIMU_OK;
}

int check_model(const rimu_model* m) {
    if (!m) return fail(RIMU_ERR_ARG, "model is null");
    if (m->kind == RIMU_MODEL_HUBBARD_REAL_SPACE) return check_real_space(m);
    if (m->kind != RIMU_MODEL_HUBBARD_REAL_1D && m->kind != RIMU_MODEL_HUBBARD_MOM_1D)
        return fail(RIMU_ERR_ARG, "unknown model kind");
    if (m->M < 1 || m->N < 1 || m->M + m->N - 1 > 128)
        return fail(RIMU_ERR_ARG, "address must fit two 64-bit chunks: need N + M - 1 <= 128");
    if (m->kind == RIMU_MODEL_HUBBARD_MOM_1D && (m->M < 3 || m->M > 64))
        return fail(RIMU_ERR_ARG, "HubbardMom1D needs 3 <= M <= 64");
    if (m->kind == RIMU_MODEL_HUBBARD_REAL_1D && m->M < 2)
        return fail(RIMU_ERR_ARG, "HubbardReal1D needs M >= 2");
    return RIMU_OK;
}

StepArgs make_args(rimu_engine* e, int64_t step, double shift, double dtau, double boost) {
    StepArgs a{};
    a.h = e->hdev;
    a.seed = e->seed;
    a.step = step;
    a.shift = shift;
    a.shift_dev = e->shift_mode ? e->shift_state.p + 2 * e->shift_cur : nullptr;
    a.dtau = dtau;
    a.boost = boost;
    a.style = e->style.kind;
    a.compression = e->style.compression;
    a.comp_threshold = e->style.compression_threshold;
    a.proj_threshold = e->style.proj_threshold;
    a.rel_threshold = e->style.rel_threshold;
    a.abs_threshold = e->style.abs_threshold;
    a.init_rule = e->init_rule;
    a.init_threshold = e->init_threshold;
    a.mix = e->mix;
    a.spin_limit = e->lb_spin_limit;
    return a;
}

// exclusive scan of n int64 values in place (data has n entries)
int scan_inplace_i64(rimu_engine* e, int64_t* data, int64_t n) {
    const int64_t tiles = ceil_div(n, SCAN_TILE);
    HIP_TRY(e->scan_tmp.alloc(tiles + 1));
    scan_tiles_kernel<int64_t><<<tiles, NT, 0, e->stream>>>(data, data, n, e->scan_tmp.p);
    scan_sums_kernel<int64_t><<<1, 1024, 0, e->stream>>>(e->scan_tmp.p, tiles, nullptr);
    scan_add_kernel<int64_t><<<tiles, NT, 0, e->stream>>>(data, n, e->scan_tmp.p);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

int scan_u32(rimu_engine* e, const uint32_t* in, uint32_t* out, int64_t n) {
    const int64_t tiles = ceil_div(n, SCAN_TILE);
    HIP_TRY(e->rtmp.alloc(tiles + 1));
    scan_tiles_kernel<uint32_t><<<tiles, NT, 0, e->stream>>>(in, out, n, e->rtmp.p);
    scan_sums_kernel<uint32_t><<<1, 1024, 0, e->stream>>>(e->rtmp.p, tiles, nullptr);
    scan_add_kernel<uint32_t><<<tiles, NT, 0, e->stream>>>(out, n, e->rtmp.p);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

// Stable LSD radix sort of n (key, value) records on key bits [0, nbits), skipping passes
// whose digit is the same for every key.  Bits at and above nbits are carried, not sorted
// (the initiator class bit).  Input in buffer `in`; returns the buffer index holding the
// result.
template <class K, class V>
int radix_sort_lb(rimu_engine* e, int64_t n, int nbits, u128 varying, int in, int* result, int lo,
                  const int64_t* n_dev);

template <class K, class V>
int radix_sort(rimu_engine* e, int64_t n, int nbits, u128 varying, int in, int* result, int lo = 0,
               const int64_t* n_dev = nullptr) {
    int src = in;
    if (n <= 1 && !n_dev) { *result = src; return RIMU_OK; }
    if (e->radix_mode != 0 || n_dev) return radix_sort_lb<K, V>(e, n, nbits, varying, in, result, lo, n_dev);
    const int64_t tiles = ceil_div(n, sort_tile<K>::value);
    const int64_t nchunks = ceil_div(tiles, RADIX_CHUNK);
    HIP_TRY(e->rcounts.alloc((size_t)tiles * RADIX));
    HIP_TRY(e->rchunks.alloc((size_t)nchunks * RADIX));
    for (int shift = lo; shift < nbits; shift += RADIX_BITS) {
        if ((uint32_t)((varying >> shift) & (RADIX - 1)) == 0) continue;  // digit constant: pass is identity
        const uint32_t dmask = shift + RADIX_BITS > nbits ? (1u << (nbits - shift)) - 1u : (uint32_t)(RADIX - 1);
        hipEvent_t ev{};
        e->timer.begin("radix_count", e->stream, &ev);
        radix_count_kernel<K><<<tiles, NT, 0, e->stream>>>((const K*)e->skeys[src].p, n, shift, dmask,
                                                           e->rcounts.p, tiles);
        e->timer.end(e->stream);
        e->timer.begin("radix_scan", e->stream, &ev);
        radix_chunk_sums_kernel<<<nchunks, RADIX, 0, e->stream>>>(e->rcounts.p, tiles, nchunks, e->rchunks.p);
        if (int rc = scan_u32(e, e->rchunks.p, e->rchunks.p, nchunks * RADIX)) return rc;
        radix_chunk_offsets_kernel<<<nchunks, RADIX, 0, e->stream>>>(e->rcounts.p, tiles, nchunks, e->rchunks.p);
        e->timer.end(e->stream);
        e->timer.begin("radix_scatter", e->stream, &ev);
        radix_scatter_kernel<K, V><<<tiles, NT, 0, e->stream>>>(
            (const K*)e->skeys[src].p, (const V*)e->svals[src].p, n, shift, dmask, e->rcounts.p, tiles,
            (K*)e->skeys[1 - src].p, (V*)e->svals[1 - src].p);
        e->timer.end(e->stream);
        HIP_TRY(hipGetLastError());
        src = 1 - src;
    }
    *result = src;
    return RIMU_OK;
}

// The same sort with one histogram pass over the keys for every digit and one look-back
// scatter per pass (no per-pass count or scan kernels).  The output is that of the per-pass
// path: the ranks within a tile and the digit offsets are the same numbers.  FLAG_LOOKBACK
// in the step's flags reports a look-back that did not complete.  With n_dev the record
// count is in device memory and n is the most the grids are sized for.
template <class K, class V>
int radix_sort_lb(rimu_engine* e, int64_t n, int nbits, u128 varying, int in, int* result, int lo,
                  const int64_t* n_dev) {
    int src = in;
    PassSet ps{};
    for (int shift = lo; shift < nbits; shift += RADIX_BITS) {
        if ((uint32_t)((varying >> shift) & (RADIX - 1)) == 0) continue;  // digit constant: pass is identity
        if (ps.n == MAX_PASSES) return fail(RIMU_ERR_INTERNAL, "too many radix passes");
        ps.shift[ps.n] = shift;
        ps.dmask[ps.n] = shift + RADIX_BITS > nbits ? (1u << (nbits - shift)) - 1u : (uint32_t)(RADIX - 1);
        ++ps.n;
    }
    if (ps.n == 0) { *result = src; return RIMU_OK; }
    if (n > (int64_t)RLB_VALUE) return fail(RIMU_ERR_CAPACITY, "radix sort of 2^32 or more records");
    const int64_t tiles = ceil_div(n, sort_tile<K>::value);
    const size_t need = (size_t)ps.n * tiles * RADIX;
    HIP_TRY(e->lb_hist.alloc((size_t)ps.n * RADIX + MAX_PASSES));
    uint32_t* counters = e->lb_hist.p + (int64_t)ps.n * RADIX;  // tile ids (two-word keys)
    int* flags = &((rimu_engine::Misc*)e->misc.p)->flags;
    hipEvent_t ev{};
    if (need > e->lb_status.n || e->lb_epoch >= RLB_EPOCH_MAX) {  // epoch 0: cleared words
        HIP_TRY(e->lb_status.alloc(need));
        HIP_TRY(hipMemsetAsync(e->lb_status.p, 0, sizeof(uint64_t) * e->lb_status.n, e->stream));
        e->lb_epoch = 0;
    }
    const uint32_t epoch = ++e->lb_epoch;
    e->timer.begin("radix_count", e->stream, &ev);
    HIP_TRY(hipMemsetAsync(counters, 0, sizeof(uint32_t) * MAX_PASSES, e->stream));
    const int copies = hist_all_copies(ps.n);
    const int hblocks = (int)std::min<int64_t>(tiles, HIST_ALL_BLOCKS);
    HIP_TRY(e->lb_part.alloc((size_t)HIST_ROWS * ps.n * RADIX));
    HIP_TRY(hipMemsetAsync(e->lb_part.p, 0, sizeof(uint32_t) * HIST_ROWS * ps.n * RADIX, e->stream));
    radix_hist_all_kernel<K><<<hblocks, NT, sizeof(uint32_t) * ps.n * copies * HIST_STRIDE, e->stream>>>(
        (const K*)e->skeys[src].p, Count{n, n_dev, n}, flags, ps, copies, e->lb_part.p);
    radix_hist_base_kernel<<<ps.n, RADIX, 0, e->stream>>>(e->lb_part.p, ps.n, e->lb_hist.p);
    e->timer.end(e->stream);
    for (int p = 0; p < ps.n; ++p) {
        e->timer.begin("radix_scatter", e->stream, &ev);
        radix_scatter_lb_kernel<K, V><<<tiles, NT, 0, e->stream>>>(
            (const K*)e->skeys[src].p, (const V*)e->svals[src].p, Count{n, n_dev, n}, ps.shift[p], ps.dmask[p],
            e->lb_hist.p + (int64_t)p * RADIX, e->lb_status.p + (int64_t)p * tiles * RADIX, epoch, counters + p,
            flags, LookbackCtl{e->lb_spin_limit, e->lb_withhold},
            (K*)e->skeys[1 - src].p, (V*)e->svals[1 - src].p);
        e->timer.end(e->stream);
        HIP_TRY(hipGetLastError());
        src = 1 - src;
    }
    *result = src;
    return RIMU_OK;
}

__global__ void clear_flags_kernel(int* flags, int bits) { atomicAnd(flags, ~bits); }

// Sort of mixed keys: radix passes on the top MIX_SORT_BITS bits, then the runs that agree
// there but not below are put in order (fixup kernels).  If a run is too long or there are
// too many, the passes below the top bits and the top passes again make the order complete
// (stable LSD passes compose, so equal keys keep their list order either way).
template <class K, class V>
int sort_mixed(rimu_engine* e, int64_t n, int nbits, u128 varying, bool init, rimu_engine::Misc* dm, int* result,
               const int64_t* n_dev = nullptr) {
    const int lo = nbits - e->sort_bits;
    int sorted = 0;
    if (int rc = radix_sort<K, V>(e, n, nbits, varying, 0, &sorted, lo, n_dev)) return rc;
    const K kmask = init ? (K)~class_bit<K>() : (K)~(K)0;
    HIP_TRY(e->fix_runs.alloc(FIXUP_LIST));
    hipEvent_t ev{};
    e->timer.begin("sort_fixup", e->stream, &ev);
    fixup_find_kernel<K><<<ceil_div(n, NT), NT, 0, e->stream>>>((const K*)e->skeys[sorted].p, Count{n, n_dev, n},
                                                                lo, kmask, e->fix_runs.p, &dm->fix_count, &dm->flags);
    fixup_sort_kernel<K, V><<<64, NT, 0, e->stream>>>((K*)e->skeys[sorted].p, (V*)e->svals[sorted].p,
                                                      Count{n, n_dev, n}, lo, kmask, e->fix_runs.p,
                                                      &dm->fix_count, &dm->flags);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    if (n_dev) {  // counts in device memory: a fallback is taken by repeating the step
        *result = sorted;
        return RIMU_OK;
    }
    HIP_TRY(hipMemcpyAsync(e->h_misc, dm, sizeof(rimu_engine::Misc), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->h_misc->flags & FLAG_FIXUP_FALLBACK) {
        // the full sort below orders the runs: the merge must not read the flag as a failure
        clear_flags_kernel<<<1, 1, 0, e->stream>>>(&dm->flags, FLAG_FIXUP_FALLBACK);
        HIP_TRY(hipGetLastError());
        int low = 0;
        if (int rc = radix_sort<K, V>(e, n, lo, varying, sorted, &low)) return rc;
        if (int rc = radix_sort<K, V>(e, n, nbits, varying, low, &sorted, lo)) return rc;
    }
    *result = sorted;
    return RIMU_OK;
}

// Merge the sorted spawn records (ak, av; S) with the vector keys and diagonal values
// (bk, bv; L) into (ok, ov) of capacity cap: tile splits, merge into per-tile slots, scan
// of the tile counts, compaction.  The new length lands in dm->L_new; `tiles` receives the
// number of merge tiles (rows of merge_p).  With an initiator rule (a.init_rule != 0) the
// spawn keys carry the class bit and bo holds the vector's values before the step.
template <class K, class V>
int merge_into(rimu_engine* e, const StepArgs& a, const K* ak, const V* av, int64_t S,
               const K* bk, const V* bv, int64_t L, K* ok, V* ov, int64_t cap,
               rimu_engine::Misc* dm, int64_t* tiles, const V* bo = nullptr, const int64_t* S_dev = nullptr) {
    constexpr int KW = sizeof(K) / 8;
    const bool init = a.init_rule != 0;
    if (init && !bo) return fail(RIMU_ERR_INTERNAL, "initiator merge without the old values");
    constexpr int64_t TILE = NT * merge_ipt<K>::value;
    const int64_t nt = std::max<int64_t>(ceil_div(S + L, TILE), 1);
    *tiles = nt;
    HIP_TRY(e->merge_p.alloc((size_t)nt * 3));
    HIP_TRY(e->splits.alloc(nt + 1));
    HIP_TRY(e->prevk.alloc((size_t)(nt + 1) * KW));
    HIP_TRY(e->tile_cnt.alloc(nt + 1));
    HIP_TRY(e->slot_k.alloc((size_t)nt * TILE * KW));
    HIP_TRY(e->slot_v.alloc((size_t)nt * TILE));
    HIP_TRY(hipMemsetAsync(e->tile_cnt.p + nt, 0, sizeof(int64_t), e->stream));
    merge_split_kernel<K><<<ceil_div(nt + 1, NT), NT, 0, e->stream>>>(
        ak, Count{S, S_dev, S}, bk, L, TILE, nt, init ? (K)~class_bit<K>() : (K)~(K)0, e->splits.p, (K*)e->prevk.p);
    hipEvent_t ev{};
    e->timer.begin("merge", e->stream, &ev);
    if (init)
        merge_kernel<K, V, true><<<nt, NT, 0, e->stream>>>(a, ak, av, Count{S, S_dev, S}, bk, bv, L, (K*)e->slot_k.p,
                                                           (V*)e->slot_v.p, e->tile_cnt.p, e->merge_p.p,
                                                           e->splits.p, (const K*)e->prevk.p, bo, &dm->flags);
    else
        merge_kernel<K, V, false><<<nt, NT, 0, e->stream>>>(a, ak, av, Count{S, S_dev, S}, bk, bv, L, (K*)e->slot_k.p,
                                                            (V*)e->slot_v.p, e->tile_cnt.p, e->merge_p.p,
                                                            e->splits.p, (const K*)e->prevk.p, nullptr,
                                                            &dm->flags);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    e->timer.begin("compact", e->stream, &ev);
    if (int rc = scan_inplace_i64(e, e->tile_cnt.p, nt + 1)) return rc;
    compact_kernel<K, V, merge_ipt<K>::value, FLAG_VECTOR_OVERFLOW><<<nt, NT, 0, e->stream>>>(
        (const K*)e->slot_k.p, (const V*)e->slot_v.p, e->tile_cnt.p, nt, ok, ov, cap, &dm->L_new, &dm->flags);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

// Stored keys from addresses (forward) or addresses from stored keys, in place.
__global__ __launch_bounds__(NT) void mix_keys_kernel(uint64_t* __restrict__ keys, int64_t n, KeyMix m,
                                                      int forward) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    keys[i] = forward ? mix_bits<uint64_t>(m, keys[i]) : unmix_bits<uint64_t>(m, keys[i]);
}

// Vectors of this many entries or more are reordered between address order and stored-key
// order on the device (one-word keys, when the spawn buffers can hold them).
constexpr int64_t DEVICE_REORDER_MIN = 1 << 16;

bool device_reorder(const rimu_engine* e, int64_t n) {
    return e->mix.on && e->kw == 1 && n >= DEVICE_REORDER_MIN && n <= e->max_spawns;
}

// n (key, value) pairs that fill() puts into spawn buffer 0: turn the keys into stored keys
// (forward) or back into addresses, sort the pairs on the key and leave them in spawn buffer
// *result.  A radix look-back that gives up (the sort buffers then hold no list) makes it
// fill and sort again with the per-pass sort.
template <class Fill>
int reorder_on_device(rimu_engine* e, int64_t n, int forward, int* result, Fill&& fill) {
    const int saved = e->radix_mode;
    int rc = RIMU_OK;
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((rc = fill())) break;
        mix_keys_kernel<<<ceil_div(n, NT), NT, 0, e->stream>>>(e->skeys[0].p, n, e->mix, forward);
        if (hipGetLastError() != hipSuccess) { rc = fail(RIMU_ERR_HIP, "mix_keys_kernel launch"); break; }
        int* flags = &((rimu_engine::Misc*)e->misc.p)->flags;
        if (hipMemsetAsync(flags, 0, sizeof(int), e->stream) != hipSuccess) { rc = fail(RIMU_ERR_HIP, "memset"); break; }
        if ((rc = radix_sort<uint64_t, uint64_t>(e, n, e->hdev.B, ~(u128)0, 0, result))) break;
        int h_flags = 0;
        if (hipMemcpyAsync(&h_flags, flags, sizeof(int), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess) { rc = fail(RIMU_ERR_HIP, "flags"); break; }
        if (!(h_flags & FLAG_LOOKBACK)) break;
        if (e->radix_mode == 0) { rc = fail(RIMU_ERR_INTERNAL, "radix look-back did not complete"); break; }
        e->radix_mode = 0;
    }
    e->radix_mode = saved;
    return rc;
}

// Bucket merge (bucket_kernels.hpp) of S spawn records sorted on their top T key bits with
// the vector (bk, bv: keys and diagonal survivors, L entries; bo: the old values, which give
// the entries' initiator class when an initiator rule is set) into (ok, ov) of capacity cap.
// `tiles` receives the number of buckets (rows of merge_p).
constexpr double BK_FILL = 1800;    // mean items (records + entries) per bucket aimed at
constexpr double BK_FILL_MAX = 2050; // highest mean kept to save a radix pass (LDS holds BK_CAP)
constexpr int64_t BK_MIN_RECORDS = 4096;
constexpr int BK_MAX_LOWBITS = 48;   // packed words: key bits below the bucket bits + a 12-bit item index
constexpr int BK_COOLDOWN = 16;      // steps on the merge path after a bucket overflowed

// Top bits that split S records and L entries into buckets of about BK_FILL items; one bit
// fewer when that saves a radix pass (8 bits each) and keeps the mean below BK_FILL_MAX.
int bucket_bits(int64_t S, int64_t L, int B) {
    const double items = (double)S + (double)L;
    int T = (int)std::ceil(std::log2(std::max(items / BK_FILL, 2.0)));
    if (T > 1 && T % RADIX_BITS == 1 && items / std::ldexp(1.0, T - 1) <= BK_FILL_MAX) --T;
    return std::min(std::max(T, 1), std::min(B, 24));
}

template <class V>
int bucket_merge_into(rimu_engine* e, const StepArgs& a, const uint64_t* ak, const V* av, int64_t S,
                      const uint64_t* bk, const V* bv, const V* bo, int64_t L, int T, uint64_t* ok, V* ov,
                      int64_t cap, rimu_engine::Misc* dm, int64_t* tiles, const int64_t* S_dev = nullptr) {
    const int B = e->hdev.B;
    const int64_t nbk = (int64_t)1 << T;
    *tiles = nbk;
    HIP_TRY(e->bk_rstart.alloc(nbk + 1));
    HIP_TRY(e->bk_vstart.alloc(nbk + 1));
    HIP_TRY(e->tile_cnt.alloc(nbk + 1));
    HIP_TRY(e->merge_p.alloc((size_t)nbk * 3));
    HIP_TRY(e->slot_k.alloc((size_t)(S + L)));
    HIP_TRY(e->slot_v.alloc((size_t)(S + L)));
    hipEvent_t ev{};
    e->timer.begin("bucket_bounds", e->stream, &ev);
    bucket_bounds_kernel<<<ceil_div(2 * nbk + 2, NT), NT, 0, e->stream>>>(ak, Count{S, S_dev, S}, bk, L, B - T, nbk,
                                                                          e->bk_rstart.p, e->bk_vstart.p);
    e->timer.end(e->stream);
    HIP_TRY(hipMemsetAsync(e->tile_cnt.p + nbk, 0, sizeof(int64_t), e->stream));
    if (e->n_cu == 0) {
        HIP_TRY(hipDeviceGetAttribute(&e->n_cu, hipDeviceAttributeMultiprocessorCount, e->device));
        if (e->n_cu <= 0) e->n_cu = 1;
    }
    const int64_t grid = std::min(nbk, (int64_t)e->n_cu * BK_WG_PER_CU);
    e->timer.begin("bucket_merge", e->stream, &ev);
    if (a.init_rule != 0)
        bucket_merge_kernel<V, true><<<grid, BK_NT, 0, e->stream>>>(
            a, ak, av, bk, bv, bo, e->bk_rstart.p, e->bk_vstart.p, nbk, B - T, e->slot_k.p, (V*)e->slot_v.p,
            e->tile_cnt.p, e->merge_p.p, &dm->flags);
    else
        bucket_merge_kernel<V, false><<<grid, BK_NT, 0, e->stream>>>(
            a, ak, av, bk, bv, nullptr, e->bk_rstart.p, e->bk_vstart.p, nbk, B - T, e->slot_k.p, (V*)e->slot_v.p,
            e->tile_cnt.p, e->merge_p.p, &dm->flags);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    e->timer.begin("compact", e->stream, &ev);
    if (int rc = scan_inplace_i64(e, e->tile_cnt.p, nbk + 1)) return rc;
    bucket_compact_kernel<V><<<nbk, NT, 0, e->stream>>>(e->slot_k.p, (const V*)e->slot_v.p, e->bk_rstart.p,
                                                        e->bk_vstart.p, e->tile_cnt.p, nbk, ok, ov, cap,
                                                        &dm->L_new, &dm->flags);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    return RIMU_OK;
}

// A HubbardRealSpace model has off-diagonal elements of value zero (which the reference
// never deposits) when a component is fermionic (Pauli blocking) or a site has a hard wall
// in some direction.
bool rs_zero_channels(const ModelDev& h) {
    if (h.fermi) return true;
    for (int s = 0; s < h.M; ++s)
        for (int d = 0; d < h.ndir; ++d)
            if (h.nbr[s][d] == 255) return true;
    return false;
}

template <class K, class V>
int step_impl(rimu_engine* e, int64_t step, double shift, double dtau, double boost,
              rimu_step_stats* out) {
    constexpr int KW = sizeof(K) / 8;
    std::memset(out, 0, sizeof(*out));
    const StepArgs a = make_args(e, step, shift, dtau, boost);
    const int64_t L = e->len;
    const K* vk = (const K*)e->vkeys[e->cur].p;
    const V* vv = (const V*)e->vvals[e->cur].p;
    rimu_engine::Misc* dm = (rimu_engine::Misc*)e->misc.p;
    hipEvent_t ev{};

    HIP_TRY(hipMemsetAsync(dm, 0, sizeof(rimu_engine::Misc), e->stream));

    // Grids of the spawn, sort and merge kernels: sized from the counts of this step, read back
    // after the plan and after the spawn (synchronous step), or -- once a step has run -- from
    // the last step's counts with headroom, the kernels reading this step's counts from device
    // memory (one host synchronisation per step).  A count beyond its grid raises FLAG_SPEC
    // and the step is repeated synchronously.  The sorts with a look-back are needed for it.
    const bool rs = e->hdev.kind == RIMU_MODEL_HUBBARD_REAL_SPACE;
    const bool dense = !e->int_vals && e->style.proj_threshold == 0 && (!rs || !rs_zero_channels(e->hdev));
    const int nbits = e->hdev.B;
    const bool bucket_ok = KW == 1 && e->radix_mode == 2 && e->mix.on && e->bucket_cooldown == 0;
    bool spec = e->spec_ok && e->radix_mode != 0 && e->W_prev > 0;
    int T_spec = 0;
    if (spec) {  // the path the counts of the last step choose; sorts with a host check stay synchronous
        if (bucket_ok && e->S_prev >= BK_MIN_RECORDS) {
            T_spec = bucket_bits(e->S_prev, L, nbits);
            if (nbits - T_spec > BK_MAX_LOWBITS) T_spec = 0;
        }
    }
    const int64_t W_cap = spec ? e->W_prev + e->W_prev / 4 + 16 * SPAWN_TILE : 0;

    // 1. plan: diagonal step, attempt offsets and the first entry of each spawn tile (as
    // many as the tile_first buffer holds; the rest below if W needs more)
    const int64_t plan_tiles = ceil_div(L + 1, PLAN_TILE);
    HIP_TRY(e->plan_p.alloc((size_t)plan_tiles * 6));
    HIP_TRY(e->plan_status.alloc(plan_tiles));
    HIP_TRY(hipMemsetAsync(e->plan_status.p, 0, sizeof(uint64_t) * plan_tiles, e->stream));
    if (spec) HIP_TRY(e->tile_first.alloc(ceil_div(W_cap, SPAWN_TILE) + 1));
    const int64_t tf_cap = (int64_t)e->tile_first.n;
    e->timer.begin("plan", e->stream, &ev);
    if (rs)
        plan_kernel<K, V, true><<<plan_tiles, NT, 0, e->stream>>>(a, vk, vv, L, (V*)e->dvals.p, e->woff.p,
                                                               e->tile_first.p, tf_cap, e->plan_status.p,
                                                               &dm->plan_counter, &dm->flags, e->plan_p.p, &dm->W);
    else
        plan_kernel<K, V, false><<<plan_tiles, NT, 0, e->stream>>>(a, vk, vv, L, (V*)e->dvals.p, e->woff.p,
                                                                e->tile_first.p, tf_cap, e->plan_status.p,
                                                                &dm->plan_counter, &dm->flags, e->plan_p.p, &dm->W);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    int64_t W = W_cap;  // synchronous: the count; speculative: the most the grids are sized for
    if (!spec) {
        HIP_TRY(hipMemcpyAsync(e->h_misc, dm, sizeof(rimu_engine::Misc), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (e->h_misc->flags & FLAG_LOOKBACK) return fail(RIMU_ERR_INTERNAL, "plan look-back did not complete");
        W = e->h_misc->W;
    }

    // 2. spawn (W = number of attempts).  Dense spawning (every attempt yields a record)
    // writes the record of attempt a at position a; otherwise tiles write to their own slot
    // ranges and a scan of the tile counts plus a compaction place the records.
    int64_t S = 0;
    int64_t spawn_tiles = 0;
    u128 kor = 0, kand = ~(u128)0;
    const int64_t* S_dev = nullptr;  // speculative: the record count in device memory
    if (W > 0) {
        spawn_tiles = ceil_div(W, SPAWN_TILE);
        // every attempt yields a record when values are floating and deposits unprojected,
        // unless the model has zero-valued channels (hard walls, Pauli blocking: the
        // reference never deposits their spawns)
        if (!spec && dense && W > e->max_spawns) {
            out->overflow = 1;
            return fail(RIMU_ERR_CAPACITY, "spawn buffer capacity exceeded");
        }
        HIP_TRY(e->spawn_p.alloc(spawn_tiles));
        HIP_TRY(e->tile_or.alloc((size_t)spawn_tiles * KW));
        HIP_TRY(e->tile_and.alloc((size_t)spawn_tiles * KW));
        if (spawn_tiles > tf_cap) {  // plan filled only the first tf_cap tile heads
            HIP_TRY(e->tile_first.alloc(spawn_tiles));
            tile_first_kernel<<<ceil_div(L, NT), NT, 0, e->stream>>>(e->woff.p, L, e->tile_first.p);
        }
        K* sk = (K*)e->skeys[0].p;
        V* sv = (V*)e->svals[0].p;
        if (!dense) {
            HIP_TRY(e->slot_k.alloc((size_t)spawn_tiles * SPAWN_TILE * KW));
            HIP_TRY(e->slot_v.alloc((size_t)spawn_tiles * SPAWN_TILE));
            HIP_TRY(e->tile_cnt.alloc(spawn_tiles + 1));
            HIP_TRY(hipMemsetAsync(e->tile_cnt.p + spawn_tiles, 0, sizeof(int64_t), e->stream));
            sk = (K*)e->slot_k.p;
            sv = (V*)e->slot_v.p;
        }
        const Count cA{W, spec ? &dm->W : nullptr, spawn_tiles * SPAWN_TILE};
        e->timer.begin("spawn", e->stream, &ev);
        if (dense && rs)
            spawn_kernel<K, V, true, true><<<spawn_tiles, NT, 0, e->stream>>>(
                a, vk, vv, L, e->woff.p, cA, e->max_spawns, &dm->flags, sk, sv, e->tile_cnt.p, e->spawn_p.p,
                e->tile_first.p, (K*)e->tile_or.p, (K*)e->tile_and.p);
        else if (dense)
            spawn_kernel<K, V, true, false><<<spawn_tiles, NT, 0, e->stream>>>(
                a, vk, vv, L, e->woff.p, cA, e->max_spawns, &dm->flags, sk, sv, e->tile_cnt.p, e->spawn_p.p,
                e->tile_first.p, (K*)e->tile_or.p, (K*)e->tile_and.p);
        else if (rs)
            spawn_kernel<K, V, false, true><<<spawn_tiles, NT, 0, e->stream>>>(
                a, vk, vv, L, e->woff.p, cA, e->max_spawns, &dm->flags, sk, sv, e->tile_cnt.p, e->spawn_p.p,
                e->tile_first.p, (K*)e->tile_or.p, (K*)e->tile_and.p);
        else
            spawn_kernel<K, V, false, false><<<spawn_tiles, NT, 0, e->stream>>>(
                a, vk, vv, L, e->woff.p, cA, e->max_spawns, &dm->flags, sk, sv, e->tile_cnt.p, e->spawn_p.p,
                e->tile_first.p, (K*)e->tile_or.p, (K*)e->tile_and.p);
        e->timer.end(e->stream);
        HIP_TRY(hipGetLastError());
        if (!spec)  // a speculative step sorts every digit and reads no OR/AND back
            key_bits_kernel<K><<<1, 1024, 0, e->stream>>>((const K*)e->tile_or.p, (const K*)e->tile_and.p,
                                                          spawn_tiles, (K*)&dm->key_or, (K*)&dm->key_and);
        if (!dense) {
            e->timer.begin("spawn_compact", e->stream, &ev);
            if (int rc = scan_inplace_i64(e, e->tile_cnt.p, spawn_tiles + 1)) return rc;
            compact_kernel<K, V, SPAWN_ITEMS, FLAG_SPAWN_OVERFLOW><<<spawn_tiles, NT, 0, e->stream>>>(
                (const K*)e->slot_k.p, (const V*)e->slot_v.p, e->tile_cnt.p, spawn_tiles, (K*)e->skeys[0].p,
                (V*)e->svals[0].p, e->max_spawns, &dm->S, &dm->flags);
            e->timer.end(e->stream);
        }
        HIP_TRY(hipGetLastError());
        if (spec) {
            // at most W_cap records (one per attempt), at most max_spawns in the buffer; every
            // digit over the bucket or key bits is sorted (no OR/AND read back)
            S = std::min(W, e->max_spawns);
            S_dev = dense ? &dm->W : &dm->S;
            kor = nbits >= 128 ? ~(u128)0 : (((u128)1 << nbits) - 1);
            kand = 0;
        } else {
            HIP_TRY(hipMemcpyAsync(e->h_misc, dm, sizeof(rimu_engine::Misc), hipMemcpyDeviceToHost, e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
            if (e->h_misc->flags & FLAG_SPAWN_OVERFLOW) {
                e->timer.collect();
                out->overflow = 1;
                return fail(RIMU_ERR_CAPACITY, "spawn buffer capacity exceeded");
            }
            S = dense ? W : e->h_misc->S;
            kor = e->h_misc->key_or;
            kand = e->h_misc->key_and;
            if (KW == 1) { kor = (uint64_t)kor; kand = (uint64_t)kand; }
        }
    }

    // 3-4. sort the spawn records, merge them with the diagonal survivors, fold, compress,
    // count.  Eligible steps sort on the top T bits only and merge bucket by bucket; if a
    // bucket outgrows its LDS tile, the records are sorted on the remaining bits and merged
    // along the merge path instead (same result).
    const int nxt = 1 - e->cur;
    int64_t merge_tiles = 0;
    int sorted = 0;
    int T = T_spec;
    if (!spec && bucket_ok && S >= BK_MIN_RECORDS) {
        T = bucket_bits(S, L, nbits);
        if (nbits - T > BK_MAX_LOWBITS) T = 0;
    }
    if (e->bucket_cooldown > 0) --e->bucket_cooldown;
    const int64_t* merge_S_dev = S_dev;  // the merge path after the fallback sort reads S from the host
    auto merge_path = [&]() -> int {
        return merge_into<K, V>(e, a, (const K*)e->skeys[sorted].p, (const V*)e->svals[sorted].p, S, vk,
                                (const V*)e->dvals.p, L, (K*)e->vkeys[nxt].p, (V*)e->vvals[nxt].p,
                                e->max_entries, dm, &merge_tiles, vv, merge_S_dev);
    };
    if (T > 0) {
        if (int rc = radix_sort<K, V>(e, S, nbits, kor ^ kand, 0, &sorted, nbits - T, S_dev)) return rc;
        if (int rc = bucket_merge_into<V>(e, a, (const uint64_t*)e->skeys[sorted].p, (const V*)e->svals[sorted].p,
                                          S, (const uint64_t*)vk, (const V*)e->dvals.p, vv, L, T,
                                          (uint64_t*)e->vkeys[nxt].p, (V*)e->vvals[nxt].p, e->max_entries, dm,
                                          &merge_tiles, S_dev)) return rc;
    } else {
        if (S > 1 || S_dev) {
            if (!e->mix.on || nbits <= e->sort_bits) {
                if (int rc = radix_sort<K, V>(e, S, nbits, kor ^ kand, 0, &sorted, 0, S_dev)) return rc;
            } else if (int rc = sort_mixed<K, V>(e, S, nbits, kor ^ kand, a.init_rule != 0, dm, &sorted, S_dev)) {
                return rc;
            }
        }
        if (int rc = merge_path()) return rc;
    }

    // 5. projections and statistics
    auto finish = [&]() -> int {
        if (e->np + e->nh > 0) {
            project_kernel<K, V><<<ceil_div(e->np + e->nh, NT), NT, 0, e->stream>>>(
                (const K*)e->vkeys[nxt].p, (const V*)e->vvals[nxt].p, &dm->L_new, (const K*)e->pkeys.p, e->pcoef.p,
                e->np + e->nh, e->terms.p);
            HIP_TRY(hipGetLastError());
        }
        PartialSet ps{};
        // plan partials rows: deaths, clones, zombies, exact, inexact, attempts
        const int plan_map[6] = {4, 5, 6, 1, 2, 0};
        for (int r = 0; r < 6; ++r) {
            ps.p[plan_map[r]] = e->plan_p.p + (int64_t)r * plan_tiles;
            ps.n[plan_map[r]] = plan_tiles;
        }
        ps.p[3] = e->spawn_p.p;
        ps.n[3] = spawn_tiles;
        for (int r = 0; r < 3; ++r) {
            ps.p[7 + r] = e->merge_p.p + (int64_t)r * merge_tiles;
            ps.n[7 + r] = merge_tiles;
        }
        ps.p[10] = e->terms.p;
        ps.n[10] = e->np;
        ps.p[11] = e->terms.p + e->np;
        ps.n[11] = e->nh;
        partial_sums_kernel<<<dim3(PARTIAL_BLOCKS, NSTAT), NT, 0, e->stream>>>(ps, e->chunk_sums.p);
        ShiftUpdate su{};
        if (e->shift_mode) {
            su.in = a.shift_dev;
            su.out = e->shift_state.p + 2 * (1 - e->shift_cur);
            su.dtau = dtau;
            su.target = e->dlu_target;
            su.zeta = e->dlu_zeta;
            su.xi = e->dlu_xi;
        }
        finalize_kernel<<<1, 64, 0, e->stream>>>(e->chunk_sums.p, PARTIAL_BLOCKS, e->stats_dev.p, su);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(e->h_stats, e->stats_dev.p, (NSTAT + 2) * sizeof(double), hipMemcpyDeviceToHost,
                               e->stream));
        HIP_TRY(hipMemcpyAsync(e->h_misc, dm, sizeof(rimu_engine::Misc), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        e->timer.collect();
        return RIMU_OK;
    };
    if (int rc = finish()) return rc;
    if (spec && (e->h_misc->flags & (FLAG_SPEC | FLAG_FIXUP_FALLBACK))) {
        // a count outgrew the grids sized from the last step, or the runs left after a sort on
        // the top bits need the full sort: repeat the step synchronously (the vector is
        // unchanged; same result)
        e->timer.begin("spec_repeat", e->stream, &ev);
        e->timer.end(e->stream);
        e->spec_ok = false;
        return step_impl<K, V>(e, step, shift, dtau, boost, out);
    }
    if (e->h_misc->flags & FLAG_SPAWN_OVERFLOW) {
        out->overflow = 1;
        e->spec_ok = false;
        return fail(RIMU_ERR_CAPACITY, "spawn buffer capacity exceeded");
    }
    if ((e->h_misc->flags & FLAG_LOOKBACK) && e->radix_mode != 0) {
        // a radix look-back wait gave up (only forced from the host: every tile waited on has
        // started): the sort buffers hold no list any more, the vector is unchanged; repeat the
        // whole step with the per-pass sort, which waits on no other tile (same result)
        e->timer.begin("lookback_fallback", e->stream, &ev);
        e->timer.end(e->stream);
        const int saved = e->radix_mode;
        e->radix_mode = 0;
        const int rc = step_impl<K, V>(e, step, shift, dtau, boost, out);
        e->radix_mode = saved;
        return rc;
    }
    if (e->h_misc->flags & FLAG_LOOKBACK) return fail(RIMU_ERR_INTERNAL, "radix look-back did not complete");
    if (e->h_misc->flags & FLAG_BUCKET_OVERFLOW) {
        // sort the bucketed records on every bit (stable passes keep their list order within
        // an address) and merge along the merge path
        e->bucket_cooldown = BK_COOLDOWN;
        e->timer.begin("bucket_fallback", e->stream, &ev);
        e->timer.end(e->stream);
        HIP_TRY(hipMemsetAsync(&dm->flags, 0, sizeof(int), e->stream));
        if (spec) {  // the counts are known now
            S = dense ? e->h_misc->W : e->h_misc->S;
            merge_S_dev = nullptr;
        }
        int full = sorted;
        if (int rc = radix_sort<K, V>(e, S, nbits, kor ^ kand, sorted, &full)) return rc;
        sorted = full;
        if (int rc = merge_path()) return rc;
        if (int rc = finish()) return rc;
        if (e->h_misc->flags & FLAG_LOOKBACK) return fail(RIMU_ERR_INTERNAL, "radix look-back did not complete");
    }
    if (e->h_misc->flags & FLAG_VECTOR_OVERFLOW) {
        out->overflow = 1;
        return fail(RIMU_ERR_CAPACITY, "vector capacity (max_entries) exceeded");
    }
    const double* s = e->h_stats;
    out->spawn_attempts = s[0];
    out->exact_steps = s[1];
    out->inexact_steps = s[2];
    out->spawns = s[3];
    out->deaths = s[4];
    out->clones = s[5];
    out->zombies = s[6];
    out->len_before = s[7];
    out->len = s[8];
    out->norm = s[9];
    out->vproj = s[10];
    out->hproj = s[11];
    const int64_t L_new = e->h_misc->L_new;
    e->cur = nxt;
    e->len = L_new;
    if (e->shift_mode) {
        e->shift_cur = 1 - e->shift_cur;
        e->h_shift = s[NSTAT];
        e->h_pnorm = s[NSTAT + 1];
    }
    // the counts that size the next step's grids
    e->W_prev = e->h_misc->W;
    e->S_prev = dense ? e->h_misc->W : e->h_misc->S;
    e->spec_ok = true;
    // algorithmic bytes: read v, write S records, read S records into the merge, write v'
    const double rec = 8.0 * KW + (double)e->vsize;
    const int64_t S_act = dense ? e->h_misc->W : e->h_misc->S;  // records this step emitted
    const double records = (double)(S_act + L);  // diagonal survivors included
    e->last_records = records;
    e->last_alg_bytes = (double)L * rec + 2.0 * records * rec + (double)L_new * rec;
    return RIMU_OK;
}

template <class K>
int overlaps_impl(rimu_engine* left, rimu_engine* right, const rimu_operator* ops, int32_t nops,
                  double* out);

}  // namespace

extern "C" {

const char* rimu_last_error(void) { return g_last_error.c_str(); }

int rimu_device_count(int32_t* count) {
    int n = 0;
    hipError_t err = hipGetDeviceCount(&n);
    if (err != hipSuccess) n = 0;
    if (count) *count = n;
    return RIMU_OK;
}

int rimu_engine_create(const rimu_model* model, const rimu_style* style, uint64_t seed,
                       int64_t max_entries, int64_t max_spawns, rimu_engine** out) {
    if (!out) return fail(RIMU_ERR_ARG, "out is null");
    *out = nullptr;
    if (int rc = check_model(model)) return rc;
    if (!style) return fail(RIMU_ERR_ARG, "style is null");
    if (style->kind < 1 || style->kind > 3) return fail(RIMU_ERR_ARG, "unknown style kind");
    if (style->kind == RIMU_STYLE_STOCHASTIC_INTEGER && style->proj_threshold != 0)
        return fail(RIMU_ERR_ARG, "Thresholding not supported for integer spawns");
    if (max_entries < 1 || max_spawns < 1) return fail(RIMU_ERR_ARG, "capacities must be positive");
    if (max_spawns >= (int64_t)UINT32_MAX) return fail(RIMU_ERR_ARG, "max_spawns must be below 2^32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RIMU_ERR_NODEVICE, "no HIP device");
    rimu_engine* e = new rimu_engine();
    e->model = *model;
    e->style = *style;
    e->hdev = to_dev(*model);
    e->mix = make_key_mix(e->hdev.B);
    e->seed = seed;
    e->int_vals = style->kind == RIMU_STYLE_STOCHASTIC_INTEGER;
    e->kw = e->hdev.B > 64 ? 2 : 1;
    e->vsize = 8;
    e->max_entries = max_entries;
    e->max_spawns = max_spawns;
    auto bad = [&](const std::string& what, hipError_t err) {
        std::string m = what + ": " + hipGetErrorString(err);
        rimu_engine_destroy(e);
        return fail(RIMU_ERR_HIP, m);
    };
    hipError_t err;
    if ((err = hipGetDevice(&e->device)) != hipSuccess) return bad("hipGetDevice", err);
    if ((err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) return bad("stream", err);
    for (int b = 0; b < 2; ++b) {
        if ((err = e->vkeys[b].alloc(max_entries * e->kw)) != hipSuccess) return bad("vector keys", err);
        if ((err = e->vvals[b].alloc(max_entries * 8)) != hipSuccess) return bad("vector values", err);
        if ((err = e->skeys[b].alloc(max_spawns * e->kw)) != hipSuccess) return bad("spawn keys", err);
        if ((err = e->svals[b].alloc(max_spawns * 8)) != hipSuccess) return bad("spawn values", err);
    }
    if ((err = e->dvals.alloc(max_entries * 8)) != hipSuccess) return bad("diagonal values", err);
    if ((err = e->woff.alloc(max_entries + 1)) != hipSuccess) return bad("work offsets", err);
    if ((err = e->misc.alloc(sizeof(rimu_engine::Misc))) != hipSuccess) return bad("misc", err);
    if ((err = e->stats_dev.alloc(16)) != hipSuccess) return bad("stats", err);
    if ((err = e->terms.alloc(16)) != hipSuccess) return bad("terms", err);
    if ((err = e->chunk_sums.alloc(PARTIAL_BLOCKS * NSTAT)) != hipSuccess) return bad("chunk sums", err);
    if ((err = e->spawn_p.alloc(16)) != hipSuccess) return bad("partials", err);
    if ((err = hipHostMalloc((void**)&e->h_misc, sizeof(rimu_engine::Misc), hipHostMallocDefault)) != hipSuccess)
        return bad("pinned misc", err);
    if ((err = hipHostMalloc((void**)&e->h_stats, 16 * sizeof(double), hipHostMallocDefault)) != hipSuccess)
        return bad("pinned stats", err);
    *out = e;
    return RIMU_OK;
}

int rimu_engine_destroy(rimu_engine* e) {
    if (!e) return RIMU_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (int b = 0; b < 2; ++b) {
        e->vkeys[b].release();
        e->vvals[b].release();
        e->skeys[b].release();
        e->svals[b].release();
    }
    e->tile_first.release(); e->splits.release(); e->tile_cnt.release();
    e->prevk.release(); e->slot_k.release(); e->slot_v.release();
    e->dvals.release(); e->woff.release(); e->scan_tmp.release(); e->rcounts.release(); e->rchunks.release();
    e->rtmp.release(); e->plan_status.release(); e->tile_or.release(); e->tile_and.release(); e->plan_p.release(); e->spawn_p.release();
    e->merge_p.release(); e->terms.release(); e->chunk_sums.release(); e->stats_dev.release(); e->shift_state.release(); e->misc.release();
    e->pkeys.release(); e->pcoef.release();
    e->ov_ops.release(); e->ov_part.release(); e->ov_out.release(); e->ov_ranges.release(); e->ov_index.release();
    if (e->h_misc) (void)hipHostFree(e->h_misc);
    if (e->h_stats) (void)hipHostFree(e->h_stats);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return RIMU_OK;
}

int rimu_vector_upload(rimu_engine* e, const uint64_t* keys, const void* vals, int64_t n) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (n < 0 || (n > 0 && (!keys || !vals))) return fail(RIMU_ERR_ARG, "bad vector arguments");
    const int kw = e->kw;
    auto key_at = [&](int64_t i) -> u128 {
        return kw == 1 ? (u128)keys[i] : (((u128)keys[2 * i + 1] << 64) | keys[2 * i]);
    };
    std::vector<uint64_t> k;
    std::vector<int64_t> vi;
    std::vector<double> vd;
    std::vector<u128> stored;  // mix(address) of the kept entries
    const bool on_device = device_reorder(e, n);  // mix and sort on the device
    k.reserve((size_t)n * kw);
    u128 prev = 0;
    for (int64_t i = 0; i < n; ++i) {
        const u128 key = key_at(i);
        if (i > 0 && key <= prev) return fail(RIMU_ERR_ARG, "keys must be strictly ascending");
        prev = key;
        if (!address_ok(e->hdev, key))
            return fail(RIMU_ERR_ARG, "key is not an address of the model (bits or particle numbers)");
        if (e->int_vals) {
            const int64_t v = ((const int64_t*)vals)[i];
            if (v == 0) continue;
            vi.push_back(v);
        } else {
            const double v = ((const double*)vals)[i];
            if (v == 0) continue;
            vd.push_back(v);
        }
        if (e->mix.on && !on_device) {
            stored.push_back(kw == 1 ? (u128)mix_bits<uint64_t>(e->mix, (uint64_t)key) : mix_bits<u128>(e->mix, key));
        } else {
            for (int w = 0; w < kw; ++w) k.push_back(keys[(size_t)kw * i + w]);
        }
    }
    if (e->mix.on && !on_device) {  // the device vector is kept in the order of the stored keys
        std::vector<int64_t> ord(stored.size());
        for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int64_t)i;
        std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return stored[a] < stored[b]; });
        std::vector<int64_t> vi2;
        std::vector<double> vd2;
        for (int64_t i : ord) {
            k.push_back((uint64_t)stored[i]);
            if (kw == 2) k.push_back((uint64_t)(stored[i] >> 64));
            if (e->int_vals) vi2.push_back(vi[i]); else vd2.push_back(vd[i]);
        }
        vi.swap(vi2);
        vd.swap(vd2);
    }
    const int64_t m = (int64_t)(k.size() / kw);
    if (m > e->max_entries) return fail(RIMU_ERR_CAPACITY, "vector longer than max_entries");
    if (on_device && m > 0) {
        int sorted = 0;
        if (int rc = reorder_on_device(e, m, 1, &sorted, [&]() -> int {
                HIP_TRY(hipMemcpyAsync(e->skeys[0].p, k.data(), m * 8, hipMemcpyHostToDevice, e->stream));
                HIP_TRY(hipMemcpyAsync(e->svals[0].p, e->int_vals ? (const void*)vi.data() : (const void*)vd.data(),
                                       m * 8, hipMemcpyHostToDevice, e->stream));
                return RIMU_OK;
            })) return rc;
        HIP_TRY(hipMemcpyAsync(e->vkeys[e->cur].p, e->skeys[sorted].p, m * 8, hipMemcpyDeviceToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->vvals[e->cur].p, e->svals[sorted].p, m * 8, hipMemcpyDeviceToDevice, e->stream));
    } else if (m > 0) {
        HIP_TRY(hipMemcpyAsync(e->vkeys[e->cur].p, k.data(), m * 8 * kw, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->vvals[e->cur].p, e->int_vals ? (const void*)vi.data() : (const void*)vd.data(),
                               m * 8, hipMemcpyHostToDevice, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->len = m;
    return RIMU_OK;
}

int rimu_vector_length(rimu_engine* e, int64_t* n) {
    if (!e || !n) return fail(RIMU_ERR_ARG, "null argument");
    *n = e->len;
    return RIMU_OK;
}

int rimu_vector_download(rimu_engine* e, uint64_t* keys, void* vals, int64_t cap, int64_t* n) {
    if (!e || !n) return fail(RIMU_ERR_ARG, "null argument");
    *n = e->len;
    if (e->len > cap) return fail(RIMU_ERR_CAPACITY, "download capacity too small");
    if (device_reorder(e, e->len)) {  // addresses back and address order on the device
        int sorted = 0;
        if (int rc = reorder_on_device(e, e->len, 0, &sorted, [&]() -> int {
                HIP_TRY(hipMemcpyAsync(e->skeys[0].p, e->vkeys[e->cur].p, e->len * 8, hipMemcpyDeviceToDevice,
                                       e->stream));
                HIP_TRY(hipMemcpyAsync(e->svals[0].p, e->vvals[e->cur].p, e->len * 8, hipMemcpyDeviceToDevice,
                                       e->stream));
                return RIMU_OK;
            })) return rc;
        HIP_TRY(hipMemcpyAsync(keys, e->skeys[sorted].p, e->len * 8, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(vals, e->svals[sorted].p, e->len * 8, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        return RIMU_OK;
    }
    if (e->len > 0) {
        HIP_TRY(hipMemcpyAsync(keys, e->vkeys[e->cur].p, e->len * 8 * e->kw, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(vals, e->vvals[e->cur].p, e->len * 8, hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->mix.on && e->len > 0) unmix_and_sort(e, keys, vals, e->len);
    return RIMU_OK;
}

int rimu_set_projector(rimu_engine* e, const uint64_t* pk, const double* pc, int64_t n,
                       const uint64_t* hk, const double* hc, int64_t nh) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (n < 0 || nh < 0) return fail(RIMU_ERR_ARG, "negative projector length");
    const int64_t tot = n + nh;
    HIP_TRY(e->pkeys.alloc(std::max<int64_t>(tot, 1) * e->kw));
    HIP_TRY(e->pcoef.alloc(std::max<int64_t>(tot, 1)));
    HIP_TRY(e->terms.alloc(std::max<int64_t>(tot, 16)));
    std::vector<uint64_t> mk;  // projector keys in the stored form
    if (e->mix.on) {
        mk.resize((size_t)tot * e->kw);
        for (int64_t i = 0; i < tot; ++i) {
            const uint64_t* src = i < n ? pk + i * e->kw : hk + (i - n) * e->kw;
            if (e->kw == 1) {
                mk[i] = mix_bits<uint64_t>(e->mix, src[0]);
            } else {
                const u128 y = mix_bits<u128>(e->mix, ((u128)src[1] << 64) | src[0]);
                mk[2 * i] = (uint64_t)y;
                mk[2 * i + 1] = (uint64_t)(y >> 64);
            }
        }
        pk = mk.data();
        hk = mk.data() + n * e->kw;
    }
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(e->pkeys.p, pk, n * 8 * e->kw, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->pcoef.p, pc, n * 8, hipMemcpyHostToDevice, e->stream));
    }
    if (nh > 0) {
        HIP_TRY(hipMemcpyAsync(e->pkeys.p + n * e->kw, hk, nh * 8 * e->kw, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->pcoef.p + n, hc, nh * 8, hipMemcpyHostToDevice, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->np = n;
    e->nh = nh;
    return RIMU_OK;
}

int rimu_step(rimu_engine* e, int64_t step, double shift, double dtau, double boost,
              rimu_step_stats* stats) {
    if (!e || !stats) return fail(RIMU_ERR_ARG, "null argument");
    if (step < 0 || step >= (1ll << 30)) return fail(RIMU_ERR_ARG, "step must be in [0, 2^30)");
    if (!(dtau > 0) || !(boost > 0)) return fail(RIMU_ERR_ARG, "dtau and boost must be positive");
    if (e->len == 0) {
        std::memset(stats, 0, sizeof(*stats));
        return RIMU_OK;
    }
    if (e->len > e->max_entries) return fail(RIMU_ERR_CAPACITY, "vector longer than max_entries");
    if (e->kw == 2)
        return e->int_vals ? step_impl<u128, int64_t>(e, step, shift, dtau, boost, stats)
                           : step_impl<u128, double>(e, step, shift, dtau, boost, stats);
    return e->int_vals ? step_impl<uint64_t, int64_t>(e, step, shift, dtau, boost, stats)
                       : step_impl<uint64_t, double>(e, step, shift, dtau, boost, stats);
}

int rimu_overlaps(rimu_engine* left, rimu_engine* right, const rimu_operator* ops, int32_t nops,
                  double* out) {
    if (!left || !right || !out || (nops > 0 && !ops)) return fail(RIMU_ERR_ARG, "null argument");
    if (nops < 0 || nops > OV_MAX_OPS) return fail(RIMU_ERR_ARG, "between 0 and 8 operators");
    if (!same_space(left->hdev, right->hdev))
        return fail(RIMU_ERR_ARG, "vectors of different address spaces");
    if (right->kw == 2) return overlaps_impl<u128>(left, right, ops, nops, out);
    return overlaps_impl<uint64_t>(left, right, ops, nops, out);
}

}  // extern "C"

namespace {
template <class K>
int overlaps_impl(rimu_engine* left, rimu_engine* right, const rimu_operator* ops, int32_t nops,
                  double* out) {
    if (left->device != right->device) return fail(RIMU_ERR_ARG, "vectors on different devices");
    const int M = right->hdev.M;
    std::vector<OpDev> dops((size_t)std::max(nops, 1));
    for (int o = 0; o < nops; ++o) {
        OpDev d{};
        d.kind = ops[o].kind;
        d.param = ops[o].param;
        d.mix = right->mix;
        if (d.kind == RIMU_OP_HAMILTONIAN) {
            if (int rc = check_model(&ops[o].model)) return rc;
            d.h = to_dev(ops[o].model);
            if (!same_space(d.h, right->hdev))
                return fail(RIMU_ERR_ARG, "operator Hamiltonian on another address space");
        } else if (right->hdev.kind == RIMU_MODEL_HUBBARD_REAL_SPACE) {
            return fail(RIMU_ERR_ARG, "G2RealCorrelator and DensityMatrixDiagonal act on one-component 1D BoseFS vectors");
        } else if (d.kind == RIMU_OP_G2_REAL) {
            if (d.param <= -M || d.param >= M) return fail(RIMU_ERR_ARG, "G2RealCorrelator needs |d| < M");
        } else if (d.kind == RIMU_OP_DENSITY_DIAGONAL) {
            if (d.param < 1 || d.param > M) return fail(RIMU_ERR_ARG, "DensityMatrixDiagonal mode outside 1..M");
        } else {
            return fail(RIMU_ERR_ARG, "unknown operator kind");
        }
        dops[o] = d;
    }
    for (int o = 0; o <= nops; ++o) out[o] = 0.0;
    const int64_t Lt = left->len, Lu = right->len;
    if (Lt == 0 || Lu == 0) return RIMU_OK;
    rimu_engine* e = right;
    const int64_t nb = ceil_div(Lu, OV_TILE);
    HIP_TRY(e->ov_ops.alloc(sizeof(OpDev) * dops.size()));
    HIP_TRY(e->ov_part.alloc((size_t)nb * (nops + 1)));
    HIP_TRY(e->ov_out.alloc(OV_MAX_OPS + 1));
    HIP_TRY(e->ov_ranges.alloc(nb + 1));
    HIP_TRY(hipMemcpyAsync(e->ov_ops.p, dops.data(), sizeof(OpDev) * dops.size(), hipMemcpyHostToDevice,
                           e->stream));
    const K* tk = (const K*)left->vkeys[left->cur].p;
    const void* tv = left->vvals[left->cur].p;
    const K* uk = (const K*)right->vkeys[right->cur].p;
    const void* uv = right->vvals[right->cur].p;
    const OpDev* dp = (const OpDev*)e->ov_ops.p;
    int need_split = 0;
    for (int o = 0; o < nops; ++o) need_split |= dops[o].kind == RIMU_OP_HAMILTONIAN;
    // mixed keys are spread evenly over their B bits: neighbours are looked up through an
    // index of t by the top P bits, about 16 keys per index entry
    const int64_t* tidx = nullptr;
    int tsh = 0;
    hipEvent_t ev{};
    e->timer.begin("overlap", e->stream, &ev);
    if (need_split && right->mix.on && Lt >= OV_INDEX_MIN) {
        const int B = right->hdev.B;
        int P = (int)std::floor(std::log2((double)Lt / 16.0));
        P = std::max(1, std::min({P, OV_INDEX_MAX_BITS, B}));
        tsh = B - P;
        HIP_TRY(e->ov_index.alloc(((size_t)1 << P) + 1));
        overlap_index_kernel<K><<<ceil_div(Lt + 1, NT), NT, 0, e->stream>>>(tk, Lt, B, P, e->ov_index.p);
        tidx = e->ov_index.p;
    }
    overlap_ranges_kernel<K><<<ceil_div(nb + 1, NT), NT, 0, e->stream>>>(tk, Lt, uk, nb, e->ov_ranges.p);
    if (left->int_vals && right->int_vals)
        overlap_kernel<K, int64_t, int64_t><<<nb, NT, 0, e->stream>>>(tk, (const int64_t*)tv, Lt, uk,
            (const int64_t*)uv, Lu, dp, nops, M, e->ov_ranges.p, need_split, tidx, tsh, e->ov_part.p);
    else if (left->int_vals)
        overlap_kernel<K, int64_t, double><<<nb, NT, 0, e->stream>>>(tk, (const int64_t*)tv, Lt, uk,
            (const double*)uv, Lu, dp, nops, M, e->ov_ranges.p, need_split, tidx, tsh, e->ov_part.p);
    else if (right->int_vals)
        overlap_kernel<K, double, int64_t><<<nb, NT, 0, e->stream>>>(tk, (const double*)tv, Lt, uk,
            (const int64_t*)uv, Lu, dp, nops, M, e->ov_ranges.p, need_split, tidx, tsh, e->ov_part.p);
    else
        overlap_kernel<K, double, double><<<nb, NT, 0, e->stream>>>(tk, (const double*)tv, Lt, uk,
            (const double*)uv, Lu, dp, nops, M, e->ov_ranges.p, need_split, tidx, tsh, e->ov_part.p);
    e->timer.end(e->stream);
    HIP_TRY(hipGetLastError());
    overlap_final_kernel<<<nops + 1, NT, 0, e->stream>>>(e->ov_part.p, nb, e->ov_out.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, e->ov_out.p, sizeof(double) * (nops + 1), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->timer.collect();
    return RIMU_OK;
}
}  // namespace

extern "C" {

int rimu_engine_set_initiator(rimu_engine* e, int32_t rule, double threshold) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (rule < RIMU_INITIATOR_NONE || rule > RIMU_INITIATOR_COHERENT)
        return fail(RIMU_ERR_ARG, "unknown initiator rule");
    if (std::isnan(threshold)) return fail(RIMU_ERR_ARG, "initiator threshold is NaN");
    if (rule != RIMU_INITIATOR_NONE && e->hdev.B > 64 * e->kw - 1)
        return fail(RIMU_ERR_ARG, "initiator rules need N + M - 1 <= 63 or 65..127 (the top key bit holds the class)");
    e->init_rule = rule;
    e->init_threshold = rule == RIMU_INITIATOR_NONE ? 0.0 : threshold;
    return RIMU_OK;
}

int rimu_engine_set_shift_update(rimu_engine* e, int32_t mode, double shift, double pnorm,
                                 double target_walkers, double zeta, double xi) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (mode != RIMU_SHIFT_HOST && mode != RIMU_SHIFT_DOUBLE_LOG) return fail(RIMU_ERR_ARG, "unknown shift update");
    if (mode == RIMU_SHIFT_HOST) {
        e->shift_mode = mode;
        return RIMU_OK;
    }
    if (!std::isfinite(shift) || !(pnorm > 0) || !std::isfinite(pnorm))
        return fail(RIMU_ERR_ARG, "shift must be finite and pnorm positive");
    if (!(target_walkers > 0) || !std::isfinite(target_walkers) || !std::isfinite(zeta) || !std::isfinite(xi))
        return fail(RIMU_ERR_ARG, "target_walkers must be positive, zeta and xi finite");
    HIP_TRY(e->shift_state.alloc(4));
    const double init[2] = {shift, pnorm};
    HIP_TRY(hipMemcpyAsync(e->shift_state.p, init, sizeof(init), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->shift_cur = 0;
    e->h_shift = shift;
    e->h_pnorm = pnorm;
    e->dlu_target = target_walkers;
    e->dlu_zeta = zeta;
    e->dlu_xi = xi;
    e->shift_mode = mode;
    return RIMU_OK;
}

int rimu_engine_get_shift(rimu_engine* e, double* shift, double* pnorm) {
    if (!e || !shift || !pnorm) return fail(RIMU_ERR_ARG, "null argument");
    if (e->shift_mode == RIMU_SHIFT_HOST) return fail(RIMU_ERR_ARG, "the engine has no device shift update");
    *shift = e->h_shift;
    *pnorm = e->h_pnorm;
    return RIMU_OK;
}

int rimu_engine_set_sort_bits(rimu_engine* e, int32_t bits) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (bits < 8 || bits > 128) return fail(RIMU_ERR_ARG, "sort bits must be in 8..128");
    e->sort_bits = bits;
    return RIMU_OK;
}

int rimu_engine_set_radix_mode(rimu_engine* e, int32_t mode) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (mode < 0 || mode > 2) return fail(RIMU_ERR_ARG, "radix mode must be 0, 1 or 2");
    e->radix_mode = mode;
    return RIMU_OK;
}

int rimu_engine_set_lookback(rimu_engine* e, uint32_t spin_limit, int32_t withhold_first) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    e->lb_spin_limit = spin_limit;
    e->lb_withhold = withhold_first != 0;
    return RIMU_OK;
}

int rimu_engine_set_timing(rimu_engine* e, int32_t enabled) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    e->timer.enabled = enabled != 0;
    e->timer.totals.clear();
    return RIMU_OK;
}

int rimu_engine_kernel_times(rimu_engine* e, char* names, int32_t name_len, double* ms,
                             int64_t* launches, int32_t cap, int32_t* n) {
    if (!e || !n) return fail(RIMU_ERR_ARG, "null argument");
    int i = 0;
    for (auto& kv : e->timer.totals) {
        if (i >= cap) break;
        if (names && name_len > 0) {
            std::strncpy(names + (size_t)i * name_len, kv.first.c_str(), name_len - 1);
            names[(size_t)i * name_len + name_len - 1] = 0;
        }
        if (ms) ms[i] = kv.second.ms;
        if (launches) launches[i] = kv.second.launches;
        ++i;
    }
    *n = i;
    return RIMU_OK;
}

int rimu_engine_last_traffic(rimu_engine* e, double* alg_bytes, double* records) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (alg_bytes) *alg_bytes = e->last_alg_bytes;
    if (records) *records = e->last_records;
    return RIMU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Stand-alone annihilation and Hamiltonian evaluation
// ---------------------------------------------------------------------------------------
namespace {

template <class K>
__global__ __launch_bounds__(NT) void ham_eval_kernel(ModelDev h, const K* __restrict__ addrs,
                                                      const int64_t* __restrict__ chosen, int64_t n,
                                                      double* __restrict__ diag, int64_t* __restrict__ nod,
                                                      K* __restrict__ naddr,
                                                      double* __restrict__ oval) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    const K k = addrs[i];
    const bool rs = h.kind == RIMU_MODEL_HUBBARD_REAL_SPACE;
    diag[i] = rs ? model_diagonal_element<true>(h, k) : model_diagonal_element<false>(h, k);
    const int64_t c = rs ? model_num_offdiagonals<true>(h, k) : model_num_offdiagonals<false>(h, k);
    nod[i] = c;
    const int64_t ch = chosen[i];
    if (ch >= 1 && ch <= c) {
        double v;
        naddr[i] = rs ? model_offdiagonal<true>(h, k, ch, &v) : model_offdiagonal<false>(h, k, ch, &v);
        oval[i] = v;
    } else {
        naddr[i] = k;
        oval[i] = 0.0;
    }
}

template <class V>
int annihilate_impl(rimu_engine* e, const uint64_t* keys, const void* vals, int64_t n, int nbits,
                    uint64_t varying, uint64_t* out_keys, void* out_vals, int64_t cap, int64_t* n_out) {
    rimu_engine::Misc* dm = (rimu_engine::Misc*)e->misc.p;
    HIP_TRY(hipMemcpyAsync(e->skeys[0].p, keys, n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->svals[0].p, vals, n * 8, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(dm, 0, sizeof(rimu_engine::Misc), e->stream));
    int sorted = 0;
    if (int rc = radix_sort<uint64_t, V>(e, n, nbits, varying, 0, &sorted)) return rc;
    StepArgs a = make_args(e, 0, 0.0, 1.0, 1.0);
    a.compression = 0;
    int64_t tiles = 0;
    if (int rc = merge_into<uint64_t, V>(e, a, e->skeys[sorted].p, (const V*)e->svals[sorted].p, n, e->vkeys[0].p,
                               (const V*)e->dvals.p, 0, e->vkeys[1].p, (V*)e->vvals[1].p, e->max_entries,
                               dm, &tiles)) return rc;
    HIP_TRY(hipMemcpyAsync(e->h_misc, dm, sizeof(rimu_engine::Misc), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const int64_t m = e->h_misc->L_new;
    *n_out = m;
    if (e->h_misc->flags & FLAG_LOOKBACK) return fail(RIMU_ERR_INTERNAL, "radix look-back did not complete");
    if (e->h_misc->flags & FLAG_VECTOR_OVERFLOW) return fail(RIMU_ERR_CAPACITY, "annihilation buffer too small");
    if (m > cap) return fail(RIMU_ERR_CAPACITY, "output capacity too small");
    if (m > 0) {
        HIP_TRY(hipMemcpyAsync(out_keys, e->vkeys[1].p, m * 8, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(out_vals, e->vvals[1].p, m * 8, hipMemcpyDeviceToHost, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RIMU_OK;
}

}  // namespace

extern "C" {

int rimu_annihilate(int32_t vtype, const uint64_t* keys, const void* vals, int64_t n,
                    uint64_t* out_keys, void* out_vals, int64_t cap, int64_t* n_out) {
    if (!n_out) return fail(RIMU_ERR_ARG, "n_out is null");
    *n_out = 0;
    if (vtype != 0 && vtype != 1) return fail(RIMU_ERR_ARG, "vtype must be 0 (int64) or 1 (double)");
    if (n < 0 || (n > 0 && (!keys || !vals))) return fail(RIMU_ERR_ARG, "bad list arguments");
    if (n == 0) return RIMU_OK;
    if (n >= (int64_t)UINT32_MAX) return fail(RIMU_ERR_ARG, "list longer than 2^32 - 2");
    uint64_t kor = 0, kand = ~0ull;
    for (int64_t i = 0; i < n; ++i) { kor |= keys[i]; kand &= keys[i]; }
    const int nbits = kor ? 64 - __builtin_clzll(kor) : 1;
    rimu_model m{};
    m.kind = RIMU_MODEL_HUBBARD_REAL_1D;
    m.M = 2;
    m.N = 1;
    rimu_style s{};
    s.kind = vtype == 0 ? RIMU_STYLE_STOCHASTIC_INTEGER : RIMU_STYLE_DETERMINISTIC;
    rimu_engine* e = nullptr;
    if (int rc = rimu_engine_create(&m, &s, 0, n, n, &e)) return rc;
    const int rc = vtype == 0
        ? annihilate_impl<int64_t>(e, keys, vals, n, nbits, kor ^ kand, out_keys, out_vals, cap, n_out)
        : annihilate_impl<double>(e, keys, vals, n, nbits, kor ^ kand, out_keys, out_vals, cap, n_out);
    std::string err = g_last_error;
    rimu_engine_destroy(e);
    g_last_error = err;
    return rc;
}

int rimu_hamiltonian_eval(const rimu_model* model, const uint64_t* addrs, const int64_t* chosen,
                          int64_t n, double* diag, int64_t* num_offdiag, uint64_t* new_addr,
                          double* offdiag_value) {
    if (int rc = check_model(model)) return rc;
    if (n < 0) return fail(RIMU_ERR_ARG, "negative count");
    if (n == 0) return RIMU_OK;
    if (!addrs || !chosen || !diag || !num_offdiag || !new_addr || !offdiag_value)
        return fail(RIMU_ERR_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RIMU_ERR_NODEVICE, "no HIP device");
    const ModelDev h = to_dev(*model);
    const int kw = h.B > 64 ? 2 : 1;
    for (int64_t i = 0; i < n; ++i) {
        const u128 key = kw == 1 ? (u128)addrs[i] : (((u128)addrs[2 * i + 1] << 64) | addrs[2 * i]);
        if (!address_ok(h, key))
            return fail(RIMU_ERR_ARG, "address is not an address of the model (bits or particle numbers)");
    }
    DevBuf<uint64_t> a, na;
    DevBuf<int64_t> c, nod;
    DevBuf<double> d, ov;
    auto cleanup = [&]() { a.release(); na.release(); c.release(); nod.release(); d.release(); ov.release(); };
    if (a.alloc(n * kw) || na.alloc(n * kw) || c.alloc(n) || nod.alloc(n) || d.alloc(n) || ov.alloc(n)) {
        cleanup();
        return fail(RIMU_ERR_HIP, "allocation failed");
    }
    hipError_t err = hipMemcpy(a.p, addrs, n * 8 * kw, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(c.p, chosen, n * 8, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        if (kw == 2)
            ham_eval_kernel<u128><<<ceil_div(n, NT), NT>>>(h, (const u128*)a.p, c.p, n, d.p, nod.p,
                                                          (u128*)na.p, ov.p);
        else
            ham_eval_kernel<uint64_t><<<ceil_div(n, NT), NT>>>(h, a.p, c.p, n, d.p, nod.p, na.p, ov.p);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpy(diag, d.p, n * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(num_offdiag, nod.p, n * 8, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(new_addr, na.p, n * 8 * kw, hipMemcpyDeviceToHost);
    if (err == hipSuccess) err = hipMemcpy(offdiag_value, ov.p, n * 8, hipMemcpyDeviceToHost);
    cleanup();
    if (err != hipSuccess) return fail(RIMU_ERR_HIP, std::string("hamiltonian_eval: ") + hipGetErrorString(err));
    return RIMU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Annihilation without device allocation (rimu_annihilate_noalloc): the sort and merge of
// rimu_annihilate on buffers carved from one file-scope device array.  Layout, in records
// of NOALLOC_MAX (8-byte keys, 8-byte values) and the sort/merge bookkeeping for that many:
//   keys[2], vals[2]   ping-pong buffers of the sort; after the sort the free pair holds the
//                      merge's per-tile slots and the sorted pair receives the result
//   status             look-back status words, passes x tiles x RADIX (cleared every call)
//   part, hist         digit counts and bases of the histogram pass
//   merge bookkeeping  partial sums, splits, preceding keys, tile counts, scan sums, flags
// ---------------------------------------------------------------------------------------
namespace {
constexpr int64_t NOALLOC_MAX = 1 << 20;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr int NOALLOC_PASSES = 8;  // 8-bit digits of a 64-bit key
constexpr int64_t NOALLOC_SORT_TILES = (NOALLOC_MAX + 4095) / 4096;
constexpr int64_t NOALLOC_MERGE_TILES = (NOALLOC_MAX + NT * MERGE_IPT - 1) / (NT * MERGE_IPT);
constexpr size_t NA_REC = align256(NOALLOC_MAX * 8);
constexpr size_t NA_KEYS0 = 0, NA_KEYS1 = NA_REC, NA_VALS0 = 2 * NA_REC, NA_VALS1 = 3 * NA_REC;
constexpr size_t NA_STATUS = 4 * NA_REC;
constexpr size_t NA_STATUS_BYTES = align256((size_t)NOALLOC_PASSES * NOALLOC_SORT_TILES * RADIX * 8);
constexpr size_t NA_PART = NA_STATUS + NA_STATUS_BYTES;
constexpr size_t NA_PART_BYTES = align256((size_t)HIST_ROWS * NOALLOC_PASSES * RADIX * 4);
constexpr size_t NA_HIST = NA_PART + NA_PART_BYTES;
constexpr size_t NA_HIST_BYTES = align256((size_t)(NOALLOC_PASSES * RADIX + MAX_PASSES) * 4);
constexpr size_t NA_MERGE_P = NA_HIST + NA_HIST_BYTES;
constexpr size_t NA_MERGE_P_BYTES = align256((size_t)NOALLOC_MERGE_TILES * 3 * 8);
constexpr size_t NA_SPLITS = NA_MERGE_P + NA_MERGE_P_BYTES;
constexpr size_t NA_TILE_BYTES = align256((size_t)(NOALLOC_MERGE_TILES + 1) * 8);
constexpr size_t NA_PREVK = NA_SPLITS + NA_TILE_BYTES;
constexpr size_t NA_TILE_CNT = NA_PREVK + NA_TILE_BYTES;
constexpr size_t NA_SCAN = NA_TILE_CNT + NA_TILE_BYTES;
constexpr size_t NA_SCAN_BYTES = align256((size_t)(NOALLOC_MERGE_TILES / SCAN_TILE + 2) * 8);
constexpr size_t NA_MISC = NA_SCAN + NA_SCAN_BYTES;
constexpr size_t NA_BYTES = NA_MISC + align256(sizeof(rimu_engine::Misc));
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.  Host
// code uses only this address: the array's name in host code is not a device address.
int scratch_address(unsigned char** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(RIMU_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return fail(RIMU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
        g_scratch[dev] = (unsigned char*)p;
    }
    *out = g_scratch[dev];
    return RIMU_OK;
}

template <class V>
int annihilate_noalloc_impl(unsigned char* base, const uint64_t* keys, const void* vals, int64_t n, int nbits,
                            uint64_t varying, uint64_t* out_keys, void* out_vals, int64_t cap, int64_t* n_out) {
    rimu_engine e;  // views of the scratch array only; its stream is the null stream
    e.stream = nullptr;
    e.max_entries = NOALLOC_MAX;
    e.max_spawns = NOALLOC_MAX;
    e.skeys[0].borrow((uint64_t*)(base + NA_KEYS0), NOALLOC_MAX);
    e.skeys[1].borrow((uint64_t*)(base + NA_KEYS1), NOALLOC_MAX);
    e.svals[0].borrow(base + NA_VALS0, NOALLOC_MAX * 8);
    e.svals[1].borrow(base + NA_VALS1, NOALLOC_MAX * 8);
    e.lb_status.borrow((uint64_t*)(base + NA_STATUS), NA_STATUS_BYTES / 8);
    e.lb_part.borrow((uint32_t*)(base + NA_PART), NA_PART_BYTES / 4);
    e.lb_hist.borrow((uint32_t*)(base + NA_HIST), NA_HIST_BYTES / 4);
    e.merge_p.borrow((double*)(base + NA_MERGE_P), NA_MERGE_P_BYTES / 8);
    e.splits.borrow((int64_t*)(base + NA_SPLITS), NA_TILE_BYTES / 8);
    e.prevk.borrow((uint64_t*)(base + NA_PREVK), NA_TILE_BYTES / 8);
    e.tile_cnt.borrow((int64_t*)(base + NA_TILE_CNT), NA_TILE_BYTES / 8);
    e.scan_tmp.borrow((int64_t*)(base + NA_SCAN), NA_SCAN_BYTES / 8);
    e.misc.borrow(base + NA_MISC, sizeof(rimu_engine::Misc));
    rimu_engine::Misc host_misc{};
    e.h_misc = &host_misc;
    rimu_engine::Misc* dm = (rimu_engine::Misc*)e.misc.p;

    HIP_TRY(hipMemcpy(e.skeys[0].p, keys, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e.svals[0].p, vals, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(dm, 0, sizeof(rimu_engine::Misc), e.stream));
    // status words of an earlier call carry the same epoch as this call's: clear them
    HIP_TRY(hipMemsetAsync(e.lb_status.p, 0, NA_STATUS_BYTES, e.stream));
    e.lb_epoch = 0;
    int sorted = 0;
    if (int rc = radix_sort<uint64_t, V>(&e, n, nbits, varying, 0, &sorted)) return rc;
    // the free pair of sort buffers holds the merge's tile slots; the result replaces the
    // sorted records once the merge has read them (compaction runs after the merge)
    e.slot_k.borrow(e.skeys[1 - sorted].p, NOALLOC_MAX);
    e.slot_v.borrow((uint64_t*)e.svals[1 - sorted].p, NOALLOC_MAX);
    StepArgs a = make_args(&e, 0, 0.0, 1.0, 1.0);
    a.compression = 0;
    int64_t tiles = 0;
    uint64_t* ok = e.skeys[sorted].p;
    V* ov = (V*)e.svals[sorted].p;
    if (int rc = merge_into<uint64_t, V>(&e, a, e.skeys[sorted].p, (const V*)e.svals[sorted].p, n, ok, ov, 0, ok,
                                         ov, NOALLOC_MAX, dm, &tiles)) return rc;
    HIP_TRY(hipMemcpyAsync(e.h_misc, dm, sizeof(rimu_engine::Misc), hipMemcpyDeviceToHost, e.stream));
    HIP_TRY(hipStreamSynchronize(e.stream));
    const int64_t m = host_misc.L_new;
    *n_out = m;
    if (host_misc.flags & FLAG_LOOKBACK) return fail(RIMU_ERR_INTERNAL, "radix look-back did not complete");
    if (host_misc.flags & FLAG_VECTOR_OVERFLOW) return fail(RIMU_ERR_INTERNAL, "annihilation scratch overflow");
    if (m > cap) return fail(RIMU_ERR_CAPACITY, "output capacity too small");
    if (m > 0) {
        HIP_TRY(hipMemcpy(out_keys, ok, m * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(out_vals, ov, m * 8, hipMemcpyDeviceToHost));
    }
    return RIMU_OK;
}
}  // namespace

extern "C" int rimu_annihilate_noalloc(int32_t vtype, const uint64_t* keys, const void* vals, int64_t n,
                                       uint64_t* out_keys, void* out_vals, int64_t cap, int64_t* n_out) {
    if (!n_out) return fail(RIMU_ERR_ARG, "n_out is null");
    *n_out = 0;
    if (vtype != 0 && vtype != 1) return fail(RIMU_ERR_ARG, "vtype must be 0 (int64) or 1 (double)");
    if (n < 0 || (n > 0 && (!keys || !vals))) return fail(RIMU_ERR_ARG, "bad list arguments");
    if (n > NOALLOC_MAX)
        return fail(RIMU_ERR_CAPACITY, "rimu_annihilate_noalloc takes at most 2^20 records; use rimu_annihilate");
    if (n == 0) return RIMU_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RIMU_ERR_NODEVICE, "no HIP device");
    unsigned char* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    uint64_t kor = 0, kand = ~0ull;
    for (int64_t i = 0; i < n; ++i) { kor |= keys[i]; kand &= keys[i]; }
    const int nbits = kor ? 64 - __builtin_clzll(kor) : 1;
    return vtype == 0
        ? annihilate_noalloc_impl<int64_t>(base, keys, vals, n, nbits, kor ^ kand, out_keys, out_vals, cap, n_out)
        : annihilate_noalloc_impl<double>(base, keys, vals, n, nbits, kor ^ kand, out_keys, out_vals, cap, n_out);
}

extern "C" int rimu_engine_reserve(rimu_engine* e, int64_t max_entries, int64_t max_spawns) {
    if (!e) return fail(RIMU_ERR_ARG, "engine is null");
    if (max_spawns >= (int64_t)UINT32_MAX) return fail(RIMU_ERR_ARG, "max_spawns must be below 2^32");
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (max_spawns > e->max_spawns) {
        for (int b = 0; b < 2; ++b) {
            HIP_TRY(e->skeys[b].alloc(max_spawns * e->kw));
            HIP_TRY(e->svals[b].alloc(max_spawns * 8));
        }
        e->max_spawns = max_spawns;
    }
    if (max_entries > e->max_entries) {
        for (int b = 0; b < 2; ++b) {
            DevBuf<uint64_t> nk;
            DevBuf<unsigned char> nv;
            HIP_TRY(nk.alloc(max_entries * e->kw));
            HIP_TRY(nv.alloc(max_entries * 8));
            if (b == e->cur && e->len > 0) {
                HIP_TRY(hipMemcpyAsync(nk.p, e->vkeys[b].p, e->len * 8 * e->kw, hipMemcpyDeviceToDevice, e->stream));
                HIP_TRY(hipMemcpyAsync(nv.p, e->vvals[b].p, e->len * 8, hipMemcpyDeviceToDevice, e->stream));
                HIP_TRY(hipStreamSynchronize(e->stream));
            }
            e->vkeys[b] = std::move(nk);
            e->vvals[b] = std::move(nv);
        }
        HIP_TRY(e->dvals.alloc(max_entries * 8));
        HIP_TRY(e->woff.alloc(max_entries + 1));
        e->max_entries = max_entries;
    }
    return RIMU_OK;
}
