// rimu_oracle.cpp -- CPU restatement of Rimu.jl's FCIQMC step path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU engine is checked
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product library (librimu_hip.so) never links or calls it.
//
// Each function cites the reference file:line it restates (paths relative to the
// Rimu.jl v0.14.0 checkout).  Where the reference draws from Julia's task-local
// Xoshiro stream (spawning.jl:20,40; compression.jl:22; Interfaces/hamiltonians.jl:363)
// this restatement draws from counter-based Philox4x32-10 streams instead, keyed by the
// run seed and addressed by (source address, attempt, step, purpose).  The GPU engine
// uses the same streams, so a single step can be compared record for record.
//
// Addresses are BoseFS on one BitString chunk (B = N + M - 1 <= 64 bits) or two (B <= 128,
// held as one 128-bit integer whose low 64 bits are the last chunk), layout of
// bitstring.jl:464-508: mode 1 occupies the least significant bits, each mode is written as
// n_i one-bits followed by a zero separator.  HubbardRealSpace also takes FermiFS (site j
// at bit j-1) and two-component keys (component 1 above component 2).  Across the C
// interface a two-chunk address is two uint64 words, the low word first.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11).  Round function and constants as published there.
// ---------------------------------------------------------------------------------
struct P4 { uint32_t v[4]; };

inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

inline P4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                        uint32_t k0, uint32_t k1) {
    uint32_t c[4] = {c0, c1, c2, c3};
    uint32_t k[2] = {k0, k1};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    P4 out;
    std::memcpy(out.v, c, sizeof(c));
    return out;
}

// Stream purposes, stored in the top two bits of counter word 3.
enum { TAG_SPAWN = 0u, TAG_DIAG = 1u, TAG_COMPRESS = 2u };

inline P4 draw(uint64_t seed, uint64_t addr, uint32_t attempt, int64_t step, uint32_t tag) {
    const uint32_t c3 = ((uint32_t)step & 0x3FFFFFFFu) | (tag << 30);
    return philox4x32_10((uint32_t)addr, (uint32_t)(addr >> 32), attempt, c3,
                         (uint32_t)seed, (uint32_t)(seed >> 32));
}

typedef unsigned __int128 u128;

// Two-chunk addresses: the high chunk selects the key (seed ^ hi * 0x9E3779B97F4A7C15, a
// bijection of hi for a given seed), the low chunk the counter; hi = 0 gives the one-chunk
// streams.
inline P4 draw(uint64_t seed, u128 addr, uint32_t attempt, int64_t step, uint32_t tag) {
    const uint64_t hi = (uint64_t)(addr >> 64);
    return draw(seed ^ (hi * 0x9E3779B97F4A7C15ull), (uint64_t)addr, attempt, step, tag);
}

// 53-bit uniform double in [0, 1) from two 32-bit words.
inline double u53(uint32_t hi, uint32_t lo) {
    const uint64_t x = ((((uint64_t)hi) << 32) | lo) >> 11;
    return (double)x * 0x1.0p-53;
}

// Uniform index in 1:n from one 32-bit word (multiply-shift; bias <= n / 2^32).
inline int64_t pick1(uint32_t r, int64_t n) {
    return 1 + (int64_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

// ---------------------------------------------------------------------------------
// BoseFS on one chunk.
// ---------------------------------------------------------------------------------
struct BoseIndex { int occnum, mode, offset; };  // BoseFSIndex, bosefs.jl / fockaddress.jl

template <class A> constexpr int WIDTH = 8 * (int)sizeof(A);
inline int tzA(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }
inline int tzA(u128 x) {
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    return lo ? __builtin_ctzll(lo) : (hi ? 64 + __builtin_ctzll(hi) : 128);
}
template <class A> inline int toA(A x) { return tzA((A)~x); }
inline int tz64(uint64_t x) { return tzA(x); }
inline int to64(uint64_t x) { return toA(x); }

// occupied_modes iterator, bitstring.jl:636-655.
template <class A>
inline int occupied_modes(A chunk, BoseIndex* out) {
    constexpr int W = WIDTH<A>;
    int n = 0;
    int empty = tzA(chunk);
    if (empty >= W) return 0;
    chunk >>= empty;
    int bit = empty, mode = 1 + empty;
    while (chunk) {
        int bosons = toA(chunk);
        chunk = bosons >= W ? (A)0 : (A)(chunk >> bosons);
        int e = chunk ? tzA(chunk) : 0;
        chunk = e >= W ? (A)0 : (A)(chunk >> e);
        out[n++] = BoseIndex{bosons, mode, bit};
        bit += bosons + e;
        mode += e;
    }
    return n;
}

// from_bose_onr, bitstring.jl:464-508 (the chunks read as one integer).
template <class A>
inline A from_onr_t(const int* onr, int M) {
    constexpr int W = WIDTH<A>;
    A r = 0;
    for (int i = M - 1; i >= 0; --i) {
        const int n = onr[i];
        r = (n + 1 >= W) ? (A)0 : (A)(r << (n + 1));
        r |= (n >= W ? (A)~(A)0 : (A)(((A)1 << n) - 1));
    }
    return r;
}
inline uint64_t from_onr(const int* onr, int M) { return from_onr_t<uint64_t>(onr, M); }

// to_bose_onr, bitstring.jl:511-520.
template <class A>
inline void to_onr(A bs, int M, int* onr) {
    constexpr int W = WIDTH<A>;
    for (int i = 0; i < M; ++i) onr[i] = 0;
    for (int mode = 0; mode < M; ++mode) {
        int b = toA(bs);
        onr[mode] = b;
        bs = (b + 1 >= W) ? (A)0 : (A)(bs >> (b + 1));
        if (!bs) break;
    }
}

// find_mode, bosefs.jl:188-203.
inline BoseIndex find_mode(const BoseIndex* occ, int nocc, int index) {
    int last_occnum = 0, last_mode = 0, last_offset = 0;
    for (int i = 0; i < nocc; ++i) {
        const int dist = index - occ[i].mode;
        if (dist == 0) return BoseIndex{occ[i].occnum, index, occ[i].offset};
        if (dist < 0) return BoseIndex{0, index, occ[i].offset + dist};
        last_occnum = occ[i].occnum; last_mode = occ[i].mode; last_offset = occ[i].offset;
    }
    return BoseIndex{0, index, last_offset + last_occnum + index - last_mode};
}

// bose_excitation_value, fockaddress.jl:544-567: destructions are applied first (in the
// order given), then creations; each operator sees the occupation numbers left by the
// operators before it.  Arguments are already reversed as in bitstring.jl:586-587.
inline int64_t excitation_value(const BoseIndex* cre, int nc, const BoseIndex* des, int nd) {
    int cocc[4], docc[4];
    for (int i = 0; i < nc; ++i) cocc[i] = cre[i].occnum;
    for (int i = 0; i < nd; ++i) docc[i] = des[i].occnum;
    int64_t value = 1;
    for (int a = 0; a < nd; ++a) {
        value *= docc[a];
        for (int i = 0; i < nc; ++i) cocc[i] -= (des[a].mode == cre[i].mode);
        for (int b = a + 1; b < nd; ++b) docc[b] -= (des[a].mode == des[b].mode);
    }
    for (int a = 0; a < nc; ++a) {
        value *= (cocc[a] + 1);
        for (int b = a + 1; b < nc; ++b) cocc[b] += (cre[a].mode == cre[b].mode);
    }
    return value;
}

// excitation(b, creations, destructions), bosefs.jl:252-255 + bitstring.jl:581-595.
// The particle moves are restated on the occupation-number representation: the resulting
// address is the one with the destroyed modes decremented and created modes incremented.
template <class A>
inline A excitation(A bs, int M, const BoseIndex* cre, const BoseIndex* des, int n, double* val) {
    BoseIndex cre_rev[4], des_rev[4];
    for (int i = 0; i < n; ++i) { cre_rev[i] = cre[n - 1 - i]; des_rev[i] = des[n - 1 - i]; }
    const int64_t v = excitation_value(cre_rev, n, des_rev, n);
    if (v == 0) { *val = 0.0; return bs; }
    int onr[128];
    to_onr(bs, M, onr);
    for (int i = 0; i < n; ++i) onr[des[i].mode - 1] -= 1;
    for (int i = 0; i < n; ++i) onr[cre[i].mode - 1] += 1;
    *val = std::sqrt((double)v);
    return from_onr_t<A>(onr, M);
}

inline int mod1(int x, int m) { int r = ((x - 1) % m + m) % m; return r + 1; }
inline void fldmod1(int64_t x, int64_t y, int64_t* f, int64_t* m) {
    // Julia fldmod1 for positive x, y
    *m = (x - 1) % y + 1;
    *f = (x - *m) / y + 1;
}

// bose_hubbard_interaction, bosefs.jl:407-413.
inline int64_t bose_hubbard_interaction(const BoseIndex* occ, int nocc) {
    int64_t r = 0;
    for (int i = 0; i < nocc; ++i) r += (int64_t)occ[i].occnum * (occ[i].occnum - 1);
    return r;
}

}  // namespace

extern "C" {

typedef struct {
    int32_t model;  // 1 = HubbardReal1D, 2 = HubbardMom1D, 3 = HubbardRealSpace
    int32_t M, N;   // modes (sites); particles (HubbardRealSpace: of component 1)
    double u, t;    // HubbardRealSpace: u[1,1], t[1]
    double kes[64];  // HubbardMom1D dispersion table, mode j at index j-1; HubbardRealSpace:
                     // trap potential of component 1 per site
    // HubbardRealSpace only
    int32_t ncomp;   // 1 or 2 components
    int32_t N2;      // particles of component 2
    int32_t fermi;   // bit c: component c+1 is a FermiFS
    int32_t ndir;    // 2D neighbour directions
    int32_t has_u;   // u is not the zero matrix
    int32_t has_pot; // trap strengths are nonzero
    double t2, u12, u22;
    double pot2[64];
    uint8_t nbr[64][6];  // neighbor_site(geometry, s+1, d+1) - 1, 255 = none
} or_model;

typedef struct {
    int32_t style;     // 1 IsStochasticInteger, 2 IsDeterministic, 3 IsDynamicSemistochastic
    int32_t compression;  // 0 NoCompression, 1 ThresholdCompression
    double compression_threshold;
    double proj_threshold;  // projection threshold of diagonal/spawn deposits
    double rel_threshold;   // DynamicSemistochastic rel_threshold
    double abs_threshold;   // DynamicSemistochastic abs_threshold
} or_style;

void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    P4 r = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    std::memcpy(out, r.v, sizeof(r.v));
}

uint64_t or_from_onr(const int32_t* onr, int32_t M) {
    int tmp[64];
    for (int i = 0; i < M; ++i) tmp[i] = onr[i];
    return from_onr(tmp, M);
}

void or_to_onr(uint64_t bs, int32_t M, int32_t* onr) {
    int tmp[64];
    to_onr(bs, M, tmp);
    for (int i = 0; i < M; ++i) onr[i] = tmp[i];
}

int32_t or_occupied_modes(uint64_t bs, int32_t* occnum, int32_t* mode, int32_t* offset) {
    BoseIndex occ[64];
    const int n = occupied_modes(bs, occ);
    for (int i = 0; i < n; ++i) { occnum[i] = occ[i].occnum; mode[i] = occ[i].mode; offset[i] = occ[i].offset; }
    return n;
}

int64_t or_bose_hubbard_interaction(uint64_t bs) {
    BoseIndex occ[64];
    const int n = occupied_modes(bs, occ);
    return bose_hubbard_interaction(occ, n);
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------------
// HubbardRealSpace (Hamiltonians/HubbardRealSpace.jl) on a BoseFS, a FermiFS or a
// CompositeFS of two of them.  A two-component key holds component 1 above component 2
// (CompositeFS isless compares components in order, multicomponent.jl:79); the geometry
// arrives as the neighbour table nbr (neighbor_site, geometry.jl:236-240).
// ---------------------------------------------------------------------------------
inline bool rs_fermi(const or_model* h, int c) { return (h->fermi >> c) & 1; }
inline int rs_width(const or_model* h, int c) {
    return rs_fermi(h, c) ? h->M : (c == 0 ? h->N : h->N2) + h->M - 1;
}
inline int model_bits(const or_model* h) {
    if (h->model != 3) return h->M + h->N - 1;
    return rs_width(h, 0) + (h->ncomp == 2 ? rs_width(h, 1) : 0);
}
template <class A> inline A low_bits(int n) { return n >= WIDTH<A> ? (A)~(A)0 : (A)(((A)1 << n) - 1); }
template <class A> inline A rs_component(const or_model* h, A key, int c) {
    if (h->ncomp == 1) return key;
    const int w2 = rs_width(h, 1);
    return c == 0 ? (A)(key >> w2) : (A)(key & low_bits<A>(w2));
}
// update_component, multicomponent.jl:70-77
template <class A> inline A rs_update_component(const or_model* h, A key, int c, A v) {
    if (h->ncomp == 1) return v;
    const int w2 = rs_width(h, 1);
    return c == 0 ? (A)((v << w2) | (key & low_bits<A>(w2))) : (A)((key & ~low_bits<A>(w2)) | v);
}

// Occupied modes of one component as BoseFSIndex-like records (occnum, mode); a FermiFS
// lists its set bits in ascending order with occnum 1 (bitstring.jl:794-831).
template <class A> inline int rs_occupied(const or_model* h, A b, int c, BoseIndex* out) {
    if (!rs_fermi(h, c)) return occupied_modes(b, out);
    int n = 0;
    for (int i = 0; i < WIDTH<A>; ++i)
        if ((b >> i) & 1) out[n++] = BoseIndex{1, i + 1, i};
    return n;
}

// dot(occupied_modes(a), occupied_modes(b)), fockaddress.jl:692-719; for two FermiFS
// bitstrings count_ones(a & b), bitstring.jl:833-838.
template <class A> inline int64_t rs_dot(const or_model* h, A a, A b) {
    if (rs_fermi(h, 0) && rs_fermi(h, 1)) {
        int64_t n = 0;
        for (int i = 0; i < WIDTH<A>; ++i) n += ((a & b) >> i) & 1;
        return n;
    }
    BoseIndex oa[128], ob[128];
    const int na = rs_occupied(h, a, 0, oa), nb = rs_occupied(h, b, 1, ob);
    int64_t acc = 0;
    int i = 0, j = 0;
    while (i < na && j < nb) {
        if (oa[i].mode > ob[j].mode) ++j;
        else if (oa[i].mode < ob[j].mode) ++i;
        else { acc += (int64_t)oa[i].occnum * ob[j].occnum; ++i; ++j; }
    }
    return acc;
}

// local_interaction(b, u) = u * bose_hubbard_interaction(b) / 2 for a BoseFS, 0 for a
// FermiFS (HubbardRealSpace.jl:18-19).
template <class A> inline double rs_self(const or_model* h, A b, int c, double u) {
    if (rs_fermi(h, c)) return 0.0;
    BoseIndex occ[128];
    const int n = occupied_modes(b, occ);
    return u * (double)bose_hubbard_interaction(occ, n) / 2;
}

// external_potential, HubbardRealSpace.jl:98-106: sum over occupied modes of n * pot[i].
template <class A> inline double rs_pot(const or_model* h, A b, int c, const double* pot) {
    BoseIndex occ[128];
    const int n = rs_occupied(h, b, c, occ);
    double pe = 0.0;
    for (int i = 0; i < n; ++i) pe += occ[i].occnum * pot[occ[i].mode - 1];
    return pe;
}

// diagonal_element, HubbardRealSpace.jl:264-279, with _interactions (:57-73) unrolled for
// two components: self(1) + cross(1,2), then + self(2).
template <class A> inline double rs_diagonal(const or_model* h, A key) {
    const A a = rs_component(h, key, 0);
    double intr = 0.0;
    if (h->has_u) {
        if (h->ncomp == 1) {
            intr = rs_self(h, a, 0, h->u);
        } else {
            const A b = rs_component(h, key, 1);
            const double self_a = rs_self(h, a, 0, h->u);
            const double row = h->u12 * (double)rs_dot(h, a, b);
            intr = self_a + row;
            intr = intr + rs_self(h, b, 1, h->u22);
        }
    }
    double pot = 0.0;
    if (h->has_pot) {
        if (h->ncomp == 1) {
            pot = rs_pot(h, a, 0, h->kes);
        } else {
            pot += rs_pot(h, a, 0, h->kes);
            pot += rs_pot(h, rs_component(h, key, 1), 1, h->pot2);
        }
    }
    return intr + pot;
}

template <class A> inline int64_t rs_num_offdiagonals(const or_model* h, A key) {
    BoseIndex occ[128];
    int64_t n = rs_occupied(h, rs_component(h, key, 0), 0, occ);
    if (h->ncomp == 2) n += rs_occupied(h, rs_component(h, key, 1), 1, occ);
    return n * h->ndir;
}

// fermi_excitation for one hop, bitstring.jl:741-791: the destruction, then the creation,
// each counting the ones below its bit; (unchanged, 0.0) when a flip finds the wrong bit.
template <class A> inline A fermi_hop(A bs, int dst_mode, int src_mode, double* val) {
    const A orig = bs;
    int count = 0;
    const int d = src_mode - 1, c = dst_mode - 1;
    for (int i = 0; i < d; ++i) count += (bs >> i) & 1;
    bs ^= (A)1 << d;
    if ((bs >> d) & 1) { *val = 0.0; return orig; }
    for (int i = 0; i < c; ++i) count += (bs >> i) & 1;
    bs ^= (A)1 << c;
    if (!((bs >> c) & 1)) { *val = 0.0; return orig; }
    *val = (count % 2 == 0) ? 1.0 : -1.0;
    return bs;
}

// HubbardRealSpaceCompOffdiagonals getindex (HubbardRealSpace.jl:303-316) inside the
// component walk of HubbardRealSpaceOffdiagonals (:380-391).
template <class A> inline A rs_offdiagonal(const or_model* h, A key, int64_t chosen, double* value) {
    BoseIndex occ[128];
    int c = 0;
    A b = rs_component(h, key, 0);
    int nocc = rs_occupied(h, b, 0, occ);
    if (h->ncomp == 2 && chosen > (int64_t)nocc * h->ndir) {
        chosen -= (int64_t)nocc * h->ndir;
        c = 1;
        b = rs_component(h, key, 1);
        nocc = rs_occupied(h, b, 1, occ);
    }
    int64_t particle, neigh;
    fldmod1(chosen, h->ndir, &particle, &neigh);
    const BoseIndex src = occ[particle - 1];           // find_occupied_mode
    const int nsite = h->nbr[src.mode - 1][neigh - 1];
    if (nsite == 255) { *value = 0.0; return key; }     // neighbor_site == 0
    const double t = c == 0 ? h->t : h->t2;
    double v;
    A nb;
    if (rs_fermi(h, c)) {
        nb = fermi_hop(b, nsite + 1, src.mode, &v);
    } else {
        BoseIndex all[128];
        const int nall = occupied_modes(b, all);
        const BoseIndex dst = find_mode(all, nall, nsite + 1);
        nb = excitation(b, h->M, &dst, &src, 1, &v);
    }
    *value = -t * v;
    return rs_update_component(h, key, c, nb);
}

// num_offdiagonals: HubbardReal1D.jl:51-53; HubbardMom1D.jl:131-140.
template <class A>
int64_t num_offdiagonals_t(const or_model* h, A bs) {
    if (h->model == 3) return rs_num_offdiagonals(h, bs);
    BoseIndex occ[128];
    const int s = occupied_modes(bs, occ);
    if (h->model == 1) return 2 * (int64_t)s;
    int d = 0;
    for (int i = 0; i < s; ++i) d += occ[i].occnum > 1;
    const int64_t M = h->M;
    return (int64_t)s * (s - 1) * (M - 2) + (int64_t)d * (M - 1);
}

// diagonal_element: HubbardReal1D.jl:55-57; HubbardMom1D.jl:174-177 with
// momentum_transfer_diagonal excitations.jl:126-137 and dot(kes, map) fockaddress.jl:316-321.
template <class A>
double diagonal_element_t(const or_model* h, A bs) {
    if (h->model == 3) return rs_diagonal(h, bs);
    BoseIndex occ[128];
    const int s = occupied_modes(bs, occ);
    if (h->model == 1) {
        return h->u * (double)bose_hubbard_interaction(occ, s) / 2;
    }
    double kin = 0.0;
    for (int i = 0; i < s; ++i) kin += h->kes[occ[i].mode - 1] * occ[i].occnum;
    int64_t onproduct = 0;
    for (int i = 0; i < s; ++i) {
        const int64_t oi = occ[i].occnum;
        onproduct += oi * (oi - 1);
        for (int j = 0; j < i; ++j) onproduct += 4 * oi * occ[j].occnum;
    }
    return kin + h->u / (2.0 * h->M) * (double)onproduct;
}

// hopnextneighbour generic form, bosefs.jl:347-353; value -t*sqrt(...), HubbardReal1D.jl:59-62.
template <class A>
A real1d_offdiagonal(const or_model* h, A bs, const BoseIndex* occ, int nocc, int64_t chosen,
                     double* value) {
    const BoseIndex src = occ[(chosen + 1) / 2 - 1];       // find_occupied_mode
    const int dir = (chosen & 1) ? 1 : -1;
    const BoseIndex dst = find_mode(occ, nocc, mod1(src.mode + dir, h->M));
    double v;
    const A nb = excitation(bs, h->M, &dst, &src, 1, &v);
    *value = -h->t * v;
    return nb;
}

// momentum_transfer_excitation (fold=true), excitations.jl:26-80; value u/(2M)*onproduct,
// HubbardMom1D.jl:189-194.
template <class A>
A mom1d_offdiagonal(const or_model* h, A bs, const BoseIndex* occ, int nocc, int64_t chosen,
                    double* value) {
    const int M = h->M;
    const int64_t singlies = nocc;
    int64_t dbl = chosen - singlies * (singlies - 1) * (M - 2);
    BoseIndex src[2];
    int64_t mom_change;
    if (dbl > 0) {
        int64_t d;
        fldmod1(dbl, M - 1, &d, &mom_change);
        BoseIndex idx = occ[0];
        for (int i = 0; i < nocc; ++i) {
            d -= occ[i].occnum >= 2;
            if (d == 0) { idx = occ[i]; break; }
        }
        src[0] = idx; src[1] = idx;
    } else {
        int64_t pair, fst, snd;
        fldmod1(chosen, M - 2, &pair, &mom_change);
        fldmod1(pair, singlies - 1, &fst, &snd);
        int64_t f_hole, s_hole;
        if (snd < fst) { f_hole = snd; s_hole = fst; }
        else { f_hole = fst; s_hole = snd + 1; }
        src[0] = occ[f_hole - 1]; src[1] = occ[s_hole - 1];
        if (mom_change >= src[1].mode - src[0].mode) mom_change += 1;
    }
    const int d0 = mod1((int)(src[0].mode + mom_change), M);
    const int d1 = mod1((int)(src[1].mode - mom_change), M);
    BoseIndex dst[2] = {find_mode(occ, nocc, d0), find_mode(occ, nocc, d1)};
    double v;
    const A nb = excitation(bs, M, dst, src, 2, &v);
    *value = h->u / (2 * M) * v;
    return nb;
}

template <class A>
A get_offdiagonal_t(const or_model* h, A bs, int64_t chosen, double* value) {
    if (h->model == 3) return rs_offdiagonal(h, bs, chosen, value);
    BoseIndex occ[128];
    const int nocc = occupied_modes(bs, occ);
    if (h->model == 1) return real1d_offdiagonal(h, bs, occ, nocc, chosen, value);
    return mom1d_offdiagonal(h, bs, occ, nocc, chosen, value);
}

inline u128 join(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

}  // namespace

extern "C" {

int64_t or_num_offdiagonals(const or_model* h, uint64_t bs) { return num_offdiagonals_t(h, bs); }
double or_diagonal_element(const or_model* h, uint64_t bs) { return diagonal_element_t(h, bs); }
uint64_t or_get_offdiagonal(const or_model* h, uint64_t bs, int64_t chosen, double* value) {
    return get_offdiagonal_t(h, bs, chosen, value);
}

// Two-chunk addresses (B = N + M - 1 > 64) as (lo, hi) words.
void or_from_onr2(const int32_t* onr, int32_t M, uint64_t* out) {
    int tmp[128];
    for (int i = 0; i < M; ++i) tmp[i] = onr[i];
    const u128 r = from_onr_t<u128>(tmp, M);
    out[0] = (uint64_t)r;
    out[1] = (uint64_t)(r >> 64);
}
void or_to_onr2(uint64_t lo, uint64_t hi, int32_t M, int32_t* onr) {
    int tmp[128];
    to_onr(join(lo, hi), M, tmp);
    for (int i = 0; i < M; ++i) onr[i] = tmp[i];
}
int64_t or_num_offdiagonals2(const or_model* h, uint64_t lo, uint64_t hi) {
    return num_offdiagonals_t(h, join(lo, hi));
}
double or_diagonal_element2(const or_model* h, uint64_t lo, uint64_t hi) {
    return diagonal_element_t(h, join(lo, hi));
}
void or_get_offdiagonal2(const or_model* h, uint64_t lo, uint64_t hi, int64_t chosen, double* value,
                         uint64_t* out) {
    const u128 r = get_offdiagonal_t(h, join(lo, hi), chosen, value);
    out[0] = (uint64_t)r;
    out[1] = (uint64_t)(r >> 64);
}

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------------
// One column of the transition operator T = 1 + dt(S - H), fciqmc.jl:78-112, applied
// per style (styles.jl:21-25, :101-105, :210-214) into a deposit sink.
// ---------------------------------------------------------------------------------
struct Params {
    const or_model* h;
    const or_style* s;
    uint64_t seed;
    int64_t step;
    double shift, dtau, boost;
};

struct Stats {
    double attempts = 0, exact = 0, inexact = 0, spawns = 0;
    double deaths = 0, clones = 0, zombies = 0;
};

// projected_deposit! for integer values, spawning.jl:33-45.
inline int64_t project_int(double x, double U) {
    const double sg = (x > 0) - (x < 0);
    return (int64_t)sg * (int64_t)std::floor(std::fabs(x) + U);
}
// projected_deposit! for floating values, spawning.jl:13-31.
inline double project_float(double x, double thresh, double U) {
    if (std::fabs(x) < thresh) {
        const double sg = (x > 0) - (x < 0);
        return (U < std::fabs(x) / thresh) ? sg * thresh : 0.0;
    }
    return x;
}

// clones_deaths_zombies, spawning.jl:79-93.
template <class V>
inline void cdz(V res, V val, Stats& st) {
    auto sgn = [](V x) { return (x > 0) - (x < 0); };
    if (res > val) {
        st.clones += std::fabs((double)(res - val));
    } else if (sgn(res) != sgn(val)) {
        st.deaths += std::fabs((double)val);
        st.zombies += std::fabs((double)res);
    } else {
        st.deaths += std::fabs((double)(res - val));
    }
}

// apply_column! of one entry.  emit(addr, value) receives every deposit in the order the
// reference performs them: the diagonal deposit first, then the spawns.
template <class V, class A, class Emit>
void apply_column(const Params& p, A addr, V val, Stats& st, Emit&& emit) {
    BoseIndex occ[128];
    const int nocc = occupied_modes(addr, occ);
    const or_model* h = p.h;
    const int style = p.s->style;

    // diagonal_step!, spawning.jl:73-77; diagonal_element(T, add) fciqmc.jl:93-96.
    {
        const double hd = diagonal_element_t(h, addr);
        const double d = 1 - p.dtau * (hd - p.shift);
        const double x = d * (double)val;
        if constexpr (std::is_integral<V>::value) {
            P4 r = draw(p.seed, addr, 0, p.step, TAG_DIAG);
            const V nv = (V)project_int(x, u53(r.v[0], r.v[1]));
            if (nv != 0) emit(addr, nv);
            cdz<V>(nv, val, st);
        } else {
            double nv = x;
            if (p.s->proj_threshold > 0) {
                P4 r = draw(p.seed, addr, 0, p.step, TAG_DIAG);
                nv = project_float(x, p.s->proj_threshold, u53(r.v[0], r.v[1]));
            }
            if (nv != 0) emit(addr, (V)nv);
            cdz<V>((V)nv, val, st);
        }
    }

    // offdiagonals(T, add), fciqmc.jl:104-112: matrix elements -dt * H_ba.
    const int64_t n_od = num_offdiagonals_t(h, addr);
    auto offdiag = [&](int64_t chosen, double* mel) {
        double hv;
        const A nb = (h->model == 3) ? rs_offdiagonal(h, addr, chosen, &hv)
                     : (h->model == 1) ? real1d_offdiagonal(h, addr, occ, nocc, chosen, &hv)
                                       : mom1d_offdiagonal(h, addr, occ, nocc, chosen, &hv);
        *mel = -hv * p.dtau;
        return nb;
    };
    auto deposit = [&](A nb, double x, double U) {
        if constexpr (std::is_integral<V>::value) {
            const V nv = (V)project_int(x, U);
            if (nv != 0) emit(nb, nv);
            st.spawns += std::fabs((double)nv);
        } else {
            const double nv = (p.s->proj_threshold > 0) ? project_float(x, p.s->proj_threshold, U) : x;
            if (nv != 0) emit(nb, (V)nv);
            st.spawns += std::fabs(nv);
        }
    };
    // spawn!(Exact), spawning.jl:174-182.
    auto exact = [&]() {
        for (int64_t j = 0; j < n_od; ++j) {
            double mel;
            const A nb = offdiag(j + 1, &mel);
            double U = 0;
            if (p.s->proj_threshold > 0) { P4 r = draw(p.seed, addr, (uint32_t)j, p.step, TAG_SPAWN); U = u53(r.v[1], r.v[2]); }
            deposit(nb, (double)val * mel, U);
        }
        st.attempts += (double)n_od;
    };
    // spawn!(WithReplacement), spawning.jl:232-243, random_offdiagonal Interfaces/hamiltonians.jl:361-366.
    auto with_replacement = [&]() {
        const int64_t n_att = std::max<int64_t>((int64_t)std::floor(std::fabs((double)val) * p.boost), 1);
        const double magnitude = (double)val / (double)n_att;
        const double prob = 1.0 / (double)n_od;
        for (int64_t j = 0; j < n_att; ++j) {
            P4 r = draw(p.seed, addr, (uint32_t)j, p.step, TAG_SPAWN);
            double mel;
            const A nb = offdiag(pick1(r.v[0], n_od), &mel);
            deposit(nb, mel * magnitude / prob, u53(r.v[1], r.v[2]));
        }
        st.attempts += (double)n_att;
    };

    if (n_od == 0) return;
    if (style == 1) {
        with_replacement();
    } else if (style == 2) {
        exact();
        st.exact += 1;
    } else {
        // spawn!(DynamicSemistochastic), spawning.jl:364-378.
        const double thresh = std::min(p.s->abs_threshold, (double)n_od);
        const double amount = p.boost * std::fabs((double)val) * p.s->rel_threshold;
        if (amount >= thresh) { exact(); st.exact += 1; }
        else { with_replacement(); st.inexact += 1; }
    }
}

// Hashing and the (lo, hi) word layout of addresses at the C interface.
struct KeyHash {
    size_t operator()(uint64_t x) const { return std::hash<uint64_t>()(x); }
    size_t operator()(u128 x) const {
        return std::hash<uint64_t>()((uint64_t)x ^ ((uint64_t)(x >> 64) * 0x9E3779B97F4A7C15ull));
    }
};
template <class A> inline A load_key(const uint64_t* keys, int64_t i);
template <> inline uint64_t load_key<uint64_t>(const uint64_t* keys, int64_t i) { return keys[i]; }
template <> inline u128 load_key<u128>(const uint64_t* keys, int64_t i) {
    return ((u128)keys[2 * i + 1] << 64) | keys[2 * i];
}
inline void store_key(uint64_t* out, int64_t i, uint64_t k) { out[i] = k; }
inline void store_key(uint64_t* out, int64_t i, u128 k) {
    out[2 * i] = (uint64_t)k;
    out[2 * i + 1] = (uint64_t)(k >> 64);
}
template <class A, class V> using KeyMap = std::unordered_map<A, V, KeyHash>;

// DVec storage: Dict semantics of dvec.jl:126-136 (a value that becomes zero is deleted).
template <class V, class A = uint64_t>
struct DictVec {
    KeyMap<A, V> m;
    void deposit(A k, V v) {  // deposit!, Interfaces/dictvectors.jl:49-51
        auto it = m.find(k);
        if (it == m.end()) {
            if (v != 0) m.emplace(k, V(0) + v);
        } else {
            it->second = it->second + v;
            if (it->second == 0) m.erase(it);
        }
    }
};

template <class V, class Map>
int64_t dump_sorted(const Map& m, uint64_t* keys, V* vals, int64_t cap) {
    using A = typename Map::key_type;
    std::vector<std::pair<A, V>> items(m.begin(), m.end());
    std::sort(items.begin(), items.end(), [](auto& a, auto& b) { return a.first < b.first; });
    const int64_t n = (int64_t)items.size();
    if (n > cap) return -n;
    for (int64_t i = 0; i < n; ++i) { store_key(keys, i, items[i].first); vals[i] = items[i].second; }
    return n;
}

// ThresholdCompression, compression.jl:18-33.
template <class V, class Map>
void compress(const Params& p, Map& m) {
    if (p.s->compression != 1) return;
    if constexpr (!std::is_integral<V>::value) {
        const double th = p.s->compression_threshold;
        for (auto it = m.begin(); it != m.end();) {
            const double prob = std::fabs(it->second) / th;
            if (prob < 1) {
                P4 r = draw(p.seed, it->first, 0, p.step, TAG_COMPRESS);
                const double sg = (it->second > 0) - (it->second < 0);
                it->second = (prob > u53(r.v[0], r.v[1])) ? th * sg : 0.0;
            }
            if (it->second == 0) it = m.erase(it); else ++it;
        }
    }
}

// The GPU engine's list order.  Above 20 address bits the engine stores each address x as
// y = mix(x), a bijection of [0, 2^B) (multiply by an odd constant mod 2^B, then fold the
// high half down with y ^= y >> ceil(B/2), twice), keeps its vector in ascending y and emits
// deposits entry by entry in that order.  order = 1 replays this list; at or below 20 bits
// y = x (ascending addresses).
constexpr int ENGINE_MIX_MIN_BITS = 20;
template <class A> A engine_order_key(int B, A x) {
    if (B <= ENGINE_MIX_MIN_BITS) return x;
    const u128 c1 = ((u128)0x9E3779B97F4A7C15ull << 64) | 0xF39CC0605CEDC835ull;
    const u128 c2 = ((u128)0xD1B54A32D192ED03ull << 64) | 0xABC98388FB8FAC03ull;
    const int s = (B + 1) / 2;
    const A mask = B >= WIDTH<A> ? (A)~(A)0 : (A)(((A)1 << B) - 1);
    A y = x;
    y = (A)(y * (A)c1) & mask;
    y ^= y >> s;
    y = (A)(y * (A)c2) & mask;
    y ^= y >> s;
    return y;
}

// Entry indices in the engine's list order.
template <class A> std::vector<int64_t> engine_entry_order(const Params& p, const uint64_t* keys, int64_t n) {
    std::vector<int64_t> idx((size_t)n);
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    const int B = model_bits(p.h);
    if (B > ENGINE_MIX_MIN_BITS) {
        std::vector<A> y((size_t)n);
        for (int64_t i = 0; i < n; ++i) y[i] = engine_order_key<A>(B, load_key<A>(keys, i));
        std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return y[a] < y[b]; });
    }
    return idx;
}

template <class A, class V>
int64_t step_impl(const Params& p, int order, const uint64_t* keys, const V* vals, int64_t n,
                  uint64_t* out_keys, V* out_vals, int64_t cap, double* stats) {
    DictVec<V, A> w;
    Stats st;
    if (order == 0) {
        // apply_operator!, Interfaces/dictvectors.jl:112-123: one column at a time.
        for (int64_t i = 0; i < n; ++i)
            apply_column<V>(p, load_key<A>(keys, i), vals[i], st, [&](A k, V v) { w.deposit(k, v); });
    } else {
        // Same deposits, reordered into a list: every diagonal deposit first (entry order),
        // then every spawn (entry, attempt order), entries in the engine's list order.  This
        // is the list the GPU engine merges.
        std::vector<std::pair<A, V>> diag, spawns;
        for (int64_t i : engine_entry_order<A>(p, keys, n)) {
            bool first = true;
            const A self = load_key<A>(keys, i);
            apply_column<V>(p, self, vals[i], st, [&](A k, V v) {
                if (first && k == self) diag.emplace_back(k, v); else spawns.emplace_back(k, v);
                first = false;
            });
        }
        for (auto& kv : diag) w.deposit(kv.first, kv.second);
        for (auto& kv : spawns) w.deposit(kv.first, kv.second);
    }
    const double len_before = (double)w.m.size();
    compress<V>(p, w.m);
    double norm = 0;
    for (auto& kv : w.m) norm += std::fabs((double)kv.second);
    stats[0] = st.attempts; stats[1] = st.exact; stats[2] = st.inexact; stats[3] = st.spawns;
    stats[4] = st.deaths; stats[5] = st.clones; stats[6] = st.zombies; stats[7] = len_before;
    stats[8] = (double)w.m.size(); stats[9] = norm;
    return dump_sorted<V>(w.m, out_keys, out_vals, cap);
}

// ---------------------------------------------------------------------------------
// Initiator approximation (DictVectors/initiators.jl).  The working memory holds an
// InitiatorValue (safe, unsafe, initiator) per address (initiators.jl:22-45).  A deposit
// from parent p_add => p_val lands in one component by to_initiator_value
// (initiators.jl:143-157): on the parent's own address (the diagonal deposit) in
// `initiator` when |p_val| > threshold, else in `safe`; on any other address in `safe`
// when |p_val| > threshold, else in `unsafe`.  An entry whose three components are all
// zero is deleted (deposit!, pdworkingmemory.jl:21-31; initiatordvec.jl deposit!).
// move_and_compress! (pdworkingmemory.jl:262-273) maps every entry through the rule's
// from_initiator_value and hands the values to compress!: ThresholdCompression keeps
// nonzero results only (compression.jl:14-26, _set_value! on a Dict deletes zeros) and
// counts every entry in len_before; NoCompression stores every entry, a zero result
// included (stochasticstyles.jl:116-121, plain Dict setindex!).
// ---------------------------------------------------------------------------------
template <class V>
struct IValue {
    V safe = 0, unsafe = 0, init = 0;
};

// rule: 0 NonInitiator, 1 Initiator, 2 SimpleInitiator, 3 CoherentInitiator.
template <class V>
V from_initiator_value(int rule, double thr, const IValue<V>& x) {
    switch (rule) {
        case 1:  // initiators.jl:138-140: safe + initiator + !iszero(initiator) * unsafe
            return (x.safe + x.init) + (x.init != 0 ? x.unsafe : V(0));
        case 2:  // initiators.jl:180-182: safe + initiator
            return x.safe + x.init;
        case 3:  // initiators.jl:208-214
            if (x.init != 0 || std::fabs((double)x.unsafe) > thr) return (x.init + x.safe) + x.unsafe;
            return x.init + x.safe;
        default:  // NonInitiator: a single value, kept in `safe`
            return x.safe;
    }
}

template <class V, class A>
struct InitiatorMemory {
    KeyMap<A, IValue<V>> m;
    int rule;
    double thr;
    // deposit!(w, add, val, p_add => p_val) with to_initiator_value
    void deposit(A add, V val, A p_add, V p_val) {
        IValue<V> d;
        const bool big = std::fabs((double)p_val) > thr;
        if (rule == 0) d.safe = val;
        else if (p_add == add) { if (big) d.init = val; else d.safe = val; }
        else { if (big) d.safe = val; else d.unsafe = val; }
        auto it = m.find(add);
        IValue<V> nv = it == m.end() ? IValue<V>{} : it->second;
        nv.safe = nv.safe + d.safe;
        nv.unsafe = nv.unsafe + d.unsafe;
        nv.init = nv.init + d.init;
        const bool zero = nv.safe == 0 && nv.unsafe == 0 && nv.init == 0;
        if (zero) { if (it != m.end()) m.erase(it); }
        else if (it == m.end()) m.emplace(add, nv);
        else it->second = nv;
    }
};

template <class A, class V>
int64_t step_initiator_impl(const Params& p, int order, int rule, double thr, const uint64_t* keys,
                            const V* vals, int64_t n, uint64_t* out_keys, V* out_vals, int64_t cap,
                            double* stats) {
    InitiatorMemory<V, A> w;
    w.rule = rule;
    w.thr = thr;
    Stats st;
    if (order == 0) {
        for (int64_t i = 0; i < n; ++i) {
            const A self = load_key<A>(keys, i);
            const V pv = vals[i];
            apply_column<V>(p, self, pv, st, [&](A k, V v) { w.deposit(k, v, self, pv); });
        }
    } else {
        // list order: every diagonal deposit first (entry order), then every spawn
        struct Dep { A k; V v; A pk; V pv; };
        std::vector<Dep> diag, spawns;
        for (int64_t i : engine_entry_order<A>(p, keys, n)) {
            const A self = load_key<A>(keys, i);
            const V pv = vals[i];
            apply_column<V>(p, self, pv, st, [&](A k, V v) {
                if (k == self) diag.push_back({k, v, self, pv}); else spawns.push_back({k, v, self, pv});
            });
        }
        for (auto& d : diag) w.deposit(d.k, d.v, d.pk, d.pv);
        for (auto& d : spawns) w.deposit(d.k, d.v, d.pk, d.pv);
    }
    // move_and_compress!
    const double len_before = (double)w.m.size();
    KeyMap<A, V> out;
    const bool threshold = p.s->compression == 1 && !std::is_integral<V>::value;
    for (auto& kv : w.m) {
        V val = from_initiator_value<V>(rule, thr, kv.second);
        if (threshold) {
            const double th = p.s->compression_threshold;
            const double prob = std::fabs((double)val) / th;
            if (val != 0 && prob < 1) {
                P4 r = draw(p.seed, kv.first, 0, p.step, TAG_COMPRESS);
                const double sg = (val > 0) - (val < 0);
                val = (V)((prob > u53(r.v[0], r.v[1])) ? th * sg : 0.0);
            }
            if (val == 0) continue;
        }
        out.emplace(kv.first, val);
    }
    double norm = 0;
    for (auto& kv : out) norm += std::fabs((double)kv.second);
    stats[0] = st.attempts; stats[1] = st.exact; stats[2] = st.inexact; stats[3] = st.spawns;
    stats[4] = st.deaths; stats[5] = st.clones; stats[6] = st.zombies; stats[7] = len_before;
    stats[8] = (double)out.size(); stats[9] = norm;
    return dump_sorted<V>(out, out_keys, out_vals, cap);
}

}  // namespace

extern "C" {

// One FCIQMC step of a vector with an initiator rule (rule 0 NonInitiator, 1 Initiator,
// 2 SimpleInitiator, 3 CoherentInitiator; threshold as in the rule's constructor).
// Arguments and result otherwise as or_step.  With NoCompression the output may hold
// entries whose value is zero (an address that received only discarded deposits).
int64_t or_step_initiator(const or_model* h, const or_style* s, uint64_t seed, int64_t step,
                          double shift, double dtau, double boost, int32_t order, int32_t vtype,
                          int32_t rule, double threshold, const uint64_t* keys, const void* vals,
                          int64_t n, uint64_t* out_keys, void* out_vals, int64_t cap, double* stats) {
    Params p{h, s, seed, step, shift, dtau, boost};
    if (model_bits(h) > 64) {
        if (vtype == 0)
            return step_initiator_impl<u128, int64_t>(p, order, rule, threshold, keys, (const int64_t*)vals,
                                                      n, out_keys, (int64_t*)out_vals, cap, stats);
        return step_initiator_impl<u128, double>(p, order, rule, threshold, keys, (const double*)vals, n,
                                                 out_keys, (double*)out_vals, cap, stats);
    }
    if (vtype == 0)
        return step_initiator_impl<uint64_t, int64_t>(p, order, rule, threshold, keys, (const int64_t*)vals, n,
                                                      out_keys, (int64_t*)out_vals, cap, stats);
    return step_initiator_impl<uint64_t, double>(p, order, rule, threshold, keys, (const double*)vals, n,
                                                 out_keys, (double*)out_vals, cap, stats);
}

// One FCIQMC step (apply_operator! + compress!) of the source vector (keys, vals).
// vtype 0: int64 values (IsStochasticInteger); 1: double values.  Keys: one uint64 word per
// address, or two (lo, hi) when N + M - 1 > 64 (also for out_keys).
// order 0: deposits in the reference's column order; 1: list order (diagonal first).
// stats[0..9] = attempts, exact_steps, inexact_steps, spawns, deaths, clones, zombies,
//               len_before, len, norm.  Returns the output length, or -(needed) if cap
// is too small.  Output keys are sorted ascending.
int64_t or_step(const or_model* h, const or_style* s, uint64_t seed, int64_t step,
                double shift, double dtau, double boost, int32_t order, int32_t vtype,
                const uint64_t* keys, const void* vals, int64_t n,
                uint64_t* out_keys, void* out_vals, int64_t cap, double* stats) {
    Params p{h, s, seed, step, shift, dtau, boost};
    if (model_bits(h) > 64) {   // two-chunk addresses: keys hold (lo, hi) words
        if (vtype == 0)
            return step_impl<u128, int64_t>(p, order, keys, (const int64_t*)vals, n, out_keys,
                                            (int64_t*)out_vals, cap, stats);
        return step_impl<u128, double>(p, order, keys, (const double*)vals, n, out_keys,
                                       (double*)out_vals, cap, stats);
    }
    if (vtype == 0)
        return step_impl<uint64_t, int64_t>(p, order, keys, (const int64_t*)vals, n, out_keys,
                                            (int64_t*)out_vals, cap, stats);
    return step_impl<uint64_t, double>(p, order, keys, (const double*)vals, n, out_keys,
                                       (double*)out_vals, cap, stats);
}

// Annihilation of a spawn list: a left fold of deposit! over the list, dvec.jl:126-136.
int64_t or_annihilate(int32_t vtype, const uint64_t* keys, const void* vals, int64_t n,
                      uint64_t* out_keys, void* out_vals, int64_t cap) {
    if (vtype == 0) {
        DictVec<int64_t> w;
        const int64_t* v = (const int64_t*)vals;
        for (int64_t i = 0; i < n; ++i) w.deposit(keys[i], v[i]);
        return dump_sorted<int64_t>(w.m, out_keys, (int64_t*)out_vals, cap);
    }
    DictVec<double> w;
    const double* v = (const double*)vals;
    for (int64_t i = 0; i < n; ++i) w.deposit(keys[i], v[i]);
    return dump_sorted<double>(w.m, out_keys, (double*)out_vals, cap);
}

// ---------------------------------------------------------------------------------
// Overlaps: dot(t, u) and the three-way dot(t, op, u) = dot_from_right(t, op, u),
// DictVectors/abstractdvec.jl:313-324: a sum over the entries (k, v) of u of
//   conj(t[k]) * diagonal_element(op, k) * v  +  sum over offdiagonals (k', h') of
//   conj(t[k']) * h' * v.
// Operators: kind 1 a Hamiltonian (model), kind 2 G2RealCorrelator(d)
// (correlation_functions.jl:68-92: (1/M) sum_i n_i (n_{i+d} - delta_{0d}), periodic),
// kind 3 DensityMatrixDiagonal(mode) (DensityMatrixDiagonal.jl:19-33: n_mode).
// Sums are taken in the order of u's entries as given.
// ---------------------------------------------------------------------------------
double or_g2_real(uint64_t bs, int32_t M, int32_t d) {
    int onr[64];
    to_onr(bs, M, onr);
    int64_t acc = 0;
    const int dd = ((d % M) + M) % M;
    if (dd == 0) {
        for (int i = 0; i < M; ++i) acc += (int64_t)onr[i] * (onr[i] - 1);
    } else {
        for (int i = 0; i < M; ++i) acc += (int64_t)onr[i] * onr[(i + dd) % M];
    }
    return (double)acc / (double)M;
}

double or_density_diagonal(uint64_t bs, int32_t M, int32_t mode) {
    int onr[64];
    to_onr(bs, M, onr);
    return (double)onr[mode - 1];
}

}  // extern "C"

namespace {
template <class A>
void overlaps_impl(const uint64_t* tk, const void* tv, int32_t t_vtype, int64_t nt, const uint64_t* uk,
                   const void* uv, int32_t u_vtype, int64_t nu, int32_t M, int32_t nops,
                   const int32_t* kinds, const int32_t* params, const or_model* models, double* out) {
    KeyMap<A, double> t;
    t.reserve((size_t)nt * 2 + 1);
    for (int64_t i = 0; i < nt; ++i)
        t[load_key<A>(tk, i)] = t_vtype == 0 ? (double)((const int64_t*)tv)[i] : ((const double*)tv)[i];
    auto T = [&](A k) { auto it = t.find(k); return it == t.end() ? 0.0 : it->second; };
    for (int o = 0; o <= nops; ++o) out[o] = 0;
    for (int64_t i = 0; i < nu; ++i) {
        const A k = load_key<A>(uk, i);
        const double v = u_vtype == 0 ? (double)((const int64_t*)uv)[i] : ((const double*)uv)[i];
        const double tk_ = T(k);
        out[0] += tk_ * v;
        int onr[128];
        to_onr(k, M, onr);
        for (int o = 0; o < nops; ++o) {
            double res;
            if (kinds[o] == 2) {
                const int dd = ((params[o] % M) + M) % M;
                int64_t acc = 0;
                for (int j = 0; j < M; ++j)
                    acc += (int64_t)onr[j] * (dd == 0 ? onr[j] - 1 : onr[(j + dd) % M]);
                res = tk_ * ((double)acc / (double)M) * v;
            } else if (kinds[o] == 3) {
                res = tk_ * (double)onr[params[o] - 1] * v;
            } else {
                const or_model* h = &models[o];
                res = tk_ * diagonal_element_t(h, k) * v;
                const int64_t n = num_offdiagonals_t(h, k);
                for (int64_t c = 1; c <= n; ++c) {
                    double hv;
                    const A nb = get_offdiagonal_t(h, k, c, &hv);
                    res += T(nb) * hv * v;
                }
            }
            out[1 + o] += res;
        }
    }
}
}  // namespace

extern "C" {

// out[0] = dot(t, u); out[1 + o] = dot(t, op_o, u) for the nops operators (kinds[o], params[o],
// models[o]).  tv / uv are int64 (vtype 0) or double (vtype 1) per vector; kw = words per
// address (1, or 2 for N + M - 1 > 64).  Returns 0.
int32_t or_overlaps(const uint64_t* tk, const void* tv, int32_t t_vtype, int64_t nt,
                    const uint64_t* uk, const void* uv, int32_t u_vtype, int64_t nu,
                    int32_t M, int32_t kw, int32_t nops, const int32_t* kinds, const int32_t* params,
                    const or_model* models, double* out) {
    if (kw == 2) overlaps_impl<u128>(tk, tv, t_vtype, nt, uk, uv, u_vtype, nu, M, nops, kinds, params, models, out);
    else overlaps_impl<uint64_t>(tk, tv, t_vtype, nt, uk, uv, u_vtype, nu, M, nops, kinds, params, models, out);
    return 0;
}

// DoubleLogUpdate, shiftstrategy.jl:174-181.
double or_double_log_update(double shift, double tnorm, double pnorm, double target,
                            double zeta, double xi, double dtau) {
    return shift - (xi / dtau * std::log(tnorm / target) + zeta / dtau * std::log(tnorm / pnorm));
}

// ---------------------------------------------------------------------------------
// Full basis and deterministic products (IsDeterministic path, mul! abstractdvec.jl:275-284).
// ---------------------------------------------------------------------------------

// All BoseFS addresses of N bosons in M modes, sorted ascending (build_basis restated:
// the number-conserving basis of basis_fock.jl).  Returns count or -(needed).
namespace {
void enumerate_compositions(int i, int left, int M, std::vector<int>& onr, std::vector<uint64_t>& out) {
    if (i == M - 1) { onr[i] = left; out.push_back(from_onr(onr.data(), M)); onr[i] = 0; return; }
    for (int k = 0; k <= left; ++k) { onr[i] = k; enumerate_compositions(i + 1, left - k, M, onr, out); }
    onr[i] = 0;
}
}  // namespace

int64_t or_build_basis(int32_t M, int32_t N, uint64_t* out, int64_t cap) {
    std::vector<int> onr(M, 0);
    std::vector<uint64_t> tmp;
    enumerate_compositions(0, N, M, onr, tmp);
    std::sort(tmp.begin(), tmp.end());
    const int64_t count = (int64_t)tmp.size();
    if (count > cap) return -count;
    std::memcpy(out, tmp.data(), sizeof(uint64_t) * tmp.size());
    return count;
}

// w = (a*1 + b*H) v on a sorted basis, accumulated column by column in basis order
// (the DVec mul! / IsDeterministic deposit order).  With a = 1 + dt*S, b = -dt this is
// T v (fciqmc.jl:93-112); with a = 0, b = 1 it is H v.  nrhs vectors stored row-major
// [D][nrhs], each with its own model (u, t) in hs[r].
int32_t or_apply_affine(const or_model* hs, int32_t nrhs, const uint64_t* basis, int64_t D,
                        const double* a, const double* b, const double* v, double* w) {
    std::memset(w, 0, sizeof(double) * D * nrhs);
    for (int64_t j = 0; j < D; ++j) {
        const uint64_t addr = basis[j];
        BoseIndex occ[64];
        const int nocc = occupied_modes(addr, occ);
        for (int r = 0; r < nrhs; ++r) {
            const or_model* h = &hs[r];
            const double vj = v[j * nrhs + r];
            const double tdiag = a[r] + b[r] * or_diagonal_element(h, addr);
            w[j * nrhs + r] += tdiag * vj;
            const int64_t n_od = or_num_offdiagonals(h, addr);
            for (int64_t c = 1; c <= n_od; ++c) {
                double hv;
                const uint64_t nb = (h->model == 3) ? rs_offdiagonal(h, addr, c, &hv)
                                  : (h->model == 1) ? real1d_offdiagonal(h, addr, occ, nocc, c, &hv)
                                                    : mom1d_offdiagonal(h, addr, occ, nocc, c, &hv);
                const int64_t k = std::lower_bound(basis, basis + D, nb) - basis;
                if (k >= D || basis[k] != nb) return -1;
                w[k * nrhs + r] += (b[r] * hv) * vj;
            }
        }
    }
    return 0;
}

// All nonzero matrix elements of H on a sorted basis as (row, col, value) triplets, in
// column order (duplicates of the same (row, col) are listed separately).  Returns the
// number of triplets, or -(needed) if cap is too small, or -1 - cap if a neighbour is
// outside the basis.
int64_t or_matrix_triplets(const or_model* h, const uint64_t* basis, int64_t D, int64_t* rows,
                           int64_t* cols, double* vals, int64_t cap) {
    int64_t n = 0;
    for (int64_t j = 0; j < D; ++j) {
        const uint64_t addr = basis[j];
        BoseIndex occ[64];
        const int nocc = occupied_modes(addr, occ);
        const double d = or_diagonal_element(h, addr);
        if (n < cap) { rows[n] = j; cols[n] = j; vals[n] = d; }
        ++n;
        const int64_t n_od = or_num_offdiagonals(h, addr);
        for (int64_t c = 1; c <= n_od; ++c) {
            double hv;
            const uint64_t nb = (h->model == 3) ? rs_offdiagonal(h, addr, c, &hv)
                              : (h->model == 1) ? real1d_offdiagonal(h, addr, occ, nocc, c, &hv)
                                                : mom1d_offdiagonal(h, addr, occ, nocc, c, &hv);
            const int64_t k = std::lower_bound(basis, basis + D, nb) - basis;
            if (k >= D || basis[k] != nb) return -1 - cap;
            if (n < cap) { rows[n] = k; cols[n] = j; vals[n] = hv; }
            ++n;
        }
    }
    return n <= cap ? n : -n;
}

// Colex rank of an address among all N-boson M-mode addresses: the address is an
// N-subset (positions of the one-bits) of B = N + M - 1 bits; rank = sum_j C(p_j, j+1).
// This ordering equals ascending numeric order of the bit strings.
int64_t or_rank(uint64_t bs) {
    int64_t r = 0;
    int j = 1;
    while (bs) {
        const int p = __builtin_ctzll(bs);
        // C(p, j)
        int64_t c = 1;
        if (p < j) c = 0;
        else for (int i = 1; i <= j; ++i) c = c * (p - j + i) / i;
        r += c;
        ++j;
        bs &= bs - 1;
    }
    return r;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Threaded restatement of PDVec's step (the timed CPU baseline).
// PDVec hash segments, pdvec.jl:43-50, fastrange_hash :6-9; per-thread spawn columns,
// perform_spawns! pdworkingmemory.jl:208-217; collect_local! :228-235;
// move_and_compress! :262-273.  Each thread owns one segment; each thread spawns into
// its own column of per-target-segment maps; then each thread merges column entries for
// its segment.  Hash maps are open-addressing tables (like Julia's Dict).
// ---------------------------------------------------------------------------------
namespace {

inline uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}
inline int segment_of(uint64_t key, int nseg) {
    return (int)(((unsigned __int128)mix64(key) * (unsigned __int128)nseg) >> 64);
}

template <class V>
struct OpenMap {
    std::vector<uint64_t> k;
    std::vector<V> v;
    std::vector<uint8_t> used;
    size_t mask = 0, count = 0;
    void init(size_t cap) {
        size_t c = 16;
        while (c < cap * 2) c <<= 1;
        k.assign(c, 0); v.assign(c, V(0)); used.assign(c, 0); mask = c - 1; count = 0;
    }
    void grow() {
        std::vector<uint64_t> ok; std::vector<V> ov; std::vector<uint8_t> ou;
        ok.swap(k); ov.swap(v); ou.swap(used);
        init((ou.size()));
        for (size_t i = 0; i < ou.size(); ++i) if (ou[i]) add(ok[i], ov[i]);
    }
    inline void add(uint64_t key, V val) {
        if ((count + 1) * 2 > mask + 1) grow();
        size_t i = mix64(key) & mask;
        while (used[i]) {
            if (k[i] == key) { v[i] = v[i] + val; return; }
            i = (i + 1) & mask;
        }
        used[i] = 1; k[i] = key; v[i] = V(0) + val; ++count;
    }
};

}  // namespace

extern "C" {

// One step of the threaded PDVec-style restatement, with nthreads segments.  Input and
// output are flat arrays (the output is not sorted).  Returns output length or -needed.
int64_t or_step_threaded(const or_model* h, const or_style* s, uint64_t seed, int64_t step,
                         double shift, double dtau, double boost, int32_t vtype, int32_t nthreads,
                         const uint64_t* keys, const void* vals, int64_t n,
                         uint64_t* out_keys, void* out_vals, int64_t cap, double* stats) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    Params p{h, s, seed, step, shift, dtau, boost};
    int64_t result = 0;
    auto run = [&](auto tag) {
        using V = decltype(tag);
        const V* vv = (const V*)vals;
        V* ov = (V*)out_vals;
        const int T = nthreads;
        std::vector<OpenMap<V>> cols((size_t)T * T);
        std::vector<Stats> sts(T);
        const int64_t per = n / T + 1;
        for (auto& c : cols) c.init((size_t)(2 * per / T + 64));
#pragma omp parallel num_threads(T)
        {
#ifdef _OPENMP
            const int t = omp_get_thread_num();
#else
            const int t = 0;
#endif
            // perform_spawns!: every thread processes a contiguous share of the source
            // entries and deposits into its own column, split by target segment.
            const int64_t lo = (int64_t)t * n / T, hi = (int64_t)(t + 1) * n / T;
            for (int64_t i = lo; i < hi; ++i) {
                apply_column<V>(p, keys[i], vv[i], sts[t], [&](uint64_t k, V v) {
                    cols[(size_t)t * T + segment_of(k, T)].add(k, v);
                });
            }
#pragma omp barrier
            // collect_local!: thread t merges column entries of segment t.
            OpenMap<V>& seg = cols[(size_t)0 * T + t];
            for (int c = 1; c < T; ++c) {
                OpenMap<V>& src = cols[(size_t)c * T + t];
                for (size_t i = 0; i <= src.mask; ++i) if (src.used[i]) seg.add(src.k[i], src.v[i]);
            }
        }
        // move_and_compress! + gather (serial gather of the segments into flat output).
        Stats st;
        for (auto& x : sts) {
            st.attempts += x.attempts; st.exact += x.exact; st.inexact += x.inexact; st.spawns += x.spawns;
            st.deaths += x.deaths; st.clones += x.clones; st.zombies += x.zombies;
        }
        int64_t len_before = 0, len = 0;
        double norm = 0;
        std::vector<int64_t> seg_count(T, 0), seg_before(T, 0);
        std::vector<double> seg_norm(T, 0);
#pragma omp parallel for num_threads(T)
        for (int t = 0; t < T; ++t) {
            OpenMap<V>& seg = cols[(size_t)0 * T + t];
            for (size_t i = 0; i <= seg.mask; ++i) {
                if (!seg.used[i] || seg.v[i] == 0) { seg.used[i] = 0; continue; }
                seg_before[t]++;
                if constexpr (!std::is_integral<V>::value) {
                    if (s->compression == 1) {
                        const double th = s->compression_threshold;
                        const double prob = std::fabs(seg.v[i]) / th;
                        if (prob < 1) {
                            P4 r = draw(seed, seg.k[i], 0, step, TAG_COMPRESS);
                            const double sg = (seg.v[i] > 0) - (seg.v[i] < 0);
                            seg.v[i] = (prob > u53(r.v[0], r.v[1])) ? th * sg : 0.0;
                        }
                        if (seg.v[i] == 0) { seg.used[i] = 0; continue; }
                    }
                }
                seg_count[t]++;
                seg_norm[t] += std::fabs((double)seg.v[i]);
            }
        }
        for (int t = 0; t < T; ++t) { len_before += seg_before[t]; len += seg_count[t]; norm += seg_norm[t]; }
        stats[0] = st.attempts; stats[1] = st.exact; stats[2] = st.inexact; stats[3] = st.spawns;
        stats[4] = st.deaths; stats[5] = st.clones; stats[6] = st.zombies; stats[7] = (double)len_before;
        stats[8] = (double)len; stats[9] = norm;
        if (len > cap) { result = -len; return; }
        std::vector<int64_t> off(T + 1, 0);
        for (int t = 0; t < T; ++t) off[t + 1] = off[t] + seg_count[t];
#pragma omp parallel for num_threads(T)
        for (int t = 0; t < T; ++t) {
            OpenMap<V>& seg = cols[(size_t)0 * T + t];
            int64_t o = off[t];
            for (size_t i = 0; i <= seg.mask; ++i)
                if (seg.used[i]) { out_keys[o] = seg.k[i]; ov[o] = seg.v[i]; ++o; }
        }
        result = len;
    };
    if (vtype == 0) run((int64_t)0); else run((double)0);
    return result;
}

}  // extern "C"
