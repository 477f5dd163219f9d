// bosefs_device.hpp -- BoseFS addresses, Hubbard matrix elements and Philox streams,
// written as register-only bit arithmetic for one wavefront lane.
//
// Address layout (BitStringAddresses/bitstring.jl:464-508): B = N + M - 1 bits in one
// 64-bit chunk, or in two (B <= 128) held as one 128-bit integer with the last BitString
// chunk (the least significant bits) low; mode 1 sits in the least significant bits and
// every mode is n_i one-bits followed by a zero separator.  A "run" is the block of ones of
// one mode.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rimu {

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter (c0..c3), key (k0, k1), ten rounds.
// ---------------------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                            uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// Stream purposes live in the top two bits of counter word 3, the step in the rest.
enum : uint32_t { TAG_SPAWN = 0u, TAG_DIAG = 1u, TAG_COMPRESS = 2u };

__host__ __device__ inline U4 draw(uint64_t seed, uint64_t addr, uint32_t attempt, int64_t step,
                                   uint32_t tag) {
    const uint32_t c3 = ((uint32_t)step & 0x3FFFFFFFu) | (tag << 30);
    return philox4x32_10((uint32_t)addr, (uint32_t)(addr >> 32), attempt, c3, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
}

// Two-chunk addresses: the high chunk selects the Philox key (seed ^ hi * golden ratio, a
// bijection of hi for a given seed), the low chunk the counter as above; hi = 0 gives the
// one-chunk streams.
__host__ __device__ inline U4 draw(uint64_t seed, unsigned __int128 addr, uint32_t attempt,
                                   int64_t step, uint32_t tag) {
    const uint64_t hi = (uint64_t)(addr >> 64);
    return draw(seed ^ (hi * 0x9E3779B97F4A7C15ull), (uint64_t)addr, attempt, step, tag);
}

// 53-bit uniform double in [0, 1).
__host__ __device__ inline double u53(uint32_t hi, uint32_t lo) {
    const uint64_t x = ((((uint64_t)hi) << 32) | lo) >> 11;
    return (double)x * 0x1.0p-53;
}

// Uniform index in 1..n (multiply-shift).
__host__ __device__ inline int64_t pick1(uint32_t r, int64_t n) {
    return 1 + (int64_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

// ---------------------------------------------------------------------------------------
// Bit helpers.  An address is one unsigned integer A: uint64_t for B <= 64 bits, u128 (two
// BitString chunks, the low chunk in the low 64 bits) for 64 < B <= 128.  Both layouts are
// the same bit string: mode 1 in the least significant bits.
// ---------------------------------------------------------------------------------------
typedef unsigned __int128 u128;

template <class A> struct addr_bits { static constexpr int value = 8 * (int)sizeof(A); };

__device__ inline int popc64(uint64_t x) { return __popcll(x); }
__device__ inline int ctz64(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }
__device__ inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
__device__ inline uint64_t low_mask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }
__device__ inline uint64_t shr(uint64_t x, int n) { return n >= 64 ? 0ull : (x >> n); }
__device__ inline uint64_t shl(uint64_t x, int n) { return n >= 64 ? 0ull : (x << n); }

__device__ inline int popcA(uint64_t x) { return __popcll(x); }
__device__ inline int popcA(u128 x) { return __popcll((uint64_t)x) + __popcll((uint64_t)(x >> 64)); }
__device__ inline int ctzA(uint64_t x) { return ctz64(x); }
__device__ inline int ctzA(u128 x) {
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    return lo ? __builtin_ctzll(lo) : (hi ? 64 + __builtin_ctzll(hi) : 128);
}
__device__ inline int clzA(uint64_t x) { return clz64(x); }
__device__ inline int clzA(u128 x) {
    const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    return hi ? __builtin_clzll(hi) : (lo ? 64 + __builtin_clzll(lo) : 128);
}
template <class A> __device__ inline A shrA(A x, int n) { return n >= addr_bits<A>::value ? (A)0 : (A)(x >> n); }
template <class A> __device__ inline A shlA(A x, int n) { return n >= addr_bits<A>::value ? (A)0 : (A)(x << n); }
template <class A> __device__ inline A low_maskA(int n) {
    return n >= addr_bits<A>::value ? (A)~(A)0 : (A)(((A)1 << n) - (A)1);
}

// Position of the k-th (0-based) set bit of x; x must have more than k set bits.
__device__ inline int select_bit(uint64_t x, int k) {
    int pos = 0;
    int c = __popc((uint32_t)x);
    if (k >= c) { k -= c; x >>= 32; pos += 32; }
    uint32_t v = (uint32_t)x;
    c = __popc(v & 0xFFFFu);
    if (k >= c) { k -= c; v >>= 16; pos += 16; }
    c = __popc(v & 0xFFu);
    if (k >= c) { k -= c; v >>= 8; pos += 8; }
    c = __popc(v & 0xFu);
    if (k >= c) { k -= c; v >>= 4; pos += 4; }
    c = __popc(v & 0x3u);
    if (k >= c) { k -= c; v >>= 2; pos += 2; }
    c = (int)(v & 1u);
    if (k >= c) { pos += 1; }
    return pos;
}
__device__ inline int select_bit(u128 x, int k) {
    const uint64_t lo = (uint64_t)x;
    const int c = __popcll(lo);
    return k < c ? select_bit(lo, k) : 64 + select_bit((uint64_t)(x >> 64), k - c);
}

// Number of consecutive ones starting at bit p upwards.
template <class A> __device__ inline int ones_from(A bits, int p) { return ctzA((A)~shrA(bits, p)); }

// First bit of the run of mode m (1-based): modes are separated by zeros, so the run of
// mode m starts right after the (m-1)-th zero.
template <class A> __device__ inline int mode_start(A bits, int m) {
    return m == 1 ? 0 : select_bit((A)~bits, m - 2) + 1;
}

// Run starts: a one whose lower neighbour is zero (or bit 0).
template <class A> __device__ inline A run_starts(A bits) { return bits & (A)~(bits << 1); }

// Delete the bit at position p (higher bits move down).
template <class A> __device__ inline A delete_bit(A x, int p) {
    return (x & low_maskA<A>(p)) | shlA(shrA(x, p + 1), p);
}
// Insert a one at position p (bits at p and above move up).
template <class A> __device__ inline A insert_one(A x, int p) {
    return (x & low_maskA<A>(p)) | ((A)1 << p) | shlA(shrA(x, p), p + 1);
}

// ---------------------------------------------------------------------------------------
// Storage order of keys.  Above MIX_MIN_BITS address bits the engine stores every address
// x as y = mix(x), a bijection of [0, 2^B) (multiply by an odd constant mod 2^B, fold the
// high half down, twice).  The top bits of y then depend on every bit of x and are spread
// evenly over the addresses of a vector: the spawn records split into buckets of about equal
// size by their top bits (the bucket merge), and sorting them on the top MIX_SORT_BITS bits of
// y leaves only rare groups to order on the rest.  The vector is kept in y order; the step's
// results per address are unchanged (every draw and every matrix element is a function of x).
// At or below MIX_MIN_BITS, y = x.
// ---------------------------------------------------------------------------------------
constexpr int MIX_MIN_BITS = 20;
constexpr int MIX_SORT_BITS = 40;

struct KeyMix {
    int on, B, s1, s2;
    u128 c1, c2, c1i, c2i;  // odd multipliers mod 2^B and their inverses
};

template <class T> __host__ __device__ inline T mix_mask(int B) {
    return B >= 8 * (int)sizeof(T) ? (T)~(T)0 : (T)(((T)1 << B) - 1);
}
// T: uint64_t when B <= 64, u128 otherwise
template <class T> __host__ __device__ inline T mix_bits(const KeyMix& m, T y) {
    const T mask = mix_mask<T>(m.B);
    y = (T)(y * (T)m.c1) & mask;
    y ^= y >> m.s1;
    y = (T)(y * (T)m.c2) & mask;
    y ^= y >> m.s2;
    return y;
}
template <class T> __host__ __device__ inline T unmix_bits(const KeyMix& m, T y) {
    const T mask = mix_mask<T>(m.B);
    y ^= y >> m.s2;  // s >= B/2: the fold is its own inverse
    y = (T)(y * (T)m.c2i) & mask;
    y ^= y >> m.s1;
    y = (T)(y * (T)m.c1i) & mask;
    return y;
}
template <class K> __host__ __device__ inline K mix_key(const KeyMix& m, K x) { return m.on ? mix_bits<K>(m, x) : x; }
template <class K> __host__ __device__ inline K unmix_key(const KeyMix& m, K y) { return m.on ? unmix_bits<K>(m, y) : y; }

inline u128 inverse_mod_2_128(u128 c) {  // Newton: each round doubles the correct low bits
    u128 x = c;
    for (int i = 0; i < 7; ++i) x *= (u128)2 - c * x;
    return x;
}
inline KeyMix make_key_mix(int B) {
    KeyMix m{};
    m.B = B;
    m.on = B > MIX_MIN_BITS;
    m.s1 = (B + 1) / 2;
    m.s2 = (B + 1) / 2;
    m.c1 = ((u128)0x9E3779B97F4A7C15ull << 64) | 0xF39CC0605CEDC835ull;
    m.c2 = ((u128)0xD1B54A32D192ED03ull << 64) | 0xABC98388FB8FAC03ull;
    m.c1i = inverse_mod_2_128(m.c1);
    m.c2i = inverse_mod_2_128(m.c2);
    return m;
}

// ---------------------------------------------------------------------------------------
// Model data passed to kernels by value.
// ---------------------------------------------------------------------------------------
struct ModelDev {
    int kind;   // 1 HubbardReal1D, 2 HubbardMom1D, 3 HubbardRealSpace
    int M, N, B;  // B: bits of the whole key
    double u, t;
    double kes[64];  // HubbardMom1D: dispersion; HubbardRealSpace: trap potential of component 1
    // HubbardRealSpace
    int ncomp, N2, fermi, ndir, has_u, has_pot;
    int cbits2;  // bits of component 2 (component 1 sits above them); 0 for one component
    double t2, u12, u22;
    double pot2[64];
    uint8_t nbr[64][6];
};

// Sum over occupied modes of n(n-1) (bose_hubbard_interaction, bosefs.jl:415-428), bit
// parallel: y_d = bits & (bits >> 1) & ... & (bits >> d) holds n - d ones for every run of
// n ones (runs are separated by zeros), so sum_{d >= 1} popcount(y_d) = sum n(n-1)/2.  The
// loop runs (largest occupation - 1) times instead of once per occupied mode.
template <class A> __device__ inline int64_t bose_hubbard_interaction(A bits) {
    int64_t acc = 0;
    A s = bits >> 1;
    A y = bits & s;
    while (y) {
        acc += popcA(y);
        s >>= 1;
        y &= s;
    }
    return 2 * acc;
}

template <class A> __device__ inline int num_occupied(A bits) { return popcA(run_starts(bits)); }

// num_offdiagonals: HubbardReal1D.jl:51-53 (2 * occupied modes);
// HubbardMom1D.jl:131-140 (s(s-1)(M-2) + d(M-1)).
template <class A> __device__ inline int64_t num_offdiagonals(const ModelDev& h, A bits) {
    const A st = run_starts(bits);
    const int64_t s = popcA(st);
    if (h.kind == 1) return 2 * s;
    const int64_t d = popcA((A)(st & (bits >> 1)));  // runs of length >= 2
    return s * (s - 1) * (h.M - 2) + d * (h.M - 1);
}

// diagonal_element: HubbardReal1D.jl:55-57; HubbardMom1D.jl:174-177 (kinetic sum over
// occupied modes in ascending order, then u/(2M) * onproduct, excitations.jl:126-137).
template <class A> __device__ inline double diagonal_element(const ModelDev& h, A bits) {
    if (h.kind == 1) return h.u * (double)bose_hubbard_interaction(bits) / 2;
    double kin = 0.0;
    int64_t sumsq = 0;
    A x = bits;
    int mode = 1;
    while (x) {
        const int z = ctzA(x);
        mode += z;
        x = shrA(x, z);
        const int n = ctzA((A)~x);
        kin += h.kes[mode - 1] * n;
        sumsq += (int64_t)n * n;
        x = shrA(x, n);
        // the separator after this run is consumed by the next ctz (counts as one mode step)
    }
    const int64_t N = h.N;
    const int64_t onproduct = 2 * N * N - N - sumsq;  // sum n(n-1) + 4 sum_{i<j} n_i n_j
    return kin + h.u / (2.0 * h.M) * (double)onproduct;
}

// hopnextneighbour (bosefs.jl:297-345): the neighbour address and the occupation product
// under the square root.  chosen in 1..2*s: site (chosen+1)/2 is the site-th occupied mode;
// odd hops right.  M modes, B = N + M - 1 bits.
template <class A>
__device__ inline A real1d_hop(A bits, int64_t chosen, int M, int B, int64_t* prod_out) {
    constexpr int W = addr_bits<A>::value;
    const int site = (int)((chosen + 1) >> 1);
    const A st = run_starts(bits);
    const int p = select_bit(st, site - 1);
    const int n = ones_from(bits, p);
    const int m = p - popcA((A)(bits & low_maskA<A>(p))) + 1;  // mode = zeros below p, plus one
    const A maskB = low_maskA<A>(B);
    A nb;
    int64_t prod;
    if (chosen & 1) {  // right
        if (m == M) {
            nb = ((bits << 1) | (A)1) & maskB;
            prod = (int64_t)n * (ones_from(bits, 0) + 1);
        } else {
            const int q = p + n - 1;
            nb = bits ^ ((A)3 << q);
            prod = (int64_t)n * (ones_from(bits, q + 2) + 1);
        }
    } else {  // left
        if (m == 1) {
            nb = (bits >> 1) | ((A)1 << (B - 1));
            const int hz = W - 1 - clzA((A)(~bits & maskB));  // highest zero inside the B bits
            prod = (int64_t)n * ((B - 1 - hz) + 1);
        } else {
            nb = bits ^ ((A)3 << (p - 1));
            int nprev = 0;
            if (p >= 2) {
                const A below = bits & low_maskA<A>(p - 1);  // bits 0 .. p-2
                const A z = (A)~below & low_maskA<A>(p - 1);
                nprev = z ? (p - 2 - (W - 1 - clzA(z))) : (p - 1);
            }
            prod = (int64_t)n * (nprev + 1);
        }
    }
    *prod_out = prod;
    return nb;
}

// get_offdiagonal of HubbardReal1D: value -t*sqrt(prod) (HubbardReal1D.jl:59-62).
template <class A>
__device__ inline A real1d_offdiagonal(const ModelDev& h, A bits, int64_t chosen, double* value) {
    int64_t prod;
    const A nb = real1d_hop(bits, chosen, h.M, h.B, &prod);
    *value = -h.t * sqrt((double)prod);
    return nb;
}

__device__ inline int imod1(int x, int m) { int r = ((x - 1) % m + m) % m; return r + 1; }

// Occupation number of mode m.
template <class A> __device__ inline int occupation(A bits, int m) { return ones_from(bits, mode_start(bits, m)); }

// momentum_transfer_excitation (fold=true), excitations.jl:26-80, value u/(2M)*sqrt(...)
// (HubbardMom1D.jl:189-194).  The map of occupied modes is read from the run starts.
template <class A>
__device__ inline A mom1d_offdiagonal(const ModelDev& h, A bits, int64_t chosen, double* value) {
    // chosen <= s(s-1)(M-2) + d(M-1) < 2^31 for M <= 64: the index arithmetic runs in 32-bit
    // unsigned integers (64-bit division is a long software sequence on the GPU)
    const int M = h.M;
    const A st = run_starts(bits);
    const int s = popcA(st);
    const uint32_t c = (uint32_t)chosen;
    const uint32_t single = (uint32_t)(s * (s - 1) * (M - 2));
    int p_mode, q_mode;  // source modes (p <= q)
    int mom;
    if (c > single) {
        const uint32_t dbl = c - single;
        const uint32_t m1 = (uint32_t)(M - 1);
        mom = (int)((dbl - 1) % m1) + 1;
        const uint32_t d = (dbl - 1) / m1 + 1;
        const A dstarts = st & (bits >> 1);
        const int pos = select_bit(dstarts, (int)d - 1);
        p_mode = q_mode = pos - popcA((A)(bits & low_maskA<A>(pos))) + 1;
    } else {
        const uint32_t m2 = (uint32_t)(M - 2), s1 = (uint32_t)(s - 1);
        mom = (int)((c - 1) % m2) + 1;
        const uint32_t pair = (c - 1) / m2 + 1;
        const int snd = (int)((pair - 1) % s1) + 1;
        const int fst = (int)((pair - 1) / s1) + 1;
        int f_hole, s_hole;
        if (snd < fst) { f_hole = snd; s_hole = fst; }
        else { f_hole = fst; s_hole = snd + 1; }
        const int pf = select_bit(st, (int)f_hole - 1);
        const int ps = select_bit(st, (int)s_hole - 1);
        p_mode = pf - popcA((A)(bits & low_maskA<A>(pf))) + 1;
        q_mode = ps - popcA((A)(bits & low_maskA<A>(ps))) + 1;
        if (mom >= q_mode - p_mode) mom += 1;
    }
    const int r_mode = imod1((int)(p_mode + mom), M);
    const int s_mode = imod1((int)(q_mode - mom), M);
    // bose_excitation_value (fockaddress.jl:559-567) with reversed operator lists:
    // destroy q, destroy p, create s, create r -- each sees the previous operators.
    const int np = occupation(bits, p_mode);
    const int nq = occupation(bits, q_mode);
    const int nr = occupation(bits, r_mode);
    const int ns = occupation(bits, s_mode);
    int64_t v = nq;
    v *= (np - (p_mode == q_mode));
    v *= (ns - (s_mode == q_mode) - (s_mode == p_mode) + 1);
    v *= (nr - (r_mode == q_mode) - (r_mode == p_mode) + (r_mode == s_mode) + 1);
    // new address: remove one boson from p and q, add one to r and s
    A nb = bits;
    nb = delete_bit(nb, mode_start(nb, q_mode));
    nb = delete_bit(nb, mode_start(nb, p_mode));
    nb = insert_one(nb, mode_start(nb, s_mode));
    nb = insert_one(nb, mode_start(nb, r_mode));
    *value = h.u / (2 * M) * sqrt((double)v);
    return nb;
}

template <class A>
__device__ inline A get_offdiagonal(const ModelDev& h, A bits, int64_t chosen, double* value) {
    return h.kind == 1 ? real1d_offdiagonal(h, bits, chosen, value)
                       : mom1d_offdiagonal(h, bits, chosen, value);
}

// ---------------------------------------------------------------------------------------
// HubbardRealSpace (HubbardRealSpace.jl) on one or two components.  Component 1 of a
// two-component key is the high part (key >> cbits2), component 2 the low cbits2 bits.
// A BoseFS component is a run-length bit string as above; a FermiFS component holds site j
// at bit j-1 (bitstring.jl:713-733).
// ---------------------------------------------------------------------------------------
template <class A> __device__ inline A rs_comp(const ModelDev& h, A key, int c) {
    if (h.ncomp == 1) return key;
    return c == 0 ? shrA(key, h.cbits2) : (A)(key & low_maskA<A>(h.cbits2));
}
template <class A> __device__ inline A rs_with_comp(const ModelDev& h, A key, int c, A v) {
    if (h.ncomp == 1) return v;
    const A low = low_maskA<A>(h.cbits2);
    return c == 0 ? (A)(shlA(v, h.cbits2) | (key & low)) : (A)((key & ~low) | v);
}
__device__ inline bool rs_fermi(const ModelDev& h, int c) { return (h.fermi >> c) & 1; }

// num_occupied_modes of a component: N for a FermiFS (fermifs.jl:129), runs for a BoseFS.
template <class A> __device__ inline int rs_occupied(const ModelDev& h, A b, int c) {
    return rs_fermi(h, c) ? popcA(b) : popcA(run_starts(b));
}

// sum over sites of n_i * pot[i] in ascending site order (external_potential,
// HubbardRealSpace.jl:98-106)
template <class A> __device__ inline double rs_potential(const ModelDev& h, A b, int c, const double* pot) {
    double pe = 0.0;
    if (rs_fermi(h, c)) {
        A x = b;
        while (x) {
            const int p = ctzA(x);
            pe += 1 * pot[p];
            x &= x - 1;
        }
    } else {
        A x = b;
        int mode = 1;
        while (x) {
            const int z = ctzA(x);
            mode += z;
            x = shrA(x, z);
            const int n = ctzA((A)~x);
            pe += n * pot[mode - 1];
            x = shrA(x, n);
        }
    }
    return pe;
}

// dot(occupied_modes(a), occupied_modes(b)) = sum_i n_ai n_bi (fockaddress.jl:692-719,
// bitstring.jl:833-838), an integer
template <class A> __device__ inline int64_t rs_overlap(const ModelDev& h, A a, A b) {
    const bool fa = rs_fermi(h, 0), fb = rs_fermi(h, 1);
    if (fa && fb) return popcA((A)(a & b));
    int64_t acc = 0;
    if (fa) {  // FermiFS a: n_ai = 1 on its set bits
        A x = a;
        while (x) {
            const int p = ctzA(x);
            acc += occupation(b, p + 1);
            x &= x - 1;
        }
        return acc;
    }
    A x = a;
    int mode = 1;
    while (x) {
        const int z = ctzA(x);
        mode += z;
        x = shrA(x, z);
        const int n = ctzA((A)~x);
        const int m = fb ? (int)((b >> (mode - 1)) & 1) : occupation(b, mode);
        acc += (int64_t)n * m;
        x = shrA(x, n);
    }
    return acc;
}

// num_offdiagonals: num_occupied_modes * 2D per component, summed (HubbardRealSpace.jl:284-290,
// :346-349).
template <class A> __device__ inline int64_t rs_num_offdiagonals(const ModelDev& h, A key) {
    int64_t n = rs_occupied(h, rs_comp(h, key, 0), 0);
    if (h.ncomp == 2) n += rs_occupied(h, rs_comp(h, key, 1), 1);
    return n * h.ndir;
}

// diagonal_element (HubbardRealSpace.jl:264-279): local interactions (:18-25 and the
// recursion _interactions :57-73: u11 self term, u12 cross term, u22 self term, in that
// order of addition), then the trap potential.  A FermiFS has no self interaction.
template <class A> __device__ inline double rs_diagonal_element(const ModelDev& h, A key) {
    const A a = rs_comp(h, key, 0);
    double intr = 0.0;
    if (h.has_u) {
        const double self_a = rs_fermi(h, 0) ? 0.0 : h.u * (double)bose_hubbard_interaction(a) / 2;
        if (h.ncomp == 1) {
            intr = self_a;
        } else {
            const A b = rs_comp(h, key, 1);
            const double row = h.u12 * (double)rs_overlap(h, a, b);
            const double self_b = rs_fermi(h, 1) ? 0.0 : h.u22 * (double)bose_hubbard_interaction(b) / 2;
            intr = (self_a + row) + self_b;
        }
    }
    double pot = 0.0;
    if (h.has_pot) {
        if (h.ncomp == 1) {
            pot = rs_potential(h, a, 0, h.kes);
        } else {
            pot += rs_potential(h, a, 0, h.kes);
            pot += rs_potential(h, rs_comp(h, key, 1), 1, h.pot2);
        }
    }
    return intr + pot;
}

// get_offdiagonal (HubbardRealSpace.jl:303-316, :369-391): channels of component 1 first;
// within a component chosen = (particle - 1) * 2D + direction.  The particle-th occupied
// site hops to its neighbour in that direction: (key, 0.0) past a hard wall; a BoseFS
// component gives -t sqrt(n_src (n_dst + 1)) (bose_excitation, bitstring.jl:581-595,
// fockaddress.jl:559-567); a FermiFS component gives -t * (+-1) with the sign of the ones
// passed (fermi_excitation, bitstring.jl:773-791), or -t * 0.0 and the unchanged key when
// the target site is occupied.
template <class A>
__device__ inline A rs_offdiagonal(const ModelDev& h, A key, int64_t chosen, double* value) {
    int c = 0;
    A b = rs_comp(h, key, 0);
    int64_t ch = chosen;
    if (h.ncomp == 2) {
        const int64_t len0 = (int64_t)rs_occupied(h, b, 0) * h.ndir;
        if (ch > len0) {
            c = 1;
            ch -= len0;
            b = rs_comp(h, key, 1);
        }
    }
    const int particle = (int)((ch - 1) / h.ndir) + 1;
    const int dir = (int)((ch - 1) % h.ndir);
    const double t = c == 0 ? h.t : h.t2;
    int src_pos, src_mode, n_src = 1;
    if (rs_fermi(h, c)) {
        src_pos = select_bit(b, particle - 1);
        src_mode = src_pos + 1;
    } else {
        const A st = run_starts(b);
        src_pos = select_bit(st, particle - 1);
        n_src = ones_from(b, src_pos);
        src_mode = src_pos - popcA((A)(b & low_maskA<A>(src_pos))) + 1;  // zeros below, plus one
    }
    const int dst = h.nbr[src_mode - 1][dir];
    if (dst == 255) {
        *value = 0.0;
        return key;
    }
    A nb;
    double v;
    if (rs_fermi(h, c)) {
        const int before_src = popcA((A)(b & low_maskA<A>(src_pos)));
        const A removed = b ^ ((A)1 << src_pos);
        if ((removed >> dst) & 1) {
            *value = -t * 0.0;
            return key;
        }
        const int before_dst = popcA((A)(removed & low_maskA<A>(dst)));
        nb = removed | ((A)1 << dst);
        v = ((before_src + before_dst) & 1) ? -1.0 : 1.0;
    } else {
        const int n_dst = occupation(b, dst + 1);
        const A removed = delete_bit(b, src_pos);
        nb = insert_one(removed, mode_start(removed, dst + 1));
        v = sqrt((double)((int64_t)n_src * (n_dst + 1)));
    }
    *value = -t * v;
    return rs_with_comp(h, key, c, nb);
}

// Model dispatch with the family fixed at compile time: RS selects HubbardRealSpace
// (kind 3), otherwise the 1D models (kinds 1 and 2, told apart at run time).  The step
// kernels are instantiated per family, so the 1D models' register allocation does not
// carry the real-space code.
template <bool RS, class A> __device__ inline int64_t model_num_offdiagonals(const ModelDev& h, A bits) {
    if constexpr (RS) return rs_num_offdiagonals(h, bits);
    else return num_offdiagonals(h, bits);
}
template <bool RS, class A> __device__ inline double model_diagonal_element(const ModelDev& h, A bits) {
    if constexpr (RS) return rs_diagonal_element(h, bits);
    else return diagonal_element(h, bits);
}
template <bool RS, class A>
__device__ inline A model_offdiagonal(const ModelDev& h, A bits, int64_t chosen, double* value) {
    if constexpr (RS) return rs_offdiagonal(h, bits, chosen, value);
    else return get_offdiagonal(h, bits, chosen, value);
}

}  // namespace rimu
