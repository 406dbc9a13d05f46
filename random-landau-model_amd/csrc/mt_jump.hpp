// MT19937 jump-ahead by polynomial arithmetic over F2 (Haramoto, Matsumoto,
// Nishimura, Panneton, L'Ecuyer, "Efficient jump ahead for F2-linear random number
// generators", INFORMS J. Computing 20 (2008)).
//
// The generator's state is the window W_k = (x_k, ..., x_{k+623}) of its word
// sequence x_{k+624} = x_{k+397} ^ twist(upper(x_k) | lower(x_{k+1})); only the top
// bit of x_k matters, so the state space has dimension 19937. The transition T
// (W_k -> W_{k+1}) has minimal polynomial phi of degree 19937 (found here once by
// Berlekamp-Massey on one output bit). Skipping J words is W_{k+J} = g(T) W_k with
// g = t^J mod phi, and g(T) W_k = XOR over i with g_i = 1 of W_{k+i}: word j of the
// jumped window is XOR_i g_i x_{k+i+j}, a correlation of g's bits with the next
// 19937 + 623 words of the sequence. The cost does not depend on J beyond the
// O(log J) squarings of t^J mod phi.
//
// vslSkipAheadStream(stream, J) of Intel MKL's VSL_BRNG_MT19937, which the reference
// calls per sample (rlm/disorder_potential.cpp:6), skips J outputs of this sequence:
// after init_by_array the state window is W_0 = mt[0..623] and output j is the
// tempered x_{624+j}, so skipping J outputs starts the generator from W_J.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

namespace rlm {
namespace mtj {

constexpr int kN = 624, kM = 397;
constexpr int kDeg = 19937;                   // degree of phi
constexpr int kPolyWords = (kDeg + 64) / 64;  // 312 words hold coefficients t^0 .. t^19967
using Poly = std::vector<uint64_t>;           // kPolyWords words, bit i = coefficient of t^i

__host__ __device__ inline uint32_t next_word(uint32_t xk, uint32_t xk1, uint32_t xkm) {
    const uint32_t y = (xk & 0x80000000u) | (xk1 & 0x7fffffffu);
    return xkm ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

__host__ __device__ inline uint32_t temper(uint32_t y) {
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

// init_by_array({seed}) of mt19937ar: the state window W_0 (the generator twists it
// before the first output)
__host__ __device__ inline void init_by_seed(uint32_t seed, uint32_t *mt) {
    mt[0] = 19650218u;
    for (int i = 1; i < kN; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    int i = 1;
    for (int k = kN; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + seed;
        i++;
        if (i >= kN) { mt[0] = mt[kN - 1]; i = 1; }
    }
    for (int k = kN - 1; k; k--) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= kN) { mt[0] = mt[kN - 1]; i = 1; }
    }
    mt[0] = 0x80000000u;
}

inline bool bit(const uint64_t *p, long i) { return (p[i >> 6] >> (i & 63)) & 1u; }
inline void flip(uint64_t *p, long i) { p[i >> 6] ^= (uint64_t)1 << (i & 63); }

// dst ^= src << s (bit shift; dst holds at least nw + s/64 + 1 words)
inline void xor_shifted(uint64_t *dst, const uint64_t *src, int nw, long s) {
    const long ws = s >> 6;
    const int bs = (int)(s & 63);
    if (bs == 0) {
        for (int k = 0; k < nw; k++) dst[ws + k] ^= src[k];
        return;
    }
    uint64_t carry = 0;
    for (int k = 0; k < nw; k++) {
        dst[ws + k] ^= (src[k] << bs) | carry;
        carry = src[k] >> (64 - bs);
    }
    dst[ws + nw] ^= carry;
}

// phi from Berlekamp-Massey on bit 0 of 2*19937 + 64 consecutive words of the
// sequence of an arbitrary nonzero state: phi is irreducible (the period is the
// Mersenne prime 2^19937 - 1), so every nonzero sequence has minimal polynomial phi.
inline Poly compute_charpoly() {
    const long NS = 2L * kDeg + 64;
    std::vector<uint8_t> s(NS);
    {
        std::vector<uint32_t> x(NS + kN);
        init_by_seed(5489u, x.data());
        for (long k = 0; k < NS; k++) x[k + kN] = next_word(x[k], x[k + 1], x[k + kM]);
        for (long k = 0; k < NS; k++) s[k] = x[k + kN] & 1u;
    }
    // C, B hold at most degree-19937 polynomials (deg(B t^m) <= the new L in every
    // update); the extra words keep every shifted write in bounds
    const int W = 2 * kPolyWords + 8;
    std::vector<uint64_t> C(W, 0), B(W, 0), T(W, 0);
    // reversed sequence: R bit (NS - 1 - k) = s_k, so sum_{i=0..L} C_i s_{n-i} is the
    // parity of C AND (R >> (NS - 1 - n))
    const long RW = (NS + 63) / 64 + 3;
    std::vector<uint64_t> R(RW, 0);
    for (long k = 0; k < NS; k++)
        if (s[k]) R[(NS - 1 - k) >> 6] |= (uint64_t)1 << ((NS - 1 - k) & 63);
    C[0] = B[0] = 1;
    long L = 0, LB = 0, m = 1;
    for (long n = 0; n < NS; n++) {
        const long off = NS - 1 - n;
        const long ow = off >> 6;
        const int ob = (int)(off & 63);
        uint64_t acc = 0;
        const long nw = (L >> 6) + 1;
        for (long k = 0; k < nw && ow + k < RW; k++) {
            uint64_t r = R[ow + k] >> ob;
            if (ob && ow + k + 1 < RW) r |= R[ow + k + 1] << (64 - ob);
            acc ^= C[k] & r;
        }
        const int d = __builtin_popcountll(acc) & 1;
        if (!d) {
            m++;
            continue;
        }
        const int bw = (int)(LB >> 6) + 1;
        if ((m >> 6) + bw + 1 >= W) return Poly();
        if (2 * L <= n) {
            T = C;
            xor_shifted(C.data(), B.data(), bw, m);
            LB = L;
            L = n + 1 - L;
            B = T;
            m = 1;
        } else {
            xor_shifted(C.data(), B.data(), bw, m);
            m++;
        }
        if (L > kDeg) return Poly();
    }
    // phi(t) = t^L C(1/t)
    Poly phi(kPolyWords, 0);
    if (L != kDeg) return Poly();
    for (long i = 0; i <= L; i++)
        if (bit(C.data(), i)) flip(phi.data(), L - i);
    return phi;
}

inline const Poly &charpoly() {
    static Poly phi;
    static std::once_flag once;
    std::call_once(once, [] { phi = compute_charpoly(); });
    return phi;
}

// p (bits 0 .. top) reduced mod phi in place; p has at least top/64 + 2 words
inline void reduce(uint64_t *p, long top) {
    const Poly &phi = charpoly();
    for (long i = top; i >= kDeg; i--)
        if (bit(p, i)) xor_shifted(p, phi.data(), kPolyWords, i - kDeg);
}

inline Poly sqrmod(const Poly &a) {
    std::vector<uint64_t> p(2 * kPolyWords + 2, 0);
    for (int k = 0; k < kPolyWords; k++) {
        const uint64_t w = a[k];
        uint64_t lo = 0, hi = 0;
        for (int b = 0; b < 32; b++) {
            lo |= ((w >> b) & 1u) << (2 * b);
            hi |= ((w >> (b + 32)) & 1u) << (2 * b);
        }
        p[2 * k] = lo;
        p[2 * k + 1] = hi;
    }
    reduce(p.data(), 2L * (kDeg - 1));
    return Poly(p.begin(), p.begin() + kPolyWords);
}

inline Poly mulmod(const Poly &a, const Poly &b) {
    std::vector<uint64_t> p(2 * kPolyWords + 2, 0);
    for (long i = 0; i < kDeg; i++)
        if (bit(a.data(), i)) xor_shifted(p.data(), b.data(), kPolyWords, i);
    reduce(p.data(), 2L * (kDeg - 1));
    return Poly(p.begin(), p.begin() + kPolyWords);
}

// t^k mod phi (left-to-right binary powering)
inline Poly powmod_t(uint64_t k) {
    Poly g(kPolyWords, 0);
    g[0] = 1;
    if (k < (uint64_t)kDeg) {
        g[0] = 0;
        flip(g.data(), (long)k);
        return g;
    }
    int top = 63;
    while (!((k >> top) & 1u)) top--;
    for (int b = top; b >= 0; b--) {
        g = sqrmod(g);
        if ((k >> b) & 1u) {
            // g *= t
            uint64_t carry = 0;
            for (int w = 0; w < kPolyWords; w++) {
                const uint64_t nc = g[w] >> 63;
                g[w] = (g[w] << 1) | carry;
                carry = nc;
            }
            if (bit(g.data(), kDeg)) {
                const Poly &phi = charpoly();
                for (int w = 0; w < kPolyWords; w++) g[w] ^= phi[w];
            }
        }
    }
    return g;
}

// out = g(T) w: the window reached from w after skipping the words that g = t^J mod
// phi encodes (out may alias w)
inline void jump_window(const uint32_t *w, const Poly &g, uint32_t *out) {
    std::vector<uint32_t> x(kDeg + kN);
    memcpy(x.data(), w, sizeof(uint32_t) * kN);
    for (long k = 0; k + kN < (long)x.size(); k++) x[k + kN] = next_word(x[k], x[k + 1], x[k + kM]);
    uint32_t acc[kN];
    memset(acc, 0, sizeof(acc));
    for (long i = 0; i < kDeg; i++)
        if (bit(g.data(), i))
            for (int j = 0; j < kN; j++) acc[j] ^= x[i + j];
    memcpy(out, acc, sizeof(acc));
}

// The window after skipping `skip` outputs from seed's initial window.
inline void seed_window(uint32_t seed, uint64_t skip, uint32_t *out) {
    init_by_seed(seed, out);
    if (skip == 0) return;
    if (skip < 4u * kDeg) {
        // short skips: step the window (cheaper than the correlation)
        std::vector<uint32_t> x(skip + kN);
        memcpy(x.data(), out, sizeof(uint32_t) * kN);
        for (uint64_t k = 0; k < skip; k++) x[k + kN] = next_word(x[k], x[k + 1], x[k + kM]);
        memcpy(out, x.data() + skip, sizeof(uint32_t) * kN);
        return;
    }
    jump_window(out, powmod_t(skip), out);
}

// Generator over a window: the outputs following it, 624 at a time.
class WindowGen {
  public:
    explicit WindowGen(const uint32_t *w) { memcpy(mt_, w, sizeof(mt_)); }
    uint32_t next() {
        if (i_ >= kN) {
            for (int k = 0; k < kN; k++) mt_[k] = next_word(mt_[k], mt_[(k + 1) % kN], mt_[(k + kM) % kN]);
            i_ = 0;
        }
        return temper(mt_[i_++]);
    }

  private:
    uint32_t mt_[kN];
    int i_ = kN;
};

}  // namespace mtj
}  // namespace rlm
