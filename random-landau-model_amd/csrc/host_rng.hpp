// Host-side tables and disorder streams of the rlm/hlm drivers.
//
// MT19937: Matsumoto & Nishimura's generator. Intel MKL's
// vslNewStream(&s, VSL_BRNG_MT19937, seed), used by the reference
// (rlm/disorder_potential.cpp:5, hlm/main_hlm_cpu.cpp:67), seeds it with
// init_by_array({seed mod 2^32}); vdRngUniform(UNIFORM_STD, a, b) maps one
// 32-bit word w to a + (b-a) * w * 2^-32; viRngUniformBits returns the raw
// words; vslSkipAheadStream(n) skips n words.
#pragma once
#include <cmath>
#include <cstdint>

namespace rlm {

class Mt19937 {
  public:
    explicit Mt19937(uint32_t seed) { init_by_array(&seed, 1); }

    uint32_t next() {
        if (mti_ >= kN) twist();
        uint32_t y = mt_[mti_++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= (y >> 18);
        return y;
    }
    void discard(uint64_t n) {
        // whole blocks of 624 words are skipped by regenerating the state
        while (n > 0) {
            if (mti_ >= kN) twist();
            const uint64_t avail = (uint64_t)(kN - mti_);
            const uint64_t take = n < avail ? n : avail;
            mti_ += (int)take;
            n -= take;
        }
    }
    double uniform(double a, double b) { return a + (b - a) * ((double)next() * 0x1p-32); }

  private:
    static constexpr int kN = 624, kM = 397;
    uint32_t mt_[kN];
    int mti_ = kN + 1;

    void init_genrand(uint32_t s) {
        mt_[0] = s;
        for (int i = 1; i < kN; i++) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
        mti_ = kN;
    }
    void init_by_array(const uint32_t *key, int len) {
        init_genrand(19650218u);
        int i = 1, j = 0;
        for (int k = (kN > len ? kN : len); k; k--) {
            mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
            i++;
            j++;
            if (i >= kN) { mt_[0] = mt_[kN - 1]; i = 1; }
            if (j >= len) j = 0;
        }
        for (int k = kN - 1; k; k--) {
            mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
            i++;
            if (i >= kN) { mt_[0] = mt_[kN - 1]; i = 1; }
        }
        mt_[0] = 0x80000000u;
        mti_ = kN;
    }
    void twist() {
        static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
        int kk;
        uint32_t y;
        for (kk = 0; kk < kN - kM; kk++) {
            y = (mt_[kk] & 0x80000000u) | (mt_[kk + 1] & 0x7fffffffu);
            mt_[kk] = mt_[kk + kM] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < kN - 1; kk++) {
            y = (mt_[kk] & 0x80000000u) | (mt_[kk + 1] & 0x7fffffffu);
            mt_[kk] = mt_[kk + (kM - kN)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (mt_[kN - 1] & 0x80000000u) | (mt_[0] & 0x7fffffffu);
        mt_[kN - 1] = mt_[kM - 1] ^ (y >> 1) ^ mag01[y & 1u];
        mti_ = 0;
    }
};

// rlm/disorder_potential.cpp:3-21 (one sample of the seed chain).
inline void host_rlm_disorder(double *x, double *y, double *s, long n_imp, double L1, double L2, uint32_t seed,
                              long sample) {
    Mt19937 g(seed);
    g.discard((uint64_t)n_imp * 3u * (uint64_t)sample);
    for (long i = 0; i < n_imp; i++) x[i] = g.uniform(0.0, L1);
    for (long i = 0; i < n_imp; i++) y[i] = g.uniform(0.0, L2);
    for (long i = 0; i < n_imp; i++) s[i] = (g.next() % 2u == 0u) ? 1.0 : -1.0;
}

// rlm/matrix_coefficients.cpp:3-25 prepare_coeff (double-precision build), with
// the single-precision expf and 1.0f/n_phi of :11 kept as the reference has them.
inline void host_prepare_coeff(double *coeff_mn, double *coeff_m_theta, double *coeff_jm, int n_phi, int off, int d,
                               int n_mesh, double L1, double L2) {
    const double PI2 = M_PI * 2.0;
    for (int m = 0; m < d; m++)
        for (int n = 0; n < d; n++) {
            const double ph = M_PI * (m - off) * (n - off) / n_phi;
            double sn, cs;
            ::sincos(ph, &sn, &cs);   // glibc sincos, so the tables do not depend on compiler merging
            const double cr = cs, ci = -sn;
            const double arg = -M_PI * ((m - off) * (m - off) * L2 / L1 + (n - off) * (n - off) * L1 / L2) / 2.0 / n_phi;
            const float e = ::expf((float)arg);
            const float r = 1.0f / (float)n_phi;
            coeff_mn[2 * (m * d + n)] = (cr * (double)e) * (double)r;
            coeff_mn[2 * (m * d + n) + 1] = (ci * (double)e) * (double)r;
        }
    for (int t = 0; t <= n_mesh; t++) {
        const double dtheta = t * PI2 / n_mesh;
        for (int m = 0; m < d; m++) {
            const double a = (m - off) * dtheta / n_phi;
            double sn, cs;
            ::sincos(a, &sn, &cs);
            coeff_m_theta[2 * (t * d + m)] = cs;
            coeff_m_theta[2 * (t * d + m) + 1] = sn;
        }
    }
    for (int j = 0; j < n_phi; j++)
        for (int m = 0; m < d; m++) {
            const double a = PI2 * (m - off) * (j + 1) / n_phi;
            double sn, cs;
            ::sincos(a, &sn, &cs);
            coeff_jm[2 * (j * d + m)] = cs;
            coeff_jm[2 * (j * d + m) + 1] = -sn;
        }
}

}  // namespace rlm
