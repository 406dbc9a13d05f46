// Host-side tables and disorder streams of the rlm/hlm drivers.
//
// MT19937: Matsumoto & Nishimura's generator. Intel MKL's
// vslNewStream(&s, VSL_BRNG_MT19937, seed), used by the reference
// (rlm/disorder_potential.cpp:5, hlm/main_hlm_cpu.cpp:67), seeds it with
// init_by_array({seed mod 2^32}); vdRngUniform(UNIFORM_STD, a, b) maps one
// 32-bit word w to a + (b-a) * w * 2^-32; viRngUniformBits returns the raw
// words; vslSkipAheadStream(n) skips n words (here by the jump-ahead of
// mt_jump.hpp, whose cost grows with log n).
#pragma once
#include <cmath>
#include <cstdint>

#include "mt_jump.hpp"

namespace rlm {

// vdRngUniform(VSL_RNG_METHOD_UNIFORM_STD, a, b) of one word
inline double mt_uniform(mtj::WindowGen &g, double a, double b) { return a + (b - a) * ((double)g.next() * 0x1p-32); }

// rlm/disorder_potential.cpp:3-21 (samples sample .. sample + count - 1 of the seed
// chain; sample k + 1 starts where sample k ends, 3 * n_imp words later, so one
// generator serves them all)
inline void host_rlm_disorder(double *x, double *y, double *s, long n_imp, double L1, double L2, uint32_t seed,
                              long sample, long count = 1) {
    uint32_t w[mtj::kN];
    mtj::seed_window(seed, (uint64_t)n_imp * 3u * (uint64_t)sample, w);
    mtj::WindowGen g(w);
    for (long k = 0; k < count; k++) {
        double *xk = x + k * n_imp, *yk = y + k * n_imp, *sk = s + k * n_imp;
        for (long i = 0; i < n_imp; i++) xk[i] = mt_uniform(g, 0.0, L1);
        for (long i = 0; i < n_imp; i++) yk[i] = mt_uniform(g, 0.0, L2);
        for (long i = 0; i < n_imp; i++) sk[i] = (g.next() % 2u == 0u) ? 1.0 : -1.0;
    }
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

// rlm/matrix_coefficients.cpp:3-25 as the reference's single-precision GPU builds
// compile it (dtype float: complex<float> tables; the double cos/sin of coeff_mn and
// coeff_jm rounded to float; coeff_m_theta from float cos/sin of a float argument;
// float L1, L2). Values written as doubles (exactly the floats).
inline void host_prepare_coeff_f32(double *coeff_mn, double *coeff_m_theta, double *coeff_jm, int n_phi, int off,
                                   int d, int n_mesh, float L1, float L2) {
    const double PI2 = M_PI * 2.0;
    for (int m = 0; m < d; m++)
        for (int n = 0; n < d; n++) {
            const double ph = M_PI * (m - off) * (n - off) / n_phi;
            double sn, cs;
            ::sincos(ph, &sn, &cs);
            // (m-o)^2 L2 / L1 + (n-o)^2 L1 / L2 in float, times -pi, / 2.0 / n_phi in double
            const float q1 = (float)((m - off) * (m - off)) * L2 / L1;
            const float q2 = (float)((n - off) * (n - off)) * L1 / L2;
            const double arg = -M_PI * (double)(q1 + q2) / 2.0 / n_phi;
            const float e = ::expf((float)arg);
            const float r = 1.0f / (float)n_phi;
            const float cr = (float)cs, ci = (float)(-sn);
            coeff_mn[2 * (m * d + n)] = (double)((cr * e) * r);
            coeff_mn[2 * (m * d + n) + 1] = (double)((ci * e) * r);
        }
    for (int t = 0; t <= n_mesh; t++) {
        const float dtheta = (float)(t * PI2 / n_mesh);
        for (int m = 0; m < d; m++) {
            const float a = (float)(m - off) * dtheta / (float)n_phi;
            coeff_m_theta[2 * (t * d + m)] = (double)::cosf(a);
            coeff_m_theta[2 * (t * d + m) + 1] = (double)::sinf(a);
        }
    }
    for (int j = 0; j < n_phi; j++)
        for (int m = 0; m < d; m++) {
            const double a = PI2 * (m - off) * (j + 1) / n_phi;
            double sn, cs;
            ::sincos(a, &sn, &cs);
            coeff_jm[2 * (j * d + m)] = (double)(float)cs;
            coeff_jm[2 * (j * d + m) + 1] = (double)(float)(-sn);
        }
}

}  // namespace rlm
