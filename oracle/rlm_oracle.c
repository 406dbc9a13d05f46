/*
 * CPU oracle for the rlm/hlm hot path (test infrastructure only).
 *
 * This file restates, in plain C and in the reference's own order of
 * floating-point operations, the parts of qiongz/Random-Landau-Model's
 * double-precision CPU build (rlm_cpu_dp / hlm_cpu_dp) whose bits matter:
 *   - the MKL VSL MT19937 stream (vslNewStream(VSL_BRNG_MT19937, seed) is
 *     Matsumoto-Nishimura init_by_array({seed}); one 32-bit word per double,
 *     u = a + (b-a) * w * 2^-32; viRngUniformBits returns raw words;
 *     vslSkipAheadStream counts words),
 *   - prepare_coeff (with its single-precision expf and 1.0f/n_phi island),
 *   - get_potential_coeff,
 *   - the projected-Hamiltonian assembly loop,
 *   - the per-line FHS Chern reduction (peer_cal_Chern).
 * The Hermitian eigensolve is LAPACK zheevd, called from oracle.py.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this library, and only as the checker. The product path never links it.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* MT19937 (Matsumoto & Nishimura, mt19937ar.c, 2002 edition).         */
/* MKL's VSL_BRNG_MT19937 seeded with one 32-bit word s is             */
/* init_by_array({s}) of that algorithm.                               */
/* ------------------------------------------------------------------ */
#define MT_N 624
#define MT_M 397

typedef struct {
    uint32_t mt[MT_N];
    int mti;
} orc_mt_state;

static void mt_init_genrand(orc_mt_state *st, uint32_t s) {
    st->mt[0] = s;
    for (int i = 1; i < MT_N; i++)
        st->mt[i] = 1812433253u * (st->mt[i - 1] ^ (st->mt[i - 1] >> 30)) + (uint32_t)i;
    st->mti = MT_N;
}

void orc_mt_init_by_array(orc_mt_state *st, const uint32_t *key, int len) {
    mt_init_genrand(st, 19650218u);
    int i = 1, j = 0;
    int k = (MT_N > len ? MT_N : len);
    for (; k; k--) {
        st->mt[i] = (st->mt[i] ^ ((st->mt[i - 1] ^ (st->mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        i++;
        j++;
        if (i >= MT_N) { st->mt[0] = st->mt[MT_N - 1]; i = 1; }
        if (j >= len) j = 0;
    }
    for (k = MT_N - 1; k; k--) {
        st->mt[i] = (st->mt[i] ^ ((st->mt[i - 1] ^ (st->mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= MT_N) { st->mt[0] = st->mt[MT_N - 1]; i = 1; }
    }
    st->mt[0] = 0x80000000u;
    st->mti = MT_N;
}

uint32_t orc_mt_next(orc_mt_state *st) {
    static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
    uint32_t y;
    if (st->mti >= MT_N) {
        int kk;
        for (kk = 0; kk < MT_N - MT_M; kk++) {
            y = (st->mt[kk] & 0x80000000u) | (st->mt[kk + 1] & 0x7fffffffu);
            st->mt[kk] = st->mt[kk + MT_M] ^ (y >> 1) ^ mag01[y & 1u];
        }
        for (; kk < MT_N - 1; kk++) {
            y = (st->mt[kk] & 0x80000000u) | (st->mt[kk + 1] & 0x7fffffffu);
            st->mt[kk] = st->mt[kk + (MT_M - MT_N)] ^ (y >> 1) ^ mag01[y & 1u];
        }
        y = (st->mt[MT_N - 1] & 0x80000000u) | (st->mt[0] & 0x7fffffffu);
        st->mt[MT_N - 1] = st->mt[MT_M - 1] ^ (y >> 1) ^ mag01[y & 1u];
        st->mti = 0;
    }
    y = st->mt[st->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

size_t orc_mt_state_size(void) { return sizeof(orc_mt_state); }

/* Raw words of the stream seeded as vslNewStream(VSL_BRNG_MT19937, seed),
 * after skipping `skip` words. */
void orc_mt_words(uint32_t seed, uint64_t skip, uint64_t count, uint32_t *out) {
    orc_mt_state st;
    orc_mt_init_by_array(&st, &seed, 1);
    for (uint64_t i = 0; i < skip; i++) (void)orc_mt_next(&st);
    for (uint64_t i = 0; i < count; i++) out[i] = orc_mt_next(&st);
}

/* vdRngUniform(VSL_RNG_METHOD_UNIFORM_STD, a, b) on one word. */
static double mt_uniform(uint32_t w, double a, double b) {
    return a + (b - a) * ((double)w * 0x1p-32);
}

/* rlm/disorder_potential.cpp:3-21 generate_disorder_potential.
 * Stream (seed), skip impurity_num*3*sample words (:6), x ~ U[0,L1) (:9),
 * y ~ U[0,L2) (:10), intensity = +1 for even word, -1 for odd (:12-17). */
void orc_rlm_disorder(uint32_t seed, long sample, long n_imp, double L1, double L2,
                      double *x, double *y, double *s) {
    orc_mt_state st;
    orc_mt_init_by_array(&st, &seed, 1);
    uint64_t skip = (uint64_t)n_imp * 3u * (uint64_t)sample;
    for (uint64_t i = 0; i < skip; i++) (void)orc_mt_next(&st);
    for (long i = 0; i < n_imp; i++) x[i] = mt_uniform(orc_mt_next(&st), 0.0, L1);
    for (long i = 0; i < n_imp; i++) y[i] = mt_uniform(orc_mt_next(&st), 0.0, L2);
    for (long i = 0; i < n_imp; i++) s[i] = (orc_mt_next(&st) % 2u == 0u) ? 1.0 : -1.0;
}

/* hlm/main_hlm_cpu.cpp:66-67,79: one stream for the run; sample n draws
 * n_sites values of U[-1,1) in sequence. out has n_samples*n_sites values. */
void orc_hlm_disorder(uint32_t seed, long n_sites, long n_samples, double *out) {
    orc_mt_state st;
    orc_mt_init_by_array(&st, &seed, 1);
    for (long i = 0; i < n_sites * n_samples; i++) out[i] = mt_uniform(orc_mt_next(&st), -1.0, 1.0);
}

/* rlm/matrix_coefficients.cpp:3-25 prepare_coeff (DP build).
 * Complex arrays are interleaved (re, im) doubles. */
void orc_prepare_coeff(int n_phi, int off, int d, int n_mesh, double L1, double L2,
                       double *coeff_mn, double *coeff_m_theta, double *coeff_jm) {
    for (int m = 0; m < d; m++)
        for (int n = 0; n < d; n++) {
            double ph = M_PI * (m - off) * (n - off) / n_phi;
            double sn, cs;
            sincos(ph, &sn, &cs);              /* glibc sincos: one call for cos and sin */
            double cr = cs, ci = -sn;
            double arg = -M_PI * ((m - off) * (m - off) * L2 / L1 + (n - off) * (n - off) * L1 / L2) / 2.0 / n_phi;
            float e = expf((float)arg);        /* single-precision exp (:11) */
            float r = 1.0f / (float)n_phi;     /* single-precision 1/n_phi (:11) */
            coeff_mn[2 * (m * d + n)] = (cr * (double)e) * (double)r;
            coeff_mn[2 * (m * d + n) + 1] = (ci * (double)e) * (double)r;
        }
    for (int t = 0; t <= n_mesh; t++) {
        double dtheta = t * (M_PI * 2.0) / n_mesh;
        for (int m = 0; m < d; m++) {
            double a = (m - off) * dtheta / n_phi, sn, cs;
            sincos(a, &sn, &cs);
            coeff_m_theta[2 * (t * d + m)] = cs;
            coeff_m_theta[2 * (t * d + m) + 1] = sn;
        }
    }
    for (int j = 0; j < n_phi; j++)
        for (int m = 0; m < d; m++) {
            double a = (M_PI * 2.0) * (m - off) * (j + 1) / n_phi, sn, cs;
            sincos(a, &sn, &cs);
            coeff_jm[2 * (j * d + m)] = cs;
            coeff_jm[2 * (j * d + m) + 1] = -sn;
        }
}

/* rlm/matrix_coefficients.cpp:27-44 get_potential_coeff: sequential sum over
 * impurities of e^{-i phase} * s_i, then times coeff_mn. */
void orc_potential(const double *x, const double *y, const double *s, long n_imp,
                   const double *coeff_mn, int d, int off, double *V) {
    for (int m = 0; m < d; m++) {
        double kx = (M_PI * 2.0) * (m - off);
        for (int n = 0; n < d; n++) {
            double ky = (M_PI * 2.0) * (n - off);
            double complex acc = 0.0;
            for (long i = 0; i < n_imp; i++) {
                double phase = kx * x[i] + ky * y[i];
                acc = acc + CMPLX(cos(phase) * s[i], -sin(phase) * s[i]);
            }
            double complex c = CMPLX(coeff_mn[2 * (m * d + n)], coeff_mn[2 * (m * d + n) + 1]);
            double complex v = CMPLX(creal(acc) * creal(c) - cimag(acc) * cimag(c),
                                     creal(acc) * cimag(c) + cimag(acc) * creal(c));
            V[2 * (m * d + n)] = creal(v);
            V[2 * (m * d + n) + 1] = cimag(v);
        }
    }
}

static inline double complex cmul(double complex a, double complex b) {
    return CMPLX(creal(a) * creal(b) - cimag(a) * cimag(b), creal(a) * cimag(b) + cimag(a) * creal(b));
}

/* rlm/hamiltonian_rlm_cpu.cpp:14-48 (peer_solve_projected, assembly part) for
 * one twist point (theta_1, theta_2). Writes the n_phi*n_phi buffer exactly as
 * the reference does: [j*n_phi+k] = h_jk for k >= j and [k*n_phi+j] =
 * conj(h_jk). LAPACK then reads that buffer column-major. */
void orc_assemble(int n_phi, int d, int off, int theta_1, int theta_2, const double *V,
                  const double *coeff_m_theta, const double *coeff_jm, double *wf) {
    const double complex *v = (const double complex *)V;
    const double complex *cmt = (const double complex *)coeff_m_theta;
    const double complex *cjm = (const double complex *)coeff_jm;
    double complex *w = (double complex *)wf;
    int dim_n = d, dim_m = d;
    long idx_kx = (long)theta_1 * dim_n;
    for (int j = 0; j < n_phi; j++)
        for (int k = j; k < n_phi; k++) {
            double complex h = 0.0;
            for (int m = 0; m < dim_m; m++) {
                long idx_m = (long)m * dim_n;
                double complex ctj = cmul(conj(cmt[(long)theta_2 * dim_m + m]), cjm[(long)j * dim_m + m]);
                double complex sum = 0.0;
                if (k - j <= off) {
                    int n = off + k - j;
                    sum = sum + cmul(v[idx_m + n], cmt[idx_kx + n]);
                }
                if ((n_phi + j - k) <= off) {
                    int n = off + k - j - n_phi;
                    sum = sum + cmul(v[idx_m + n], cmt[idx_kx + n]);
                }
                h = h + cmul(sum, ctj);
            }
            w[(long)j * n_phi + k] = h;
            w[(long)k * n_phi + j] = conj(h);
        }
}

/* diag_wrappers/chern.cpp:37-75 (peer_cal_Chern) for all theta_2 of one
 * theta_1 line. wfs is the 2-slot ring [slot][theta_2 0..n_mesh][state][dim_vec]
 * (interleaved complex). Output chern_theta[m + n*n_mesh] = -arg(prod)/2pi. */
void orc_chern_line(const double *wfs, int n_phi, int dim_vec, int n_mesh, int theta_1,
                    double *chern_theta) {
    const double complex *w = (const double complex *)wfs;
    long dki[4], dkj[4];
    dki[2] = (theta_1 % 2 == 0 ? 1 : -1);
    dki[0] = -dki[2];
    dki[1] = dki[3] = dkj[0] = dkj[2] = 0;
    dkj[3] = -1;
    dkj[1] = 1;
    unsigned long dim_wfs = (unsigned long)n_phi * dim_vec;
    long n_mesh_pbc = n_mesh + 1;
    double complex *prod = (double complex *)malloc(sizeof(double complex) * n_phi);
    for (long m = 0; m < n_mesh; m++) {
        long lki = 1 - theta_1 % 2, lkj = m;
        for (int n = 0; n < n_phi; n++) prod[n] = 1.0;
        for (int l = 0; l < 4; l++) {
            long rki = lki + dki[l], rkj = lkj + dkj[l];
            for (int n = 0; n < n_phi; n++) {
                const double complex *a = w + (lki * n_mesh_pbc + lkj) * dim_wfs + (unsigned long)dim_vec * n;
                const double complex *b = w + (rki * n_mesh_pbc + rkj) * dim_wfs + (unsigned long)dim_vec * n;
                double complex sum = 0.0;
                for (int s = 0; s < dim_vec; s++) sum = sum + cmul(conj(a[s]), b[s]);
                prod[n] = cmul(prod[n], sum);
            }
            lki = rki;
            lkj = rkj;
        }
        for (int n = 0; n < n_phi; n++) chern_theta[m + (long)n * n_mesh] = -carg(prod[n]) / (M_PI * 2.0);
    }
    free(prod);
}

/* float expf as the DP build calls it, exposed so tests can compare tables. */
float orc_expf(float x) { return expf(x); }
/* libm's single-precision cosf/sinf: coeff_m_theta of the reference's FP32 builds takes
 * cos/sin of a float argument (rlm/matrix_coefficients.cpp:17 with dtype float). */
float orc_cosf(float x) { return cosf(x); }
float orc_sinf(float x) { return sinf(x); }
