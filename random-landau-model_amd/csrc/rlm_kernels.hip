// rlm model kernels: disorder Fourier components, per-line band tables and the
// batched projected-Hamiltonian assembly (FP64).
//
// Reference: rlm/matrix_coefficients.cpp:27-44 (get_potential_coeff),
// rlm/hamiltonian_rlm_cpu.cpp:14-48 (assembly loop), and their FP32 GPU twins
// rlm/potential_coeff.cu:3-26 and rlm/hamiltonian_rlm_gpu.cu:9-50, which this
// file replaces (FP64, batched over samples and the whole twist mesh).
//
// The assembly uses the factorisation of the reference loop
//   h_jk = sum_m E[theta_2][m][j] * W[theta_1][m][k-j],
//   E[theta_2][m][j] = conj(coeff_m_theta[theta_2][m]) * coeff_jm[j][m]      (coeff_theta_jm)
//   W[theta_1][m][D] = [D<=o] V[m][o+D] c1[o+D] + [D>=N-o] V[m][o+D-N] c1[o+D-N]  (`sum`)
// so E is built once per run, W once per (sample, theta_1), and every
// k-point only runs the d-term dot products of its band entries.
#include "common.hpp"
#include "rlm_kernels.hpp"

namespace rlm {

// V[s][m][n] = coeff_mn[m][n] * sum_i s_i e^{-i 2pi((m-o) x_i + (n-o) y_i)}
// grid (ceil(d/64), d, S), block 256 = 64 n-values x 4 impurity slices.
__global__ __launch_bounds__(256) void potential_kernel(const double *__restrict__ X, const double *__restrict__ Y,
                                                        const double *__restrict__ Sg, long n_imp,
                                                        const double *__restrict__ cmn, int d, int off,
                                                        double *__restrict__ V) {
    __shared__ double red[2 * 256];
    const int s = blockIdx.z, m = blockIdx.y;
    const int nl = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + nl;
    const double PI2 = 6.283185307179586232;   // (M_PI*2.0) as a double
    const double kx = PI2 * (double)(m - off);
    const double ky = PI2 * (double)(n - off);
    const double *x = X + s * n_imp, *y = Y + s * n_imp, *sg = Sg + s * n_imp;
    double re = 0.0, im = 0.0;
    if (n < d) {
        for (long i = slice; i < n_imp; i += 4) {
            const double phase = __dadd_rn(__dmul_rn(kx, x[i]), __dmul_rn(ky, y[i]));
            double sn, cs;
            sincos(phase, &sn, &cs);
            re = fma(cs, sg[i], re);
            im = fma(-sn, sg[i], im);
        }
    }
    red[2 * threadIdx.x] = re;
    red[2 * threadIdx.x + 1] = im;
    __syncthreads();
    if (slice == 0 && n < d) {
        for (int k = 1; k < 4; k++) {
            re += red[2 * (threadIdx.x + 64 * k)];
            im += red[2 * (threadIdx.x + 64 * k) + 1];
        }
        const cd c = ld(cmn, (long)m * d + n);
        st(V, (long)s * d * d + (long)m * d + n, cmul(cd{re, im}, c));
    }
}

// E[t][m][j] = conj(cmt[t][m]) * cjm[j][m]; grid ((N+255)/256, d, g+1)
__global__ void etable_kernel(const double *__restrict__ cmt, const double *__restrict__ cjm, int n_phi, int d,
                              double *__restrict__ E) {
    const int t = blockIdx.z, m = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_phi) return;
    const cd a = ld(cmt, (long)t * d + m), b = ld(cjm, (long)j * d + m);
    st(E, ((long)t * d + m) * n_phi + j, cmul(cconj(a), b));
}

// W[s][t1][m][D], D in [0,N); grid ((N+255)/256, d, S*(g+1)) with z = s*(g+1)+t1
__global__ void wtable_kernel(const double *__restrict__ V, const double *__restrict__ cmt, int n_phi, int d, int off,
                              int n_mesh, double *__restrict__ Wt) {
    const int z = blockIdx.z, m = blockIdx.y;
    const int s = z / (n_mesh + 1), t1 = z % (n_mesh + 1);
    const int D = blockIdx.x * blockDim.x + threadIdx.x;
    if (D >= n_phi) return;
    const double *Vs = V + 2L * s * d * d;
    cd sum{0.0, 0.0};
    if (D <= off) {
        const int nn = off + D;
        sum = cadd(sum, cmul(ld(Vs, (long)m * d + nn), ld(cmt, (long)t1 * d + nn)));
    }
    if (n_phi - D <= off) {
        const int nn = off + D - n_phi;
        sum = cadd(sum, cmul(ld(Vs, (long)m * d + nn), ld(cmt, (long)t1 * d + nn)));
    }
    st(Wt, ((long)z * d + m) * n_phi + D, sum);
}

// Assembly: one workgroup per k-point. kp -> (sample, theta_1, theta_2) through
// kp0 + blockIdx.x over the full (g+1)x(g+1) mesh of each sample.
// A(r,c) = conj(h_{r,c}) for c >= r (D = c-r), h_{c,r} for c < r (D = r-c);
// zero outside the band D <= o || D >= N-o. Column-major, lda = N.
__global__ __launch_bounds__(256) void assemble_kernel(const double *__restrict__ E, const double *__restrict__ Wt,
                                                       int n_phi, int d, int off, int n_mesh, long kp0,
                                                       double *__restrict__ Aout, long strideA,
                                                       float *__restrict__ A32) {
    const long kp = kp0 + blockIdx.x;
    const int g1 = n_mesh + 1;
    const long s = kp / ((long)g1 * g1);
    const int rem = (int)(kp % ((long)g1 * g1));
    const int t1 = rem / g1, t2 = rem % g1;
    const double *Et = E + 2L * t2 * d * n_phi;
    const double *Wl = Wt + 2L * (s * g1 + t1) * d * n_phi;
    double *A = Aout + 2L * blockIdx.x * strideA;
    float *Af = A32 ? A32 + 2L * blockIdx.x * strideA : nullptr;
    const int N = n_phi;
    for (int c = 0; c < N; c++) {
        for (int r = threadIdx.x; r < N; r += blockDim.x) {
            int D, j;
            bool upper = (c >= r);
            if (upper) { D = c - r; j = r; } else { D = r - c; j = c; }
            cd h{0.0, 0.0};
            if (D <= off || D >= N - off) {
                for (int m = 0; m < d; m++) {
                    const cd w = ld(Wl, (long)m * N + D);
                    const cd e = ld(Et, (long)m * N + j);
                    cfma(h, w, e);
                }
            }
            cd a = upper ? cconj(h) : h;
            if (r == c) a.y = 0.0;
            st(A, (long)c * N + r, a);
            if (Af) stx(Af, (long)c * N + r, cx<float>{(float)a.x, (float)a.y});
        }
    }
}

int launch_potential(const double *X, const double *Y, const double *S, long n_imp, const double *cmn, int d, int off,
                     double *V, int n_samples, hipStream_t stream) {
    dim3 grid((d + 63) / 64, d, n_samples);
    hipLaunchKernelGGL(potential_kernel, grid, dim3(256), 0, stream, X, Y, S, n_imp, cmn, d, off, V);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_etable(const double *cmt, const double *cjm, int n_phi, int d, int n_mesh, double *E, hipStream_t stream) {
    dim3 grid((n_phi + 255) / 256, d, n_mesh + 1);
    hipLaunchKernelGGL(etable_kernel, grid, dim3(256), 0, stream, cmt, cjm, n_phi, d, E);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_wtable(const double *V, const double *cmt, int n_phi, int d, int off, int n_mesh, int n_samples, double *Wt,
                  hipStream_t stream) {
    dim3 grid((n_phi + 255) / 256, d, n_samples * (n_mesh + 1));
    hipLaunchKernelGGL(wtable_kernel, grid, dim3(256), 0, stream, V, cmt, n_phi, d, off, n_mesh, Wt);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_assemble(const double *E, const double *Wt, int n_phi, int d, int off, int n_mesh, long kp0, long nkp,
                    double *A, long strideA, hipStream_t stream, float *A32) {
    if (nkp <= 0) return 0;
    hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)nkp), dim3(256), 0, stream, E, Wt, n_phi, d, off, n_mesh, kp0,
                       A, strideA, A32);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

// Mixed-precision refinement: FP64 Rayleigh quotients of the FP32 eigenvectors
// against the FP64 matrix, lambda_j = Re(u^H A u) / (u^H u), replacing the
// eigenvalues of the FP32-reduced tridiagonal. One workgroup per matrix, one
// thread per state; A(r,c) is zero outside the circular band |c-r| <= bw (bw >= n/2:
// dense). resid[kp] = max_j ||A u - lambda u|| / ||u|| (diagnostic).
__global__ __launch_bounds__(256) void rayleigh_kernel(const double *__restrict__ A, const float *__restrict__ U,
                                                       int n, int bw, double *__restrict__ W,
                                                       double *__restrict__ resid) {
    __shared__ double red[8];
    const long kp = blockIdx.x;
    const double *Ak = A + 2L * kp * n * n;
    const float *Uk = U + 2L * kp * n * n;
    const bool dense = 2 * bw + 1 >= n;
    double worst = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        cd q{0.0, 0.0};
        double uu = 0.0, yy = 0.0;
        for (int r = 0; r < n; r++) {
            cd y{0.0, 0.0};
            const int c0 = dense ? 0 : r - bw, c1 = dense ? n - 1 : r + bw;
            for (int cc = c0; cc <= c1; cc++) {
                const int c = cc < 0 ? cc + n : (cc >= n ? cc - n : cc);
                const cx<float> u = ldx(Uk, (long)c * n + j);
                cfma(y, ld(Ak, (long)c * n + r), cd{(double)u.x, (double)u.y});
            }
            const cx<float> ur = ldx(Uk, (long)r * n + j);
            const cd u{(double)ur.x, (double)ur.y};
            cfmac(q, u, y);
            uu += cabs2(u);
            yy += cabs2(y);
        }
        const double lam = q.x / uu;
        W[kp * n + j] = lam;
        const double r2 = fmax(yy / uu - lam * lam, 0.0);
        worst = fmax(worst, sqrt(r2));
    }
    // max over states
    for (int off = kWave / 2; off > 0; off >>= 1) worst = fmax(worst, __shfl_xor(worst, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = worst;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) m = fmax(m, red[w]);
        if (resid) resid[kp] = m;
    }
}

int launch_rayleigh(const double *A, const float *U, int n, int bw, long nkp, double *W, double *resid,
                    hipStream_t stream) {
    if (nkp <= 0) return 0;
    hipLaunchKernelGGL(rayleigh_kernel, dim3((unsigned)nkp), dim3(256), 0, stream, A, U, n, bw, W, resid);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace rlm
