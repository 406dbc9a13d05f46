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

#include <algorithm>

namespace rlm {

// V[s][m][n] = coeff_mn[m][n] * sum_i s_i e^{-i 2pi((m-o) x_i + (n-o) y_i)}
// grid (ceil(npair / 64), S), block 256 = 64 (m, n) pairs x 4 impurity slices; the pairs
// computed are numbered p = m d + n (rows m < o in full, row o up to n = o with mirror,
// every row without), so that every lane of a wave has a pair (64-column blocks of the
// d = 73 columns at n_phi = 256 left 43 % of the lanes idle: 3.4 -> 2.0 ms).
// mirror (d = 2o + 1): the phases of (m, n) and (2o - m, 2o - n) are exact negatives
// (both products and their sum negate exactly), so the second sum is the conjugate of the
// first (to the last bit of sincos's odd / even symmetry): rows m <= o are computed (row
// o up to n = o) and each sum is also written, conjugated, to its mirror; half the sincos
// (n_phi = 3000, 4 samples: 26 -> 14 ms).
__global__ __launch_bounds__(256) void potential_kernel(const double *__restrict__ X, const double *__restrict__ Y,
                                                        const double *__restrict__ Sg, long n_imp,
                                                        const double *__restrict__ cmn, int d, int off,
                                                        double *__restrict__ V, int mirror, int npair) {
    __shared__ double red[2 * 256];
    const int s = blockIdx.y;
    const int nl = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int pr = blockIdx.x * 64 + nl;
    const int m = pr / d, n = pr - m * d;
    const double PI2 = 6.283185307179586232;   // (M_PI*2.0) as a double
    const double kx = PI2 * (double)(m - off);
    const double ky = PI2 * (double)(n - off);
    const double *x = X + s * n_imp, *y = Y + s * n_imp, *sg = Sg + s * n_imp;
    const bool act = pr < npair;
    double re = 0.0, im = 0.0;
    if (act) {
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
    if (slice == 0 && act) {
        for (int k = 1; k < 4; k++) {
            re += red[2 * (threadIdx.x + 64 * k)];
            im += red[2 * (threadIdx.x + 64 * k) + 1];
        }
        const cd c = ld(cmn, (long)m * d + n);
        st(V, (long)s * d * d + (long)m * d + n, cmul(cd{re, im}, c));
        if (mirror && !(m == off && n == off)) {
            const int m2 = 2 * off - m, n2 = 2 * off - n;
            const cd c2 = ld(cmn, (long)m2 * d + n2);
            st(V, (long)s * d * d + (long)m2 * d + n2, cmul(cd{re, -im}, c2));
        }
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
    const int mirror = d == 2 * off + 1 ? 1 : 0;
    const int npair = mirror ? off * d + off + 1 : d * d;
    dim3 grid((npair + 63) / 64, n_samples);
    hipLaunchKernelGGL(potential_kernel, grid, dim3(256), 0, stream, X, Y, S, n_imp, cmn, d, off, V, mirror, npair);
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

// ---------------------------------------------------------------- band form
// The projected Hamiltonian of one k-point is non-zero only on the circular band
// D = |c - r| in B = {0..o} U {N-o..N-1} (all of [0, N) when 2o+1 >= N). The band
// form keeps H[j][b] = h_{j, j+D(b)} = sum_m E[t2][m][j] W[t1][m][D(b)], b < nb =
// |B|, from which A(r, c) = conj(H[r][b(c-r)]) for c >= r and H[c][b(r-c)] below
// (real diagonal), the matrix the dense assembly writes.
int band_width(int N, int off) { return band_nb(N, off); }

// Wb[z][m][b] = W[z][m][D(b)] (the dense table's non-zero columns); grid
// ((nb+63)/64, d, S*(g+1)), z = s*(g+1)+t1.
__global__ void wband_kernel(const double *__restrict__ V, const double *__restrict__ cmt, int n_phi, int d, int off,
                             int n_mesh, int nb, double *__restrict__ Wb) {
    const int z = blockIdx.z, m = blockIdx.y;
    const int s = z / (n_mesh + 1), t1 = z % (n_mesh + 1);
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int D = band_D(b, n_phi, off);
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
    st(Wb, ((long)z * d + m) * nb + b, sum);
}

// Row tiles of 16 that band column tile cb of H needs: the readers take A(r, c), r > c, as
// h_{c, r} = H[c][b(r - c)] and A(r, c), r <= c, as conj(H[r][b(c - r)]) (band_entry,
// the register-resident reduction, the Rayleigh kernels), so they read H[j][b] only where
// j + D(b) < N, never the entries whose column j + D(b) wraps past N (h's lower triangle,
// about half of the band). D increases with b, so tile (j0, b0) holds such an entry iff
// j0 + D(b0) < N.
__host__ __device__ inline int hband_rows(int N, int off, int nrb, int cb) {
    const int b0 = cb * 16;
    const int D = (2 * off + 1 >= N || b0 <= off) ? b0 : b0 + N - 2 * off - 1;
    const int r = (N - D + 15) / 16;
    return r < 0 ? 0 : (r > nrb ? nrb : r);
}

// H[kp][j][b] = sum_m E[t2][m][j] Wb[s,t1][m][b] on the FP64 matrix cores: nsplit
// workgroups per k-point (several where the k-points are too few to fill the chip),
// 16x16 output tiles dealt to their waves, K = m in steps of 4. Output kp is local
// (blockIdx.x / nsplit); the k-point index is kp0 + that over a mesh of pl points per
// line (theta_1 = 0..pl-1 and theta_2 = 0..pl-1 per sample; pl = g for the solved
// points, g+1 for the whole mesh). Wb is laid out per (s, theta_1) with g+1 lines per
// sample.
__global__ __launch_bounds__(256) void hband_kernel(const double *__restrict__ E, const double *__restrict__ Wb, int N,
                                                    int d, int nb, int off, int n_mesh, int pl, long kp0,
                                                    double *__restrict__ H, int nsplit) {
    using M = Mf<double>;
    using v4 = M::v4;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const long kpl = blockIdx.x / nsplit;
    const int part = (int)(blockIdx.x % nsplit);
    const long kp = kp0 + kpl;
    const int g1 = n_mesh + 1;
    const long s = kp / ((long)pl * pl);
    const int rem = (int)(kp % ((long)pl * pl));
    const int t1 = rem / pl, t2 = rem % pl;
    const double *Et = E + 2L * t2 * d * N;
    const double *Wz = Wb + 2L * (s * g1 + t1) * d * nb;
    double *Hk = H + 2L * kpl * N * nb;
    const int nrb = (N + 15) / 16, ncb = (nb + 15) / 16;
    int ntile = 0;
    for (int cb = 0; cb < ncb; cb++) ntile += hband_rows(N, off, nrb, cb);
    for (int tile = part * 4 + w; tile < ntile; tile += 4 * nsplit) {
        int cb = 0, rt = tile;                              // the needed tiles, column tile by column tile
        while (rt >= hband_rows(N, off, nrb, cb)) rt -= hband_rows(N, off, nrb, cb++);
        const int j0 = rt * 16, b0 = cb * 16;
        // three real products per complex one: P1 = (er + ei) xr, P2 = er (xi - xr),
        // P3 = ei (xr + xi); Re = P1 - P3, Im = P1 + P2
        v4 p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
        for (int m0 = 0; m0 < d; m0 += 4) {
            const int m = m0 + lq;
            const cd e = (m < d && j0 + li < N) ? ld(Et, (long)m * N + j0 + li) : cd{0.0, 0.0};    // E[m][j]
            const cd x = (m < d && b0 + li < nb) ? ld(Wz, (long)m * nb + b0 + li) : cd{0.0, 0.0}; // Wb[m][b]
            p1 = M::mma(e.x + e.y, x.x, p1);
            p2 = M::mma(e.x, x.y - x.x, p2);
            p3 = M::mma(e.y, x.x + x.y, p3);
        }
        const int b = b0 + li;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int j = j0 + M::row(lane, v);
            if (j < N && b < nb) st(Hk, (long)j * nb + b, cd{p1[v] - p3[v], b == 0 ? 0.0 : p1[v] + p2[v]});
        }
    }
}

// Dense column-major A of each k-point from its band form; blockIdx.x = kp * nch +
// chunk, the nch chunks of a k-point striding over its N*N entries.
__global__ __launch_bounds__(256) void expand_band_kernel(const double *__restrict__ H, int N, int off, int nb,
                                                          double *__restrict__ A, long strideA, int nch) {
    const long kp = blockIdx.x / nch;
    const double *Hk = H + 2L * kp * N * nb;
    double *Ak = A + 2L * kp * strideA;
    const int step = nch * (int)blockDim.x;
    for_2d((int)(blockIdx.x % nch) * (int)blockDim.x + (int)threadIdx.x, step, N, N, [&](int c, int r) {
        cd v{0.0, 0.0};
        if (r <= c) {
            const int b = band_b(c - r, N, off);
            if (b >= 0) v = cconj(ld(Hk, (long)r * nb + b));
        } else {
            const int b = band_b(r - c, N, off);
            if (b >= 0) v = ld(Hk, (long)c * nb + b);
        }
        st(Ak, (long)c * N + r, v);
    });
}

int launch_wband(const double *V, const double *cmt, int n_phi, int d, int off, int n_mesh, int n_samples, double *Wb,
                 hipStream_t stream) {
    const int nb = band_width(n_phi, off);
    dim3 grid((nb + 63) / 64, d, n_samples * (n_mesh + 1));
    hipLaunchKernelGGL(wband_kernel, grid, dim3(64), 0, stream, V, cmt, n_phi, d, off, n_mesh, nb, Wb);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_hband(const double *E, const double *Wb, int n_phi, int d, int off, int n_mesh, int pl, long kp0,
                 long nkp, double *H, hipStream_t stream) {
    if (nkp <= 0 || pl < 1 || pl > n_mesh + 1) return nkp <= 0 ? 0 : -22;
    // enough workgroups for two per CU (n_phi = 3000, 64 k-points: 19.5 ms with one per k-point)
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const int nrb = (n_phi + 15) / 16, ncb = (band_width(n_phi, off) + 15) / 16;
    int ntile = 0;
    for (int cb = 0; cb < ncb; cb++) ntile += hband_rows(n_phi, off, nrb, cb);
    const long want = (2L * cus + nkp - 1) / nkp;
    const int nsplit = (int)std::max(1L, std::min<long>(want, (ntile + 3) / 4));
    hipLaunchKernelGGL(hband_kernel, dim3((unsigned)(nkp * nsplit)), dim3(256), 0, stream, E, Wb, n_phi, d,
                       band_width(n_phi, off), off, n_mesh, pl, kp0, H, nsplit);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_expand_band(const double *H, int n_phi, int off, long nkp, double *A, long strideA, hipStream_t stream) {
    if (nkp <= 0) return 0;
    const long nn = (long)n_phi * n_phi;
    const int nch = (int)std::min<long>(256, std::max<long>(1, nn / (256L * 64)));
    hipLaunchKernelGGL(expand_band_kernel, dim3((unsigned)(nkp * nch)), dim3(256), 0, stream, H, n_phi, off,
                       band_width(n_phi, off), A, strideA, nch);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

// Mixed-precision refinement: FP64 Rayleigh quotients of the FP32 eigenvectors
// against the FP64 matrix, lambda_j = Re(u^H A u) / (u^H u), replacing the
// eigenvalues of the FP32-reduced tridiagonal. The product runs on the FP64 matrix
// cores: one wave per 16 states, looping over 16-row blocks; for row block r0 only
// the columns of the circular band window enter (A is zero outside |c - r| <= bw mod
// n; the window is the whole row when it would wrap onto itself). The FP32 vectors
// are widened exactly. Four waves per workgroup, the workgroups of one matrix on one
// XCD (blockIdx % 8). With Hb (band form, off = bw) the matrix entries come from it
// instead of A.
// HALF (eigenvalues only): A = B + B^H with B the upper half of the band (the
// entries at circular offsets 1..bw of each row and half the diagonal), so
// u^H A u = 2 Re(u^H B u) and the window of row block r0 is [r0, r0 + 15 + bw]
// (16 + bw columns instead of 16 + 2 bw; the whole upper triangle when dense).
// !HALF: the full product A u, eigenvalues (if W) and resid[kp*nsb + sb] = max over
// the 16 states of block sb of ||A u - lambda u|| / ||u|| (diagnostic).
template <bool HALF, typename TU>
__global__ __launch_bounds__(256) void rayleigh_kernel(const double *__restrict__ A, const double *__restrict__ Hb,
                                                       const TU *__restrict__ U, int n, int bw, long nkp,
                                                       int wg_per_kp, double *__restrict__ W,
                                                       double *__restrict__ resid) {
    using M = Mf<double>;
    using v4 = M::v4;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const long q = blockIdx.x >> 3;
    const long kp = (q / wg_per_kp) * 8 + (blockIdx.x & 7);
    const int sb = (int)(q % wg_per_kp) * 4 + w;           // state block of this wave
    const int nsb = (n + 15) / 16;
    if (kp >= nkp || sb >= nsb) return;                     // no barriers in this kernel
    const int nb = band_nb(n, bw);
    const double *Ak = Hb ? nullptr : A + 2 * kp * (long)n * n;
    const double *Hk = Hb ? Hb + 2 * kp * (long)n * nb : nullptr;
    const TU *Uk = U + 2 * kp * (long)n * n;
    const int j = sb * 16 + li;
    const bool jin = j < n;
    const bool dense = 16 + 2 * bw + 3 >= n;
    double num = 0.0, yy = 0.0, uu = 0.0;
    for (int r0 = 0; r0 < n; r0 += 16) {
        v4 yr = {0.0, 0.0, 0.0, 0.0}, yi = {0.0, 0.0, 0.0, 0.0};
        int cs, kw;
        if (HALF) {
            cs = r0;
            kw = dense ? n - r0 : 16 + bw;
        } else {
            cs = dense ? 0 : r0 - bw;
            kw = dense ? n : 16 + 2 * bw;
        }
        const int r = r0 + li;
        for (int k0 = 0; k0 < kw; k0 += 4) {
            const int kk = k0 + lq;
            int c = cs + kk;
            c = c < 0 ? c + n : (c >= n ? c - n : c);
            const bool kin = kk < kw;
            cd a{0.0, 0.0};                                                               // A(r, c) or B(r, c)
            if (kin && r < n) {
                if (!HALF) {
                    a = band_entry(Ak, Hk, n, bw, nb, r, c);
                } else {
                    const int dl = kk - li;                                               // c - r along the window
                    if (dl == 0) a = cd{0.5 * band_entry(Ak, Hk, n, bw, nb, r, c).x, 0.0};
                    else if (dl > 0) a = band_entry(Ak, Hk, n, bw, nb, r, c);
                }
            }
            const cx<TU> u = (kin && jin) ? ldx(Uk, (long)c * n + j) : cx<TU>{(TU)0, (TU)0};   // U[c][j]
            const double ur = u.x, ui = u.y;
            yr = M::mma(a.x, ur, yr);
            yr = M::mma(-a.y, ui, yr);
            yi = M::mma(a.x, ui, yi);
            yi = M::mma(a.y, ur, yi);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int rr = r0 + M::row(lane, v);
            if (rr < n && jin) {
                const cx<TU> u = ldx(Uk, (long)rr * n + j);
                const double ur = u.x, ui = u.y;
                num += ur * yr[v] + ui * yi[v];
                if (!HALF) yy += yr[v] * yr[v] + yi[v] * yi[v];
                uu += ur * ur + ui * ui;
            }
        }
    }
    // sum over the four lanes holding state j (lane bits 4-5)
    for (int off = 16; off < kWave; off <<= 1) {
        num += __shfl_xor(num, off, kWave);
        if (!HALF) yy += __shfl_xor(yy, off, kWave);
        uu += __shfl_xor(uu, off, kWave);
    }
    if (HALF) num *= 2.0;
    double worst = 0.0;
    if (jin && uu > 0.0) {
        const double lam = num / uu;
        if (W && lq == 0) W[kp * n + j] = lam;
        if (!HALF) worst = sqrt(fmax(yy / uu - lam * lam, 0.0));
    } else if (jin && W && lq == 0) {
        W[kp * n + j] = 0.0;
    }
    if (!HALF) {
        for (int off = 1; off < 16; off <<= 1) worst = fmax(worst, __shfl_xor(worst, off, kWave));
        if (resid && lane == 0) resid[kp * nsb + sb] = worst;
    }
}

// Eigenvalues-only Rayleigh quotients for a banded window (16 + 2 bw + 3 < n): the
// HALF product 2 Re(u^H B u) of rayleigh_kernel with two state blocks (32 states) per
// wave, so that each gathered band entry feeds both, the row blocks and their K-steps
// run as one sequence whose operands are loaded one step ahead, and each complex
// product takes three real MFMAs instead of four: with a = ar + i ai, u = ur + i ui,
// P1 = (ar + ai) ur, P2 = ar (ui - ur), P3 = ai (ur + ui), Re(a u) = P1 - P3 and
// Im(a u) = P1 + P2 (absolute error a few ulps of |a||u|, as for the direct product).
// Row r of block r0 meets columns r .. r + bw only (offsets kk - li in [0, bw] of the
// window [r0, r0 + 15 + bw]); the window may wrap past column n - 1.
template <typename TU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void rayleigh_half_band_kernel(const double *__restrict__ A,
                                                                 const double *__restrict__ Hb,
                                                                 const TU *__restrict__ U, int n, int bw, long nkp,
                                                                 int wg_per_kp, double *__restrict__ W) {
    using M = Mf<double>;
    using v4 = M::v4;
    constexpr int SB = 2;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const long q = blockIdx.x >> 3;
    const long kp = (q / wg_per_kp) * 8 + (blockIdx.x & 7);
    const int sb0 = ((int)(q % wg_per_kp) * 4 + w) * SB;   // first state block of this wave
    const int nsb = (n + 15) / 16;
    if (kp >= nkp || sb0 >= nsb) return;                    // no barriers in this kernel
    const int nb = band_nb(n, bw);
    const double *Ak = Hb ? nullptr : A + 2 * kp * (long)n * n;
    const double *Hk = Hb ? Hb + 2 * kp * (long)n * nb : nullptr;
    const TU *Uk = U + 2 * kp * (long)n * n;
    int jj[SB];
    bool jin[SB];
#pragma unroll
    for (int b = 0; b < SB; b++) {
        jj[b] = (sb0 + b) * 16 + li;
        jin[b] = jj[b] < n;
    }
    const int kw = 16 + bw, nks = (kw + 3) / 4;
    // operands of K-step k0 of row block r0: this lane's A element, loaded raw with a
    // mode applied when it is used (0: outside the half band, 1: as stored, 2: conjugate,
    // 3: half the real part, the diagonal), so that no load is waited for before the
    // current step's MFMAs issue; U[c][j] of each block. Band form: A(r, c) =
    // conj(H[r][c - r]) for c >= r, H[c][2 bw + 1 - (kk - li)] where the window wrapped.
    auto fetch = [&](int r0, int k0, cd &araw, int &mode, cx<TU> (&u)[SB]) {
        const int kk = k0 + lq, r = r0 + li, dl = kk - li;
        int c = r0 + kk;
        c = c >= n ? c - n : c;
        long off;
        int md;
        if (Hk) {
            const bool wrap = c < r;
            off = wrap ? (long)c * nb + (2 * bw + 1 - dl) : (long)r * nb + dl;
            md = wrap ? 1 : 2;
        } else {
            off = (long)c * n + r;
            md = 1;
        }
        if (dl == 0) md = 3;
        if (!(r < n && dl >= 0 && dl <= bw)) {
            off = 0;
            md = 0;
        }
        araw = ld(Hk ? Hk : Ak, off);
        mode = md;
        const bool kin = kk < kw;
#pragma unroll
        for (int b = 0; b < SB; b++) u[b] = (kin && jin[b]) ? ldx(Uk, (long)c * n + jj[b]) : cx<TU>{(TU)0, (TU)0};
    };
    double num[SB], uu[SB];
    v4 p1[SB], p2[SB], p3[SB];
#pragma unroll
    for (int b = 0; b < SB; b++) {
        num[b] = 0.0;
        uu[b] = 0.0;
        p1[b] = p2[b] = p3[b] = v4{0.0, 0.0, 0.0, 0.0};
    }
    cd a_n;
    int mode_n;
    cx<TU> u_n[SB], eu[SB][4];
    fetch(0, 0, a_n, mode_n, u_n);
    int r0 = 0, ks = 0;
    while (r0 < n) {
        const double ar = mode_n == 0 ? 0.0 : (mode_n == 3 ? 0.5 * a_n.x : a_n.x);
        const double ai = (mode_n == 0 || mode_n == 3) ? 0.0 : (mode_n == 2 ? -a_n.y : a_n.y);
        cx<TU> u[SB];
#pragma unroll
        for (int b = 0; b < SB; b++) u[b] = u_n[b];
        if (ks == 0) {      // this row block's U rows for the final dot products, loaded early
#pragma unroll
            for (int b = 0; b < SB; b++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int rr = r0 + M::row(lane, v);
                    eu[b][v] = (rr < n && jin[b]) ? ldx(Uk, (long)rr * n + jj[b]) : cx<TU>{(TU)0, (TU)0};
                }
        }
        const bool last_k = ks + 1 == nks;
        const int r0n = last_k ? r0 + 16 : r0, ksn = last_k ? 0 : ks + 1;
        if (r0n < n) fetch(r0n, 4 * ksn, a_n, mode_n, u_n);
        const double apb = ar + ai;
#pragma unroll
        for (int b = 0; b < SB; b++) {
            const double ur = u[b].x, ui = u[b].y;
            p1[b] = M::mma(apb, ur, p1[b]);
            p2[b] = M::mma(ar, ui - ur, p2[b]);
            p3[b] = M::mma(ai, ur + ui, p3[b]);
        }
        if (last_k) {
#pragma unroll
            for (int b = 0; b < SB; b++) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const double ur = eu[b][v].x, ui = eu[b][v].y;     // zero outside the matrix
                    num[b] += ur * (p1[b][v] - p3[b][v]) + ui * (p1[b][v] + p2[b][v]);
                    uu[b] += ur * ur + ui * ui;
                }
                p1[b] = p2[b] = p3[b] = v4{0.0, 0.0, 0.0, 0.0};
            }
        }
        r0 = r0n;
        ks = ksn;
    }
#pragma unroll
    for (int b = 0; b < SB; b++) {
        // sum over the four lanes holding state j (lane bits 4-5)
        for (int off = 16; off < kWave; off <<= 1) {
            num[b] += __shfl_xor(num[b], off, kWave);
            uu[b] += __shfl_xor(uu[b], off, kWave);
        }
        if (jin[b] && W && lq == 0) W[kp * n + jj[b]] = uu[b] > 0.0 ? 2.0 * num[b] / uu[b] : 0.0;
    }
}

int rayleigh_resid_slots(int n) { return (n + 15) / 16; }

template <typename TU>
static int launch_rayleigh_t(const double *A, const TU *U, int n, int bw, long nkp, double *W, double *resid,
                             hipStream_t stream, const double *Hb) {
    if (nkp <= 0) return 0;
    const int nsb = (n + 15) / 16, wg = (nsb + 3) / 4;
    const long groups = (nkp + 7) / 8;
    if (resid)
        hipLaunchKernelGGL((rayleigh_kernel<false, TU>), dim3((unsigned)(groups * 8 * wg)), dim3(256), 0, stream, A,
                           Hb, U, n, bw, nkp, wg, W, resid);
    else if (16 + 2 * bw + 3 < n) {
        const int wg2 = (nsb + 7) / 8;   // 4 waves x 2 state blocks
        hipLaunchKernelGGL((rayleigh_half_band_kernel<TU>), dim3((unsigned)(groups * 8 * wg2)), dim3(256), 0, stream,
                           A, Hb, U, n, bw, nkp, wg2, W);
    } else
        hipLaunchKernelGGL((rayleigh_kernel<true, TU>), dim3((unsigned)(groups * 8 * wg)), dim3(256), 0, stream, A,
                           Hb, U, n, bw, nkp, wg, W, resid);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}
int launch_rayleigh(const double *A, const float *U, int n, int bw, long nkp, double *W, double *resid,
                    hipStream_t stream, const double *Hb) {
    return launch_rayleigh_t<float>(A, U, n, bw, nkp, W, resid, stream, Hb);
}
int launch_rayleigh(const double *A, const double *U, int n, int bw, long nkp, double *W, double *resid,
                    hipStream_t stream, const double *Hb) {
    return launch_rayleigh_t<double>(A, U, n, bw, nkp, W, resid, stream, Hb);
}

}  // namespace rlm
