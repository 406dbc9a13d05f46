// Batched Hermitian eigensolver (FP64) for gfx950.
//
// Replaces the per-k-point LAPACK zheevd('V') of diag_wrappers/mkl_diag.cpp:2-9
// (and the cusolverDnCheevd / magma_cheevd_gpu calls of
// diag_wrappers/cusolver_diag.cu:3-22, diag_wrappers/magma_diag.cu:2-49) with
// three persistent kernels that each take one workgroup per matrix:
//
//   1. zhetrd_kernel: Householder reduction A = Q T Q^H (LAPACK zhetd2
//      convention, lower storage of the reflectors), with the rank-2 update of
//      step i fused with the Hermitian matrix-vector product of step i+1 so the
//      trailing block is read and written once per column.
//   2. tail_kernel, stage 1: divide & conquer on the real tridiagonal T
//      (Cuppen tear with leaf size 1, LAPACK dlaed2-style deflation,
//      rational-model secular solver, Gu-Eisenstat eigenvector recomputation),
//      level-synchronous so every tree level runs its nodes in parallel.
//   3. tail_kernel, stage 2: back-transform U = Q Z with compact-WY blocks of
//      reflectors (I - V T V^H), one thread per eigenvector column.
//
// Matrices are complex128, column-major with full Hermitian storage on input.
// Eigenvectors are returned component-major ("row-major": U[r*n + j] is
// component r of state j) so that later per-state kernels read coalesced.
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "eig.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace rlm {

// ============================================================================
// 1. Householder tridiagonalisation
// ============================================================================
template <typename T, int NT, int MAXR>
__global__ __launch_bounds__(NT) void zhetrd_kernel(T *__restrict__ Abase, long strideA, int lda, int n,
                                                    double *__restrict__ Dout, double *__restrict__ Eout,
                                                    T *__restrict__ TAUout, long strideV, int batch) {
    using C = cx<T>;
    extern __shared__ double smem_d[];
    // LDS carve: v[2][n], w[2][n] (complex), part[NT/64][n] complex, red[2*NT/64], scal[8]
    C *vbuf = reinterpret_cast<C *>(smem_d);          // 2*n
    C *wbuf = vbuf + 2 * n;                           // 2*n
    C *part = wbuf + 2 * n;                           // (NT/64)*n
    T *red = reinterpret_cast<T *>(part + (NT / kWave) * n);
    T *sc = red + 2 * (NT / kWave);                   // 8 scalars
    const int t = threadIdx.x;

    for (long b = blockIdx.x; b < batch; b += gridDim.x) {
        T *A = Abase + 2 * b * strideA;                // complex index = 2*...
        double *Dv = Dout + b * strideV;
        double *Ev = Eout + b * strideV;
        // strideV: per-matrix stride in doubles of D, E and TAU alike (TAU holds T values)
        T *Tv = TAUout + b * strideV * (long)(sizeof(double) / sizeof(T));
        auto Aidx = [&](int r, int c) -> long { return (long)c * lda + r; };

        int cur = 0;          // vbuf/wbuf slot holding the pending update
        bool pending = false;
        for (int i = 0; i < n - 1; i++) {
            C *vold = vbuf + cur * n, *wold = wbuf + cur * n;    // indexed from row i
            C *vnew = vbuf + (cur ^ 1) * n;                      // indexed from row i+1
            C *wnew = wbuf + (cur ^ 1) * n;
            // (a) bring column i (rows >= i) up to date, collect ||x||^2, alpha, d_i
            T nrm2 = 0;
            for (int r = i + t; r < n; r += NT) {
                C a = ldx(A, Aidx(r, i));
                if (pending) {
                    const C vr = vold[r - i], wr = wold[r - i];
                    const C vc = vold[0], wc = wold[0];
                    // a -= v_r conj(w_c) + w_r conj(v_c)
                    a = a - (vr * conjx(wc) + wr * conjx(vc));
                    stx(A, Aidx(r, i), a);
                }
                if (r == i) sc[0] = a.x;
                else if (r == i + 1) { sc[1] = a.x; sc[2] = a.y; }
                else nrm2 += abs2x(a);
            }
            __syncthreads();
            nrm2 = block_sum_t(nrm2, red);
            if (t == 0) {
                Dv[i] = (double)sc[0];
                const T ar = sc[1], ai = sc[2];
                T beta, taur, taui, sr, si;
                if (nrm2 == (T)0 && ai == (T)0) {
                    beta = ar; taur = 0; taui = 0; sr = 0; si = 0;
                } else {
                    // zlarfg: beta = -sign(alphr) * ||(alpha, x)||, tau = (beta-alpha)/beta, x *= 1/(alpha-beta)
                    const T nrm = sqrt(ar * ar + ai * ai + nrm2);
                    beta = (ar >= (T)0) ? -nrm : nrm;
                    taur = (beta - ar) / beta;
                    taui = -ai / beta;
                    const T dr = ar - beta, di = ai;
                    const T inv = (T)1 / (dr * dr + di * di);
                    sr = dr * inv; si = -di * inv;
                }
                Ev[i] = (double)beta;
                Tv[2 * i] = taur;
                Tv[2 * i + 1] = taui;
                sc[3] = taur; sc[4] = taui; sc[5] = sr; sc[6] = si; sc[7] = beta;
            }
            __syncthreads();
            const C tau{sc[3], sc[4]};
            const C scal{sc[5], sc[6]};
            // (b) v = [1; x*scal], stored back into A(i+2:n, i); A(i+1,i) = beta
            for (int r = i + 1 + t; r < n; r += NT) {
                C v;
                if (r == i + 1) {
                    v = C{(T)1, (T)0};
                    stx(A, Aidx(r, i), C{sc[7], (T)0});
                } else {
                    v = ldx(A, Aidx(r, i)) * scal;
                    stx(A, Aidx(r, i), v);
                }
                vnew[r - (i + 1)] = v;
            }
            __syncthreads();
            // (c) fused pass over the trailing block rows/cols >= i+1:
            //     apply the pending update of step i-1, accumulate acc = A22 * vnew
            const int nr = n - (i + 1);
            const int R = (nr >= NT) ? NT : ((nr + kWave - 1) / kWave) * kWave;
            const int G = NT / R;
            const int rl = t % R, cg = t / R;
            C acc[MAXR];
#pragma unroll
            for (int k = 0; k < MAXR; k++) acc[k] = C{(T)0, (T)0};
            if (cg < G) {
                for (int c = i + 1 + cg; c < n; c += G) {
                    const C vn = vnew[c - (i + 1)];
                    C voc{(T)0, (T)0}, woc{(T)0, (T)0};
                    if (pending) { voc = conjx(vold[c - i]); woc = conjx(wold[c - i]); }
#pragma unroll
                    for (int k = 0; k < MAXR; k++) {
                        const int r = i + 1 + rl + k * R;
                        if (r < n) {
                            C a = ldx(A, Aidx(r, c));
                            if (pending) {
                                const C vr = vold[r - i], wr = wold[r - i];
                                a = a - (vr * woc + wr * voc);
                                if (r == c) a.y = (T)0;   // Hermitian diagonal stays real (zher2)
                                stx(A, Aidx(r, c), a);
                            }
                            fmax_(acc[k], a, vn);
                        }
                    }
                }
            }
            // reduce the G column groups
#pragma unroll
            for (int k = 0; k < MAXR; k++) {
                const int r = i + 1 + rl + k * R;
                if (r < n) part[cg * n + (r - (i + 1))] = acc[k];
            }
            __syncthreads();
            C dotl{(T)0, (T)0};
            for (int r = i + 1 + t; r < n; r += NT) {
                C p{(T)0, (T)0};
                for (int g = 0; g < G; g++) p = p + part[g * n + (r - (i + 1))];
                p = tau * p;
                wnew[r - (i + 1)] = p;                        // p for now
                fmacx(dotl, p, vnew[r - (i + 1)]);             // conj(p) v
            }
            const C dot = block_sum_cx(dotl, red);
            // alpha2 = -1/2 tau (p^H v); w = p + alpha2 v
            const C alpha2 = (tau * dot) * (T)(-0.5);
            for (int r = i + 1 + t; r < n; r += NT) {
                const C p = wnew[r - (i + 1)];
                wnew[r - (i + 1)] = p + alpha2 * vnew[r - (i + 1)];
            }
            __syncthreads();
            pending = (tau.x != (T)0 || tau.y != (T)0);
            cur ^= 1;
        }
        // last diagonal element
        if (t == 0) {
            C a = ldx(A, Aidx(n - 1, n - 1));
            if (pending && n >= 2) {
                const C v = vbuf[cur * n], w = wbuf[cur * n];
                a = a - (v * conjx(w) + w * conjx(v));
            }
            Dv[n - 1] = (double)a.x;
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// Blocked variant (LAPACK zlatrd/zhetrd structure): panels of NB columns. Inside a
// panel the trailing block is only read (Hermitian matrix-vector product against
// the panel-start matrix, corrected with the panel's V and W); after the panel
// one rank-2*NB update A22 -= V W^H + W V^H writes it once. Per column the
// trailing block is read once instead of read and written.
// V(r,l) lives in A (column k0+l, unit at row k0+l+1); W is kept in LDS.
template <typename T, int NT, int NB, int MAXR>
__global__ __launch_bounds__(NT) void zhetrd_blocked_kernel(T *__restrict__ Abase, long strideA, int lda, int n,
                                                            double *__restrict__ Dout, double *__restrict__ Eout,
                                                            T *__restrict__ TAUout, long strideV, int batch) {
    using C = cx<T>;
    constexpr int NQ = 2 * NB;              // dot products per column (W^H v and V^H v)
    constexpr int RG = NT / NQ;             // row groups for those dot products
    extern __shared__ double smem_d[];
    C *Wp = reinterpret_cast<C *>(smem_d);           // n x NB, row-major
    C *vb = Wp + (long)n * NB;                       // n (v of the current column, from row i+1)
    C *part = vb + n;                                // (NT/64) x n
    C *dpart = part + (NT / kWave) * n;              // RG x NQ
    C *dots = dpart + RG * NQ;                       // NQ
    T *red = reinterpret_cast<T *>(dots + NQ);
    T *sc = red + 2 * (NT / kWave);
    const int t = threadIdx.x;

    for (long b = blockIdx.x; b < batch; b += gridDim.x) {
        T *A = Abase + 2 * b * strideA;
        double *Dv = Dout + b * strideV;
        double *Ev = Eout + b * strideV;
        T *Tv = TAUout + b * strideV * (long)(sizeof(double) / sizeof(T));
        auto Aidx = [&](int r, int c) -> long { return (long)c * lda + r; };
        // V(r, l) of the current panel (reflector of column k0 + l)
        auto Vget = [&](int k0, int r, int l) -> C {
            const int j = k0 + l;
            if (r < j + 1) return C{(T)0, (T)0};
            if (r == j + 1) return C{(T)1, (T)0};
            return ldx(A, Aidx(r, j));
        };
        for (int k0 = 0; k0 < n - 1; k0 += NB) {
            const int nbk = min(NB, n - 1 - k0);
            for (int j = 0; j < nbk; j++) {
                const int i = k0 + j;
                // (1) column i, rows >= i: apply the panel's earlier updates
                T nrm2 = 0;
                for (int r = i + t; r < n; r += NT) {
                    C a = ldx(A, Aidx(r, i));
                    if (j > 0) {
                        for (int l = 0; l < j; l++) {
                            const C vr = Vget(k0, r, l), wr = Wp[(long)r * NB + l];
                            const C vi = Vget(k0, i, l), wi = Wp[(long)i * NB + l];
                            a = a - (vr * conjx(wi) + wr * conjx(vi));
                        }
                        if (r == i) a.y = (T)0;
                        stx(A, Aidx(r, i), a);
                    }
                    if (r == i) sc[0] = a.x;
                    else if (r == i + 1) { sc[1] = a.x; sc[2] = a.y; }
                    else nrm2 += abs2x(a);
                }
                __syncthreads();
                nrm2 = block_sum_t(nrm2, red);
                if (t == 0) {
                    Dv[i] = (double)sc[0];
                    const T ar = sc[1], ai = sc[2];
                    T beta, taur, taui, sr, si;
                    if (nrm2 == (T)0 && ai == (T)0) {
                        beta = ar; taur = 0; taui = 0; sr = 0; si = 0;
                    } else {
                        const T nrm = sqrt(ar * ar + ai * ai + nrm2);
                        beta = (ar >= (T)0) ? -nrm : nrm;
                        taur = (beta - ar) / beta;
                        taui = -ai / beta;
                        const T dr = ar - beta, di = ai;
                        const T inv = (T)1 / (dr * dr + di * di);
                        sr = dr * inv; si = -di * inv;
                    }
                    Ev[i] = (double)beta;
                    Tv[2 * i] = taur;
                    Tv[2 * i + 1] = taui;
                    sc[3] = taur; sc[4] = taui; sc[5] = sr; sc[6] = si; sc[7] = beta;
                }
                __syncthreads();
                const C tau{sc[3], sc[4]};
                const C scal{sc[5], sc[6]};
                // (2) v = [1; x*scal] into A(i+1:, i) (A(i+1,i) keeps beta) and LDS
                for (int r = i + 1 + t; r < n; r += NT) {
                    C v;
                    if (r == i + 1) {
                        v = C{(T)1, (T)0};
                        stx(A, Aidx(r, i), C{sc[7], (T)0});
                    } else {
                        v = ldx(A, Aidx(r, i)) * scal;
                        stx(A, Aidx(r, i), v);
                    }
                    vb[r - (i + 1)] = v;
                }
                __syncthreads();
                // (3) p = A22 v over rows/cols >= i+1 (A22 as at the panel start)
                const int nr = n - (i + 1);
                const int R = (nr >= NT) ? NT : ((nr + kWave - 1) / kWave) * kWave;
                const int G = NT / R;
                const int rl = t % R, cg = t / R;
                C acc[MAXR];
#pragma unroll
                for (int k = 0; k < MAXR; k++) acc[k] = C{(T)0, (T)0};
                if (cg < G) {
                    for (int c = i + 1 + cg; c < n; c += G) {
                        const C vn = vb[c - (i + 1)];
#pragma unroll
                        for (int k = 0; k < MAXR; k++) {
                            const int r = i + 1 + rl + k * R;
                            if (r < n) fmax_(acc[k], ldx(A, Aidx(r, c)), vn);
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < MAXR; k++) {
                    const int r = i + 1 + rl + k * R;
                    if (r < n) part[cg * n + (r - (i + 1))] = acc[k];
                }
                // (4) dot products x_l = W(:,l)^H v, y_l = V(:,l)^H v (rows >= i+1, l < j)
                {
                    const int q = t % NQ, rg = t / NQ;
                    const int l = q >> 1, kind = q & 1;
                    C s{(T)0, (T)0};
                    if (l < j) {
                        for (int r = i + 1 + rg; r < n; r += RG) {
                            const C x = kind == 0 ? Wp[(long)r * NB + l] : Vget(k0, r, l);
                            fmacx(s, x, vb[r - (i + 1)]);
                        }
                    }
                    dpart[rg * NQ + q] = s;
                }
                __syncthreads();
                if (t < NQ) {
                    C s{(T)0, (T)0};
                    for (int g = 0; g < RG; g++) s = s + dpart[g * NQ + t];
                    dots[t] = s;
                }
                __syncthreads();
                // (5) p -= V x + W y; p *= tau; w = p - 1/2 tau (p^H v) v
                C dotl{(T)0, (T)0};
                for (int r = i + 1 + t; r < n; r += NT) {
                    C p{(T)0, (T)0};
                    for (int g = 0; g < G; g++) p = p + part[g * n + (r - (i + 1))];
                    for (int l = 0; l < j; l++) {
                        p = p - (Vget(k0, r, l) * dots[2 * l] + Wp[(long)r * NB + l] * dots[2 * l + 1]);
                    }
                    p = tau * p;
                    part[r - (i + 1)] = p;               // row slot of group 0, read back by this thread only
                    fmacx(dotl, p, vb[r - (i + 1)]);
                }
                const C dot = block_sum_cx(dotl, red);
                const C alpha2 = (tau * dot) * (T)(-0.5);
                for (int r = i + 1 + t; r < n; r += NT)
                    Wp[(long)r * NB + j] = part[r - (i + 1)] + alpha2 * vb[r - (i + 1)];
                __syncthreads();
            }
            // rank-2*nbk update of the trailing block rows/cols >= k1 = k0 + nbk
            const int k1 = k0 + nbk;
            const int nr = n - k1;
            if (nr > 0) {
                const int R = (nr >= NT) ? NT : ((nr + kWave - 1) / kWave) * kWave;
                const int G = NT / R;
                const int rl = t % R, cg = t / R;
                if (cg < G) {
#pragma unroll
                    for (int k = 0; k < MAXR; k++) {
                        const int r = k1 + rl + k * R;
                        if (r >= n) continue;
                        C vr[NB], wr[NB];
#pragma unroll
                        for (int l = 0; l < NB; l++) {
                            vr[l] = l < nbk ? Vget(k0, r, l) : C{(T)0, (T)0};
                            wr[l] = l < nbk ? Wp[(long)r * NB + l] : C{(T)0, (T)0};
                        }
                        for (int c = k1 + cg; c < n; c += G) {
                            C a = ldx(A, Aidx(r, c));
#pragma unroll
                            for (int l = 0; l < NB; l++) {
                                if (l < nbk) {
                                    const C vc = Vget(k0, c, l), wc = Wp[(long)c * NB + l];
                                    a = a - (vr[l] * conjx(wc) + wr[l] * conjx(vc));
                                }
                            }
                            if (r == c) a.y = (T)0;
                            stx(A, Aidx(r, c), a);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (t == 0) Dv[n - 1] = (double)ldx(A, Aidx(n - 1, n - 1)).x;
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// Lower-storage blocked variant: as zhetrd_blocked_kernel, but only the lower
// triangle of A is ever read or written. The Hermitian product A22*v walks the
// lower triangle in 32x32 tiles, one wave per tile: each element read once feeds
// both y_r += A(r,c) v_c and y_c += conj(A(r,c)) v_r (column sums by wave
// shuffles). Each wave accumulates into its own LDS vector; the 4 vectors are
// summed in a fixed order, so results do not depend on scheduling. Halves the
// trailing-matrix traffic of the full-storage kernels.
template <typename T, int NT, int NB, int MAXR>
// ncols: columns to reduce (n - 1 for the whole matrix; fewer leaves the updated
// trailing block A(ncols:n, ncols:n) for another kernel to finish).
__global__ __launch_bounds__(NT) void zhetrd_lower_kernel(T *__restrict__ Abase, long strideA, int lda, int n,
                                                          double *__restrict__ Dout, double *__restrict__ Eout,
                                                          T *__restrict__ TAUout, long strideV, int batch,
                                                          int ncols) {
    using C = cx<T>;
    constexpr int NW = NT / kWave;
    constexpr int NQ = 2 * NB;
    constexpr int RG = NT / NQ;
    constexpr int TS = 32;
    extern __shared__ double smem_d[];
    C *Wp = reinterpret_cast<C *>(smem_d);           // n x NB, row-major
    C *vb = Wp + (long)n * NB;                       // n
    C *ypart = vb + n;                               // NW x n (one vector per wave)
    C *dpart = ypart + NW * n;                       // RG x NQ
    C *dots = dpart + RG * NQ;                       // NQ
    T *red = reinterpret_cast<T *>(dots + NQ);
    T *sc = red + 2 * NW;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1), wid = t / kWave;

    for (long b = blockIdx.x; b < batch; b += gridDim.x) {
        T *A = Abase + 2 * b * strideA;
        double *Dv = Dout + b * strideV;
        double *Ev = Eout + b * strideV;
        T *Tv = TAUout + b * strideV * (long)(sizeof(double) / sizeof(T));
        auto Aidx = [&](int r, int c) -> long { return (long)c * lda + r; };
        auto Vget = [&](int k0, int r, int l) -> C {
            const int j = k0 + l;
            if (r < j + 1) return C{(T)0, (T)0};
            if (r == j + 1) return C{(T)1, (T)0};
            return ldx(A, Aidx(r, j));
        };
        for (int k0 = 0; k0 < ncols; k0 += NB) {
            const int nbk = min(NB, ncols - k0);
            for (int j = 0; j < nbk; j++) {
                const int i = k0 + j;
                // (1) column i, rows >= i
                T nrm2 = 0;
                for (int r = i + t; r < n; r += NT) {
                    C a = ldx(A, Aidx(r, i));
                    if (j > 0) {
                        for (int l = 0; l < j; l++) {
                            const C vr = Vget(k0, r, l), wr = Wp[(long)r * NB + l];
                            const C vi = Vget(k0, i, l), wi = Wp[(long)i * NB + l];
                            a = a - (vr * conjx(wi) + wr * conjx(vi));
                        }
                        if (r == i) a.y = (T)0;
                        stx(A, Aidx(r, i), a);
                    }
                    if (r == i) sc[0] = a.x;
                    else if (r == i + 1) { sc[1] = a.x; sc[2] = a.y; }
                    else nrm2 += abs2x(a);
                }
                __syncthreads();
                nrm2 = block_sum_t(nrm2, red);
                if (t == 0) {
                    Dv[i] = (double)sc[0];
                    const T ar = sc[1], ai = sc[2];
                    T beta, taur, taui, sr, si;
                    if (nrm2 == (T)0 && ai == (T)0) {
                        beta = ar; taur = 0; taui = 0; sr = 0; si = 0;
                    } else {
                        const T nrm = sqrt(ar * ar + ai * ai + nrm2);
                        beta = (ar >= (T)0) ? -nrm : nrm;
                        taur = (beta - ar) / beta;
                        taui = -ai / beta;
                        const T dr = ar - beta, di = ai;
                        const T inv = (T)1 / (dr * dr + di * di);
                        sr = dr * inv; si = -di * inv;
                    }
                    Ev[i] = (double)beta;
                    Tv[2 * i] = taur;
                    Tv[2 * i + 1] = taui;
                    sc[3] = taur; sc[4] = taui; sc[5] = sr; sc[6] = si; sc[7] = beta;
                }
                __syncthreads();
                const C tau{sc[3], sc[4]};
                const C scal{sc[5], sc[6]};
                // (2) v, and clear the per-wave accumulators
                for (int r = i + 1 + t; r < n; r += NT) {
                    C v;
                    if (r == i + 1) {
                        v = C{(T)1, (T)0};
                        stx(A, Aidx(r, i), C{sc[7], (T)0});
                    } else {
                        v = ldx(A, Aidx(r, i)) * scal;
                        stx(A, Aidx(r, i), v);
                    }
                    vb[r - (i + 1)] = v;
                }
                for (int e = t; e < NW * n; e += NT) ypart[e] = C{(T)0, (T)0};
                __syncthreads();
                // (3) y = A22 v from the lower triangle, 32x32 tiles, one wave per tile.
                // Lane = (rg, cg) owns the 4x4 sub-block rows 4rg.., cols 4cg.. of the tile:
                // row sums reduce over cg (lane bits 3-5), column sums over rg (lane bits 0-2).
                const int m = n - (i + 1);
                const int Tn = (m + TS - 1) / TS;
                const int ntiles = Tn * (Tn + 1) / 2;
                C *yw = ypart + wid * n;
                const int lrg = lane & 7, lcg = lane >> 3;
                // The 16 entries of a sub-block load unconditionally (indices clamped into the
                // trailing block, masked afterwards) and the next tile's loads are issued before
                // the current tile's arithmetic, so memory latency overlaps the work.
                auto tile_of = [&](int tk, int &R, int &Cc) {
                    R = 0;
                    while ((R + 1) * (R + 2) / 2 <= tk) R++;
                    Cc = tk - R * (R + 1) / 2;
                };
                auto load_tile = [&](int tk, C (&av)[16]) {
                    int R, Cc;
                    tile_of(tk, R, Cc);
                    const int rb = R * TS + 4 * lrg, cb = Cc * TS + 4 * lcg;
#pragma unroll
                    for (int cu = 0; cu < 4; cu++)
#pragma unroll
                        for (int ru = 0; ru < 4; ru++) {
                            const int rr = min(rb + ru, m - 1), cc = min(cb + cu, m - 1);
                            av[ru * 4 + cu] = ldx(A, Aidx(i + 1 + rr, i + 1 + cc));
                        }
                };
                C cur[16], nxt[16];
                if (wid < ntiles) load_tile(wid, cur);
                for (int tk = wid; tk < ntiles; tk += NW) {
                    if (tk + NW < ntiles) load_tile(tk + NW, nxt);
                    int R, Cc;
                    tile_of(tk, R, Cc);
                    const int rb = R * TS + 4 * lrg;          // first local row of the sub-block
                    const int cb = Cc * TS + 4 * lcg;         // first local column
                    C vrow[4], vcol[4], rs[4], cs[4];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        vrow[u] = (rb + u < m) ? vb[rb + u] : C{(T)0, (T)0};
                        vcol[u] = (cb + u < m) ? vb[cb + u] : C{(T)0, (T)0};
                        rs[u] = C{(T)0, (T)0};
                        cs[u] = C{(T)0, (T)0};
                    }
#pragma unroll
                    for (int cu = 0; cu < 4; cu++) {
                        const int cloc = cb + cu;
#pragma unroll
                        for (int ru = 0; ru < 4; ru++) {
                            const int rloc = rb + ru;
                            const bool use = rloc < m && cloc < m && (R > Cc || rloc >= cloc);
                            const C a = use ? cur[ru * 4 + cu] : C{(T)0, (T)0};
                            fmax_(rs[ru], a, vcol[cu]);                      // y_r += A(r,c) v_c
                            if (rloc != cloc) fmacx(cs[cu], a, vrow[ru]);   // y_c += conj(A(r,c)) v_r
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 16; e++) cur[e] = nxt[e];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
#pragma unroll
                        for (int off = 8; off < 64; off <<= 1) {     // sum over the 8 column groups
                            rs[u].x += __shfl_xor(rs[u].x, off, kWave);
                            rs[u].y += __shfl_xor(rs[u].y, off, kWave);
                        }
#pragma unroll
                        for (int off = 1; off < 8; off <<= 1) {      // sum over the 8 row groups
                            cs[u].x += __shfl_xor(cs[u].x, off, kWave);
                            cs[u].y += __shfl_xor(cs[u].y, off, kWave);
                        }
                    }
                    if (lcg == 0) {
#pragma unroll
                        for (int u = 0; u < 4; u++)
                            if (rb + u < m) yw[rb + u] = yw[rb + u] + rs[u];
                    }
                    if (lrg == 0) {
#pragma unroll
                        for (int u = 0; u < 4; u++)
                            if (cb + u < m) yw[cb + u] = yw[cb + u] + cs[u];
                    }
                }
                // (4) dot products x_l = W(:,l)^H v, y_l = V(:,l)^H v (rows >= i+1, l < j)
                {
                    const int q = t % NQ, rg = t / NQ;
                    const int l = q >> 1, kind = q & 1;
                    C s2{(T)0, (T)0};
                    if (l < j) {
                        for (int r = i + 1 + rg; r < n; r += RG) {
                            const C x = kind == 0 ? Wp[(long)r * NB + l] : Vget(k0, r, l);
                            fmacx(s2, x, vb[r - (i + 1)]);
                        }
                    }
                    dpart[rg * NQ + q] = s2;
                }
                __syncthreads();
                if (t < NQ) {
                    C s2{(T)0, (T)0};
                    for (int g = 0; g < RG; g++) s2 = s2 + dpart[g * NQ + t];
                    dots[t] = s2;
                }
                __syncthreads();
                // (5) p = sum of the wave vectors - V x - W y; p *= tau; w = p - 1/2 tau (p^H v) v
                C dotl{(T)0, (T)0};
                for (int r = i + 1 + t; r < n; r += NT) {
                    const int rr = r - (i + 1);
                    C p{(T)0, (T)0};
                    for (int w = 0; w < NW; w++) p = p + ypart[w * n + rr];
                    for (int l = 0; l < j; l++)
                        p = p - (Vget(k0, r, l) * dots[2 * l] + Wp[(long)r * NB + l] * dots[2 * l + 1]);
                    p = tau * p;
                    ypart[rr] = p;                    // slot of wave 0, read back by this thread only
                    fmacx(dotl, p, vb[rr]);
                }
                const C dot = block_sum_cx(dotl, red);
                const C alpha2 = (tau * dot) * (T)(-0.5);
                for (int r = i + 1 + t; r < n; r += NT)
                    Wp[(long)r * NB + j] = ypart[r - (i + 1)] + alpha2 * vb[r - (i + 1)];
                __syncthreads();
            }
            // rank-2*nbk update of the lower triangle of the trailing block (rows >= cols >= k1)
            // on the matrix cores: A(row, col) -= sum_k X[row][k] conj(Y[col][k]) with
            // X = [V W], Y = [W V] (K = 2*NB), per 16x16 tile of the lower triangle. The
            // product is formed transposed, M[col][row], so that element v of lane l is
            // A(row = R + l%16, col = Cb + Mf::row(l, v)): 16 consecutive rows per column.
            const int k1 = k0 + nbk;
            const int nr = n - k1;
            if (nr > 0) {
                using M = Mf<T>;
                using v4 = typename M::v4;
                const int li = lane & 15, lq = lane >> 4;
                const int nt16 = (nr + 15) / 16, ntl = nt16 * (nt16 + 1) / 2;
                for (int tl = wid; tl < ntl; tl += NW) {
                    int tr = 0;
                    while ((tr + 1) * (tr + 2) / 2 <= tl) tr++;
                    const int tc = tl - tr * (tr + 1) / 2;
                    const int R = k1 + tr * 16, Cb = k1 + tc * 16;
                    v4 accr, acci;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = R + li, col = Cb + M::row(lane, v);
                        const C x = (row < n && col < n && row >= col) ? ldx(A, Aidx(row, col)) : C{(T)0, (T)0};
                        accr[v] = x.x;
                        acci[v] = x.y;
                    }
                    const int rowq = R + li, colq = Cb + li;      // operand rows of X and Y
#pragma unroll
                    for (int s4 = 0; s4 < 2 * NB / 4; s4++) {
                        const int kk = 4 * s4 + lq;               // 0..2NB-1: [V | W] for X, [W | V] for Y
                        const int l = kk < NB ? kk : kk - NB;
                        C xv{(T)0, (T)0}, yv{(T)0, (T)0};
                        if (l < nbk) {
                            if (rowq < n) xv = kk < NB ? Vget(k0, rowq, l) : Wp[(long)rowq * NB + l];
                            if (colq < n) yv = kk < NB ? Wp[(long)colq * NB + l] : Vget(k0, colq, l);
                        }
                        // A-operand P[col][k] = conj(Y[col][k]), B-operand Q[k][row] = X[row][k]
                        accr = M::mma(-yv.x, xv.x, accr);
                        accr = M::mma(-yv.y, xv.y, accr);
                        acci = M::mma(-yv.x, xv.y, acci);
                        acci = M::mma(yv.y, xv.x, acci);
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = R + li, col = Cb + M::row(lane, v);
                        if (row < n && col < n && row >= col)
                            stx(A, Aidx(row, col), C{accr[v], row == col ? (T)0 : acci[v]});
                    }
                }
            }
            __syncthreads();
        }
        if (t == 0 && ncols == n - 1) Dv[n - 1] = (double)ldx(A, Aidx(n - 1, n - 1)).x;
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// Register-resident variant: the lower triangle of A is read into VGPRs once and
// stays there for the whole reduction, so a matrix costs one read of its lower
// triangle plus the writes of d, e, tau and the reflectors (the kernels above
// stream the trailing block through memory once per column). One workgroup of
// 512 threads (8 waves) per matrix. The lower triangle is cut into 32x32 tiles
// (diagonal tiles held as full Hermitian blocks) dealt to the waves by
// resident_deal, and lane (lrg, lcg) = (lane & 7, lane >> 3) of a wave holds the
// 4x4 sub-block at tile rows 4*lrg.., columns 4*lcg.. of each of its tiles. Step i
// (reflector v_i known):
//   all waves: apply the pending rank-2 update of step i-1 to the trailing block,
//              accumulate y = A22 v_i from the lower triangle (row sums and
//              conjugate column sums, one LDS vector per wave) and copy column
//              i+1 to LDS;
//   wave 0:    w_i = tau_i y - 1/2 tau_i^2 (y^H v_i) v_i; column i+1 -= v_i w_i(0)^*
//              + w_i v_i(0)^*; reflector of step i+1 from that column (zlarfg).
// Two barriers per column. NS = tiles per wave held in registers; up to NL tiles
// of tile column 0 that find no register slot are kept in LDS (they are live for
// the first 32 columns only). n <= 256.

// Cross-lane moves for the tile reductions: DPP within a 16-lane row, ds_swizzle
// within 32 lanes (no LDS traffic), bpermute only across the two wave halves.
// (bound_ctrl set: every source lane of the patterns used here lies inside its row,
// so it changes no value, and it lets the compiler fold the move into the add)
template <int CTRL>
__device__ inline float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ inline double dpp_mov(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffLL), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int PAT>
__device__ inline float swz_mov(float v) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), PAT));
}
template <int PAT>
__device__ inline double swz_mov(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(u & 0xffffffffLL), PAT);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(u >> 32), PAT);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum over lane bits 0-2 (xor 1, xor 2 by quad_perm, then the half-row mirror,
// which pairs quad 0 with quad 1 of each 8-lane group)
template <typename T>
__device__ inline T sum_lanes_012(T v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    v += dpp_mov<0x141>(v);
    return v;
}
// sum over lane bits 3-5 (row rotate by 8 = xor 8, swizzle xor 16, bpermute xor 32)
template <typename T>
__device__ inline T sum_lanes_345(T v) {
    v += dpp_mov<0x128>(v);
    v += swz_mov<0x401F>(v);
    v += __shfl_xor(v, 32, kWave);
    return v;
}

// Lane-swap instructions of gfx950 (VALU, no LDS round trip). Swapping a register
// with itself, v_permlane16_swap returns rows (0,0,2,2) and (1,1,3,3) of the 16-lane
// rows, v_permlane32_swap the low half twice and the high half twice: the sum of
// the two results is the xor-16 resp. xor-32 partner sum, in every lane.
__device__ inline float sum_xor16(float v) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ inline float sum_xor32(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ inline double sum_xor16(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    return __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0])) +
           __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ inline double sum_xor32(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    return __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0])) +
           __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
// sum over all 64 lanes in VALU only (the latency-bound single-wave phase)
template <typename T>
__device__ inline T sum_wave_valu(T v) {
    v = sum_lanes_012(v);
    v += dpp_mov<0x128>(v);
    return sum_xor32(sum_xor16(v));
}
// value of lane l (wave-uniform) in every lane, through a scalar register
__device__ inline float lane_bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double lane_bcast(double v, int l) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One tile of the fused pass of step i (see above). The vectors are indexed by
// absolute row and are zero outside their support (v_i, w_i: rows i+1..n-1), so the
// pending rank-2 update leaves every entry outside the trailing block unchanged and
// the stale entries left of it add nothing to y = A v_i: no per-entry masks.
// Storage of one lane's 4x4 sub-block: registers, or LDS (entry k of lane l at
// p[k * 64], p = tile base + l, conflict-free).
template <typename T>
struct RegSub {
    cx<T> (&a)[16];
    __device__ cx<T> get(int k) const { return a[k]; }
    __device__ void set(int k, cx<T> x) { a[k] = x; }
};
template <typename T>
struct LdsSub {
    cx<T> *p;
    __device__ cx<T> get(int k) const { return p[k * kWave]; }
    __device__ void set(int k, cx<T> x) { p[k * kWave] = x; }
};

// x -= a * b
template <typename T>
__device__ __forceinline__ void fms_(cx<T> &x, cx<T> a, cx<T> b) {
    x.x = fma(-a.x, b.x, x.x);
    x.x = fma(a.y, b.y, x.x);
    x.y = fma(-a.x, b.y, x.y);
    x.y = fma(-a.y, b.x, x.y);
}

// Tile strictly below the diagonal: A(r,c) -= v_r w_c^* + w_r v_c^* (step i-1), then
// y_r += A(r,c) v_c and y_c += A(r,c)^* v_r (step i) into the wave's partial y.
template <typename T, typename Sub>
__device__ __forceinline__ void resident_tile_off(Sub a, int rb, int cb, int lrg, int lcg, const cx<T> *vn,
                                                  const cx<T> *vo, const cx<T> *wo, cx<T> *yw) {
    using CT = cx<T>;
    const CT z{(T)0, (T)0};
    CT vc[4], poc[4], woc[4], cs[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        vc[u] = vn[cb + u];
        poc[u] = conjx(vo[cb + u]);
        woc[u] = conjx(wo[cb + u]);
        cs[u] = z;
    }
#pragma unroll
    for (int ru = 0; ru < 4; ru++) {
        const int r = rb + ru;
        const CT vr = vn[r], por = vo[r], wor = wo[r];
        const T nvrx = -vr.x;
        CT rs = z;
#pragma unroll
        for (int cu = 0; cu < 4; cu++) {
            CT x = a.get(ru * 4 + cu);
            fms_(x, por, woc[cu]);
            fms_(x, wor, poc[cu]);
            a.set(ru * 4 + cu, x);
            fmax_(rs, x, vc[cu]);                            // y_r += A(r,c) v_c
            cs[cu].x = fma(x.x, vr.x, cs[cu].x);             // y_c += conj(A(r,c)) v_r
            cs[cu].x = fma(x.y, vr.y, cs[cu].x);
            cs[cu].y = fma(x.x, vr.y, cs[cu].y);
            cs[cu].y = fma(x.y, nvrx, cs[cu].y);
        }
        rs.x = sum_lanes_345(rs.x);
        rs.y = sum_lanes_345(rs.y);
        yw[r] = yw[r] + rs;            // the 8 lanes of a row hold the same sum: same-value writes
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        cs[u].x = sum_lanes_012(cs[u].x);
        cs[u].y = sum_lanes_012(cs[u].y);
        yw[cb + u] = yw[cb + u] + cs[u];
    }
}

// Diagonal tile, held as the full Hermitian 32x32 block (both triangles, real
// diagonal): the same update, then y_r += sum_c A(r,c) v_c over the whole block.
template <typename T>
__device__ __forceinline__ void resident_tile_diag(cx<T> (&a)[16], int rb, int cb, int lrg, int lcg,
                                                   const cx<T> *vn, const cx<T> *vo, const cx<T> *wo, cx<T> *yw) {
    using CT = cx<T>;
    const CT z{(T)0, (T)0};
    const bool dsub = lrg == lcg;
    CT vc[4], poc[4], woc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        vc[u] = vn[cb + u];
        poc[u] = conjx(vo[cb + u]);
        woc[u] = conjx(wo[cb + u]);
    }
#pragma unroll
    for (int ru = 0; ru < 4; ru++) {
        const int r = rb + ru;
        const CT por = vo[r], wor = wo[r];
        CT rs = z;
#pragma unroll
        for (int cu = 0; cu < 4; cu++) {
            CT x = a[ru * 4 + cu];
            fms_(x, por, woc[cu]);
            fms_(x, wor, poc[cu]);
            if (ru == cu) x.y = dsub ? (T)0 : x.y;
            a[ru * 4 + cu] = x;
            fmax_(rs, x, vc[cu]);
        }
        rs.x = sum_lanes_345(rs.x);
        rs.y = sum_lanes_345(rs.y);
        yw[r] = yw[r] + rs;
    }
}

// column c (lane group lcg == (c % 32) / 4, entry c % 4, wave-uniform) of one tile to colb
// (0/1 weights rather than branches or selects, which the compiler folds into a
// dynamically indexed tile and moves the register tiles to scratch)
template <typename T, typename Sub>
__device__ __forceinline__ void resident_capture(Sub a, int rb, int cu, cx<T> *colb) {
    T m[4];
#pragma unroll
    for (int k = 0; k < 4; k++) m[k] = (cu == k) ? (T)1 : (T)0;
#pragma unroll
    for (int ru = 0; ru < 4; ru++) {
        cx<T> x{(T)0, (T)0};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const cx<T> e = a.get(ru * 4 + k);
            x.x = fma(m[k], e.x, x.x);
            x.y = fma(m[k], e.y, x.y);
        }
        colb[rb + ru] = x;
    }
}

// one of NQ per-lane values by a wave-uniform index (no dynamic register indexing)
template <typename T, int NQ>
__device__ __forceinline__ cx<T> pick_q(const cx<T> (&v)[NQ], int q) {
    cx<T> r = v[0];
#pragma unroll
    for (int k = 1; k < NQ; k++) r = (q == k) ? v[k] : r;
    return r;
}

// Tile ownership of the register-resident kernel: wave w holds tile (r[w][s], c[w][s])
// in register slot s (-1: empty) and at most one LDS tile (lr[w], LDS buffer lq[w])
// of tile column 0.
struct ResidentDeal {
    signed char r[8][4], c[8][4], lr[8], lq[8];
};

// Deals the TN(TN+1)/2 tiles to 8 waves x NS slots. A tile of tile column c is live
// for the first 32(c+1) columns, so the tiles are placed from the last tile column
// back, each on the SIMD (waves w and w+4 share one) and then the wave with the
// least work so far: at every column the live work is spread evenly. Work units:
// diagonal tile 2 (row sums only), off-diagonal 3. DIAG0: slot 0 holds diagonal
// tiles only (its code path is fixed at compile time), slots 1.. off-diagonal ones.
// Off-diagonal tiles of column 0 that find no register slot go to LDS (at most NL,
// one per wave). Returns false if the tiles do not fit.
static bool resident_deal(int TN, int NS, bool diag0, int NL, int nw, ResidentDeal &d) {
    memset(&d, -1, sizeof(d));
    int load[8] = {0, 0, 0, 0, 0, 0, 0, 0}, used[8] = {0, 0, 0, 0, 0, 0, 0, 0}, nlds = 0;
    bool slot0[8] = {false, false, false, false, false, false, false, false};
    auto better = [&](int a, int b) {   // is wave a less loaded than wave b
        if (b < 0) return true;
        const int sa = load[a % 4] + load[a % 4 + 4], sb = load[b % 4] + load[b % 4 + 4];
        if (sa != sb) return sa < sb;
        if (load[a] != load[b]) return load[a] < load[b];
        return a < b;
    };
    for (int c = TN - 1; c >= 0; c--) {
        for (int r = c; r < TN; r++) {
            const bool diag = r == c;
            int best = -1;
            for (int w = 0; w < nw; w++) {
                const bool fits = diag0 ? (diag ? !slot0[w] : used[w] < NS - 1) : used[w] < NS;
                if (fits && better(w, best)) best = w;
            }
            if (best >= 0) {
                int s;
                if (diag0 && diag) { s = 0; slot0[best] = true; }
                else s = (diag0 ? 1 : 0) + used[best]++;
                d.r[best][s] = (signed char)r;
                d.c[best][s] = (signed char)c;
                load[best] += diag ? 2 : 3;
                continue;
            }
            if (diag || c != 0 || nlds >= NL) return false;
            for (int w = 0; w < nw; w++)
                if (d.lr[w] < 0 && better(w, best)) best = w;
            if (best < 0) return false;
            d.lr[best] = (signed char)r;
            d.lq[best] = (signed char)nlds++;
            load[best] += 3;
        }
    }
    return true;
}

// Tile layouts of the register-resident kernel (its TL parameter). Each gives the
// register storage of one 32x32 tile (R, C) and, for it: the load; the pending rank-2
// update of step i-1 (update; a no-op where product applies it itself); the product
// pass into the wave's partial y, which also captures column i1 = i + 1 into colb; the
// same for a tile of tile column 0 held in LDS (load_lds, col0_lds, step_lds: entry k of
// lane l at p[k * 64], p = tile base + l).

// 4x4 sub-blocks: lane (lrg, lcg) = (lane & 7, lane >> 3) holds tile rows 4 lrg..,
// columns 4 lcg..; update and products in one VALU pass.
template <typename T>
struct ResidentSub4 {
    using CT = cx<T>;
    struct Tile {
        CT a[16];
    };
    template <typename F>
    static __device__ __forceinline__ void load(Tile &t, int R, int C, int lane, int n, F entry) {
        const int rb = R * 32 + 4 * (lane & 7), cb = C * 32 + 4 * (lane >> 3);
#pragma unroll
        for (int cu = 0; cu < 4; cu++)
#pragma unroll
            for (int ru = 0; ru < 4; ru++) {
                const int r = rb + ru, c = cb + cu;
                CT x{(T)0, (T)0};
                if (R >= 0 && r < n && c < n) x = r >= c ? entry(r, c) : conjx(entry(c, r));
                if (r == c) x.y = (T)0;
                t.a[ru * 4 + cu] = x;
            }
    }
    template <typename F>
    static __device__ __forceinline__ void load_lds(CT *l, int R, int lane, int n, F entry) {
        const int rb = R * 32 + 4 * (lane & 7), cb = 4 * (lane >> 3);
#pragma unroll
        for (int cu = 0; cu < 4; cu++)
#pragma unroll
            for (int ru = 0; ru < 4; ru++) {
                const int r = rb + ru, c = cb + cu;
                l[(ru * 4 + cu) * kWave] = r < n ? entry(r, c) : CT{(T)0, (T)0};
            }
    }
    static __device__ __forceinline__ void col0(const Tile &t, int R, int lane, int n, CT *colb) {
        if ((lane >> 3) != 0) return;
#pragma unroll
        for (int ru = 0; ru < 4; ru++) {
            const int r = R * 32 + 4 * (lane & 7) + ru;
            if (r < n) colb[r] = t.a[ru * 4];
        }
    }
    static __device__ __forceinline__ void col0_lds(const CT *l, int R, int lane, int n, CT *colb) {
        if ((lane >> 3) != 0) return;
#pragma unroll
        for (int ru = 0; ru < 4; ru++) {
            const int r = R * 32 + 4 * (lane & 7) + ru;
            if (r < n) colb[r] = l[(ru * 4) * kWave];
        }
    }
    static __device__ __forceinline__ void update(Tile &, bool, int, int, int, int, const CT *, const CT *) {}
    static __device__ __forceinline__ void product(Tile &t, bool diag, int R, int C, int lane, int i1, const CT *vn,
                                                   const CT *vo, const CT *wo, CT *yw, CT *colb) {
        const int lrg = lane & 7, lcg = lane >> 3;
        int rb = R * 32 + 4 * lrg, cb = C * 32 + 4 * lcg;
        asm volatile("" : "+v"(rb), "+v"(cb));       // keep per-slot indices out of loop-invariant hoisting
        if (diag)
            resident_tile_diag<T>(t.a, rb, cb, lrg, lcg, vn, vo, wo, yw);
        else
            resident_tile_off<T>(RegSub<T>{t.a}, rb, cb, lrg, lcg, vn, vo, wo, yw);
        if (C == i1 / 32 && lcg == (i1 % 32) / 4) resident_capture<T>(RegSub<T>{t.a}, rb, i1 & 3, colb);
    }
    static __device__ __forceinline__ void step_lds(CT *l, int R, int lane, int i1, const CT *vn, const CT *vo,
                                                    const CT *wo, CT *yw, CT *colb) {
        const int lrg = lane & 7, lcg = lane >> 3;
        const int rb = R * 32 + 4 * lrg, cb = 4 * lcg;
        resident_tile_off<T>(LdsSub<T>{l}, rb, cb, lrg, lcg, vn, vo, wo, yw);
        if (i1 / 32 == 0 && lcg == (i1 % 32) / 4) resident_capture<T>(LdsSub<T>{l}, rb, i1 & 3, colb);
    }
};

// Sum over the 16 lanes of each DPP row (lane bits 0-3), scattered: lane l returns the
// sum of x[l & 15]. Each step halves the set of values a lane keeps and adds its
// partner's copy of that set: bit 3 by a row rotation of 8 (xor 8), bit 2 by the
// half-row mirror (xor 7), bits 1 and 0 by quad permutations (xor 2, xor 1). The four
// partner sets together reach all 16 lanes of the row.
// The first two steps split the row at 4-lane bank boundaries (bit 3: lanes 0-7 keep
// x[k], lanes 8-15 x[8 + k]; bit 2: banks 0 and 2 keep y[k], banks 1 and 3 y[4 + k]), so
// each kept value is one DPP add restricted by bank_mask (both halves of y[k] from two
// adds) instead of two selects and an add: own + partner in the same order as the select
// form, all lanes active. One s_nop 1 first covers the DPP read-after-VALU-write hazard on
// x; the second step's DPP reads of y come at least eight instructions after their writes.
__device__ __forceinline__ void reduce_scatter16_b32(const float (&x)[16], float (&z)[4]) {
    float y[8];
    asm volatile(
        "s_nop 1\n\t"
        "v_add_f32_dpp %0, %12, %12 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %0, %20, %20 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %1, %13, %13 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %1, %21, %21 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %2, %14, %14 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %2, %22, %22 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %3, %15, %15 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %3, %23, %23 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %4, %16, %16 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %4, %24, %24 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %5, %17, %17 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %5, %25, %25 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %6, %18, %18 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %6, %26, %26 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %7, %19, %19 row_ror:8 row_mask:0xf bank_mask:0x3\n\t"
        "v_add_f32_dpp %7, %27, %27 row_ror:8 row_mask:0xf bank_mask:0xc\n\t"
        "v_add_f32_dpp %8, %0, %0 row_half_mirror row_mask:0xf bank_mask:0x5\n\t"
        "v_add_f32_dpp %8, %4, %4 row_half_mirror row_mask:0xf bank_mask:0xa\n\t"
        "v_add_f32_dpp %9, %1, %1 row_half_mirror row_mask:0xf bank_mask:0x5\n\t"
        "v_add_f32_dpp %9, %5, %5 row_half_mirror row_mask:0xf bank_mask:0xa\n\t"
        "v_add_f32_dpp %10, %2, %2 row_half_mirror row_mask:0xf bank_mask:0x5\n\t"
        "v_add_f32_dpp %10, %6, %6 row_half_mirror row_mask:0xf bank_mask:0xa\n\t"
        "v_add_f32_dpp %11, %3, %3 row_half_mirror row_mask:0xf bank_mask:0x5\n\t"
        "v_add_f32_dpp %11, %7, %7 row_half_mirror row_mask:0xf bank_mask:0xa"
        : "=&v"(y[0]), "=&v"(y[1]), "=&v"(y[2]), "=&v"(y[3]), "=&v"(y[4]), "=&v"(y[5]), "=&v"(y[6]),
          "=&v"(y[7]), "=&v"(z[0]), "=&v"(z[1]), "=&v"(z[2]), "=&v"(z[3])
        : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(x[8]),
          "v"(x[9]), "v"(x[10]), "v"(x[11]), "v"(x[12]), "v"(x[13]), "v"(x[14]), "v"(x[15]));
}
__device__ __forceinline__ float reduce_scatter16(const float (&x)[16], int li) {
    const bool b1 = (li & 2) != 0, b0 = (li & 1) != 0;
    float z[4], w[2];
    reduce_scatter16_b32(x, z);
#pragma unroll
    for (int k = 0; k < 2; k++) w[k] = (b1 ? z[2 + k] : z[k]) + dpp_mov<0x4E>(b1 ? z[k] : z[2 + k]);
    return (b0 ? w[1] : w[0]) + dpp_mov<0xB1>(b0 ? w[0] : w[1]);
}

// Matrix-core layout (complex float): the tile is four 16x16 blocks (p, q) in the
// accumulator layout of v_mfma_f32_16x16x4_f32, real and imaginary parts apart: lane
// (li, lq) = (lane & 15, lane >> 4) holds rows 16 p + 4 lq + u (u = 0..3) of column
// 16 q + li. The pending rank-2 update A -= v w^H + w v^H is a real rank-4 update of
// each part,
//   Re A -= [vr vi wr wi]_r [wr wi vr vi]_c^T,   Im A -= [vi -vr wi -wr]_r [wr wi vr vi]_c^T,
// one 16x16x4 product per block and part, ahead of the VALU products of the same tile.
// Column sums reduce over lq (two lane swaps), row sums over li (reduce_scatter16).
// Diagonal tiles (full Hermitian blocks) add their part of y as column sums,
// y_c = sum_r conj(A(r, c)) v_r. Column blocks q = 0 wholly left of the trailing block
// are skipped: their vectors are zero.
struct ResidentMfma {
    using CT = cx<float>;
    typedef float v4 __attribute__((ext_vector_type(4)));
    struct Tile {
        v4 re[2][2], im[2][2];
    };
    template <typename F>
    static __device__ __forceinline__ void load(Tile &t, int R, int C, int lane, int n, F entry) {
        const int li = lane & 15, lq = lane >> 4;
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 2; q++)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int r = R * 32 + 16 * p + 4 * lq + u, c = C * 32 + 16 * q + li;
                    CT x{0.f, 0.f};
                    if (R >= 0 && r < n && c < n) x = r >= c ? entry(r, c) : conjx(entry(c, r));
                    if (r == c) x.y = 0.f;
                    t.re[p][q][u] = x.x;
                    t.im[p][q][u] = x.y;
                }
    }
    static __device__ __forceinline__ void to_lds(const Tile &t, CT *l) {
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 2; q++)
#pragma unroll
                for (int u = 0; u < 4; u++) l[((p * 2 + q) * 4 + u) * kWave] = CT{t.re[p][q][u], t.im[p][q][u]};
    }
    static __device__ __forceinline__ void from_lds(Tile &t, const CT *l) {
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 2; q++)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const CT x = l[((p * 2 + q) * 4 + u) * kWave];
                    t.re[p][q][u] = x.x;
                    t.im[p][q][u] = x.y;
                }
    }
    template <typename F>
    static __device__ __forceinline__ void load_lds(CT *l, int R, int lane, int n, F entry) {
        Tile t;
        load(t, R, 0, lane, n, entry);
        to_lds(t, l);
    }
    // column c1 (lanes li == c1 % 16 of tile column c1 / 32) to colb; 0/1 weights pick the
    // column block (selects would index the tile dynamically and move it to scratch)
    static __device__ __forceinline__ void capture(const Tile &t, int R, int C, int c1, int lane, CT *colb) {
        if (C != c1 / 32 || (lane & 15) != c1 % 16) return;
        const int lq = lane >> 4;
        const float m1 = (c1 & 16) ? 1.f : 0.f, m0 = 1.f - m1;
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int u = 0; u < 4; u++)
                colb[R * 32 + 16 * p + 4 * lq + u] = CT{fmaf(m1, t.re[p][1][u], m0 * t.re[p][0][u]),
                                                        fmaf(m1, t.im[p][1][u], m0 * t.im[p][0][u])};
    }
    static __device__ __forceinline__ void col0(const Tile &t, int R, int lane, int, CT *colb) {
        capture(t, R, 0, 0, lane, colb);
    }
    static __device__ __forceinline__ void col0_lds(const CT *l, int R, int lane, int n, CT *colb) {
        Tile t;
        from_lds(t, l);
        col0(t, R, lane, n, colb);
    }
    template <int Q0, bool DG>
    static __device__ __forceinline__ void update_q(Tile &t, const float (&ar)[2], const float (&ai)[2],
                                                    const float (&bc)[2]) {
#pragma unroll
        for (int p = DG ? Q0 : 0; p < 2; p++) {
#pragma unroll
            for (int q = Q0; q < 2; q++) {
                t.re[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[p], bc[q], t.re[p][q], 0, 0, 0);
                t.im[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[p], bc[q], t.im[p][q], 0, 0, 0);
            }
        }
    }
    static __device__ __forceinline__ void update(Tile &t, bool diag, int R, int C, int lane, int i1, const CT *vo,
                                                  const CT *wo) {
        const int li = lane & 15, lq = lane >> 4;
        const bool sw = lq >= 2, sy = (lq & 1) != 0;
        float ar[2], ai[2], bc[2];          // operands: k = lq of [vr vi wr wi]_r (negated), of the
#pragma unroll                              // Im row operand, of [wr wi vr vi]_c
        for (int p = 0; p < 2; p++) {
            const int r = R * 32 + 16 * p + li;
            const CT s = sw ? wo[r] : vo[r];
            ar[p] = sy ? -s.y : -s.x;
            ai[p] = sy ? s.x : -s.y;
        }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int c = C * 32 + 16 * q + li;
            const CT s = sw ? vo[c] : wo[c];
            bc[q] = sy ? s.y : s.x;
        }
        // (diag and the column-block skip as template arguments: loop bounds that
        // index the tile must be compile-time constants)
        const bool half = i1 > C * 32 + 15;
        if (diag) {
            if (half) update_q<1, true>(t, ar, ai, bc);
            else update_q<0, true>(t, ar, ai, bc);
        } else {
            if (half) update_q<1, false>(t, ar, ai, bc);
            else update_q<0, false>(t, ar, ai, bc);
        }
    }
    template <int Q0, bool DG>
    static __device__ __forceinline__ void product_q(Tile &t, int R, int C, int lane, const CT *vn, CT *yw) {
        const int li = lane & 15, lq = lane >> 4;
        // the partial-y entries this tile adds to, read before any of its stores: the row
        // band R and the column band C of an off-diagonal tile are disjoint, so both reads
        // issue at once instead of the column read waiting for the row store
        float *yf = reinterpret_cast<float *>(yw);
        const int e = 2 * (R * 32 + 16 * (li >> 3) + 4 * lq + ((li >> 1) & 3)) + (li & 1);
        const float yr0 = DG ? 0.f : yf[e];
        // the column sums end reduce-scattered: lane row lq holds component lq & 1 of column
        // half Q0 + (lq >> 1) (Q0 = 1: rows 2 and 3 repeat rows 0 and 1, entries and values)
        const int ec = 2 * (C * 32 + 16 * (Q0 + ((lq >> 1) & (1 - Q0))) + li) + (lq & 1);
        const float yc0 = yf[ec];
        CT vr[2][4];
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int u = 0; u < 4; u++) vr[p][u] = vn[R * 32 + 16 * p + 4 * lq + u];
        if constexpr (DG) {
            const int dv = li - 4 * lq;       // entry u of block (p, p) is on the diagonal
#pragma unroll
            for (int p = Q0; p < 2; p++)
#pragma unroll
                for (int u = 0; u < 4; u++) t.im[p][p][u] = dv == u ? 0.f : t.im[p][p][u];
        } else {
            float x[16];
#pragma unroll
            for (int p = 0; p < 2; p++)
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    CT rs{0.f, 0.f};
#pragma unroll
                    for (int q = Q0; q < 2; q++) fmax_(rs, CT{t.re[p][q][u], t.im[p][q][u]}, vn[C * 32 + 16 * q + li]);
                    x[(p * 4 + u) * 2] = rs.x;
                    x[(p * 4 + u) * 2 + 1] = rs.y;
                }
            yf[e] = yr0 + reduce_scatter16(x, li);
        }
        CT cs[2];
#pragma unroll
        for (int q = Q0; q < 2; q++) {
            cs[q] = CT{0.f, 0.f};
#pragma unroll
            for (int p = DG ? Q0 : 0; p < 2; p++)
#pragma unroll
                for (int u = 0; u < 4; u++) fmacx(cs[q], CT{t.re[p][q][u], t.im[p][q][u]}, vr[p][u]);
        }
        // sums over the four 16-lane rows, two values per swap: permlane16_swap(x, y) leaves
        // row pairs (x0 + x1, y0 + y1, x2 + x3, y2 + y3), permlane32_swap(a, b) then
        // (a0 + a2, a1 + a3, b0 + b2, b1 + b3): the additions of sum_xor32(sum_xor16(.)) in the
        // same order, with three swaps for the four components instead of eight
        auto pair16 = [](float x, float y) {
            const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
            return __uint_as_float(r[0]) + __uint_as_float(r[1]);
        };
        auto pair32 = [](float a, float b) {
            const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
            return __uint_as_float(r[0]) + __uint_as_float(r[1]);
        };
        if constexpr (Q0 == 0) {
            yf[ec] = yc0 + pair32(pair16(cs[0].x, cs[0].y), pair16(cs[1].x, cs[1].y));
        } else {
            // every lane stores (rows 2 and 3 the same values to the same entries as rows 0
            // and 1): a store under a lane condition would let the swap sink into a branch
            // whose exec mask hides the lanes it reads
            const float s = pair16(cs[1].x, cs[1].y);
            yf[ec] = yc0 + pair32(s, s);
        }
    }
    static __device__ __forceinline__ void product(Tile &t, bool diag, int R, int C, int lane, int i1, const CT *vn,
                                                   const CT *, const CT *, CT *yw, CT *colb) {
        const bool half = i1 > C * 32 + 15;
        if (diag) {
            if (half) product_q<1, true>(t, R, C, lane, vn, yw);
            else product_q<0, true>(t, R, C, lane, vn, yw);
        } else {
            if (half) product_q<1, false>(t, R, C, lane, vn, yw);
            else product_q<0, false>(t, R, C, lane, vn, yw);
        }
        capture(t, R, C, i1, lane, colb);
    }
    static __device__ __forceinline__ void step_lds(CT *l, int R, int lane, int i1, const CT *vn, const CT *vo,
                                                    const CT *wo, CT *yw, CT *colb) {
        Tile t;
        from_lds(t, l);
        update(t, false, R, 0, lane, i1, vo, wo);
        to_lds(t, l);
        product(t, false, R, 0, lane, i1, vn, vo, wo, yw, colb);
    }
};

// Hb (optional): the matrices in the band form of offset hb_off (complex double,
// n x band_nb(n, hb_off) per matrix) instead of A; A then only receives the reflectors.
template <typename T, int NS, int NL, bool DIAG0, int NW, int NMAX, typename TL>
__global__ __launch_bounds__(512) void zhetrd_resident_kernel(T *__restrict__ Abase, long strideA, int lda, int n,
                                                             double *__restrict__ Dout, double *__restrict__ Eout,
                                                             T *__restrict__ TAUout, long strideV, int batch,
                                                             const double *__restrict__ Hb, int hb_off,
                                                             const ResidentDeal deal) {
    using CT = cx<T>;
    constexpr int NT = NW * kWave, TS = 32, NQ = NMAX / kWave;
    constexpr int NLS = NL > 0 ? NL : 1;
    __shared__ CT vbuf[2][NMAX], wbuf[2][NMAX], colb[NMAX], yp[NW][NMAX];
    __shared__ CT ltile[NLS][NL > 0 ? 16 * kWave : 1];
    __shared__ T sd[NMAX], se[NMAX];      // d, e and tau, written to memory at the end
    __shared__ CT stau[NMAX];
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(t / kWave);
    // tiles of this wave (host-dealt, see resident_deal)
    int tR[NS], tC[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        tR[s] = deal.r[wid][s];
        tC[s] = deal.c[wid][s];
    }
    const int lR = NL > 0 ? deal.lr[wid] : -1, lq = lR >= 0 ? deal.lq[wid] : 0;
    CT *lsub = NL > 0 ? &ltile[lq][lane] : nullptr;

    {
        const long b = blockIdx.x;            // one workgroup per matrix (grid = batch)
        if (b >= batch) return;
        T *A = Abase + 2 * b * strideA;
        double *Dv = Dout + b * strideV;
        double *Ev = Eout + b * strideV;
        T *Tv = TAUout + b * strideV * (long)(sizeof(double) / sizeof(T));
        if (n == 1) {
            if (t == 0) Dv[0] = Hb ? Hb[2 * b] : (double)ldx(A, 0).x;
            return;
        }
        const int hnb = Hb ? band_nb(n, hb_off) : 0;
        const double *Hk = Hb ? Hb + 2 * b * (long)n * hnb : nullptr;
        // A(r, c), r >= c: dense, or H[c][b(r-c)] (real on the diagonal)
        auto lower_entry = [&](int r, int c) -> CT {
            if (!Hk) return ldx(A, (long)c * lda + r);
            const int bb = band_b(r - c, n, hb_off);
            if (bb < 0) return CT{(T)0, (T)0};
            const cd h = ld(Hk, (long)c * hnb + bb);
            return CT{(T)h.x, r == c ? (T)0 : (T)h.y};
        };
        // diagonal tiles hold both triangles: A(r, c) = conj(A(c, r)) above the diagonal
        typename TL::Tile a[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) TL::load(a[s], tR[s], tC[s], lane, n, lower_entry);
        if (NL > 0 && lR >= 0) TL::load_lds(lsub, lR, lane, n, lower_entry);
        for (int e = t; e < NMAX; e += NT) {
            vbuf[0][e] = CT{(T)0, (T)0};
            vbuf[1][e] = CT{(T)0, (T)0};
            wbuf[0][e] = CT{(T)0, (T)0};
            wbuf[1][e] = CT{(T)0, (T)0};
        }
        // column 0 to LDS
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (tC[s] == 0) TL::col0(a[s], tR[s], lane, n, colb);
        if (NL > 0 && lR >= 0) TL::col0_lds(lsub, lR, lane, n, colb);
        __syncthreads();

        // wave 0 only: reflector of step i from cq[q] = column i at row lane + 64 q
        // (updated; rows >= i valid). Writes v_i (absolute rows, zero outside i+1..n-1)
        // to vbuf[slot], beta and the scaled vector into A(i+1:n, i), d_i, e_i and tau_i.
        auto reflector = [&](int i, const CT (&cq)[NQ], int slot, CT &tau) {
            const int j1 = i + 1;
            const T dii = lane_bcast(pick_q<T, NQ>(cq, i / kWave).x, i % kWave);
            const CT a1v = pick_q<T, NQ>(cq, j1 / kWave);
            const T ar = lane_bcast(a1v.x, j1 % kWave), ai = lane_bcast(a1v.y, j1 % kWave);
            T s2 = 0;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int r = lane + kWave * q;
                if (r > j1 && r < n) s2 += abs2x(cq[q]);
            }
            const T nrm2 = sum_wave_valu(s2);                   // every lane gets the same sum
            T beta, taur, taui, sr, si;
            if (nrm2 == (T)0 && ai == (T)0) {
                beta = ar; taur = 0; taui = 0; sr = 0; si = 0;
            } else if constexpr (sizeof(T) == sizeof(float)) {
                // hardware sqrt and reciprocal (about 1 ulp; the FP32 reduction's
                // backward error is O(n eps) anyway): a short dependent chain
                const T nrm = __builtin_amdgcn_sqrtf(ar * ar + ai * ai + nrm2);
                beta = (ar >= (T)0) ? -nrm : nrm;
                const T rb = __builtin_amdgcn_rcpf(beta);
                taur = (beta - ar) * rb;
                taui = -ai * rb;
                const T dr = ar - beta, di = ai;
                const T inv = __builtin_amdgcn_rcpf(dr * dr + di * di);
                sr = dr * inv; si = -di * inv;
            } else {
                const T nrm = sqrt(ar * ar + ai * ai + nrm2);
                beta = (ar >= (T)0) ? -nrm : nrm;
                taur = (beta - ar) / beta;
                taui = -ai / beta;
                const T dr = ar - beta, di = ai;
                const T inv = (T)1 / (dr * dr + di * di);
                sr = dr * inv; si = -di * inv;
            }
            const CT scal{sr, si};
            CT *vw = vbuf[slot];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int r = lane + kWave * q;
                CT v{(T)0, (T)0};
                if (r == j1) v = CT{(T)1, (T)0};
                else if (r > j1 && r < n) v = cq[q] * scal;
                vw[r] = v;
            }
            if (lane == 0) {   // the reflector goes to memory from the next pass (off this wave's path)
                sd[i] = dii;
                se[i] = beta;
                stau[i] = CT{taur, taui};
            }
            tau = CT{taur, taui};
        };

        CT tau{(T)0, (T)0};   // tau of the current step (wave 0)
        if (wid == 0) {
            CT cq[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int r = lane + kWave * q;
                cq[q] = r < n ? colb[r] : CT{(T)0, (T)0};
            }
            reflector(0, cq, 0, tau);
        }
        __syncthreads();

        for (int i = 0; i < n - 1; i++) {
            const int cur = i & 1, i1 = i + 1;
            const CT *vn = vbuf[cur];                            // v_i (absolute rows)
            const CT *vo = vbuf[cur ^ 1], *wo = wbuf[cur ^ 1];   // v_{i-1}, w_{i-1}
            CT *yw = yp[wid];
            for (int e = lane; e < NMAX; e += kWave) yw[e] = CT{(T)0, (T)0};
            if (wid == (i + 1) % NW) {        // reflector i into A(i+1:n, i) (beta at i+1), a wave in turn
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int r = lane + kWave * q;
                    if (r > i && r < n) stx(A, (long)i * lda + r, r == i1 ? CT{se[i], (T)0} : vn[r]);
                }
            }
            // per live tile, its update then its products (one basic block per tile: the other
            // wave of the SIMD fills the VALU while this one's update runs; 111.3 -> 107.8 ms at
            // n = 256 against issuing every tile's update first)
#pragma unroll
            for (int s = 0; s < NS; s++) {
                if (tR[s] < 0 || tC[s] * TS + TS - 1 < i1) continue;   // wave-uniform
                const bool dg = DIAG0 ? s == 0 : tR[s] == tC[s];
                TL::update(a[s], dg, tR[s], tC[s], lane, i1, vo, wo);
                TL::product(a[s], dg, tR[s], tC[s], lane, i1, vn, vo, wo, yw, colb);
            }
            if (NL > 0 && lR >= 0 && i1 < TS) TL::step_lds(lsub, lR, lane, i1, vn, vo, wo, yw, colb);
            __syncthreads();
            if (wid == 0) {
                // every LDS read up front and unconditional (all rows < NMAX), then
                // masks by select: one LDS round trip instead of one per branch
                CT p[NQ], cq[NQ], vv[NQ], cb[NQ];
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int r = lane + kWave * q;
                    CT y = yp[0][r];
#pragma unroll
                    for (int w = 1; w < NW; w++) y = y + yp[w][r];
                    p[q] = y;
                    vv[q] = vn[r];
                    cb[q] = colb[r];
                }
                CT dl{(T)0, (T)0};
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int r = lane + kWave * q;
                    p[q] = (r >= i1 && r < n) ? tau * p[q] : CT{(T)0, (T)0};
                    fmacx(dl, p[q], vv[q]);
                }
                dl.x = sum_wave_valu(dl.x);
                dl.y = sum_wave_valu(dl.y);
                const CT alpha2 = (tau * dl) * (T)(-0.5);
                CT *wn = wbuf[cur];
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int r = lane + kWave * q;
                    p[q] = (r >= i1 && r < n) ? p[q] + alpha2 * vv[q] : CT{(T)0, (T)0};   // w_i
                    wn[r] = p[q];
                }
                const CT w0s = pick_q<T, NQ>(p, i1 / kWave);
                const CT w0{lane_bcast(w0s.x, i1 % kWave), lane_bcast(w0s.y, i1 % kWave)};
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int r = lane + kWave * q;
                    CT x = cb[q] - (vv[q] * conjx(w0) + p[q]);          // v_i(i+1) = 1
                    if (r == i1) x.y = (T)0;
                    cq[q] = (r >= i1 && r < n) ? x : CT{(T)0, (T)0};
                }
                if (i1 < n - 1) reflector(i1, cq, cur ^ 1, tau);
                else if (lane == (n - 1) % kWave) sd[n - 1] = pick_q<T, NQ>(cq, (n - 1) / kWave).x;
            }
            __syncthreads();
        }
        for (int e = t; e < n; e += NT) {
            Dv[e] = (double)sd[e];
            if (e < n - 1) {
                Ev[e] = (double)se[e];
                Tv[2 * e] = stau[e].x;
                Tv[2 * e + 1] = stau[e].y;
            }
        }
    }
}

// Small matrices (n <= 32 RPT, n <= 8 CPT), register-resident in full storage: one
// 256-thread workgroup per matrix (persistent over the batch), thread t holding A(r, c)
// for rows r = t % 32 + 32 k and columns c = t / 32 + 8 q, so that several matrices share a
// CU (the fused kernel keeps them in L2 and streams the trailing block twice per column).
// Per column i: the column's owners write it to LDS; wave 0 builds reflector i (zlarfg,
// as the other kernels); every thread forms its part of y = A_i v_i (row sums over its
// columns, then over the 8 column groups); wave 0 forms w_i = tau y - 1/2 tau (tau y)^H v v;
// every thread applies A -= v w^H + w v^H to its entries. Same outputs as zhetrd_kernel.
template <typename T, int RPT, int CPT>
__global__ __launch_bounds__(256) void zhetrd_small_kernel(T *__restrict__ Abase, long strideA, int lda, int n,
                                                          double *__restrict__ Dout, double *__restrict__ Eout,
                                                          T *__restrict__ TAUout, long strideV, int batch) {
    using CT = cx<T>;
    constexpr int MAXN = 32 * RPT, NQV = (MAXN + kWave - 1) / kWave;
    __shared__ CT col[MAXN], vsh[MAXN], wsh[MAXN], ypart[4][MAXN];
    __shared__ T sc[2];
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    const int rl = t & 31, cg = t >> 5;                      // row lane, column group
    for (long b = blockIdx.x; b < batch; b += gridDim.x) {
        T *A = Abase + 2 * b * strideA;
        double *Dv = Dout + b * strideV;
        double *Ev = Eout + b * strideV;
        T *Tv = TAUout + b * strideV * (long)(sizeof(double) / sizeof(T));
        CT a[RPT][CPT];
#pragma unroll
        for (int k = 0; k < RPT; k++)
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int r = rl + 32 * k, c = cg + 8 * q;
                a[k][q] = (r < n && c < n) ? ldx(A, (long)c * lda + r) : CT{(T)0, (T)0};
                if (r == c) a[k][q].y = (T)0;
            }
        for (int i = 0; i < n - 1; i++) {
            const int i1 = i + 1;
            // (1) column i (updated through step i-1), rows >= i, to LDS; 0/1 weights rather
            // than selects, which the compiler would fold into a dynamically indexed
            // (scratch-resident) array
            if (cg == (i & 7)) {
                const int qi = i >> 3;
                T m[CPT];
#pragma unroll
                for (int q = 0; q < CPT; q++) m[q] = (q == qi) ? (T)1 : (T)0;
#pragma unroll
                for (int k = 0; k < RPT; k++) {
                    CT x{(T)0, (T)0};
#pragma unroll
                    for (int q = 0; q < CPT; q++) {
                        x.x = fma(m[q], a[k][q].x, x.x);
                        x.y = fma(m[q], a[k][q].y, x.y);
                    }
                    const int r = rl + 32 * k;
                    if (r >= i && r < n) col[r] = x;
                }
            }
            __syncthreads();
            // (2) reflector i from col[i..n-1] (wave 0); v (absolute rows, zero outside
            // i+1..n-1) to vsh, beta and the scaled vector into A(i+1:n, i)
            if (wid == 0) {
                T s2 = 0;
                for (int r = i + 2 + lane; r < n; r += kWave) s2 += abs2x(col[r]);
                const T nrm2 = sum_wave_valu(s2);
                const T dii = col[i].x;
                const T ar = col[i1].x, ai = col[i1].y;
                T beta, taur, taui, sr, si;
                if (nrm2 == (T)0 && ai == (T)0) {
                    beta = ar; taur = 0; taui = 0; sr = 0; si = 0;
                } else {
                    const T nrm = sqrt(ar * ar + ai * ai + nrm2);
                    beta = (ar >= (T)0) ? -nrm : nrm;
                    taur = (beta - ar) / beta;
                    taui = -ai / beta;
                    const T dr = ar - beta, di = ai;
                    const T inv = (T)1 / (dr * dr + di * di);
                    sr = dr * inv; si = -di * inv;
                }
                const CT scal{sr, si};
                for (int r = lane; r < MAXN; r += kWave) {
                    CT v{(T)0, (T)0};
                    if (r == i1) {
                        v = CT{(T)1, (T)0};
                        stx(A, (long)i * lda + r, CT{beta, (T)0});
                    } else if (r > i1 && r < n) {
                        v = col[r] * scal;
                        stx(A, (long)i * lda + r, v);
                    }
                    vsh[r] = v;
                }
                if (lane == 0) {
                    Dv[i] = (double)dii;
                    Ev[i] = (double)beta;
                    Tv[2 * i] = taur;
                    Tv[2 * i + 1] = taui;
                    sc[0] = taur;
                    sc[1] = taui;
                }
            }
            __syncthreads();
            // (3) y = A_i v: row sums over this thread's columns, then over the column groups
            {
                CT y[RPT];
#pragma unroll
                for (int k = 0; k < RPT; k++) y[k] = CT{(T)0, (T)0};
#pragma unroll
                for (int q = 0; q < CPT; q++) {
                    const int c = cg + 8 * q;
                    const CT vc = c < MAXN ? vsh[c] : CT{(T)0, (T)0};
#pragma unroll
                    for (int k = 0; k < RPT; k++) fmax_(y[k], a[k][q], vc);
                }
#pragma unroll
                for (int k = 0; k < RPT; k++) {
                    y[k].x += __shfl_xor(y[k].x, 32, kWave);        // the wave's two column groups
                    y[k].y += __shfl_xor(y[k].y, 32, kWave);
                    if (lane < 32) ypart[wid][rl + 32 * k] = y[k];
                }
            }
            __syncthreads();
            // (4) w_i (wave 0)
            if (wid == 0) {
                const CT tau{sc[0], sc[1]};
                CT pv[NQV], vv[NQV];
                CT dl{(T)0, (T)0};
#pragma unroll
                for (int k = 0; k < NQV; k++) {
                    const int r = lane + kWave * k;
                    CT p{(T)0, (T)0}, v{(T)0, (T)0};
                    if (r < MAXN) {
                        p = ((ypart[0][r] + ypart[1][r]) + ypart[2][r]) + ypart[3][r];
                        v = vsh[r];
                    }
                    p = (r > i && r < n) ? tau * p : CT{(T)0, (T)0};
                    fmacx(dl, p, v);
                    pv[k] = p;
                    vv[k] = v;
                }
                dl.x = sum_wave_valu(dl.x);
                dl.y = sum_wave_valu(dl.y);
                const CT alpha2 = (tau * dl) * (T)(-0.5);
#pragma unroll
                for (int k = 0; k < NQV; k++) {
                    const int r = lane + kWave * k;
                    if (r < MAXN) wsh[r] = (r > i && r < n) ? pv[k] + alpha2 * vv[k] : CT{(T)0, (T)0};
                }
            }
            __syncthreads();
            // (5) A -= v w^H + w v^H on this thread's entries (v, w vanish outside i+1..n-1)
            {
                CT vr[RPT], wr[RPT];
#pragma unroll
                for (int k = 0; k < RPT; k++) {
                    vr[k] = vsh[rl + 32 * k];
                    wr[k] = wsh[rl + 32 * k];
                }
#pragma unroll
                for (int q = 0; q < CPT; q++) {
                    const int c = cg + 8 * q;
                    const CT wc = c < MAXN ? conjx(wsh[c]) : CT{(T)0, (T)0};
                    const CT vc = c < MAXN ? conjx(vsh[c]) : CT{(T)0, (T)0};
#pragma unroll
                    for (int k = 0; k < RPT; k++) {
                        CT x = a[k][q];
                        fms_(x, vr[k], wc);
                        fms_(x, wr[k], vc);
                        if (rl + 32 * k == c) x.y = (T)0;           // Hermitian diagonal stays real
                        a[k][q] = x;
                    }
                }
            }
        }
        // last diagonal element
        {
            const int r = n - 1;
            if (rl == (r & 31) && cg == (r & 7)) {
                const int kk = r >> 5, qq = r >> 3;
                T x = 0;
#pragma unroll
                for (int k = 0; k < RPT; k++)
#pragma unroll
                    for (int q = 0; q < CPT; q++) x = fma((k == kk && q == qq) ? (T)1 : (T)0, a[k][q].x, x);
                Dv[n - 1] = (double)x;
            }
        }
        __syncthreads();
    }
}

// ============================================================================
// 2. Divide & conquer on the symmetric tridiagonal + 3. back-transform
// ============================================================================
struct DcShared {
    double *dq, *dqn, *ds, *zs, *dl, *wz, *z2, *tau, *vals, *rc, *rs, *what, *inrm;
    double *red;   // 8 doubles of reduction scratch
    int *perm, *orig, *rank, *rpj, *rnj, *ndx, *dfx;
    int *pcol, *lstL, *lstR, *rflag;  // eigenvector-update gather index, per-half column lists, rotated flags
    int *cnt;  // per node: K, ndefl, nrot, |lstL|, |lstR|  (5 ints per node slot)
};

__device__ inline double dc_eps() { return 1.1102230246251565e-16; }  // 2^-53
// Deflation threshold unit: 2^-53 for FP64 eigenvector matrices; 2^-30 for the FP32
// ones of the mixed-precision solver, whose tridiagonal carries FP32 rounding anyway and
// whose eigenvalues are replaced by FP64 Rayleigh quotients (a deflated z_i perturbs T by
// at most 8 * 2^-30 of its scale; divide & conquer at n_phi = 256: 69.2 -> 66.5 ms, at
// n_phi = 3000: 426 -> 327 ms; energies and Chern integers unchanged against FP64; 2^-26
// measured no faster).
template <typename QS>
__device__ inline double defl_eps() {
    return sizeof(QS) == sizeof(float) ? 0x1p-30 : dc_eps();
}

// 1/x from the hardware reciprocal estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r) (each step squares the relative error); x must be finite
// and non-zero, which holds for every pole distance taken below.
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
// One Newton-Raphson step: the estimate measures within 4.6e-8 (2^-24.4) relative on
// MI355X, so one step leaves ~2^-48, enough for the FP32 eigenvector matrices of the
// mixed-precision solver and for its secular roots at 2^-32.
__device__ inline double rcp_nr1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
// the reciprocal for eigenvector matrices of storage type QS
template <typename QS>
__device__ inline double rcp_q(double x) {
    return sizeof(QS) == sizeof(float) ? rcp_nr1(x) : rcp_nr(x);
}

// Root j of 1 + sum_i z2_i / (dl_i - lambda) = 0 with z2_i = rho w_i^2, poles dl
// ascending (sum_i w_i^2 = 1), to relative accuracy teps (2^-53 for FP64
// eigenvectors; 2^-32 for the FP32 eigenvector matrices of the mixed-precision
// solver, whose vectors carry FP32 accuracy and whose eigenvalues are replaced by FP64
// Rayleigh quotients: 4.3 iterations per root at 2^-53, 3.9 at 2^-40, 3.6 at 2^-32 for
// n = 256 random Hermitian matrices; the plans' mixed energies stay within 2.2e-12 of FP64).
// Returns origin index o and tau with lambda = dl[o] + tau.
// [jlo, jhi] must contain j; the callers pass the range of roots their wave solves at once,
// so the pole loops split there uniformly (see below).
__device__ void secular_root(int j, int K, const double *dl, const double *z2, double rho, int &o_out,
                             double &tau_out, double teps, int jlo, int jhi) {
    const bool one_nr = teps > dc_eps();      // see rcp_nr1
    int o;
    double lo, hi, tau;
    if (j < K - 1) {
        const double Delta = dl[j + 1] - dl[j];
        // f at the midpoint, origin dl[j]
        double f = 1.0;
        const double tm = 0.5 * Delta;
#pragma unroll 4
        for (int i = 0; i < K; i++) {
            const double del = (dl[i] - dl[j]) - tm;
            f += z2[i] * (one_nr ? rcp_nr1(del) : rcp_nr(del));
        }
        if (f >= 0.0) { o = j; lo = 0.0; hi = tm; tau = tm; }
        else { o = j + 1; lo = -tm; hi = 0.0; tau = -tm; }
    } else {
        o = j; lo = 0.0; hi = rho; tau = rho;
    }
    const double dlo = dl[o];
    const double Dj = dl[j] - dlo;
    const double Dj1 = (j < K - 1) ? (dl[j + 1] - dlo) : 0.0;
    const int jm = min(jhi, K - 1);
    for (int it = 0; it < 120; it++) {
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0, err;
        // poles at or below j feed psi, the others phi (all terms of one sign each). The
        // wave's roots lie in [jlo, jhi]: poles below jlo feed psi and poles above jhi phi
        // for every lane; in between (at most 64 poles) the sign of the pole distance
        // decides, which is exact: tau stays strictly inside (0, D/2] (origin j) or
        // [-D/2, 0) (origin j + 1), so (dl_i - dl_o) - tau < 0 exactly for i <= j and > 0
        // for i > j. The other sum then gains +0, so psi, phi and their derivatives are
        // bit-identical to splitting each lane at its own j, and the wave runs K pole
        // iterations instead of K + (jhi - jlo).
        int i = 0;
#pragma unroll 4
        for (; i < jlo; i++) {
            const double inv = one_nr ? rcp_nr1((dl[i] - dlo) - tau) : rcp_nr((dl[i] - dlo) - tau);
            const double tt = z2[i] * inv;
            psi += tt;
            dpsi = fma(tt, inv, dpsi);
        }
#pragma unroll 2
        for (; i <= jm; i++) {
            const double inv = one_nr ? rcp_nr1((dl[i] - dlo) - tau) : rcp_nr((dl[i] - dlo) - tau);
            const double tt = z2[i] * inv;
            const double tn = fmin(tt, 0.0), tp = fmax(tt, 0.0);
            psi += tn;
            phi += tp;
            dpsi = fma(tn, inv, dpsi);
            dphi = fma(tp, inv, dphi);
        }
#pragma unroll 4
        for (; i < K; i++) {
            const double inv = one_nr ? rcp_nr1((dl[i] - dlo) - tau) : rcp_nr((dl[i] - dlo) - tau);
            const double tt = z2[i] * inv;
            phi += tt;
            dphi = fma(tt, inv, dphi);
        }
        // sum of |terms|: the psi terms are all <= 0 (poles below the root), the phi terms
        // all >= 0, so it is phi - psi without a third accumulation
        err = phi - psi;
        const double delj = (dl[j] - dlo) - tau;
        const double delj1 = (j + 1 < K) ? (dl[j + 1] - dlo) - tau : 0.0;
        const double f = 1.0 + psi + phi;
        if (f > 0.0) hi = fmin(hi, tau);
        else if (f < 0.0) lo = fmax(lo, tau);
        if (f == 0.0 || fabs(f) <= 8.0 * teps * (1.0 + err)) break;
        if (hi - lo <= 4.0 * teps * fmax(fabs(lo), fabs(hi))) break;
        const double bq = dpsi * delj * delj;
        const double aq = psi - dpsi * delj;
        double nt;
        if (j < K - 1) {
            const double Bq = dphi * delj1 * delj1;
            const double Aq = phi - dphi * delj1;
            const double c = 1.0 + aq + Aq;
            if (o == j) {
                const double beta = c * Dj1 + bq + Bq;
                nt = 2.0 * bq * Dj1 / (beta + sqrt(fmax(beta * beta - 4.0 * c * bq * Dj1, 0.0)));
            } else {
                const double beta = c * Dj + bq + Bq;
                nt = 2.0 * Bq * Dj / (beta + sqrt(fmax(beta * beta - 4.0 * c * Bq * Dj, 0.0)));
            }
        } else {
            const double c = 1.0 + aq + phi;
            nt = (c > 0.0) ? bq / c : hi + 1.0;  // force bisection when the model fails
        }
        if (!(nt > lo && nt < hi)) nt = 0.5 * (lo + hi);
        tau = nt;
    }
    o_out = o;
    tau_out = tau;
}

// Per-item steps of a merge (node at lo with K secular roots, m = K + nd values), shared
// by stedc_block and the multi-workgroup secular phase (stedc_secular_kernel).
// Root j: origin pole index, tau and the origin pole rc (root j = rc[j] + tau[j]).
// [jlo, jhi]: a range that contains j (see secular_root).
template <typename QS>
__device__ inline void dc_root(int j, int K, int lo, double rho, DcShared &sh, int jlo, int jhi) {
    int o;
    double tu;
    secular_root(j, K, sh.dl + lo, sh.z2 + lo, rho, o, tu, sizeof(QS) == sizeof(float) ? 0x1p-32 : dc_eps(), jlo,
                 jhi);
    sh.orig[lo + j] = o;
    sh.tau[lo + j] = tu;
    sh.rc[lo + j] = sh.dl[lo + o];
}
// Gu-Eisenstat recomputed weight i (two partial products; the j == i factor is 1; the pole
// distances are (dl_i - rc_j) - tau_j).
template <typename QS>
__device__ inline void dc_weight(int i, int K, int lo, DcShared &sh) {
    const double *dl = sh.dl + lo, *ro = sh.rc + lo, *ta = sh.tau + lo;
    const double di = dl[i];
    double pa = (di - ro[i]) - ta[i], pb = 1.0;
    int j = 0;
    for (; j + 1 < K; j += 2) {
        const bool sa = j == i, sb = j + 1 == i;
        const double fa = ((di - ro[j]) - ta[j]) * rcp_q<QS>(sa ? 1.0 : di - dl[j]);
        const double fb = ((di - ro[j + 1]) - ta[j + 1]) * rcp_q<QS>(sb ? 1.0 : di - dl[j + 1]);
        pa *= sa ? 1.0 : fa;
        pb *= sb ? 1.0 : fb;
    }
    if (j < K && j != i) pa *= ((di - ro[j]) - ta[j]) * rcp_q<QS>(di - dl[j]);
    const double p = pa * pb;
    sh.what[lo + i] = copysign(sqrt(fmax(-p, 0.0)), sh.wz[lo + i]);
}
// Inverse norm of eigenvector j (two partial sums).
template <typename QS>
__device__ inline void dc_norm(int j, int K, int lo, DcShared &sh) {
    const double *dl = sh.dl + lo, *wh = sh.what + lo;
    const double oj = sh.rc[lo + j];
    const double tj = sh.tau[lo + j];
    double sa = 0.0, sb = 0.0;
    int i = 0;
    for (; i + 1 < K; i += 2) {
        const double ua = wh[i] * rcp_q<QS>((dl[i] - oj) - tj);
        const double ub = wh[i + 1] * rcp_q<QS>((dl[i + 1] - oj) - tj);
        sa += ua * ua;
        sb += ub * ub;
    }
    if (i < K) {
        const double ua = wh[i] * rcp_q<QS>((dl[i] - oj) - tj);
        sa += ua * ua;
    }
    sh.inrm[lo + j] = 1.0 / sqrt(sa + sb);
}
// Value q of the merged node: root q < K, deflated value q - K after them.
__device__ inline void dc_value(int q, int K, int lo, DcShared &sh) {
    sh.vals[lo + q] = q < K ? sh.rc[lo + q] + sh.tau[lo + q] : sh.ds[lo + sh.dfx[lo + q - K]];
}
// Rank of value q among the node's m values (ties by index) and its place in dqn.
__device__ inline void dc_rank(int q, int m, int lo, DcShared &sh, double *dqn) {
    const double v = sh.vals[lo + q];
    int rk = 0;
    for (int p = 0; p < m; p++) {
        const double u = sh.vals[lo + p];
        rk += (u < v) || (u == v && p < q);
    }
    sh.rank[lo + q] = rk;
    dqn[lo + rk] = v;
}

// Number of elements of the ascending array a[0..len) that are < v (strict) or <= v.
__device__ inline int count_below(const double *a, int len, double v, bool inclusive) {
    int lo = 0, hi = len;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        bool go = inclusive ? (a[mid] <= v) : (a[mid] < v);
        if (go) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Matrix-core merge GEMM of one node, Qb = Qa S (see stedc_block): task = 16 RB rows
// (RB row blocks) x 16 QT output columns (QT column tiles); each gathered Qa value
// feeds QT MFMAs, each generated S value RB. S[i][q] = what_i / (dl_i - lambda_q) *
// inrm_q is generated in FP64 per K-step and never stored; every row group of a node
// generates it anew, so RB = 4, QT = 2 halves that FP64 work against RB = 2, QT = 4
// (same accumulator registers, twice the Qa loads) where the halves have >= 64 rows. QS: storage and matrix-core type of the eigenvector matrices (double; float in
// the mixed-precision solver, whose eigenvectors carry the FP32 accuracy of its
// reflectors anyway).
// Uo (top merge only, lo = 0): the products go straight to the component-major complex
// output, Uo[r * n + rank] = (value, 0), instead of column rank of Qb.
template <typename QS, int RB, int QT>
__device__ __forceinline__ void merge_gemm_mfma(int lane, int G, int lo, int mid, int hi, int n1, int K, int KL, int KR,
                                                int n, const QS *Qa, QS *Qb, const DcShared &sh,
                                                QS *Uo = nullptr) {
    using MQ = Mf<QS>;
    using v4 = typename MQ::v4;
    constexpr int RH = 16 * RB;
    const int wv = lane / kWave, nwv = G / kWave;
    const int l = lane & (kWave - 1), li = l & 15, lq = l >> 4;
    const double *dl = sh.dl + lo;
    const int nqt = (K + 15) / 16, nqr = (nqt + QT - 1) / QT;
    const int rows1 = hi - mid;                       // right half (may be short)
    const int ng0 = (n1 + RH - 1) / RH, ng1 = (rows1 + RH - 1) / RH;
    const int ntask = (ng0 + ng1) * nqr;
    for (int task = wv; task < ntask; task += nwv) {
        const int qr = task % nqr, g = task / nqr;
        const bool right = g >= ng0;
        const int rbase = right ? mid + (g - ng0) * RH : lo + g * RH;
        const int rend = right ? hi : mid;
        const int cnt = right ? KR : KL;
        const int *lst = (right ? sh.lstR : sh.lstL) + lo;
        const int qt0 = qr * QT, nq = min(QT, nqt - qt0);     // wave-uniform
        double qo[QT], qtu[QT], qiv[QT];
#pragma unroll
        for (int u = 0; u < QT; u++) {
            const int q = (qt0 + u) * 16 + li;
            const bool ok = u < nq && q < K;
            qo[u] = ok ? sh.rc[lo + q] : 0.0;
            qtu[u] = ok ? sh.tau[lo + q] : 0.0;
            qiv[u] = ok ? sh.inrm[lo + q] : 0.0;          // 0 marks an empty column
        }
        v4 acc[RB][QT];
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int u = 0; u < QT; u++) acc[rb][u] = v4{0, 0, 0, 0};
        // K-step operands one step ahead: list entry, weights and the gathered Qa
        // values of step idx0 + 4 load while step idx0 runs on the matrix cores
        bool kin_n;
        double wi_n, di_n;
        QS av_n[RB];
        auto fetch = [&](int idx) {
            kin_n = idx < cnt;
            const int i = kin_n ? lst[idx] : 0;
            wi_n = sh.what[lo + i];
            di_n = dl[i];
            const QS *col = Qa + (long)(lo + sh.pcol[lo + i]) * n;
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const int r = rbase + rb * 16 + li;
                av_n[rb] = (kin_n && r < rend) ? col[r] : (QS)0;
            }
        };
        fetch(lq);
        for (int idx0 = 0; idx0 < cnt; idx0 += 4) {
            const bool kin = kin_n;
            const double wi = wi_n, di = di_n;
            QS av[RB];
#pragma unroll
            for (int rb = 0; rb < RB; rb++) av[rb] = av_n[rb];
            if (idx0 + 4 < cnt) fetch(idx0 + 4 + lq);
#pragma unroll
            for (int u = 0; u < QT; u++) {
                if (u < nq) {
                    const double bv0 = wi * rcp_q<QS>((di - qo[u]) - qtu[u]) * qiv[u];
                    const QS bv = (kin && qiv[u] != 0.0) ? (QS)bv0 : (QS)0;
#pragma unroll
                    for (int rb = 0; rb < RB; rb++) acc[rb][u] = MQ::mma(av[rb], bv, acc[rb][u]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < QT; u++) {
            const int q = (qt0 + u) * 16 + li;
            if (u < nq && q < K) {
                const int oc = lo + sh.rank[lo + q];
                QS *ocol = Qb + (long)oc * n;
#pragma unroll
                for (int rb = 0; rb < RB; rb++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int r = rbase + rb * 16 + MQ::row(l, v);
                        if (r < rend) {
                            if (Uo) stx(Uo, (long)r * n + oc, cx<QS>{acc[rb][u][v], (QS)0});
                            else ocol[r] = acc[rb][u][v];
                        }
                    }
            }
        }
    }
}

// Stage 1: D&C. d (n), e (n-1) in; eigenvalues ascending in sh.dq (returned via
// pointer swap) and eigenvectors of T in the returned buffer (row-major n x n).
// QS: storage type of the eigenvector matrices Qa, Qb and of S (see merge_gemm_mfma);
// every other quantity is FP64.
// lev_stop > 0: stop at merge level lev_stop once its nodes are merged, deflated and
// rotated and the secular inputs are set (nullptr returned, the scale left in sh.red[0]);
// its roots, weights, norms and ranks (stedc_secular_kernel) and its eigenvector GEMM
// (stedc_level_gemm_kernel, or stedc_top_gemm_kernel for the top level, which writes the
// outputs) then run on several workgroups per matrix.
// lev_begin > 1: resume at that level from the state a stopped call and the GEMM kernel
// left in sh, Qa / Qb (the caller passes the workspace's first and second matrix; the
// buffers and eigenvalue arrays swap after every level) and the scale in sh.red[0].
// Uout (optional): where the top merge takes the matrix-core path, its products and
// deflated columns go straight to the component-major complex output Uout[r * n + j]
// (*direct set), which saves a transposing pass over the eigenvectors.
template <int NT, typename QS>
__device__ QS *stedc_block(int n, const double *dIn, const double *eIn, QS *Qa, QS *Qb, QS *S, DcShared &sh,
                           double **dq_final, int lev_stop = 0, QS *Uout = nullptr, bool *direct = nullptr,
                           int lev_begin = 1) {
    const int t = threadIdx.x;
    double *dq = sh.dq, *dqn = sh.dqn;
    double scale;
    if (lev_begin > 1) {
        scale = sh.red[0];
        if (!(lev_begin & 1)) {
            QS *tq = Qa; Qa = Qb; Qb = tq;
            double *td = dq; dq = dqn; dqn = td;
        }
    } else {
    // scale T to unit max-norm (as LAPACK dstedc does) so that the deflation
    // tolerance is relative to the matrix, whatever its magnitude
    double amax = 0.0;
    for (int k = t; k < n; k += NT) {
        amax = fmax(amax, fabs(dIn[k]));
        if (k < n - 1) amax = fmax(amax, fabs(eIn[k]));
    }
    for (int off = kWave / 2; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off, kWave));
    if ((t & (kWave - 1)) == 0) sh.red[t / kWave] = amax;
    __syncthreads();
    amax = 0.0;
    for (int w = 0; w < NT / kWave; w++) amax = fmax(amax, sh.red[w]);
    __syncthreads();
    scale = (amax > 0.0) ? amax : 1.0;
    const double is0 = 1.0 / scale;
    // init: torn diagonal (scaled), identity / zero
    for (int k = t; k < n; k += NT) {
        double v = dIn[k] * is0;
        if (k > 0) v -= fabs(eIn[k - 1] * is0);
        if (k < n - 1) v -= fabs(eIn[k] * is0);
        sh.dq[k] = v;
    }
    // Qa, Qb: eigenvector matrices of the current level, column-major (column c =
    // eigenvector c, Q[c*n + r])
    for_2d(t, NT, n, n, [&](int c, int r) {
        const long idx = (long)c * n + r;
        Qa[idx] = (r == c) ? (QS)1 : (QS)0;
        Qb[idx] = (QS)0;
    });
    __syncthreads();
    }
    const double iscale = 1.0 / scale;
    for (int L = lev_begin; (1 << (L - 1)) < n; L++) {
        const int half = 1 << (L - 1), blk = half << 1;
        const int nnodes = (n + blk - 1) / blk;
        int G = 1;
        while (G * 2 <= NT / nnodes && G * 2 <= blk) G *= 2;
        // halves of >= 32 rows merge on the matrix cores, which take a whole wave per node:
        // fewer nodes per round rather than the VALU product
        if (half >= 32 && G < kWave) G = kWave;
        if (G > NT) G = NT;
        const int per = NT / G;  // nodes handled per round
        for (int nb0 = 0; nb0 < nnodes; nb0 += per) {
            const int k = nb0 + t / G;
            const int lane = t % G;
            const bool active = k < nnodes;
            const int lo = k * blk, mid = lo + half, hi = min(lo + blk, n);
            const bool merging = active && mid < n;
            const int m = hi - lo, n1 = mid - lo;
            const double beta = merging ? eIn[mid - 1] * iscale : 0.0;
            const double sgn = (beta >= 0.0) ? 1.0 : -1.0;
            const double rho = 2.0 * fabs(beta);
            const double isq2 = 0.70710678118654752440;
            // P1: merge the two sorted halves; gather z
            if (merging) {
                for (int i = lane; i < m; i += G) {
                    double val, zi;
                    int pos;
                    if (i < n1) {
                        val = dq[lo + i];
                        pos = i + count_below(dq + mid, hi - mid, val, false);
                        zi = Qa[(long)(lo + i) * n + mid - 1];
                    } else {
                        const int j = i - n1;
                        val = dq[mid + j];
                        pos = j + count_below(dq + lo, n1, val, true);
                        zi = sgn * Qa[(long)(mid + j) * n + mid];
                    }
                    sh.ds[lo + pos] = val;
                    sh.zs[lo + pos] = zi * isq2;
                    sh.perm[lo + pos] = i;
                    sh.rflag[lo + pos] = 0;
                }
            }
            __syncthreads();
            // P2: deflation scan. Nodes with a whole wave: a parallel pass (ballot
            // compaction) that is exact whenever no Givens rotation is needed, which it
            // detects with the same test on the same values as the sequential scan (the
            // first rotation of that scan always involves untouched entries); the
            // sequential scan (one lane) runs otherwise, and for smaller nodes.
            bool seq = merging && lane == 0;
            if (merging && G >= kWave && lane < kWave) {
                double dmx = 0.0, zmx = 0.0;
                for (int e = lane; e < m; e += kWave) {
                    dmx = fmax(dmx, fabs(sh.ds[lo + e]));
                    zmx = fmax(zmx, fabs(sh.zs[lo + e]));
                }
                for (int off = kWave / 2; off > 0; off >>= 1) {
                    dmx = fmax(dmx, __shfl_xor(dmx, off, kWave));
                    zmx = fmax(zmx, __shfl_xor(zmx, off, kWave));
                }
                const double tol = 8.0 * defl_eps<QS>() * fmax(dmx, zmx);
                const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
                int K = 0, nd = 0, KL = 0, KR = 0;
                bool have = false, conflict = false;
                double pz = 0.0, pd = 0.0;                    // last kept entry of earlier chunks
                for (int c0 = 0; c0 < m; c0 += kWave) {
                    const int e = c0 + lane;
                    const bool valid = e < m;
                    const double z = valid ? sh.zs[lo + e] : 0.0, d = valid ? sh.ds[lo + e] : 0.0;
                    const bool defl = valid && rho * fabs(z) <= tol;
                    const bool keep = valid && !defl;
                    const unsigned long long km = __ballot(keep);
                    const unsigned long long lower = km & below;
                    const int pl = lower ? 63 - __clzll(lower) : 0;
                    double qz = __shfl(z, pl, kWave), qd = __shfl(d, pl, kWave);
                    if (!lower) { qz = pz; qd = pd; }
                    if (keep && (lower != 0 || have)) {
                        double S_ = qz, C_ = z;
                        const double tz = hypot(C_, S_);
                        const double tt = d - qd;
                        C_ /= tz;
                        S_ = -S_ / tz;
                        if (fabs(tt * C_ * S_) <= tol) conflict = true;
                    }
                    const int pc = valid ? sh.perm[lo + e] : 0;
                    const unsigned long long dm = __ballot(defl);
                    const unsigned long long lm = __ballot(keep && pc < n1);
                    const unsigned long long rm = __ballot(keep && pc >= n1);
                    if (keep) {
                        const int kq = K + __popcll(lower);
                        sh.ndx[lo + kq] = e;
                        sh.pcol[lo + kq] = pc;
                        if (pc < n1) sh.lstL[lo + KL + __popcll(lm & below)] = kq;
                        else sh.lstR[lo + KR + __popcll(rm & below)] = kq;
                    }
                    if (defl) sh.dfx[lo + nd + __popcll(dm & below)] = e;
                    if (km) {
                        const int last = 63 - __clzll(km);
                        pz = __shfl(z, last, kWave);
                        pd = __shfl(d, last, kWave);
                        have = true;
                    }
                    K += __popcll(km);
                    nd += __popcll(dm);
                    KL += __popcll(lm);
                    KR += __popcll(rm);
                }
                const bool any = __ballot(conflict) != 0;
                if (!any && lane == 0) {
                    sh.cnt[5 * k + 0] = K;
                    sh.cnt[5 * k + 1] = nd;
                    sh.cnt[5 * k + 2] = 0;
                    sh.cnt[5 * k + 3] = KL;
                    sh.cnt[5 * k + 4] = KR;
                }
                seq = any && lane == 0;
            }
            if (seq) {
                double dmax = 0.0, zmax = 0.0;
                for (int i = 0; i < m; i++) {
                    dmax = fmax(dmax, fabs(sh.ds[lo + i]));
                    zmax = fmax(zmax, fabs(sh.zs[lo + i]));
                }
                const double tol = 8.0 * defl_eps<QS>() * fmax(dmax, zmax);
                int K = 0, nd = 0, nr = 0, pj = -1, KL = 0, KR = 0;
                // a non-deflated column q: record its node-local source column and which
                // halves of the node's rows it can be non-zero in (a rotation mixes halves)
                auto append = [&](int sj) {
                    const int pc = sh.perm[lo + sj];
                    sh.ndx[lo + K] = sj;
                    sh.pcol[lo + K] = pc;
                    if (sh.rflag[lo + sj] || pc < n1) sh.lstL[lo + KL++] = K;
                    if (sh.rflag[lo + sj] || pc >= n1) sh.lstR[lo + KR++] = K;
                    K++;
                };
                for (int nj = 0; nj < m; nj++) {
                    if (rho * fabs(sh.zs[lo + nj]) <= tol) {
                        sh.dfx[lo + nd++] = nj;
                        continue;
                    }
                    if (pj < 0) { pj = nj; continue; }
                    double S_ = sh.zs[lo + pj], C_ = sh.zs[lo + nj];
                    const double tz = hypot(C_, S_);
                    const double tt = sh.ds[lo + nj] - sh.ds[lo + pj];
                    C_ /= tz;
                    S_ = -S_ / tz;
                    if (fabs(tt * C_ * S_) <= tol) {
                        sh.zs[lo + nj] = tz;
                        sh.zs[lo + pj] = 0.0;
                        sh.rpj[lo + nr] = pj;
                        sh.rnj[lo + nr] = nj;
                        sh.rflag[lo + pj] = 1;
                        sh.rflag[lo + nj] = 1;
                        sh.rc[lo + nr] = C_;
                        sh.rs[lo + nr] = S_;
                        nr++;
                        const double dp = sh.ds[lo + pj], dn = sh.ds[lo + nj];
                        const double t1 = dp * C_ * C_ + dn * S_ * S_;
                        sh.ds[lo + nj] = dp * S_ * S_ + dn * C_ * C_;
                        sh.ds[lo + pj] = t1;
                        sh.dfx[lo + nd++] = pj;
                        pj = nj;
                    } else {
                        append(pj);
                        pj = nj;
                    }
                }
                if (pj >= 0) append(pj);
                sh.cnt[5 * k + 0] = K;
                sh.cnt[5 * k + 1] = nd;
                sh.cnt[5 * k + 2] = nr;
                sh.cnt[5 * k + 3] = KL;
                sh.cnt[5 * k + 4] = KR;
            }
            __syncthreads();
            // P3: apply deflation rotations to Qa's node block; secular inputs
            int K = 0, nd = 0, nr = 0;
            int KL = 0, KR = 0;
            if (merging) {
                K = sh.cnt[5 * k]; nd = sh.cnt[5 * k + 1]; nr = sh.cnt[5 * k + 2];
                KL = sh.cnt[5 * k + 3]; KR = sh.cnt[5 * k + 4];
            }
            if (merging) {
                if (nr > 0) {
                    for (int r = lo + lane; r < hi; r += G) {
                        for (int q = 0; q < nr; q++) {
                            const int cp = sh.perm[lo + sh.rpj[lo + q]], cn = sh.perm[lo + sh.rnj[lo + q]];
                            const double c = sh.rc[lo + q], s = sh.rs[lo + q];
                            QS *xp = Qa + (long)(lo + cp) * n + r, *yp = Qa + (long)(lo + cn) * n + r;
                            const double x = *xp, y = *yp;
                            *xp = (QS)(c * x + s * y);
                            *yp = (QS)(c * y - s * x);
                        }
                    }
                }
                for (int q = lane; q < K; q += G) {
                    const double zq = sh.zs[lo + sh.ndx[lo + q]];
                    sh.dl[lo + q] = sh.ds[lo + sh.ndx[lo + q]];
                    sh.wz[lo + q] = zq;
                    sh.z2[lo + q] = rho * zq * zq;
                }
            }
            __syncthreads();
            // a stopped level's roots, weights, norms and ranks and its GEMM run on several
            // workgroups per matrix (stedc_secular_kernel, stedc_level_gemm_kernel or
            // stedc_top_gemm_kernel); uniform over the workgroup
            if (L == lev_stop) continue;
            // P4: secular roots
            // a wave solves the roots j - lane % W .. + W - 1 together (W = min(G, wave):
            // node groups are aligned to G threads)
            if (merging) {
                const int W = min(G, kWave);
                for (int j = lane; j < K; j += G) dc_root<QS>(j, K, lo, rho, sh, j - lane % W, j - lane % W + W - 1);
            }
            __syncthreads();
            // P5: Gu-Eisenstat recomputed weights
            if (merging)
                for (int i = lane; i < K; i += G) dc_weight<QS>(i, K, lo, sh);
            __syncthreads();
            // P6: eigenvector norms; output values and ranks
            if (merging) {
                for (int j = lane; j < K; j += G) dc_norm<QS>(j, K, lo, sh);
                for (int q = lane; q < m; q += G) dc_value(q, K, lo, sh);
            }
            __syncthreads();
            if (merging)
                for (int q = lane; q < m; q += G) dc_rank(q, m, lo, sh, dqn);
            __syncthreads();
            // Matrix-core path for halves of >= 32 rows (whole waves per node): Qb = Qa S
            // with S[i][q] = what_i / (dl_i - lambda_q) * inrm_q generated per K-step,
            // S never stored. Below that, S goes through memory and the VALU path.
            const bool mfma = merging && G >= kWave && half >= 32;
            // 16-row halves, two nodes per wave: each wave runs its nodes' products on the
            // matrix cores one after the other (uniform over the workgroup); at n = 256 this
            // level's product went from 11 % of the divide & conquer (VALU) to 1.4 %
            const bool mfma_pair = G == kWave / 2 && half == 16;
            // P7: U (K x K) into S rows lo.., scaled by the inverse norms
            if (merging && !mfma && !mfma_pair) {
                const double *dl = sh.dl + lo;
                for_2d(lane, G, K, K, [&](int i, int q) {
                    const double del = (dl[i] - sh.rc[lo + q]) - sh.tau[lo + q];
                    S[(long)(lo + i) * n + q] = (QS)(sh.what[lo + i] * rcp_q<QS>(del) * sh.inrm[lo + q]);
                });
            }
            __syncthreads();
            // P8: Qb block = Qa block (rotated, gathered) * U; deflated columns copied.
            // Rows of the left half only meet columns of lstL (left child or rotated), rows
            // of the right half only lstR (the block-diagonal structure dlaed3 exploits).
            // the top merge on the matrix cores writes the output itself (top_direct is
            // uniform over the workgroup; Uo is set on the lanes of node 0)
            const bool top_direct = Uout && blk >= n && G >= kWave && half >= 32;
            QS *Uo = (mfma && top_direct) ? Uout : nullptr;
            if (mfma) {
                if (sizeof(QS) == sizeof(float) && half >= 64)
                    merge_gemm_mfma<QS, 4, 2>(lane, G, lo, mid, hi, n1, K, KL, KR, n, Qa, Qb, sh, Uo);
                else
                    merge_gemm_mfma<QS, 2, 4>(lane, G, lo, mid, hi, n1, K, KL, KR, n, Qa, Qb, sh, Uo);
            } else if (mfma_pair) {
                const int wv = t / kWave;
                for (int kk = 0; kk < 2; kk++) {
                    const int kn = nb0 + 2 * wv + kk;              // wave-uniform
                    const int lo2 = kn * blk, mid2 = lo2 + half;
                    if (kn >= nnodes || mid2 >= n) continue;       // no merge (copied below)
                    merge_gemm_mfma<QS, 1, 4>(t & (kWave - 1), kWave, lo2, mid2, min(lo2 + blk, n), half,
                                              sh.cnt[5 * kn], sh.cnt[5 * kn + 3], sh.cnt[5 * kn + 4], n, Qa, Qb, sh);
                }
            } else if (merging) {
                // VALU path: each task is one row and 4 consecutive output columns,
                // consecutive lanes on consecutive rows
                const int nq4 = (K + 3) >> 2;
                for (int task = lane; task < m * nq4; task += G) {
                    const int rloc = task % m, q0 = (task / m) * 4;
                    const int r = lo + rloc;
                    const bool left = rloc < n1;
                    const int *lst = left ? sh.lstL + lo : sh.lstR + lo;
                    const int cntl = left ? KL : KR;
                    const int qa = q0, qb = min(q0 + 1, K - 1), qc = min(q0 + 2, K - 1), qd = min(q0 + 3, K - 1);
                    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll 4
                    for (int idx = 0; idx < cntl; idx++) {
                        const int i = lst[idx];
                        const double x = Qa[(long)(lo + sh.pcol[lo + i]) * n + r];
                        const QS *srow = S + (long)(lo + i) * n;
                        a0 = fma(x, srow[qa], a0);
                        a1 = fma(x, srow[qb], a1);
                        a2 = fma(x, srow[qc], a2);
                        a3 = fma(x, srow[qd], a3);
                    }
                    Qb[(long)(lo + sh.rank[lo + q0]) * n + r] = (QS)a0;
                    if (q0 + 1 < K) Qb[(long)(lo + sh.rank[lo + q0 + 1]) * n + r] = (QS)a1;
                    if (q0 + 2 < K) Qb[(long)(lo + sh.rank[lo + q0 + 2]) * n + r] = (QS)a2;
                    if (q0 + 3 < K) Qb[(long)(lo + sh.rank[lo + q0 + 3]) * n + r] = (QS)a3;
                }
            }
            if (Uo) {
                // deflated columns, consecutive lanes on consecutive output columns of a row
                for_2d(lane, G, m, nd, [&](int r, int qq) {
                    stx(Uo, (long)r * n + sh.rank[K + qq], cx<QS>{Qa[(long)sh.perm[sh.dfx[qq]] * n + r], (QS)0});
                });
            } else if (merging) {
                for_2d(lane, G, nd, m, [&](int qq, int rloc) {
                    const int r = lo + rloc;
                    Qb[(long)(lo + sh.rank[lo + K + qq]) * n + r] = Qa[(long)(lo + sh.perm[lo + sh.dfx[lo + qq]]) * n + r];
                });
            } else if (active) {
                for_2d(lane, G, m, m, [&](int cl, int rl) {
                    const long e = (long)(lo + cl) * n + lo + rl;
                    Qb[e] = Qa[e];
                });
                for (int q = lane; q < m; q += G) dqn[lo + q] = dq[lo + q];
            }
            if (top_direct) *direct = true;
            __syncthreads();
        }
        if (L == lev_stop) {                        // every round's deflation done: stop
            if (t == 0) sh.red[0] = scale;
            *dq_final = dqn;
            return nullptr;
        }
        QS *tq = Qa; Qa = Qb; Qb = tq;
        double *td = dq; dq = dqn; dqn = td;
    }
    for (int k = t; k < n; k += NT) dq[k] *= scale;
    __syncthreads();
    *dq_final = dq;
    return Qa;
}

// Stage 1 kernel: divide & conquer per matrix; eigenvalues to W, eigenvectors of
// the tridiagonal to U (component-major complex, zero imaginary part).
// GSH: the per-node arrays live in the workgroup's global workspace after Qa, Qb, S
// (n > 512, where they exceed the LDS); the LDS then holds only the transpose tile.
// Node arrays of the divide & conquer at p (13 double arrays of n, 8 doubles of
// reduction scratch, 11 int arrays of n and 5 ints per node slot).
__device__ inline void dc_shared_at(double *p, int n, DcShared &sh) {
    sh.dq = p; p += n; sh.dqn = p; p += n; sh.ds = p; p += n; sh.zs = p; p += n;
    sh.dl = p; p += n; sh.wz = p; p += n; sh.z2 = p; p += n; sh.tau = p; p += n; sh.vals = p; p += n;
    sh.rc = p; p += n; sh.rs = p; p += n; sh.what = p; p += n; sh.inrm = p; p += n;
    sh.red = p; p += 8;
    int *ip = reinterpret_cast<int *>(p);
    sh.perm = ip; ip += n; sh.orig = ip; ip += n; sh.rank = ip; ip += n; sh.rpj = ip; ip += n;
    sh.rnj = ip; ip += n; sh.ndx = ip; ip += n; sh.dfx = ip; ip += n;
    sh.pcol = ip; ip += n; sh.lstL = ip; ip += n; sh.lstR = ip; ip += n; sh.rflag = ip; ip += n;
    sh.cnt = ip; ip += 5 * n;
}

// Component-major complex output U[r*n + j] = Z[j*n + r] (column-major Z) through 32x32
// LDS tiles.
template <typename T, int NT>
__device__ inline void dc_transpose_out(int n, const T *Z, T *U, double *tile) {
    using C = cx<T>;
    const int t = threadIdx.x;
    for (int tr = 0; tr < n; tr += 32)
        for (int tc = 0; tc < n; tc += 32) {
            for (int e = t; e < 1024; e += NT) {
                const int jj = e >> 5, rr = e & 31;            // read column tc+jj, row tr+rr
                if (tc + jj < n && tr + rr < n) tile[rr * 33 + jj] = Z[(long)(tc + jj) * n + tr + rr];
            }
            __syncthreads();
            for (int e = t; e < 1024; e += NT) {
                const int rr = e >> 5, jj = e & 31;            // write row tr+rr, column tc+jj
                if (tc + jj < n && tr + rr < n) stx(U, (long)(tr + rr) * n + tc + jj, C{(T)tile[rr * 33 + jj], (T)0});
            }
            __syncthreads();
        }
}

// Stage 1 kernel: divide & conquer per matrix; eigenvalues to W, eigenvectors of
// the tridiagonal to U (component-major complex, zero imaginary part).
// GSH: the per-node arrays live in the workgroup's global workspace after Qa, Qb, S
// (n > 512, where they exceed the LDS); the LDS then holds only the transpose tile.
// lev_stop > 0 (GSH only): stop before merge level lev_stop's GEMM (see stedc_block);
// workgroup b then handles matrix b (grid = batch) and keeps its state in workspace slot b.
// n <= 512: registers capped for three waves per SIMD (three workgroups per CU, a few
// registers spilled): n_phi = 256, 32768 matrices 58 -> 51 ms (complex float output),
// 78 -> 69 ms (complex double); two waves per SIMD measured no faster than one.
template <typename T, int NT, bool GSH>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GSH ? 1 : 3))) void stedc_kernel(int n, const double *__restrict__ Din,
                                                   const double *__restrict__ Ein, long strideV,
                                                   double *__restrict__ Wout, long strideW, T *__restrict__ Uout,
                                                   long strideU, double *__restrict__ work, int batch,
                                                   int lev_stop) {
    extern __shared__ double smem[];
    const int t = threadIdx.x;
    // per-workgroup global workspace: Qa, Qb, S (n*n double slots each, holding the
    // eigenvector matrices in the precision T of the output) [+ the node arrays]
    const long wsz = GSH ? stedc_global_work_doubles(n) : 3L * n * n;
    double *wg = work + (long)blockIdx.x * wsz;
    T *Qa0 = reinterpret_cast<T *>(wg);
    T *Qb0 = reinterpret_cast<T *>(wg + (long)n * n);
    T *S = reinterpret_cast<T *>(wg + 2L * n * n);
    DcShared sh;
    dc_shared_at(GSH ? wg + 3L * n * n : smem, n, sh);
    for (long b = blockIdx.x; b < batch; b += gridDim.x) {
        const double *dIn = Din + b * strideV;
        const double *eIn = Ein + b * strideV;
        double *W = Wout + b * strideW;
        T *U = Uout + 2 * b * strideU;
        double *dqf;
        bool direct = false;
        T *Z = stedc_block<NT, T>(n, dIn, eIn, Qa0, Qb0, S, sh, &dqf, GSH ? lev_stop : 0, U, &direct);
        if (!Z) return;                    // split: the top merge continues in other kernels
        for (int k = t; k < n; k += NT) W[k] = dqf[k];
        __syncthreads();
        // the D&C arrays are free now; the launch sizes LDS for the tile
        if (!direct) dc_transpose_out<T, NT>(n, Z, U, smem);
    }
}

// Level of the top merge of stedc_block (blk = 2^L >= n), and which buffer holds its
// input eigenvector matrix (Qa0 at odd levels, Qb0 at even ones: the buffers swap after
// every level) and its eigenvalues (dqn: sh.dqn at odd levels, sh.dq at even ones).
__device__ inline int dc_top_level(int n) {
    int L = 1;
    while ((1 << L) < n) L++;
    return L;
}

// A stopped divide & conquer resumed at merge level lev (its secular roots, weights,
// norms and ranks) after stedc_level_gemm_kernel of level lev - 1; stops again before the
// level's GEMM. One workgroup per matrix of the round (workspace slot = blockIdx.x).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void stedc_resume_kernel(int n, const double *__restrict__ Din,
                                                          const double *__restrict__ Ein, long strideV,
                                                          double *__restrict__ work, int lev) {
    const long b = blockIdx.x;
    double *wg = work + b * stedc_global_work_doubles(n);
    DcShared sh;
    dc_shared_at(wg + 3L * n * n, n, sh);
    T *Qa0 = reinterpret_cast<T *>(wg), *Qb0 = reinterpret_cast<T *>(wg + (long)n * n);
    T *S = reinterpret_cast<T *>(wg + 2L * n * n);
    double *dqf;
    (void)stedc_block<NT, T>(n, Din + b * strideV, Ein + b * strideV, Qa0, Qb0, S, sh, &dqf, lev, nullptr, nullptr,
                             lev);
}

// The secular phase of a stopped merge level L, nwg workgroups per matrix, the items of
// all merging nodes dealt over them: phase 0 the roots (dc_root), phase 1 the weights and
// the output values (dc_weight, dc_value), phase 2 the norms and the ranks (dc_norm,
// dc_rank), one launch each. Same arithmetic as stedc_block's single-workgroup phases.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void stedc_secular_kernel(int n, const double *__restrict__ Ein, long strideV,
                                                           double *__restrict__ work, int nwg, int L, int phase) {
    const long b = blockIdx.x / nwg;
    const int part = blockIdx.x % nwg;
    double *wg = work + b * stedc_global_work_doubles(n);
    DcShared sh;
    dc_shared_at(wg + 3L * n * n, n, sh);
    const double *eIn = Ein + b * strideV;
    double *dqn = (L & 1) ? sh.dqn : sh.dq;
    const double iscale = 1.0 / sh.red[0];
    const int half = 1 << (L - 1), blk = half << 1, nnodes = (n + blk - 1) / blk;
    const int G = nwg * NT;
    // item idx of the level: per merging node K items (phase 0) or m items (phases 1, 2)
    for (int idx = part * NT + (int)threadIdx.x;; idx += G) {
        int k = 0, it = idx, lo = 0, mid = 0, m = 0, K = 0;
        for (; k < nnodes; k++) {
            lo = k * blk;
            mid = lo + half;
            if (mid >= n) continue;
            m = min(lo + blk, n) - lo;
            K = sh.cnt[5 * k];
            const int cnt = phase == 0 ? K : m;
            if (it < cnt) break;
            it -= cnt;
        }
        if (k >= nnodes) break;
        if (phase == 0) {
            // the wave's 64 consecutive items lie in it - lane .. it - lane + 63 when the wave
            // does not cross a node boundary (a valid range for each lane either way)
            const int wl = (int)(threadIdx.x & (kWave - 1));
            dc_root<T>(it, K, lo, 2.0 * fabs(eIn[mid - 1] * iscale), sh, it - wl, it - wl + kWave - 1);
        } else if (phase == 1) {
            if (it < K) dc_weight<T>(it, K, lo, sh);
            dc_value(it, K, lo, sh);
        } else {
            if (it < K) dc_norm<T>(it, K, lo, sh);
            dc_rank(it, m, lo, sh, dqn);
        }
    }
}

// The eigenvector products Qb = Qa S of every node of merge level L (below the top), nwg
// workgroups per matrix (each node's wave tasks dealt over all of them), with the
// nodes' deflated columns, and the copies of the nodes that do not merge at this level.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void stedc_level_gemm_kernel(int n, double *__restrict__ work, int nwg, int L) {
    const long b = blockIdx.x / nwg;
    const int part = blockIdx.x % nwg;
    double *wg = work + b * stedc_global_work_doubles(n);
    DcShared sh;
    dc_shared_at(wg + 3L * n * n, n, sh);
    T *Qa0 = reinterpret_cast<T *>(wg), *Qb0 = reinterpret_cast<T *>(wg + (long)n * n);
    const T *Qa = (L & 1) ? Qa0 : Qb0;
    T *Qb = (L & 1) ? Qb0 : Qa0;
    const double *dq = (L & 1) ? sh.dq : sh.dqn;
    double *dqn = (L & 1) ? sh.dqn : sh.dq;
    const int half = 1 << (L - 1), blk = half << 1, nnodes = (n + blk - 1) / blk;
    const int lane = part * NT + threadIdx.x, G = nwg * NT;
    for (int k = 0; k < nnodes; k++) {
        const int lo = k * blk, mid = lo + half, hi = min(lo + blk, n), m = hi - lo;
        if (mid < n) {
            const int K = sh.cnt[5 * k], nd = sh.cnt[5 * k + 1], KL = sh.cnt[5 * k + 3], KR = sh.cnt[5 * k + 4];
            if (sizeof(T) == sizeof(float) && half >= 64)
                merge_gemm_mfma<T, 4, 2>(lane, G, lo, mid, hi, half, K, KL, KR, n, Qa, Qb, sh);
            else
                merge_gemm_mfma<T, 2, 4>(lane, G, lo, mid, hi, half, K, KL, KR, n, Qa, Qb, sh);
            for_2d(lane, G, nd, m, [&](int qq, int rloc) {
                const int r = lo + rloc;
                Qb[(long)(lo + sh.rank[lo + K + qq]) * n + r] = Qa[(long)(lo + sh.perm[lo + sh.dfx[lo + qq]]) * n + r];
            });
        } else {
            for_2d(lane, G, m, m, [&](int cl, int rl) {
                const long e = (long)(lo + cl) * n + lo + rl;
                Qb[e] = Qa[e];
            });
            for (int q = lane; q < m; q += G) dqn[lo + q] = dq[lo + q];
        }
    }
}

// The top merge's product of stedc_block, nwg workgroups per matrix (the wave tasks of
// merge_gemm_mfma dealt over all of them), written with its deflated columns straight to
// the component-major output U; the eigenvalues (scaled back) to W.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void stedc_top_gemm_kernel(int n, double *__restrict__ work, int nwg,
                                                            double *__restrict__ Wout, long strideW,
                                                            T *__restrict__ Uout, long strideU) {
    const long b = blockIdx.x / nwg;
    const int part = blockIdx.x % nwg;
    double *wg = work + b * stedc_global_work_doubles(n);
    DcShared sh;
    dc_shared_at(wg + 3L * n * n, n, sh);
    const int L = dc_top_level(n);
    T *Qa0 = reinterpret_cast<T *>(wg), *Qb0 = reinterpret_cast<T *>(wg + (long)n * n);
    const T *Qa = (L & 1) ? Qa0 : Qb0;
    T *Qb = (L & 1) ? Qb0 : Qa0;
    T *U = Uout + 2 * b * strideU;
    const int half = 1 << (L - 1), lo = 0, mid = half, hi = n, n1 = half, m = n;
    const int K = sh.cnt[0], nd = sh.cnt[1], KL = sh.cnt[3], KR = sh.cnt[4];
    const int lane = part * NT + threadIdx.x, G = nwg * NT;
    if (sizeof(T) == sizeof(float) && half >= 64)
        merge_gemm_mfma<T, 4, 2>(lane, G, lo, mid, hi, n1, K, KL, KR, n, Qa, Qb, sh, U);
    else
        merge_gemm_mfma<T, 2, 4>(lane, G, lo, mid, hi, n1, K, KL, KR, n, Qa, Qb, sh, U);
    for_2d(lane, G, m, nd, [&](int r, int qq) {
        stx(U, (long)r * n + sh.rank[K + qq], cx<T>{Qa[(long)sh.perm[sh.dfx[qq]] * n + r], (T)0});
    });
    if (part == 0) {
        const double *dqn = (L & 1) ? sh.dqn : sh.dq;
        const double scale = sh.red[0];
        double *W = Wout + b * strideW;
        for (int k = threadIdx.x; k < n; k += NT) W[k] = dqn[k] * scale;
    }
}

// Stage 2 kernel: back-transform U = Q U with compact-WY blocks of nb_wy reflectors
// (I - V T V^H), last block first, one thread per eigenvector column.
template <typename T, int NT, int NBW>
__global__ __launch_bounds__(NT) void backtransform_kernel(const T *__restrict__ Abase, long strideA, int lda, int n,
                                                           const T *__restrict__ TAUin, long strideV,
                                                           T *__restrict__ Uout, long strideU, int nb_wy,
                                                           int batch) {
    using C = cx<T>;
    extern __shared__ double smem[];
    const int t = threadIdx.x;
    for (long b = blockIdx.x; b < batch; b += gridDim.x) {
        const T *A = Abase + 2 * b * strideA;
        const T *tauv = TAUin + b * strideV * (long)(sizeof(double) / sizeof(T));
        T *U = Uout + 2 * b * strideU;
        // back-transform with blocks of reflectors, last block first
        C *Vl = reinterpret_cast<C *>(smem);          // rows x nb
        const int nref = n - 1;
        const int nblk = (nref + nb_wy - 1) / nb_wy;
        C *Tl = Vl + (long)n * nb_wy;                 // nb x nb
        for (int bk = nblk - 1; bk >= 0; bk--) {
            const int j0 = bk * nb_wy;
            const int nbk = min(nb_wy, nref - j0);
            const int r0 = j0 + 1;
            const int nrow = n - r0;
            // V block: V[(r - r0)*nb + k] for reflector j0+k
            for (int idx = t; idx < nrow * nb_wy; idx += NT) {
                const int rr = idx / nb_wy, k = idx % nb_wy;
                const int r = r0 + rr, j = j0 + k;
                C v{(T)0, (T)0};
                if (k < nbk) {
                    if (r == j + 1) v = C{(T)1, (T)0};
                    else if (r > j + 1) v = ldx(A, (long)j * lda + r);
                }
                Vl[idx] = v;
            }
            __syncthreads();
            // T (upper triangular): T[i][k] stored Tl[i*nb+k]; first Gram products V_i^H V_k
            for (int idx = t; idx < nb_wy * nb_wy; idx += NT) {
                const int i = idx / nb_wy, k = idx % nb_wy;
                C g{(T)0, (T)0};
                if (i < k && k < nbk) {
                    for (int rr = 0; rr < nrow; rr++) fmacx(g, Vl[rr * nb_wy + i], Vl[rr * nb_wy + k]);
                }
                Tl[idx] = g;
            }
            __syncthreads();
            if (t == 0) {
                // T[0:k,k] = -tau_k * T[0:k,0:k] * G[0:k,k]; T[k][k] = tau_k (zlarft, forward, columnwise)
                C *tmp = Tl + nb_wy * nb_wy;   // nb scratch entries in LDS
                for (int k = 0; k < nbk; k++) {
                    const C tk{tauv[2 * (j0 + k)], tauv[2 * (j0 + k) + 1]};
                    for (int i = 0; i < k; i++) {
                        C s{(T)0, (T)0};
                        for (int l = i; l < k; l++) fmax_(s, Tl[i * nb_wy + l], Tl[l * nb_wy + k]);
                        tmp[i] = s;
                    }
                    for (int i = 0; i < k; i++) Tl[i * nb_wy + k] = (tk * tmp[i]) * (T)(-1);
                    Tl[k * nb_wy + k] = tk;
                }
                for (int k = nbk; k < nb_wy; k++)
                    for (int i = 0; i < nb_wy; i++) Tl[i * nb_wy + k] = C{(T)0, (T)0};
            }
            __syncthreads();
            // apply to every column: w = V^H u ; w = T w ; u -= V w
            for (int c = t; c < n; c += NT) {
                C w[NBW];
#pragma unroll
                for (int k = 0; k < NBW; k++) w[k] = C{(T)0, (T)0};
                for (int rr = 0; rr < nrow; rr++) {
                    const C u = ldx(U, (long)(r0 + rr) * n + c);
#pragma unroll
                    for (int k = 0; k < NBW; k++)
                        if (k < nb_wy) fmacx(w[k], Vl[rr * nb_wy + k], u);
                }
                C w2[NBW];
#pragma unroll
                for (int i = 0; i < NBW; i++) {
                    C s{(T)0, (T)0};
#pragma unroll
                    for (int k = 0; k < NBW; k++)
                        if (i < nb_wy && k >= i && k < nb_wy) fmax_(s, Tl[i * nb_wy + k], w[k]);
                    w2[i] = s * (T)(-1);
                }
                for (int rr = 0; rr < nrow; rr++) {
                    C u = ldx(U, (long)(r0 + rr) * n + c);
#pragma unroll
                    for (int k = 0; k < NBW; k++)
                        if (k < nb_wy) fmax_(u, Vl[rr * nb_wy + k], w2[k]);
                    stx(U, (long)(r0 + rr) * n + c, u);
                }
            }
            __syncthreads();
        }
    }
}

// ----------------------------------------------------------------------------
// Back-transform on the matrix cores (n <= 256; complex float n <= 512), with the MFMA helpers Mf<T> of
// common.hpp.
// Element (row, k) of the reflector block starting at reflector j0 (zhetrd lower
// storage: unit at row j+1, A(j+2:n, j) below, zero above; zero for j >= n-1).
// Branch-free: every lane loads from a clamped in-bounds address (a 32-bit offset: the
// callers have n, lda <= 4096) and selects zero above the reflector, the unit entry on it;
// per-lane branches here cost an exec-mask save, branch and restore per load.
// refl_raw: the (clamped) stored entry of reflector j at row; refl_sel: V(row, j) from it.
template <typename T>
__device__ inline cx<T> refl_raw(const T *A, int lda, int n, int j, int row) {
    const int jc = min(j, max(n - 2, 0)), rc = min(max(row, 0), n - 1);
    return ldx(A, jc * lda + rc);
}
template <typename T>
__device__ inline cx<T> refl_sel(cx<T> x, int n, int j, int row) {
    const bool live = j < n - 1;
    const T re = (live && row > j + 1 && row < n) ? x.x : (live && row == j + 1 ? (T)1 : (T)0);
    const T im = (live && row > j + 1 && row < n) ? x.y : (T)0;
    return cx<T>{re, im};
}
template <typename T>
__device__ inline cx<T> refl_get(const T *A, int lda, int n, int j0, int row, int k) {
    return refl_sel(refl_raw(A, lda, n, j0 + k, row), n, j0 + k, row);
}
// The same by per-lane branches (exec-masked loads): for a register prefetch, whose value
// is not used until later and which a select after the load would make wait for it.
template <typename T>
__device__ inline cx<T> refl_get_masked(const T *A, int lda, int n, int j0, int row, int k) {
    const int j = j0 + k;
    if (j >= n - 1 || row <= j || row >= n) return cx<T>{(T)0, (T)0};
    if (row == j + 1) return cx<T>{(T)1, (T)0};
    return ldx(A, j * lda + row);
}

// Triangular factors of the compact-WY blocks (zlarft, forward, columnwise):
// block bk of matrix b covers reflectors 16bk..16bk+15; T[k][k] = tau_k,
// T[0:k, k] = -tau_k T[0:k, 0:k] G[0:k, k] with G = V^H V (one MFMA product per
// block). Stored column-major per block: Tout[b*strideT + (bk*256 + k2*16 + k1)]
// (complex). One wave per block, 4 waves per workgroup, wpm workgroups per matrix
// (workgroup q of a matrix takes blocks 4q + w, 4q + w + 4 wpm, ...).
template <typename T>
__global__ __launch_bounds__(256) void larft_kernel(const T *__restrict__ Abase, long strideA, int lda, int n,
                                                    const T *__restrict__ TAUin, long strideV, T *__restrict__ Tout,
                                                    long strideT, int batch, int wpm) {
    using M = Mf<T>;
    using v4 = typename M::v4;
    using CT = cx<T>;
    __shared__ CT gs[4][16 * 17];
    __shared__ CT ts[4][16 * 17];
    const long b = blockIdx.x / wpm;
    const int wq = (int)(blockIdx.x % wpm);
    if (b >= batch) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const T *A = Abase + 2 * b * strideA;
    const T *tau = TAUin + b * strideV * (long)(sizeof(double) / sizeof(T));
    CT *Tb = reinterpret_cast<CT *>(Tout + 2 * b * strideT);
    const int nref = n - 1, nblk = (nref + 15) / 16;
    for (int bk = 4 * wq + w; bk < nblk; bk += 4 * wpm) {
        const int j0 = bk * 16;
        v4 gr = {0, 0, 0, 0}, gi = {0, 0, 0, 0};
        // G[k1][k2] = sum_r conj(V[r][k1]) V[r][k2]; K-steps over rows 4s.. (rows >= j0+1),
        // four at a time so that their four (branch-free) loads are in flight together
        for (int r16 = j0 + 1 - ((j0 + 1) & 3); r16 < n; r16 += 16) {
            CT x[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                x[u] = refl_get(A, lda, n, j0, r16 + 4 * u + lq, li);   // zero at rows >= n
#pragma unroll
            for (int u = 0; u < 4; u++) {   // A operand: conj(V)[k1=li][row]; B: V[row][k2=li]
                gr = M::mma(x[u].x, x[u].x, gr);
                gr = M::mma(x[u].y, x[u].y, gr);
                gi = M::mma(x[u].x, x[u].y, gi);
                gi = M::mma(-x[u].y, x[u].x, gi);
            }
        }
        // G to LDS (D layout: element v of lane l is G[row(l, v)][l%16])
#pragma unroll
        for (int v = 0; v < 4; v++) gs[w][M::row(lane, v) * 17 + li] = CT{gr[v], gi[v]};
        __builtin_amdgcn_wave_barrier();
        // recurrence over columns k; lane i < 16 owns row i of T
        for (int k = 0; k < 16; k++) {
            const int j = j0 + k;
            const CT tk = (j < nref) ? CT{tau[2 * j], tau[2 * j + 1]} : CT{(T)0, (T)0};
            CT val{(T)0, (T)0};
            if (lane < k) {
                CT sacc{(T)0, (T)0};
                for (int l2 = lane; l2 < k; l2++) fmax_(sacc, ts[w][lane * 17 + l2], gs[w][l2 * 17 + k]);
                val = (tk * sacc) * (T)(-1);
            } else if (lane == k) {
                val = tk;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < 16) ts[w][lane * 17 + k] = val;
            __builtin_amdgcn_wave_barrier();
        }
        for (int e = lane; e < 256; e += kWave) {
            const int k2 = e >> 4, k1 = e & 15;
            Tb[bk * 256 + e] = ts[w][k1 * 17 + k2];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// U = Q U for a slab of CW eigenvector columns per workgroup (NWV waves; wave w owns
// the 16-row blocks w, w + NWV, w + 2 NWV, w + 3 NWV: n <= 64 NWV). Per reflector block, last first:
// W = V^H U (each wave over its rows, summed through LDS in a fixed order),
// W2 = T W, U -= V W2. Workgroups are mapped so that the slabs of one matrix share
// blockIdx % 8, i.e. one XCD and its L2, which then serves V and T to all of them.
template <typename T, int CW, int NWV, int WPE = 2>
__global__ __launch_bounds__(NWV > 4 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(WPE))) void backtransform_mfma_kernel(const T *__restrict__ Abase, long strideA, int lda,
                                                                 int n, const T *__restrict__ Tin, long strideT,
                                                                 T *__restrict__ Uout, long strideU, int nslab,
                                                                 int batch) {
    using M = Mf<T>;
    using v4 = typename M::v4;
    using CT = cx<T>;
    constexpr int NCB = CW / 16, RBW = 4, VP = 17;            // VP: padded row stride of the V stage
    // complex float: each wave stages all its V rows once per block and prefetches the
    // next block's; complex double stages one 16-row block at a time and reads V again
    // in the update (LDS and registers then allow two workgroups per CU)
    constexpr bool STAGE_ALL = sizeof(T) == sizeof(float);
    constexpr bool PREFETCH = STAGE_ALL;
    __shared__ CT vst[NWV][STAGE_ALL ? RBW : 1][16 * VP];        // per wave: V(rows of row block i, k)
    __shared__ T wred[2][NWV][2 * NCB * 4 * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const long q = blockIdx.x >> 3;
    const long b = (q / nslab) * 8 + (blockIdx.x & 7);
    const int slab = (int)(q % nslab);
    if (b >= batch) return;
    const T *A = Abase + 2 * b * strideA;
    const CT *Tb = reinterpret_cast<const CT *>(Tin + 2 * b * strideT);
    T *U = Uout + 2 * b * strideU;
    const int c0 = slab * CW;
    v4 ur[RBW][NCB], ui[RBW][NCB];
#pragma unroll
    for (int i = 0; i < RBW; i++) {
        const int R = (w + NWV * i) * 16;
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int col = c0 + cb * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = R + M::row(lane, v);
                const CT x = (row < n && col < n) ? ldx(U, (long)row * n + col) : CT{(T)0, (T)0};
                ur[i][cb][v] = x.x;
                ui[i][cb][v] = x.y;
            }
        }
    }
    const int nref = n - 1, nblk = (nref + 15) / 16;
    // V(R + li, k) for k = lq + 4kk of this wave's row blocks (coalesced: 16 rows per k)
    CT vpre[STAGE_ALL ? RBW : 1][4];
    auto load_v = [&](int bk) {
        if (!STAGE_ALL) return;
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int R = (w + NWV * i) * 16;
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
                vpre[i][kk] = (R + 15 >= bk * 16 + 1) ? refl_get_masked(A, lda, n, bk * 16, R + li, lq + 4 * kk)
                                                      : CT{(T)0, (T)0};
        }
    };
    if (PREFETCH) load_v(nblk - 1);
    // complex float: T[k1 = li][k2 = row(l, s)] of a block is loaded one block ahead (n_phi =
    // 256, 32768 matrices: 60.5 -> 58.7 ms; complex double, whose extra registers cost it a
    // wave per SIMD, 159 -> 199 ms)
    constexpr bool TPF = sizeof(T) == sizeof(float);
    CT tnext[4];
    auto load_t = [&](int bk) {
#pragma unroll
        for (int s = 0; s < 4; s++) tnext[s] = Tb[bk * 256 + M::row(lane, s) * 16 + li];
    };
    if (TPF) load_t(nblk - 1);
    for (int bk = nblk - 1; bk >= 0; bk--) {
        const int j0 = bk * 16, r0 = j0 + 1;
        if (STAGE_ALL) {
            // stage this wave's V rows (read back transposed by W = V^H U, directly by U -= V W2)
#pragma unroll
            for (int i = 0; i < (STAGE_ALL ? RBW : 1); i++)
#pragma unroll
                for (int kk = 0; kk < 4; kk++) vst[w][i][(lq + 4 * kk) * VP + li] = vpre[i][kk];
            __builtin_amdgcn_wave_barrier();
        }
        // complex products from three real MFMAs: a b with P1 = (a_r + a_i) b_r,
        // P2 = a_r (b_i - b_r), P3 = a_i (b_r + b_i), Re = P1 - P3, Im = P1 + P2 (P1 in q1,
        // -P3 and P2 straight into the real / imaginary accumulators)
        v4 wr[NCB], wi[NCB], q1[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            wr[cb] = v4{0, 0, 0, 0};
            wi[cb] = v4{0, 0, 0, 0};
            q1[cb] = v4{0, 0, 0, 0};
        }
        // W = V^H U over this wave's rows (a = conj V)
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int R = (w + NWV * i) * 16;
            if (R + 15 < r0) continue;                    // wave-uniform: V is zero above row r0
            const int si = STAGE_ALL ? i : 0;
            if (!STAGE_ALL) {
#pragma unroll
                for (int kk = 0; kk < 4; kk++) vst[w][0][(lq + 4 * kk) * VP + li] = refl_get(A, lda, n, j0, R + li, lq + 4 * kk);
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const CT x = vst[w][si][li * VP + M::row(lane, v)];  // V[R + row(l, v)][k = li]
                const T xd = x.x - x.y;
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    const T br = ur[i][cb][v], bi = ui[i][cb][v];
                    q1[cb] = M::mma(xd, br, q1[cb]);
                    wr[cb] = M::mma(x.y, br + bi, wr[cb]);
                    wi[cb] = M::mma(x.x, bi - br, wi[cb]);
                }
            }
            if (!STAGE_ALL) __builtin_amdgcn_wave_barrier();   // the next row block rewrites vst
        }
        // sum W over the waves (double-buffered by block parity: one barrier per block)
        T *wb = wred[bk & 1][w];
#pragma unroll
        for (int cb = 0; cb < NCB; cb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                wb[((0 * NCB + cb) * 4 + v) * kWave + lane] = q1[cb][v] + wr[cb][v];
                wb[((1 * NCB + cb) * 4 + v) * kWave + lane] = q1[cb][v] + wi[cb][v];
            }
        __syncthreads();
        CT tcur[4];
        if (TPF) {
#pragma unroll
            for (int s = 0; s < 4; s++) tcur[s] = tnext[s];
            if (bk > 0) load_t(bk - 1);
        }
        if (PREFETCH && bk > 0) load_v(bk - 1);           // in flight during T W and the update
#pragma unroll
        for (int cb = 0; cb < NCB; cb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T sr = 0, si = 0;
#pragma unroll
                for (int ww = 0; ww < NWV; ww++) {
                    sr += wred[bk & 1][ww][((0 * NCB + cb) * 4 + v) * kWave + lane];
                    si += wred[bk & 1][ww][((1 * NCB + cb) * 4 + v) * kWave + lane];
                }
                wr[cb][v] = sr;
                wi[cb][v] = si;
            }
        // W2 = T W (K index = reflector k2 = row(l, s))
        v4 w2r[NCB], w2i[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            w2r[cb] = v4{0, 0, 0, 0};
            w2i[cb] = v4{0, 0, 0, 0};
            q1[cb] = v4{0, 0, 0, 0};
        }
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const CT tv = TPF ? tcur[s] : Tb[bk * 256 + M::row(lane, s) * 16 + li];   // T[k1 = li][k2 = row(l, s)]
            const T ts = tv.x + tv.y;
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) {
                const T br = wr[cb][s], bi = wi[cb][s];
                q1[cb] = M::mma(ts, br, q1[cb]);
                w2r[cb] = M::mma(-tv.y, br + bi, w2r[cb]);
                w2i[cb] = M::mma(tv.x, bi - br, w2i[cb]);
            }
        }
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            w2r[cb] += q1[cb];
            w2i[cb] += q1[cb];
        }
        // U -= V W2 on this wave's rows (K index = reflector row(l, s))
#pragma unroll
        for (int i = 0; i < RBW; i++) {
            const int R = (w + NWV * i) * 16;
            if (R + 15 < r0) continue;
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) q1[cb] = v4{0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const CT x = STAGE_ALL ? vst[w][STAGE_ALL ? i : 0][M::row(lane, s) * VP + li]
                                       : refl_get(A, lda, n, j0, R + li, M::row(lane, s));   // V[R + li][k = row(l, s)]
                const T xs = x.x + x.y;
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    const T br = w2r[cb][s], bi = w2i[cb][s];
                    q1[cb] = M::mma(xs, br, q1[cb]);                     // P1
                    ur[i][cb] = M::mma(x.y, br + bi, ur[i][cb]);         // U_r -= P1 - P3
                    ui[i][cb] = M::mma(-x.x, bi - br, ui[i][cb]);        // U_i -= P1 + P2
                }
            }
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) {
                ur[i][cb] -= q1[cb];
                ui[i][cb] -= q1[cb];
            }
        }
        __builtin_amdgcn_wave_barrier();                   // the next block rewrites vst
    }
#pragma unroll
    for (int i = 0; i < RBW; i++) {
        const int R = (w + NWV * i) * 16;
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int col = c0 + cb * 16 + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = R + M::row(lane, v);
                if (row < n && col < n) stx(U, (long)row * n + col, CT{ur[i][cb][v], ui[i][cb][v]});
            }
        }
    }
}

// ============================================================================
// host launchers
// ============================================================================
// divide & conquer arrays, and at least the 32 x 33 transpose tile of the output copy
size_t heevd_dc_lds_bytes(int n) {
    const size_t dc = (size_t)n * (13 * sizeof(double) + 16 * sizeof(int)) + 8 * sizeof(double);
    const size_t tile = 32 * 33 * sizeof(double);
    return dc > tile ? dc : tile;
}

long backtransform_tws_complex(int n) { return (long)((n - 1 + 15) / 16) * 256; }

static size_t backtransform_lds_bytes(int n, int nb_wy, size_t real) {
    return ((size_t)n * nb_wy + (size_t)nb_wy * (nb_wy + 1)) * 2 * real;
}

int heevd_dc_workspace_doubles(int n, int grid_tail, long *per_matrix, long *per_wg) {
    *per_matrix = 4L * n;         // d, e, tau(complex)
    *per_wg = 3L * n * n;
    (void)grid_tail;
    return 0;
}

// register-resident small-n reduction (zhetrd_small_kernel), n <= 72 (wider column
// ranges would cost the second workgroup per CU); persistent workgroups as many as the
// CUs hold at once
template <typename T, int RPT, int CPT>
static int launch_zhetrd_small_rc(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV,
                                  int batch, hipStream_t stream) {
    int dev = 0, cus = 0, per_cu = 0;
    RLM_HIP_CHECK(hipGetDevice(&dev));
    RLM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RLM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)zhetrd_small_kernel<T, RPT, CPT>,
                                                               256, 0));
    int grid = per_cu > 0 && cus > 0 ? per_cu * cus : 1024;
    if (grid > batch) grid = batch;
    hipLaunchKernelGGL((zhetrd_small_kernel<T, RPT, CPT>), dim3(grid), dim3(256), 0, stream, A, strideA, lda, n, D, E,
                       TAU, strideV, batch);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}
template <typename T>
static int launch_zhetrd_small_t(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV,
                                 int batch, hipStream_t stream) {
    if (n < 1 || n > 72) return -22;
    if (n <= 64) return launch_zhetrd_small_rc<T, 2, 8>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream);
    return launch_zhetrd_small_rc<T, 3, 9>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream);
}

template <typename T>
int launch_zhetrd_t(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV, int batch,
                    int grid, hipStream_t stream) {
    if (n < 1 || n > 512) return -22;
    constexpr int NT = 256;
    const size_t cs = 2 * sizeof(T);
    size_t lds = (size_t)(4 * n + (NT / kWave) * n) * cs + (2 * (NT / kWave) + 8) * sizeof(T);
    if (lds > 64 * 1024) RLM_HIP_CHECK(hipFuncSetAttribute((const void *)zhetrd_kernel<T, NT, 2>,
                                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // persistent workgroups: as many as the CUs hold at once (small n: several per CU)
    {
        int dev = 0, cus = 0, per_cu = 0;
        RLM_HIP_CHECK(hipGetDevice(&dev));
        RLM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        RLM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)zhetrd_kernel<T, NT, 2>, NT,
                                                                   lds));
        if (per_cu > 0 && cus > 0 && per_cu * cus > grid) grid = per_cu * cus;
    }
    if (grid > batch) grid = batch;
    hipLaunchKernelGGL((zhetrd_kernel<T, NT, 2>), dim3(grid), dim3(NT), lds, stream, A, strideA, lda, n, D, E, TAU,
                       strideV, batch);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

static size_t blocked_lds_bytes(int n, int NB, size_t real) {
    constexpr int NT = 256;
    return ((size_t)n * NB + n + (NT / kWave) * n + (NT / (2 * NB)) * 2 * NB + 2 * NB) * 2 * real +
           (2 * (NT / kWave) + 8) * real;
}

template <typename T, int NB>
int launch_zhetrd_blocked_t(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV, int batch,
                            int grid, hipStream_t stream) {
    if (n < 1 || n > 512) return -22;
    constexpr int NT = 256;
    const size_t lds = blocked_lds_bytes(n, NB, sizeof(T));
    if (lds > 160 * 1024) return -22;
    if (grid > batch) grid = batch;
    if (lds > 64 * 1024)
        RLM_HIP_CHECK(hipFuncSetAttribute((const void *)zhetrd_blocked_kernel<T, NT, NB, 2>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((zhetrd_blocked_kernel<T, NT, NB, 2>), dim3(grid), dim3(NT), lds, stream, A, strideA, lda, n, D,
                       E, TAU, strideV, batch);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

static size_t lower_lds_bytes(int n, int NB, size_t real) {
    constexpr int NT = 512;
    return ((size_t)n * NB + n + (NT / kWave) * n + (NT / (2 * NB)) * 2 * NB + 2 * NB) * 2 * real +
           (2 * (NT / kWave) + 8) * real;
}

template <typename T, int NB>
int launch_zhetrd_lower_t(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV, int batch,
                          int grid, hipStream_t stream, int ncols = -1) {
    if (ncols < 0) ncols = n - 1;
    if (n < 1 || n > 512) return -22;
    constexpr int NT = 512;
    const size_t lds = lower_lds_bytes(n, NB, sizeof(T));
    if (lds > 160 * 1024) return -22;
    if (grid > batch) grid = batch;
    if (lds > 64 * 1024)
        RLM_HIP_CHECK(hipFuncSetAttribute((const void *)zhetrd_lower_kernel<T, NT, NB, 1>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((zhetrd_lower_kernel<T, NT, NB, 1>), dim3(grid), dim3(NT), lds, stream, A, strideA, lda, n, D,
                       E, TAU, strideV, batch, ncols);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
int launch_tail_t(const T *A, long strideA, int lda, int n, const double *D, const double *E, const T *TAU,
                  long strideV, double *W, long strideW, T *U, long strideU, double *work, T *tws, int batch, int grid,
                  hipStream_t stream, hipEvent_t mid) {
    if (n < 1 || n > 512) return -22;
    const size_t lds_dc = heevd_dc_lds_bytes(n);
    if (lds_dc > 160 * 1024) return -22;
    if (grid > batch) grid = batch;
    if (n <= 128) {
        // two waves per matrix: more matrices per CU where the nodes are small
        hipLaunchKernelGGL((stedc_kernel<T, 128, false>), dim3(grid), dim3(128), lds_dc, stream, n, D, E, strideV, W,
                           strideW, U, strideU, work, batch, 0);
    } else {
        constexpr int NT = 256;
        if (lds_dc > 64 * 1024)
            RLM_HIP_CHECK(hipFuncSetAttribute((const void *)stedc_kernel<T, NT, false>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_dc));
        hipLaunchKernelGGL((stedc_kernel<T, NT, false>), dim3(grid), dim3(NT), lds_dc, stream, n, D, E, strideV, W,
                           strideW, U, strideU, work, batch, 0);
    }
    RLM_HIP_CHECK(hipGetLastError());
    if (mid) RLM_HIP_CHECK(hipEventRecord(mid, stream));
    if (n == 1) return 0;
    if (n <= 256 && tws) {
        // matrix-core path: triangular factors, then slabs of eigenvector columns
        const long strideT = backtransform_tws_complex(n);
        hipLaunchKernelGGL((larft_kernel<T>), dim3(batch), dim3(256), 0, stream, A, strideA, lda, n, TAU, strideV, tws,
                           strideT, batch, 1);
        RLM_HIP_CHECK(hipGetLastError());
        // n <= 128: two waves per slab (twice the workgroups per CU, a 2-way W reduction),
        // slabs of 32 columns for complex float and 16 for complex double. Above: slabs of
        // 16 columns at three waves per SIMD (166 VGPRs complex float, 148 complex double;
        // three 51 KB workgroups per CU). Complex float n = 256, 32768 matrices: 58.4 ms with
        // 32-column slabs at two waves per SIMD (241 VGPRs, 67.5 KB of LDS: a cap at three
        // waves leaves it at one), 70.0 ms with 16-column slabs at two, 57.2 ms at three.
        if (n <= 128) {
            constexpr int CW = sizeof(T) == sizeof(float) ? 32 : 16;
            const int nslab = (n + CW - 1) / CW;
            const long groups = (batch + 7) / 8;
            hipLaunchKernelGGL((backtransform_mfma_kernel<T, CW, 2>), dim3((unsigned)(groups * 8 * nslab)), dim3(128),
                               0, stream, A, strideA, lda, n, tws, strideT, U, strideU, nslab, batch);
        } else {
            const int nslab = (n + 15) / 16;
            const long groups = (batch + 7) / 8;
            hipLaunchKernelGGL((backtransform_mfma_kernel<T, 16, 4, 3>), dim3((unsigned)(groups * 8 * nslab)),
                               dim3(256), 0, stream, A, strideA, lda, n, tws, strideT, U, strideU, nslab, batch);
        }
        RLM_HIP_CHECK(hipGetLastError());
        return 0;
    }
    // complex float 256 < n <= 512: the same slabs with eight waves, each still holding four
    // 16-row blocks (4624 matrices, n = 384 / 512: 40.8 / 69.2 ms against 51.6 / 108.6 ms
    // streaming; complex double, at 256 VGPRs, loses: 102 / 207 against 73 / 161 ms)
    if constexpr (sizeof(T) == sizeof(float)) {
        if (n <= 512 && tws) {
            const long strideT = backtransform_tws_complex(n);
            hipLaunchKernelGGL((larft_kernel<T>), dim3(batch), dim3(256), 0, stream, A, strideA, lda, n, TAU, strideV, tws,
                               strideT, batch, 1);
            RLM_HIP_CHECK(hipGetLastError());
            const int nslab = (n + 31) / 32;
            const long groups = (batch + 7) / 8;
            hipLaunchKernelGGL((backtransform_mfma_kernel<T, 32, 8>), dim3((unsigned)(groups * 8 * nslab)), dim3(512), 0,
                               stream, A, strideA, lda, n, tws, strideT, U, strideU, nslab, batch);
            RLM_HIP_CHECK(hipGetLastError());
            return 0;
        }
    }
    if (tws && lda == n) {
        // T factors, then the streaming matrix-core back-transform (the slab no longer fits
        // the accumulators)
        const long strideT = backtransform_tws_complex(n);
        hipLaunchKernelGGL((larft_kernel<T>), dim3(batch), dim3(256), 0, stream, A, strideA, lda, n, TAU, strideV, tws,
                           strideT, batch, 1);
        RLM_HIP_CHECK(hipGetLastError());
        return launch_backtransform_stream(A, strideA, n, tws, strideT, U, strideU, batch, stream);
    }
    // reflectors per compact-WY block: the V block (n x NBW complex) lives in LDS
    constexpr int NT = 256, NBW = sizeof(T) == 4 ? 32 : 16;
    const size_t lds_bt = backtransform_lds_bytes(n, NBW, sizeof(T));
    if (lds_bt > 160 * 1024) return -22;
    if (lds_bt > 64 * 1024)
        RLM_HIP_CHECK(hipFuncSetAttribute((const void *)backtransform_kernel<T, NT, NBW>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bt));
    hipLaunchKernelGGL((backtransform_kernel<T, NT, NBW>), dim3(batch), dim3(NT), lds_bt, stream, A, strideA, lda, n,
                       TAU, strideV, U, strideU, NBW, batch);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename T, int NS, int NL, bool DIAG0, int NW = 8, int NMAX = 256, typename TL = ResidentSub4<T>>
static int launch_zhetrd_resident_t(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV,
                                    int batch, hipStream_t stream, const double *Hb, int hb_off) {
    const int tn = (n + 31) / 32;
    ResidentDeal deal;
    if (n < 1 || n > NMAX || !resident_deal(tn, NS, DIAG0, NL, NW, deal)) return -22;
    hipLaunchKernelGGL((zhetrd_resident_kernel<T, NS, NL, DIAG0, NW, NMAX, TL>), dim3(batch), dim3(64 * NW), 0, stream,
                       A, strideA, lda, n, D, E, TAU, strideV, batch, Hb, hb_off, deal);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

// Largest n of the register-resident kernel: complex float keeps 4 tiles per wave
// in registers (128 VGPRs) plus up to 4 tiles of tile column 0 in LDS (n <= 256);
// complex double keeps 1 tile per wave in registers plus up to 4 in LDS (n <= 128).
static int resident_max_n(size_t real) { return real == sizeof(float) ? 256 : 128; }

// RLM_HETRD selects the tridiagonalisation kernel: "resident" (lower triangle in
// registers), "fused" (unblocked, full storage), "blocked" (panels, full storage),
// "lower" (panels of 16, lower storage, tiled symv), "lower8" (panels of 8), "hybrid"
// (lower for the leading columns, resident for the trailing block it fits), "small"
// (complex double n <= 72, full storage in registers). Unset or "auto": the autotuner's
// choice for (n, precision) when tune_hetrd has run, else the fixed defaults below.
// Read at every launch so one process can compare them.
enum { kTrdAuto = 0, kTrdFused, kTrdBlocked, kTrdLower, kTrdResident, kTrdHybrid, kTrdSmall, kTrdLower8 };
static const char *const kTrdNames[] = {"auto", "fused", "blocked", "lower", "resident", "hybrid", "small", "lower8"};
static int hetrd_variant() {
    const char *e = getenv("RLM_HETRD");
    if (!e) return kTrdAuto;
    for (int v = 1; v < (int)(sizeof(kTrdNames) / sizeof(kTrdNames[0])); v++)
        if (strcmp(e, kTrdNames[v]) == 0) return v;
    return kTrdAuto;
}

// ---------------------------------------------------------------- autotuner
// Choices of tune_hetrd per (n, precision), per process; also kept in a text file
// (RLM_TUNE_CACHE, default $HOME/.cache/rlm_hip_tune.txt: lines "<device> <n> <fp32>
// <variant>") so that later processes on the same kind of GPU skip the timing.
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, int> g_tuned;

static int tuned_variant(int n, bool fp32) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find({n, fp32 ? 1 : 0});
    return it == g_tuned.end() ? kTrdAuto : it->second;
}

// the variant a launch uses: RLM_HETRD, else the tuned choice, else kTrdAuto (defaults)
static int effective_variant(int n, bool fp32) {
    const int v = hetrd_variant();
    return v != kTrdAuto ? v : tuned_variant(n, fp32);
}

// True when launch_chetrd_band can take band-form input (register-resident kernel).
bool chetrd_band_supported(int n) {
    const int v = effective_variant(n, true);
    return (v == kTrdAuto || v == kTrdResident) && n <= resident_max_n(sizeof(float));
}

template <typename T>
static int launch_resident_any(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV,
                               int batch, hipStream_t stream, const double *Hb, int hb_off) {
    const int tn = (n + 31) / 32;
    if constexpr (sizeof(T) == sizeof(float)) {
        // n <= 96 (at most 6 tiles): 4 waves with two register tiles each, so that two
        // matrices share a CU and one's single-wave reflector phase overlaps the other's
        // work (complex double would spill at two tiles per wave)
        // RLM_RESIDENT_VALU set: the VALU tile layout (update and products in one VALU
        // pass) instead of the matrix-core update, for comparison
        if (getenv("RLM_RESIDENT_VALU")) {
            if (tn <= 3)
                return launch_zhetrd_resident_t<T, 2, 0, false, 4, 128>(A, strideA, lda, n, D, E, TAU, strideV, batch,
                                                                        stream, Hb, hb_off);
            if (tn <= 4)
                return launch_zhetrd_resident_t<T, 1, 4, false>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream,
                                                                Hb, hb_off);
            if (tn <= 5)
                return launch_zhetrd_resident_t<T, 2, 4, false>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream,
                                                                Hb, hb_off);
            return launch_zhetrd_resident_t<T, 4, 4, true>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream, Hb,
                                                           hb_off);
        }
        if (tn <= 3)
            return launch_zhetrd_resident_t<T, 2, 0, false, 4, 128, ResidentMfma>(A, strideA, lda, n, D, E, TAU,
                                                                                  strideV, batch, stream, Hb, hb_off);
        if (tn <= 4)
            return launch_zhetrd_resident_t<T, 1, 4, false, 8, 256, ResidentMfma>(A, strideA, lda, n, D, E, TAU,
                                                                                  strideV, batch, stream, Hb, hb_off);
        if (tn <= 5)
            return launch_zhetrd_resident_t<T, 2, 4, false, 8, 256, ResidentMfma>(A, strideA, lda, n, D, E, TAU,
                                                                                  strideV, batch, stream, Hb, hb_off);
        return launch_zhetrd_resident_t<T, 4, 4, true, 8, 256, ResidentMfma>(A, strideA, lda, n, D, E, TAU, strideV,
                                                                            batch, stream, Hb, hb_off);
    } else {
        return launch_zhetrd_resident_t<T, 1, 4, false>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream, Hb,
                                                        hb_off);
    }
}

// Whether variant v can reduce complex T matrices of size n.
template <typename T>
static bool variant_valid(int v, int n) {
    const bool dbl = sizeof(T) == sizeof(double);
    switch (v) {
        case kTrdSmall: return dbl && n <= 72;
        case kTrdFused: return n <= 512;
        case kTrdResident: return n <= resident_max_n(sizeof(T));
        case kTrdHybrid: return n > resident_max_n(sizeof(T)) && n <= 512 &&
                                lower_lds_bytes(n, dbl ? 8 : 16, sizeof(T)) <= 160 * 1024;
        case kTrdLower: return n <= 512 && lower_lds_bytes(n, 16, sizeof(T)) <= 160 * 1024;
        case kTrdLower8: return n <= 512 && lower_lds_bytes(n, 8, sizeof(T)) <= 160 * 1024;
        case kTrdBlocked: return n <= 512 && blocked_lds_bytes(n, 16, sizeof(T)) <= 160 * 1024;
        default: return false;
    }
}

template <typename T>
static int launch_hetrd_any(T *A, long strideA, int lda, int n, double *D, double *E, T *TAU, long strideV, int batch,
                            int grid, hipStream_t stream) {
    int v = effective_variant(n, sizeof(T) == sizeof(float));
    const int rmax = resident_max_n(sizeof(T));
    if (v != kTrdAuto && !variant_valid<T>(v, n)) v = kTrdAuto;
    // defaults (no tuned choice): complex double n <= 72 takes zhetrd_small_kernel
    // (registers, two matrices per CU: hlm n = 72, 67584 matrices: 30.3 ms), 72 < n <= 96
    // the fused kernel (its small LDS footprint lets five matrices share a CU, where the
    // resident kernel holds one: hlm n = 72 52.7 against 111 ms), n <= 128 the resident
    // kernel (n = 128: 23.8 against 26.5 ms fused); complex float n <= 256 resident, n > 256
    // hybrid (panels for the leading n - 256 columns, then the trailing 256 x 256 block in
    // registers: n = 384: 143 -> 99 ms, n = 512: 290 -> 248 ms for 4624 matrices)
    if (v == kTrdAuto) {
        if (sizeof(T) == sizeof(double) && n <= 72) v = kTrdSmall;
        else if (sizeof(T) == sizeof(double) && n <= 96) v = kTrdFused;
        else if (n <= rmax) v = kTrdResident;
        else if (sizeof(T) == sizeof(float) && variant_valid<T>(kTrdHybrid, n)) v = kTrdHybrid;
        else if (variant_valid<T>(kTrdLower, n)) v = kTrdLower;
        else if (variant_valid<T>(kTrdLower8, n)) v = kTrdLower8;
        else v = kTrdFused;
    }
    if constexpr (sizeof(T) == sizeof(double))
        if (v == kTrdSmall) return launch_zhetrd_small_t<T>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream);
    if (v == kTrdResident && n <= rmax)
        return launch_resident_any<T>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream, nullptr, 0);
    if (v == kTrdHybrid && variant_valid<T>(kTrdHybrid, n)) {
        // columns 0..ncols-1 by panels of 16 through memory; the updated trailing block
        // A(ncols:n, ncols:n) (at most rmax) is then a Hermitian matrix whose
        // reduction continues on chip, writing d, e, tau and the reflectors in place
        // (complex double: panels of 8, as its lower variant, and a trailing 128 block: the
        // autotuner takes it over lower8 for n = 192..512; n = 256, 32768 matrices: 584 ->
        // 529 ms)
        const int ncols = ((n - rmax) + 15) / 16 * 16;
        int rc = sizeof(T) == sizeof(double)
                     ? launch_zhetrd_lower_t<T, 8>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream, ncols)
                     : launch_zhetrd_lower_t<T, 16>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream, ncols);
        if (rc) return rc;
        return launch_resident_any<T>(A + 2 * ((long)ncols * lda + ncols), strideA, lda, n - ncols, D + ncols,
                                      E + ncols, TAU + 2 * ncols, strideV, batch, stream, nullptr, 0);
    }
    // the panel kernels keep W (n x NB) in LDS: panels of 8 where 16 do not fit (complex
    // double n = 512: 148 KB; 448 against 1848 ms fused for 4624 matrices)
    if ((v == kTrdLower || v == kTrdResident || v == kTrdHybrid) && variant_valid<T>(kTrdLower, n))
        return launch_zhetrd_lower_t<T, 16>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream);
    if ((v == kTrdLower || v == kTrdLower8 || v == kTrdResident || v == kTrdHybrid) && variant_valid<T>(kTrdLower8, n))
        return launch_zhetrd_lower_t<T, 8>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream);
    if (v == kTrdBlocked && variant_valid<T>(kTrdBlocked, n))
        return launch_zhetrd_blocked_t<T, 16>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream);
    return launch_zhetrd_t<T>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream);
}

// Random Hermitian test matrices of the autotuner (full storage, values in [-1, 1)).
template <typename T>
__global__ void random_herm_kernel(T *A, long strideA, int n, unsigned seed) {
    T *M = A + 2 * blockIdx.x * strideA;
    for (long idx = threadIdx.x; idx < (long)n * n; idx += blockDim.x) {
        const int c = (int)(idx / n), r = (int)(idx % n);
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        unsigned h = seed ^ (blockIdx.x * 0x9E3779B9u) ^ ((unsigned)lo * 0x85EBCA6Bu) ^ ((unsigned)hi * 0xC2B2AE35u);
        h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
        const T re = (T)((h & 0xFFFF) * (1.0 / 32768.0) - 1.0), im = (T)((h >> 16) * (1.0 / 32768.0) - 1.0);
        M[2 * idx] = re;
        M[2 * idx + 1] = r == c ? (T)0 : (r > c ? im : -im);
    }
}

static std::string tune_cache_path() {
    const char *e = getenv("RLM_TUNE_CACHE");
    if (e && *e) return std::string(e);
    const char *home = getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/rlm_hip_tune.txt";
}

static std::string device_tag() {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string arch(prop.gcnArchName);
    for (char &c : arch)
        if (c == ' ' || c == ':') c = '_';
    return arch + "_cu" + std::to_string(prop.multiProcessorCount);
}

template <typename T>
static int tune_hetrd_t(int n, long batch_hint, int *chosen) {
    const bool fp32 = sizeof(T) == sizeof(float);
    std::vector<int> cand;
    for (int v = kTrdFused; v <= kTrdLower8; v++)
        if (variant_valid<T>(v, n)) cand.push_back(v);
    if (cand.size() <= 1) {
        *chosen = cand.empty() ? kTrdAuto : cand[0];
        return 0;
    }
    // a sample batch that fills the chip as the real launch does
    const long cap = n <= 128 ? 1024 : n <= 256 ? 512 : 256;
    const int S = (int)std::max<long>(1, std::min<long>(batch_hint, cap));
    const long strideV = 4L * n;
    T *A = nullptr;
    double *V = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        if (A) (void)hipFree(A);
        if (V) (void)hipFree(V);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    };
    if (hipMalloc(&A, sizeof(T) * 2 * (size_t)n * n * S) != hipSuccess ||
        hipMalloc(&V, sizeof(double) * (size_t)strideV * S) != hipSuccess || hipStreamCreate(&s) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return -12;
    }
    double *D = V, *E = V + n;
    T *TAU = reinterpret_cast<T *>(V + 2 * n);
    float best = 1e30f;
    int best_v = kTrdAuto;
    for (int v : cand) {
        float t_min = 1e30f;
        for (int rep = 0; rep < 3 && !rc; rep++) {   // rep 0 warms up (code object load)
            hipLaunchKernelGGL(random_herm_kernel<T>, dim3(S), dim3(256), 0, s, A, (long)n * n, n, 1234u + rep);
            if (hipEventRecord(e0, s) != hipSuccess) rc = -1;
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tuned[{n, fp32 ? 1 : 0}] = v;      // launch through the normal dispatch
            }
            if (!rc) rc = launch_hetrd_any<T>(A, (long)n * n, n, n, D, E, TAU, strideV, S, 1024, s);
            if (!rc && hipEventRecord(e1, s) != hipSuccess) rc = -1;
            if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = -1;
            float ms = 0.f;
            if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = -1;
            if (!rc && rep > 0) t_min = std::min(t_min, ms);
        }
        if (rc) break;
        if (t_min < best) {
            best = t_min;
            best_v = v;
        }
    }
    cleanup();
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned.erase({n, fp32 ? 1 : 0});
    }
    if (rc) return rc;
    *chosen = best_v;
    return 0;
}

int tune_hetrd(int n, bool fp32, long batch_hint) {
    if (n < 2 || n > 512) return 0;
    const char *en = getenv("RLM_TUNE");
    if ((en && strcmp(en, "0") == 0) || hetrd_variant() != kTrdAuto) return 0;
    if (tuned_variant(n, fp32) != kTrdAuto) return 0;
    const std::string tag = device_tag(), path = tune_cache_path();
    {
        FILE *f = fopen(path.c_str(), "r");
        if (f) {
            char dtag[128], name[32];
            int fn = 0, ff = 0, found = kTrdAuto;
            while (fscanf(f, "%127s %d %d %31s", dtag, &fn, &ff, name) == 4)
                if (tag == dtag && fn == n && ff == (fp32 ? 1 : 0))
                    for (int v = 1; v < (int)(sizeof(kTrdNames) / sizeof(kTrdNames[0])); v++)
                        if (strcmp(name, kTrdNames[v]) == 0) found = v;
            fclose(f);
            const bool ok = fp32 ? variant_valid<float>(found, n) : variant_valid<double>(found, n);
            if (found != kTrdAuto && ok) {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tuned[{n, fp32 ? 1 : 0}] = found;
                return 0;
            }
        }
    }
    int v = kTrdAuto, rc = fp32 ? tune_hetrd_t<float>(n, batch_hint, &v) : tune_hetrd_t<double>(n, batch_hint, &v);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[{n, fp32 ? 1 : 0}] = v;
    }
    if (v != kTrdAuto) {
        const size_t slash = path.rfind('/');
        if (slash != std::string::npos) {
            const std::string dir = path.substr(0, slash);
            (void)mkdir(dir.c_str(), 0755);
        }
        FILE *f = fopen(path.c_str(), "a");
        if (f) {
            fprintf(f, "%s %d %d %s\n", tag.c_str(), n, fp32 ? 1 : 0, kTrdNames[v]);
            fclose(f);
        }
    }
    return 0;
}

int hetrd_choice(int n, bool fp32) { return effective_variant(n, fp32); }
const char *hetrd_choice_name(int n, bool fp32) {
    const int v = effective_variant(n, fp32);
    return kTrdNames[v >= 0 && v < (int)(sizeof(kTrdNames) / sizeof(kTrdNames[0])) ? v : 0];
}

int launch_zhetrd(double *A, long strideA, int lda, int n, double *D, double *E, double *TAU, long strideV, int batch,
                  int grid, hipStream_t stream) {
    return launch_hetrd_any<double>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream);
}
int launch_chetrd(float *A, long strideA, int lda, int n, double *D, double *E, float *TAU, long strideV, int batch,
                  int grid, hipStream_t stream) {
    return launch_hetrd_any<float>(A, strideA, lda, n, D, E, TAU, strideV, batch, grid, stream);
}
int launch_chetrd_band(float *A, long strideA, int lda, int n, double *D, double *E, float *TAU, long strideV,
                       int batch, hipStream_t stream, const double *Hb, int hb_off) {
    if (!chetrd_band_supported(n) || !Hb) return -22;
    return launch_resident_any<float>(A, strideA, lda, n, D, E, TAU, strideV, batch, stream, Hb, hb_off);
}
int launch_tail(const double *A, long strideA, int lda, int n, const double *D, const double *E, const double *TAU,
                long strideV, double *W, long strideW, double *U, long strideU, double *work, double *tws, int batch,
                int grid, hipStream_t stream, hipEvent_t mid) {
    return launch_tail_t<double>(A, strideA, lda, n, D, E, TAU, strideV, W, strideW, U, strideU, work, tws, batch,
                                 grid, stream, mid);
}
int launch_tail_f32(const float *A, long strideA, int lda, int n, const double *D, const double *E, const float *TAU,
                    long strideV, double *W, long strideW, float *U, long strideU, double *work, float *tws, int batch,
                    int grid, hipStream_t stream, hipEvent_t mid) {
    return launch_tail_t<float>(A, strideA, lda, n, D, E, TAU, strideV, W, strideW, U, strideU, work, tws, batch,
                                grid, stream, mid);
}

// ---------------------------------------------------------------- n > 512 pieces
// Divide & conquer with the node arrays in global memory (any n); work holds
// grid * stedc_global_work_doubles(n) doubles.
template <typename T>
static int launch_stedc_global_t(int n, const double *D, const double *E, long strideV, double *W, long strideW, T *U,
                                 long strideU, double *work, int batch, int grid, hipStream_t stream) {
    // 512 threads: few, large matrices (one workgroup each), so twice the lanes per merge
    constexpr int NT = 512;
    if (n < 1 || batch < 1) return batch < 1 ? 0 : -22;
    if (grid > batch) grid = batch;
    const size_t lds = 32 * 33 * sizeof(double);
    const char *sp = getenv("RLM_DC_SPLIT");
    const bool split = n >= 1024 && !(sp && strcmp(sp, "0") == 0);
    if (!split) {
        hipLaunchKernelGGL((stedc_kernel<T, NT, true>), dim3(grid), dim3(NT), lds, stream, n, D, E, strideV, W,
                           strideW, U, strideU, work, batch, 0);
        RLM_HIP_CHECK(hipGetLastError());
        return 0;
    }
    // large n: the secular phase and the GEMMs of the merge levels with halves of >= 128
    // rows (the top GEMM alone is about 70 % of the divide & conquer's matrix-core work)
    // on several workgroups per matrix instead of one, in rounds of `grid` matrices (one
    // workspace slot each): the single-workgroup kernel runs the levels below and merges
    // and deflates the first split level, then secular, GEMM and resume kernels alternate
    // up to the top
    int top = 1;
    while ((1 << top) < n) top++;                  // level whose block covers n
    // first split level: halves of 128 rows (with the secular phase on several workgroups
    // as well, n_phi = 3000, 64 matrices: 55.4 ms against 57.7 ms from 256-row halves and
    // 55.5 ms from 64; RLM_DC_SPLIT_LEVEL overrides, >= 6)
    int first = 8;
    if (const char *e = getenv("RLM_DC_SPLIT_LEVEL")) first = std::max(6, atoi(e));
    int dev = 0, cus = 0;
    RLM_HIP_CHECK(hipGetDevice(&dev));
    RLM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    for (int b0 = 0; b0 < batch; b0 += grid) {
        const int cnt = std::min(grid, batch - b0);
        const int nwg = std::max(1, std::min(16, (2 * std::max(cus, 1) + cnt - 1) / cnt));
        const double *Db = D + (long)b0 * strideV, *Eb = E + (long)b0 * strideV;
        const int lev0 = std::min(first, top);
        hipLaunchKernelGGL((stedc_kernel<T, NT, true>), dim3(cnt), dim3(NT), lds, stream, n, Db, Eb, strideV,
                           W + (long)b0 * strideW, strideW, U + 2L * b0 * strideU, strideU, work, cnt, lev0);
        for (int L = lev0; L <= top; L++) {
            for (int phase = 0; phase < 3; phase++)
                hipLaunchKernelGGL((stedc_secular_kernel<T, NT>), dim3((unsigned)((long)cnt * nwg)), dim3(NT), 0,
                                   stream, n, Eb, strideV, work, nwg, L, phase);
            if (L == top) break;
            hipLaunchKernelGGL((stedc_level_gemm_kernel<T, NT>), dim3((unsigned)((long)cnt * nwg)), dim3(NT), 0,
                               stream, n, work, nwg, L);
            hipLaunchKernelGGL((stedc_resume_kernel<T, NT>), dim3(cnt), dim3(NT), 0, stream, n, Db, Eb, strideV, work,
                               L + 1);
        }
        hipLaunchKernelGGL((stedc_top_gemm_kernel<T, NT>), dim3((unsigned)((long)cnt * nwg)), dim3(NT), 0, stream, n,
                           work, nwg, W + (long)b0 * strideW, strideW, U + 2L * b0 * strideU, strideU);
        RLM_HIP_CHECK(hipGetLastError());
    }
    return 0;
}
int launch_stedc_global(int n, const double *D, const double *E, long strideV, double *W, long strideW, double *U,
                        long strideU, double *work, int batch, int grid, hipStream_t stream) {
    return launch_stedc_global_t<double>(n, D, E, strideV, W, strideW, U, strideU, work, batch, grid, stream);
}
int launch_stedc_global(int n, const double *D, const double *E, long strideV, double *W, long strideW, float *U,
                        long strideU, double *work, int batch, int grid, hipStream_t stream) {
    return launch_stedc_global_t<float>(n, D, E, strideV, W, strideW, U, strideU, work, batch, grid, stream);
}

// Triangular factors of every 16-reflector block, wpm workgroups per matrix.
template <typename T>
static int launch_larft_t(const T *A, long strideA, int lda, int n, const T *TAU, long strideV, T *tws, long strideT,
                          int batch, int wpm, hipStream_t stream) {
    if (batch < 1 || n < 2) return 0;
    hipLaunchKernelGGL((larft_kernel<T>), dim3((unsigned)((long)batch * wpm)), dim3(256), 0, stream, A, strideA, lda,
                       n, TAU, strideV, tws, strideT, batch, wpm);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}
int launch_larft(const double *A, long strideA, int lda, int n, const double *TAU, long strideV, double *tws,
                 long strideT, int batch, int wpm, hipStream_t stream) {
    return launch_larft_t<double>(A, strideA, lda, n, TAU, strideV, tws, strideT, batch, wpm, stream);
}
int launch_larft(const float *A, long strideA, int lda, int n, const float *TAU, long strideV, float *tws,
                 long strideT, int batch, int wpm, hipStream_t stream) {
    return launch_larft_t<float>(A, strideA, lda, n, TAU, strideV, tws, strideT, batch, wpm, stream);
}

}  // namespace rlm
