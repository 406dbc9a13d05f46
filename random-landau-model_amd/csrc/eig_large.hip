// Batched Hermitian eigensolver for 512 < n <= 4096: the n_phi = 3000 regime of
// the random Landau model (rlm/main_rlm_cpu.cpp:29-34 with -n 3000 -q 1 -i 16),
// where one matrix is 144 MB in complex double and the parallelism has to come
// from inside each matrix. Same conventions as eig.hip (LAPACK zhetrd lower,
// zlarfg reflectors kept in A, compact-WY back-transform), so it replaces the
// per-k-point zheevd of diag_wrappers/mkl_diag.cpp:2-9 at these sizes.
//
//   Tridiagonalisation (zhetrd/zlatrd structure, panels of kLgNB columns; only the
//   lower triangle is read and updated):
//     lg_col_kernel    one workgroup per matrix and column: finishes w of the
//                      previous column (w = tau y - 1/2 tau (w^H v) v), applies
//                      the newest term of the panel's pending update to column c
//                      and builds reflector c (zlarfg);
//     lg_symv_lower_kernel + lg_symv_reduce_kernel
//                      y = A22 v - V (W^H v) - W (V^H v) from the lower triangle
//                      in 32 x 32 tiles (row and column sums from one read),
//                      blocks of 4 x 16 (late columns 4 x 8) tiles per workgroup: the trailing lower
//                      triangle is streamed once per column; one extra workgroup
//                      per matrix forms W^H v, V^H v beside the tiles, and the
//                      reduction's idle half-wave applies the older pending terms
//                      to the next column, strip by strip;
//     lg_her2k_kernel  after each panel A22 -= V W^H + W V^H on the lower tiles, on
//                      the matrix cores.
//   Divide & conquer: stedc_kernel with its node arrays in global memory.
//   Back-transform: larft_kernel with several workgroups per matrix, the factors of
//     128-reflector passes (lgbt_tbig_kernel), then lgbt_slab_kernel: a slab of 64
//     eigenvector columns per workgroup through every pass, one read and one write
//     of the slab per pass (lg_backtransform2_kernel, two 16-reflector blocks per
//     read and write, remains for complex double 256 < n <= 512).
#include "common.hpp"
#include "eig.hpp"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

namespace rlm {

constexpr int kLgNB = 32;      // reflectors per panel
constexpr int kLgColNT = 512;  // threads of the column kernel
constexpr int kLgRPT = 8;      // rows per thread of the column kernel: n <= 4096

int heevd_large_max_n() { return kLgColNT * kLgRPT; }

// Launch timing of the Hermitian product kernel (lg_symv_lower_kernel, one launch per
// column): while enabled, eager launches are bracketed by HIP events from a pool.
namespace {
std::mutex g_kt_mu;
bool g_kt_on = false;
std::vector<hipEvent_t> g_kt_ev;
size_t g_kt_used = 0;

// the two events of the next timed launch (nullptr: timing off or the stream is capturing)
void kt_pair(hipStream_t s, hipEvent_t &e0, hipEvent_t &e1) {
    e0 = e1 = nullptr;
    std::lock_guard<std::mutex> lk(g_kt_mu);
    if (!g_kt_on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    while (g_kt_ev.size() < g_kt_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        g_kt_ev.push_back(e);
    }
    e0 = g_kt_ev[g_kt_used];
    e1 = g_kt_ev[g_kt_used + 1];
    g_kt_used += 2;
}
}  // namespace

int symv_timing(int enable, double *total_ms, long *launches) {
    std::lock_guard<std::mutex> lk(g_kt_mu);
    if (enable) {
        g_kt_on = true;
        g_kt_used = 0;
        return 0;
    }
    g_kt_on = false;
    double tot = 0.0;
    if (g_kt_used) RLM_HIP_CHECK(hipEventSynchronize(g_kt_ev[g_kt_used - 1]));
    for (size_t k = 0; k + 1 < g_kt_used; k += 2) {
        float ms = 0.f;
        RLM_HIP_CHECK(hipEventElapsedTime(&ms, g_kt_ev[k], g_kt_ev[k + 1]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = (long)(g_kt_used / 2);
    g_kt_used = 0;
    return 0;
}

// Per-matrix panel workspace, complex elements: Vp (kLgNB x n, column l = reflector
// k0 + l with explicit zeros above and the unit entry), Wp (kLgNB x n), P (2 kLgNB:
// W^H v then V^H v of the current column), then the Hermitian product's partial sums
// over strips of 32 rows (strip K = rows base + 32 K .., base = (c + 1) rounded down to
// a multiple of 8, so that every 8-row group of a column is one aligned segment):
// Rp (ns x njb x 32: row sums of strip K over the tiles of column block jb) and Cp
// (ns x nib x 32: column sums of tile column K over the strips of row block ib),
// ns = strips of the first column.
constexpr int kLgTS = 32;   // tile edge
constexpr int kLgTR = 4;    // strips per row block (one per wave)
constexpr int kLgTC = 16;   // tiles per column block (at most)
constexpr int kLgTCmin = 8; // tiles per column block where the trailing matrix is small
static long lg_strips(int n) { return (n + kLgTS - 1) / kLgTS; }
__host__ __device__ inline long lg_row_blocks(long ns) { return (ns + kLgTR - 1) / kLgTR; }
// column blocks of a strip's row sums as allocated (the narrowest blocks)
__host__ __device__ inline long lg_col_blocks(long ns) { return (ns + kLgTCmin - 1) / kLgTCmin; }
long heevd_large_aux_complex(int n) {
    const long ns = lg_strips(n);
    return 2L * kLgNB * n + 2L * kLgNB + ns * (lg_row_blocks(ns) + lg_col_blocks(ns)) * kLgTS;
}

// ----------------------------------------------------------------------------
// Column c = k0 + i of the panel starting at k0 (nbk reflectors). i == nbk: only
// finish w of the last panel column (and d[n-1] when c == n - 1).
template <typename T>
__global__ __launch_bounds__(kLgColNT) void lg_col_kernel(T *__restrict__ Abase, long strideA, int n,
                                                          double *__restrict__ Dout, double *__restrict__ Eout,
                                                          T *__restrict__ TAUout, long strideV, T *__restrict__ aux,
                                                          long strideAux, int k0, int i, int nbk) {
    using C = cx<T>;
    constexpr int NT = kLgColNT, NW = NT / kWave;
    __shared__ T red[2 * NW];
    __shared__ C rowc[2 * kLgNB];
    __shared__ T sc[8];
    const long b = blockIdx.x;
    const int t = threadIdx.x;
    T *A = Abase + 2 * b * strideA;
    C *Vp = reinterpret_cast<C *>(aux + 2 * b * strideAux);
    C *Wp = Vp + (long)kLgNB * n;
    double *Dv = Dout + b * strideV, *Ev = Eout + b * strideV;
    T *Tv = TAUout + b * strideV * (long)(sizeof(double) / sizeof(T));
    const int c = k0 + i;
    // (a) w of panel column i - 1 (reflector c - 1, support rows >= c)
    if (i > 0) {
        const C tau{Tv[2 * (c - 1)], Tv[2 * (c - 1) + 1]};
        C *y = Wp + (long)(i - 1) * n;
        const C *v = Vp + (long)(i - 1) * n;
        C wl[kLgRPT];
        C s{(T)0, (T)0};
#pragma unroll
        for (int q = 0; q < kLgRPT; q++) {
            const int r = c + t + q * NT;
            wl[q] = C{(T)0, (T)0};
            if (r < n) {
                wl[q] = tau * y[r];
                fmacx(s, wl[q], v[r]);                       // s += conj(w) v
            }
        }
        s = block_sum_cx(s, red);
        const C alpha = (tau * s) * (T)(-0.5);
#pragma unroll
        for (int q = 0; q < kLgRPT; q++) {
            const int r = c + t + q * NT;
            if (r < n) y[r] = wl[q] + alpha * v[r];
        }
        __syncthreads();
    }
    const bool refl = i < nbk && c < n - 1;
    const bool last = c == n - 1;
    if (!refl && !last) return;
    // (b) column c (rows >= c) with the panel's pending rank-2i update: inside a panel
    // (0 < i < nbk) lg_symv_reduce_kernel of column c - 1 has applied the terms l < i - 1
    // to A(c:n, c) already, so only l = i - 1 remains; the column after the panel's last
    // (i == nbk, the last column of the matrix) takes them all
    const int lfirst = (i > 0 && i < nbk) ? i - 1 : 0;
    if (t >= lfirst && t < i) {
        rowc[t] = Wp[(long)t * n + c];
        rowc[kLgNB + t] = Vp[(long)t * n + c];
    }
    __syncthreads();
    C a[kLgRPT];
    T nrm2 = 0;
#pragma unroll
    for (int q = 0; q < kLgRPT; q++) {
        const int r = c + t + q * NT;
        a[q] = r < n ? ldx(A, (long)c * n + r) : C{(T)0, (T)0};
    }
    // the pending update over all rows of the thread, UL panel columns per step: the
    // 2 UL kLgRPT loads of a step are independent, so their latencies overlap (the products
    // are applied in panel-column order, as one column at a time would); complex double
    // keeps one column per step (two would not fit 256 registers)
    constexpr int UL = sizeof(T) == sizeof(float) ? 2 : 1;
    int l = lfirst;
    for (; l + UL <= i; l += UL) {
        C xv[UL][kLgRPT], xw[UL][kLgRPT];
#pragma unroll
        for (int u = 0; u < UL; u++) {
            const C *vl = Vp + (long)(l + u) * n, *wl = Wp + (long)(l + u) * n;
#pragma unroll
            for (int q = 0; q < kLgRPT; q++) {
                const int r = c + t + q * NT;
                xv[u][q] = r < n ? vl[r] : C{(T)0, (T)0};
                xw[u][q] = r < n ? wl[r] : C{(T)0, (T)0};
            }
        }
#pragma unroll
        for (int u = 0; u < UL; u++) {
            const C wc = conjx(rowc[l + u]), vc = conjx(rowc[kLgNB + l + u]);
#pragma unroll
            for (int q = 0; q < kLgRPT; q++) a[q] = a[q] - (xv[u][q] * wc + xw[u][q] * vc);
        }
    }
    for (; l < i; l++) {
        const C wc = conjx(rowc[l]), vc = conjx(rowc[kLgNB + l]);
        const C *vl = Vp + (long)l * n, *wl = Wp + (long)l * n;
#pragma unroll
        for (int q = 0; q < kLgRPT; q++) {
            const int r = c + t + q * NT;
            if (r < n) a[q] = a[q] - (vl[r] * wc + wl[r] * vc);
        }
    }
#pragma unroll
    for (int q = 0; q < kLgRPT; q++) {
        const int r = c + t + q * NT;
        C x = a[q];
        if (r < n) {
            if (r == c) {
                x.y = (T)0;
                sc[0] = x.x;
            } else if (r == c + 1) {
                sc[1] = x.x;
                sc[2] = x.y;
            } else {
                nrm2 += abs2x(x);
            }
        }
        a[q] = x;
    }
    nrm2 = block_sum_t(nrm2, red);                           // its barriers publish sc[0..2]
    if (last) {
        if (t == 0) Dv[c] = (double)sc[0];
        return;
    }
    if (t == 0) {
        Dv[c] = (double)sc[0];
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
        Ev[c] = (double)beta;
        Tv[2 * c] = taur;
        Tv[2 * c + 1] = taui;
        sc[5] = sr; sc[6] = si; sc[7] = beta;
    }
    __syncthreads();
    const C scal{sc[5], sc[6]};
    // (c) v: reflector storage in A, panel column i of Vp; rows k0..c of
    //     Vp(:, i) and Wp(:, i) cleared (the product fills Wp below row c)
#pragma unroll
    for (int q = 0; q < kLgRPT; q++) {
        const int r = c + t + q * NT;
        if (r < n) {
            C v{(T)0, (T)0};
            if (r == c + 1) {
                v = C{(T)1, (T)0};
                stx(A, (long)c * n + r, C{sc[7], (T)0});
            } else if (r > c + 1) {
                v = a[q] * scal;
                stx(A, (long)c * n + r, v);
            }
            Vp[(long)i * n + r] = v;
        }
    }
    for (int r = k0 + t; r <= c; r += NT) {
        if (r < c) Vp[(long)i * n + r] = C{(T)0, (T)0};
        Wp[(long)i * n + r] = C{(T)0, (T)0};
    }
    __syncthreads();
    // (the dot products W^H v, V^H v of this reflector are formed by the first workgroups
    // of lg_symv_lower_kernel, beside the Hermitian product)
}

// ----------------------------------------------------------------------------
// Hermitian product from the lower triangle: y = A22 v with A22 = A(s:n, s:n), s = c+1,
// in 32 x 32 tiles (I, J), I >= J, of the strips of rows base + 32 I .. (rows and columns
// below s masked; base = s rounded down to a multiple of 8). Workgroup (b, ib, jb) takes
// the tiles of strips 4 ib .. 4 ib + 3 (wave w: strip 4 ib + w) and tile columns
// tc jb .. tc jb + tc - 1 (tc = 16, or 8 for the small trailing matrices of the late
// columns, where more, smaller workgroups balance the chip), so the four waves read
// neighbouring 256-byte segments of the same columns together. Lane (lrg, lcg) = (lane & 7, lane >> 3) holds the 4 x 4 sub-block at
// tile rows lrg + 8 ru (the 8 lanes of a column group read 8 consecutive rows of the
// column-major matrix per load), columns 4 lcg.., read once for both y_I += A_IJ v_J
// (row sums: kept in registers over the wave's tiles, reduced over lcg once, into
// Rp[I][jb]) and y_J += A_IJ^H v_I (column sums: reduced over lrg per tile into LDS,
// summed over the waves in a fixed order into Cp[J][ib]).
template <typename T>
__global__ __launch_bounds__(256) void lg_symv_lower_kernel(const T *__restrict__ Abase, long strideA, int n,
                                                            T *__restrict__ aux, long strideAux, int k0, int i,
                                                            int per, int ns0, int nb, int tc) {
    using C = cx<T>;
    // complex float issues the next tile's loads before this tile's sums (complex double
    // has no registers to spare for it and gains nothing)
    constexpr bool kPrefetch = sizeof(T) == sizeof(float);
    __shared__ C vrow[kLgTR * kLgTS];
    __shared__ C vcol[kLgTC * kLgTS];
    __shared__ C colp[kLgTR][kLgTC][kLgTS];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    const int c = k0 + i, s0 = c + 1, m = n - s0;
    if (blockIdx.x < (unsigned)nb) {
        // workgroup b < nb: P[l] = W(:, l)^H v, P[kLgNB + l] = V(:, l)^H v over the rows > c
        // of matrix b, one wave per product (dispatched first, they run beside the tiles)
        const long bm = blockIdx.x;
        C *Vq = reinterpret_cast<C *>(aux + 2 * bm * strideAux);
        C *Wq = Vq + (long)kLgNB * n;
        C *P = Wq + (long)kLgNB * n;
        const C *vs = Vq + (long)i * n + s0;
        for (int qd = w; qd < 2 * i; qd += 4) {
            const C *col = ((qd < i) ? Wq + (long)qd * n : Vq + (long)(qd - i) * n) + s0;
            C s{(T)0, (T)0};
            int e = lane;
            // eight products in flight per step, accumulated in the same order as one at a time
            for (; e + 7 * kWave < m; e += 8 * kWave) {
                C x[8], y[8];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    x[k] = col[e + k * kWave];
                    y[k] = vs[e + k * kWave];
                }
#pragma unroll
                for (int k = 0; k < 8; k++) fmacx(s, x[k], y[k]);
            }
            for (; e < m; e += kWave) fmacx(s, col[e], vs[e]);
            s.x = wave_sum_t(s.x);
            s.y = wave_sum_t(s.y);
            if (lane == 0) P[qd < i ? qd : kLgNB + (qd - i)] = s;
        }
        return;
    }
    const int lrg = lane & 7, lcg = lane >> 3;
    const int base = s0 & ~7, mb = n - base, lo = s0 - base;   // local rows / columns lo .. mb-1
    const int ns = (mb + kLgTS - 1) / kLgTS, nib = (ns + kLgTR - 1) / kLgTR;
    // workgroup -> (matrix b, row block ib, column block jb); row block ib has
    // floor(min(4 ib + 3, ns - 1) / 16) + 1 column blocks
    const long b = (blockIdx.x - nb) / per;
    int q = (int)((blockIdx.x - nb) % per), ib = 0;
    for (; ib < nib - 1; ib++) {
        const int nj = min(kLgTR * ib + kLgTR - 1, ns - 1) / tc + 1;
        if (q < nj) break;
        q -= nj;
    }
    const int jb = q;
    const int I = kLgTR * ib + w;
    const int J0 = jb * tc, J1 = min(J0 + tc, min(kLgTR * ib + kLgTR - 1, ns - 1) + 1);
    const T *A = Abase + 2 * b * strideA;
    C *Vp = reinterpret_cast<C *>(aux + 2 * b * strideAux);
    const C *v = Vp + (long)i * n;                             // v(r) for rows r > c
    C *Rp = Vp + 2L * kLgNB * n + 2L * kLgNB;
    C *Cp = Rp + (long)ns0 * lg_col_blocks(ns0) * kLgTS;
    const int njbm = (int)lg_col_blocks(ns0), nibm = (int)lg_row_blocks(ns0);
    for (int e = t; e < kLgTR * kLgTS; e += 256) {
        const int r = base + kLgTR * ib * kLgTS + e;
        vrow[e] = (r >= s0 && r < n) ? v[r] : C{(T)0, (T)0};
    }
    for (int e = t; e < (J1 - J0) * kLgTS; e += 256) {
        const int r = base + J0 * kLgTS + e;
        vcol[e] = (r >= s0 && r < n) ? v[r] : C{(T)0, (T)0};
    }
    __syncthreads();
    const int rb = I * kLgTS + lrg;                          // local rows rb + 8 ru of the sub-block
    C racc[4], vr[4];
#pragma unroll
    for (int ru = 0; ru < 4; ru++) {
        racc[ru] = C{(T)0, (T)0};
        vr[ru] = vrow[w * kLgTS + lrg + 8 * ru];
    }
    auto load_tile = [&](int J, C (&dst)[4][4]) {
        const int cb = J * kLgTS + 4 * lcg;
#pragma unroll
        for (int cu = 0; cu < 4; cu++) {
            const long colo = (long)(base + min(cb + cu, mb - 1)) * n + base;
#pragma unroll
            for (int ru = 0; ru < 4; ru++) dst[ru][cu] = ldx(A, colo + min(rb + 8 * ru, mb - 1));
        }
    };
    const int Je = I < ns ? min(J1, I + 1) : J0;            // this wave's tiles J0 .. Je-1
    C a[4][4];
    if (kPrefetch && J0 < Je) load_tile(J0, a);
    for (int J = J0; J < Je; J++) {
        const int cb = J * kLgTS + 4 * lcg;
        C an[4][4], vc[4];
        if (kPrefetch)
            load_tile(min(J + 1, Je - 1), an);
        else
            load_tile(J, a);
#pragma unroll
        for (int cu = 0; cu < 4; cu++) vc[cu] = vcol[(J - J0) * kLgTS + 4 * lcg + cu];
        const bool diag = J == I;
        C csum[4];
#pragma unroll
        for (int cu = 0; cu < 4; cu++) csum[cu] = C{(T)0, (T)0};
#pragma unroll
        for (int ru = 0; ru < 4; ru++)
#pragma unroll
            for (int cu = 0; cu < 4; cu++) {
                const int lr = rb + 8 * ru, lc = cb + cu;
                const bool in = lr < mb && lc < mb && lr >= lo && lc >= lo && (!diag || lc <= lr);
                const C x = in ? a[ru][cu] : C{(T)0, (T)0};
                fmax_(racc[ru], x, vc[cu]);                                  // A(r, c) v_c
                if (!diag || lc < lr) fmacx(csum[cu], x, vr[ru]);           // conj(A(r, c)) v_r
            }
        // column sums over lrg (lane bits 0-2), then lanes lrg == 0 keep them in LDS
#pragma unroll
        for (int cu = 0; cu < 4; cu++)
#pragma unroll
            for (int off = 1; off < 8; off <<= 1) {
                csum[cu].x += __shfl_xor(csum[cu].x, off, kWave);
                csum[cu].y += __shfl_xor(csum[cu].y, off, kWave);
            }
        if (lrg == 0) {
#pragma unroll
            for (int cu = 0; cu < 4; cu++) colp[w][J - J0][4 * lcg + cu] = csum[cu];
        }
        if (kPrefetch) {
#pragma unroll
            for (int ru = 0; ru < 4; ru++)
#pragma unroll
                for (int cu = 0; cu < 4; cu++) a[ru][cu] = an[ru][cu];
        }
    }
    // row sums over lcg (lane bits 3-5)
#pragma unroll
    for (int ru = 0; ru < 4; ru++)
#pragma unroll
        for (int off = 8; off < kWave; off <<= 1) {
            racc[ru].x += __shfl_xor(racc[ru].x, off, kWave);
            racc[ru].y += __shfl_xor(racc[ru].y, off, kWave);
        }
    if (I < ns && lcg == 0) {
#pragma unroll
        for (int ru = 0; ru < 4; ru++) Rp[((long)I * njbm + jb) * kLgTS + lrg + 8 * ru] = racc[ru];
    }
    __syncthreads();
    // column sums of tile column J over the strips J .. of this row block, waves in order
    for (int e = t; e < (J1 - J0) * kLgTS; e += 256) {
        const int J = J0 + e / kLgTS;
        C sacc{(T)0, (T)0};
#pragma unroll
        for (int ww = 0; ww < kLgTR; ww++)
            if (kLgTR * ib + ww >= J && kLgTR * ib + ww < ns) sacc = sacc + colp[ww][e / kLgTS][e % kLgTS];
        Cp[((long)J * nibm + ib) * kLgTS + e % kLgTS] = sacc;
    }
}

// y(r) = sum_jb Rp[K][jb][r'] + sum_{ib >= K/4} Cp[K][ib][r'] - V (W^H v) - W (V^H v) for
// the rows r = base + 32 K + r' >= c + 1 of strip K, into Wp(:, i).
// Threads 32..63 meanwhile apply to column c + 1 (rows of the strip, all >= c + 1) the
// terms l < i of the panel's pending update, when column c + 1 is in the same panel
// (i + 1 < nbk): A(r, c+1) -= V(r, l) conj(W(c+1, l)) + W(r, l) conj(V(c+1, l)), in the
// order lg_col_kernel would apply them; lg_col_kernel then adds only the term l = i.
template <typename T>
__global__ __launch_bounds__(64) void lg_symv_reduce_kernel(T *__restrict__ Abase, long strideA, T *__restrict__ aux,
                                                            long strideAux, int n, int k0, int i, int nbk, int ns0,
                                                            int tc) {
    using C = cx<T>;
    const int c = k0 + i, s0 = c + 1;
    const int base = s0 & ~7, mb = n - base;
    const int ns = (mb + kLgTS - 1) / kLgTS, nib = (ns + kLgTR - 1) / kLgTR;
    const long b = blockIdx.x / ns;
    const int K = (int)(blockIdx.x % ns);
    const int t = threadIdx.x;
    C *Vp = reinterpret_cast<C *>(aux + 2 * b * strideAux);
    C *Wp = Vp + (long)kLgNB * n;
    const C *P = Wp + (long)kLgNB * n;
    const int njbm = (int)lg_col_blocks(ns0), nibm = (int)lg_row_blocks(ns0);
    const C *Rp = Vp + 2L * kLgNB * n + 2L * kLgNB;
    const C *Cp = Rp + (long)ns0 * njbm * kLgTS;
    const int r = base + K * kLgTS + (t & (kLgTS - 1));
    if (r < s0 || r >= n) return;
    if (t >= kLgTS) {
        if (i + 1 >= nbk || i == 0) return;
        const int cn = c + 1;
        T *A = Abase + 2 * b * strideA;
        C x = ldx(A, (long)cn * n + r);
        for (int l = 0; l < i; l++)
            x = x - (Vp[(long)l * n + r] * conjx(Wp[(long)l * n + cn]) + Wp[(long)l * n + r] * conjx(Vp[(long)l * n + cn]));
        stx(A, (long)cn * n + r, x);
        return;
    }
    // loads issued ahead of the fixed-order sums
    const int njb = K / tc + 1, ib0 = K / kLgTR;
    C y{(T)0, (T)0};
    for (int jb = 0; jb < njb; jb++) y = y + Rp[((long)K * njbm + jb) * kLgTS + t];
    int ib = ib0;
    for (; ib + 8 <= nib; ib += 8) {
        C x[8];
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = Cp[((long)K * nibm + ib + k) * kLgTS + t];
#pragma unroll
        for (int k = 0; k < 8; k++) y = y + x[k];
    }
    for (; ib < nib; ib++) y = y + Cp[((long)K * nibm + ib) * kLgTS + t];
    int l = 0;
    for (; l + 4 <= i; l += 4) {
        C xv[4], xw[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            xv[k] = Vp[(long)(l + k) * n + r];
            xw[k] = Wp[(long)(l + k) * n + r];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) y = y - (xv[k] * P[l + k] + xw[k] * P[kLgNB + l + k]);
    }
    for (; l < i; l++) y = y - (Vp[(long)l * n + r] * P[l] + Wp[(long)l * n + r] * P[kLgNB + l]);
    Wp[(long)i * n + r] = y;
}

// ----------------------------------------------------------------------------
// A(r, c') -= sum_l V(r,l) conj(W(c',l)) + W(r,l) conj(V(c',l)) for r, c' >= s0 on the
// 64 x 64 tiles of the lower triangle (diagonal tiles whole; the upper triangle is never
// read again). With X = [V W], Y = [W V] (n x 2 nbk) the update is X Y^H; the
// matrix cores compute its transpose D[c'][r] = sum_k (-conj(Y(c',k))) X(r,k) so
// that the 16 lanes of an accumulator row hold 16 consecutive rows of A
// (coalesced in column-major storage). 64 x 64 tile per workgroup, 32 x 32 per wave.
template <typename T>
__global__ __launch_bounds__(256) void lg_her2k_kernel(T *__restrict__ Abase, long strideA, int n,
                                                       const T *__restrict__ aux, long strideAux, int s0, int nbk,
                                                       int nt) {
    using M = Mf<T>;
    using v4 = typename M::v4;
    using C = cx<T>;
    const int tiles = nt * (nt + 1) / 2;                    // tiles (tr, tc), tr >= tc
    const long b = blockIdx.x / tiles;
    const int tile = (int)(blockIdx.x % tiles);
    int tr = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
    if ((tr + 1) * (tr + 2) / 2 <= tile) tr++;
    if (tr * (tr + 1) / 2 > tile) tr--;
    const int tc = tile - tr * (tr + 1) / 2;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int li = lane & 15, lq = lane >> 4;
    const int R0 = s0 + 64 * tr + 32 * (w & 1), C0 = s0 + 64 * tc + 32 * (w >> 1);
    T *A = Abase + 2 * b * strideA;
    const C *Vp = reinterpret_cast<const C *>(aux + 2 * b * strideAux);
    const C *Wp = Vp + (long)kLgNB * n;
    v4 dr[2][2], di[2][2];
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            const int r = R0 + 16 * rb + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int cc = C0 + 16 * cb + M::row(lane, v);
                const C x = (r < n && cc < n) ? ldx(A, (long)cc * n + r) : C{(T)0, (T)0};
                dr[cb][rb][v] = x.x;
                di[cb][rb][v] = x.y;
            }
        }
    // D -= conj(ya) xb. Complex float: three real MFMAs per complex product (a = conj ya,
    // b = xb): P1 = (a_r + a_i) b_r in q, D_r -= P1 - P3 and D_i -= P1 + P2 with
    // P3 = a_i (b_r + b_i), P2 = a_r (b_i - b_r) (n_phi = 3000: 63.2 -> 56.8 ms). Complex
    // double keeps four: the q accumulators would cost it a wave per SIMD (128 -> 153 ms).
    constexpr bool G3 = sizeof(T) == sizeof(float);
    v4 q[2][2];
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++) q[cb][rb] = v4{0, 0, 0, 0};
    const int K = 2 * nbk;
    for (int k4 = 0; k4 < K; k4 += 4) {
        const int k = k4 + lq;
        const bool kin = k < K;
        const C *Xc = k < nbk ? Vp + (long)k * n : Wp + (long)(k - nbk) * n;
        const C *Yc = k < nbk ? Wp + (long)k * n : Vp + (long)(k - nbk) * n;
        C ya[2], xb[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int cc = C0 + 16 * u + li, r = R0 + 16 * u + li;
            ya[u] = (kin && cc < n) ? Yc[cc] : C{(T)0, (T)0};
            xb[u] = (kin && r < n) ? Xc[r] : C{(T)0, (T)0};
        }
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
            const T yd = ya[cb].x - ya[cb].y;
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                const T br = xb[rb].x, bi = xb[rb].y;
                if constexpr (G3) {
                    q[cb][rb] = M::mma(yd, br, q[cb][rb]);
                    dr[cb][rb] = M::mma(-ya[cb].y, br + bi, dr[cb][rb]);
                    di[cb][rb] = M::mma(-ya[cb].x, bi - br, di[cb][rb]);
                } else {
                    dr[cb][rb] = M::mma(-ya[cb].x, br, dr[cb][rb]);
                    dr[cb][rb] = M::mma(-ya[cb].y, bi, dr[cb][rb]);
                    di[cb][rb] = M::mma(-ya[cb].x, bi, di[cb][rb]);
                    di[cb][rb] = M::mma(ya[cb].y, br, di[cb][rb]);
                }
            }
        }
    }
    if constexpr (G3) {
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                dr[cb][rb] -= q[cb][rb];
                di[cb][rb] -= q[cb][rb];
            }
    }
#pragma unroll
    for (int cb = 0; cb < 2; cb++)
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            const int r = R0 + 16 * rb + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int cc = C0 + 16 * cb + M::row(lane, v);
                if (r < n && cc < n) stx(A, (long)cc * n + r, C{dr[cb][rb][v], di[cb][rb][v]});
            }
        }
}

// ----------------------------------------------------------------------------
// Back-transform U = Q U, Q = H_0 H_1 ... H_{n-2}: blocks of 16 reflectors
// I - V T V^H (T from larft_kernel), last block first. A workgroup owns a slab of
// 16 NCB eigenvector columns of one matrix; wave w owns the 16-row blocks w, w+4,
// ... of it in every phase, so U is only ever re-read by the wave that wrote it. The
// slabs of one matrix share blockIdx % 8 (one XCD, whose L2 serves V and T).
// Element (row, k) of reflector block j0 (unit at row j+1, A(j+2:n, j) below).
template <typename T>
__device__ inline cx<T> lg_refl(const T *A, int n, int j0, int row, int k) {
    const int j = j0 + k;
    if (j >= n - 1 || row <= j || row >= n) return cx<T>{(T)0, (T)0};
    if (row == j + 1) return cx<T>{(T)1, (T)0};
    return ldx(A, (long)j * n + row);
}

// Two reflector blocks per pass (blocks B0 = bk and B1 = bk - 1, applied in that order):
// the slab's U rows are read and written once per 32 reflectors instead of once per 16.
// Entering a pass, the waves hold partials of W0 = V_B0^H U, W1' = V_B1^H U and
// C2 = V_B0^H V_B1 (all from the U before B0); summed over the waves in LDS (fixed
// order) they give W2_0 = T_B0 W0, and W1 = W1' - C W2_0 with C = C2^H (the W of B1 after
// B0 is applied), W2_1 = T_B1 W1. The pass then applies U -= V_B0 W2_0 + V_B1 W2_1 row
// block by row block and accumulates the next pair's W0, W1' and C2 from the updated
// rows. Complex float prefetches the next row block's U rows while the current one runs
// on the matrix cores (prefetching its V operands too, or the U rows in complex double,
// would not fit 256 registers). Two waves per SIMD (at most 256 registers): the second
// workgroup per CU hides the per-pair barriers and LDS sums. n_phi = 3000, 64 matrices:
// one block per pass took 368 ms (complex float, one wave per SIMD) / 300 ms (two) and
// 527 ms (complex double); two blocks per pass take 247 / 425 ms (the U stream is the
// bound: skipping the stores alone saved 138 ms).
template <typename T, int NCB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void lg_backtransform2_kernel(
        const T *__restrict__ Abase, long strideA, int n, const T *__restrict__ Tin, long strideT,
        T *__restrict__ Uout, long strideU, int nslab, int batch) {
    using M = Mf<T>;
    using v4 = typename M::v4;
    using CT = cx<T>;
    constexpr int WS = 2 * NCB * 4 * kWave;               // one W partial (re, im) of a wave
    constexpr bool PREFU = sizeof(T) == sizeof(float);
    __shared__ T wred[4][2 * WS + 2 * 4 * kWave];          // W0, W1', C2 partials per wave
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int li = lane & 15;
    const long q = blockIdx.x >> 3;
    const long b = (q / nslab) * 8 + (blockIdx.x & 7);
    const int slab = (int)(q % nslab);
    if (b >= batch) return;                                  // whole workgroup: no barrier is skipped
    const T *A = Abase + 2 * b * strideA;
    const CT *Tb = reinterpret_cast<const CT *>(Tin + 2 * b * strideT);
    T *U = Uout + 2 * b * strideU;
    const int c0 = slab * 16 * NCB;
    const int nblk = (n - 1 + 15) / 16;
    // V operands of block bk for the row block R: apply form V[R + li][k = row(l, s)],
    // accumulate form V[R + row(l, v)][k = li]; zero for a block index < 0
    auto vapply = [&](int bk, int R, CT (&x)[4]) {
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++)
            x[s2] = bk >= 0 ? lg_refl(A, n, 16 * bk, R + li, M::row(lane, s2)) : CT{(T)0, (T)0};
    };
    auto vacc = [&](int bk, int R, CT (&x)[4]) {
#pragma unroll
        for (int v = 0; v < 4; v++)
            x[v] = bk >= 0 ? lg_refl(A, n, 16 * bk, R + M::row(lane, v), li) : CT{(T)0, (T)0};
    };
    auto loadu = [&](int R, v4 (&ur)[NCB], v4 (&ui)[NCB]) {
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int col = c0 + 16 * cb + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = R + M::row(lane, v);
                const CT x = (row < n && col < n) ? ldx(U, (long)row * n + col) : CT{(T)0, (T)0};
                ur[cb][v] = x.x;
                ui[cb][v] = x.y;
            }
        }
    };
    // acc += conj(V)^T B (three real products per complex one: W = q + (r, i))
    auto accw = [&](const CT (&xv)[4], const v4 (&br_)[NCB], const v4 (&bi_)[NCB], v4 (&r)[NCB], v4 (&i)[NCB],
                    v4 (&qq)[NCB]) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const CT x = xv[v];
            const T xd = x.x - x.y;
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) {
                const T br = br_[cb][v], bi = bi_[cb][v];
                qq[cb] = M::mma(xd, br, qq[cb]);
                r[cb] = M::mma(x.y, br + bi, r[cb]);
                i[cb] = M::mma(x.x, bi - br, i[cb]);
            }
        }
    };
    auto accc = [&](const CT (&xa)[4], const CT (&xb)[4], v4 &r, v4 &i, v4 &qq) {   // C2 += conj(Va)^T Vb
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const CT x = xa[v];
            const T xd = x.x - x.y, br = xb[v].x, bi = xb[v].y;
            qq = M::mma(xd, br, qq);
            r = M::mma(x.y, br + bi, r);
            i = M::mma(x.x, bi - br, i);
        }
    };
    const v4 z4 = v4{0, 0, 0, 0};
    v4 p0r[NCB], p0i[NCB], p0q[NCB], p1r[NCB], p1i[NCB], p1q[NCB];     // W0, W1' partials
    v4 pcr = z4, pci = z4, pcq = z4;                                    // C2 partial
#pragma unroll
    for (int cb = 0; cb < NCB; cb++) {
        p0r[cb] = p0i[cb] = p0q[cb] = z4;
        p1r[cb] = p1i[cb] = p1q[cb] = z4;
    }
    // partials of the first pair from the original U
    {
        const int B0 = nblk - 1, B1 = nblk - 2;
        const int lo = 16 * (B1 >= 0 ? B1 : B0) + 1;
        for (int rbk = w; 16 * rbk < n; rbk += 4) {
            const int R = 16 * rbk;
            if (R + 15 < lo) continue;
            v4 ur[NCB], ui[NCB];
            loadu(R, ur, ui);
            CT xa[4], xb[4];
            vacc(B0, R, xa);
            vacc(B1, R, xb);
            accw(xa, ur, ui, p0r, p0i, p0q);
            if (B1 >= 0) {
                accw(xb, ur, ui, p1r, p1i, p1q);
                accc(xa, xb, pcr, pci, pcq);
            }
        }
    }
    for (int B0 = nblk - 1; B0 >= 0; B0 -= 2) {
        const int B1 = B0 - 1, N0 = B0 - 2, N1 = B0 - 3;             // this pair, the next pair
        T *wb = wred[w];
#pragma unroll
        for (int cb = 0; cb < NCB; cb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                wb[((0 * NCB + cb) * 4 + v) * kWave + lane] = p0q[cb][v] + p0r[cb][v];
                wb[((1 * NCB + cb) * 4 + v) * kWave + lane] = p0q[cb][v] + p0i[cb][v];
                wb[WS + ((0 * NCB + cb) * 4 + v) * kWave + lane] = p1q[cb][v] + p1r[cb][v];
                wb[WS + ((1 * NCB + cb) * 4 + v) * kWave + lane] = p1q[cb][v] + p1i[cb][v];
            }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            wb[2 * WS + v * kWave + lane] = pcq[v] + pcr[v];
            wb[2 * WS + (4 + v) * kWave + lane] = pcq[v] + pci[v];
        }
        __syncthreads();
        v4 w0r[NCB], w0i[NCB], w1r[NCB], w1i[NCB];
        T c2r[4], c2i[4];
#pragma unroll
        for (int cb = 0; cb < NCB; cb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
                for (int ww = 0; ww < 4; ww++) {
                    a0 += wred[ww][((0 * NCB + cb) * 4 + v) * kWave + lane];
                    a1 += wred[ww][((1 * NCB + cb) * 4 + v) * kWave + lane];
                    a2 += wred[ww][WS + ((0 * NCB + cb) * 4 + v) * kWave + lane];
                    a3 += wred[ww][WS + ((1 * NCB + cb) * 4 + v) * kWave + lane];
                }
                w0r[cb][v] = a0;
                w0i[cb][v] = a1;
                w1r[cb][v] = a2;
                w1i[cb][v] = a3;
            }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            T a0 = 0, a1 = 0;
#pragma unroll
            for (int ww = 0; ww < 4; ww++) {
                a0 += wred[ww][2 * WS + v * kWave + lane];
                a1 += wred[ww][2 * WS + (4 + v) * kWave + lane];
            }
            c2r[v] = a0;
            c2i[v] = a1;
        }
        __syncthreads();                                     // wred is rewritten by the next pair
        // D = X Y for a 16 x 16 X given as the A operand of each K-step s (xs[s] =
        // X[li][row(l, s)]) and Y in accumulator layout (lane holds Y[row(l, s)][li])
        auto mul = [&](const CT (&xs)[4], const v4 (&yr)[NCB], const v4 (&yi)[NCB], v4 (&dr)[NCB], v4 (&di)[NCB]) {
            v4 qq[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) dr[cb] = di[cb] = qq[cb] = z4;
#pragma unroll
            for (int s2 = 0; s2 < 4; s2++) {
                const CT tv = xs[s2];
                const T ts = tv.x + tv.y;
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    const T br = yr[cb][s2], bi = yi[cb][s2];
                    qq[cb] = M::mma(ts, br, qq[cb]);
                    dr[cb] = M::mma(-tv.y, br + bi, dr[cb]);
                    di[cb] = M::mma(tv.x, bi - br, di[cb]);
                }
            }
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) {
                dr[cb] += qq[cb];
                di[cb] += qq[cb];
            }
        };
        v4 w20r[NCB], w20i[NCB], w21r[NCB], w21i[NCB];
        {
            CT t0[4];
#pragma unroll
            for (int s2 = 0; s2 < 4; s2++) t0[s2] = Tb[B0 * 256 + M::row(lane, s2) * 16 + li];   // T_B0[li][row]
            mul(t0, w0r, w0i, w20r, w20i);                                                       // W2_0 = T_B0 W0
        }
        if (B1 >= 0) {
            CT cs[4];
#pragma unroll
            for (int s2 = 0; s2 < 4; s2++) cs[s2] = CT{c2r[s2], -c2i[s2]};    // C[li][row(l, s)] = conj C2[row][li]
            v4 dr[NCB], di[NCB];
            mul(cs, w20r, w20i, dr, di);
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) {
                w1r[cb] -= dr[cb];
                w1i[cb] -= di[cb];
            }
            CT t1[4];
#pragma unroll
            for (int s2 = 0; s2 < 4; s2++) t1[s2] = Tb[B1 * 256 + M::row(lane, s2) * 16 + li];
            mul(t1, w1r, w1i, w21r, w21i);                                                       // W2_1 = T_B1 W1
        } else {
#pragma unroll
            for (int cb = 0; cb < NCB; cb++) w21r[cb] = w21i[cb] = z4;
        }
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            p0r[cb] = p0i[cb] = p0q[cb] = z4;
            p1r[cb] = p1i[cb] = p1q[cb] = z4;
        }
        pcr = pci = pcq = z4;
        // rows this pass touches: the applied blocks' and the next pair's supports
        const int lowest = N1 >= 0 ? N1 : (N0 >= 0 ? N0 : (B1 >= 0 ? B1 : B0));
        const int lo = 16 * lowest + 1;
        const int alo = 16 * (B1 >= 0 ? B1 : B0) + 1;          // first row the applied blocks change
        int rbk = w;
        if (16 * rbk + 15 < lo) rbk += 4 * ((lo - 15 - 16 * rbk + 63) / 64);
        v4 nur[NCB], nui[NCB];
        if (PREFU && 16 * rbk < n) loadu(16 * rbk, nur, nui);
        for (; 16 * rbk < n; rbk += 4) {
            const int R = 16 * rbk;
            v4 ur[NCB], ui[NCB];
            if (PREFU) {
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    ur[cb] = nur[cb];
                    ui[cb] = nui[cb];
                }
            } else {
                loadu(R, ur, ui);
            }
            CT x0[4], x1[4], y0[4], y1[4];
            vapply(B0, R, x0);
            vapply(B1, R, x1);
            vacc(N0, R, y0);
            vacc(N1, R, y1);
            if (PREFU && 16 * (rbk + 4) < n) loadu(16 * (rbk + 4), nur, nui);
            if (R + 15 >= alo) {
                // U -= V_B0 W2_0 + V_B1 W2_1 on rows R..
                v4 qu[NCB];
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) qu[cb] = z4;
#pragma unroll
                for (int s2 = 0; s2 < 4; s2++) {
                    const CT xa = x0[s2], xb = x1[s2];
                    const T xsa = xa.x + xa.y, xsb = xb.x + xb.y;
#pragma unroll
                    for (int cb = 0; cb < NCB; cb++) {
                        const T ar = w20r[cb][s2], ai = w20i[cb][s2];
                        qu[cb] = M::mma(xsa, ar, qu[cb]);
                        ur[cb] = M::mma(xa.y, ar + ai, ur[cb]);
                        ui[cb] = M::mma(-xa.x, ai - ar, ui[cb]);
                        const T br = w21r[cb][s2], bi = w21i[cb][s2];
                        qu[cb] = M::mma(xsb, br, qu[cb]);
                        ur[cb] = M::mma(xb.y, br + bi, ur[cb]);
                        ui[cb] = M::mma(-xb.x, bi - br, ui[cb]);
                    }
                }
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    ur[cb] -= qu[cb];
                    ui[cb] -= qu[cb];
                }
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    const int col = c0 + 16 * cb + li;
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = R + M::row(lane, v);
                        if (row < n && col < n) stx(U, (long)row * n + col, CT{ur[cb][v], ui[cb][v]});
                    }
                }
            }
            if (N0 >= 0) {
                accw(y0, ur, ui, p0r, p0i, p0q);
                if (N1 >= 0) {
                    accw(y1, ur, ui, p1r, p1i, p1q);
                    accc(y0, y1, pcr, pci, pcq);
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Back-transform in passes of NB reflectors (128 for complex float, 64 for complex double):
// Q_P = H_j0 ... H_{j0+NB-1} = I - V T V^H with T (NB x NB, upper triangular) combined from
// the 16-reflector factors of larft_kernel by the product rule of two compact-WY blocks,
// T_{0:l, l} = -T_{0:l,0:l} (V_{0:l}^H V_l) T_ll over the 16-column blocks l
// (lgbt_tbig_kernel); then every pass, last first, applies U -= V T (V^H U) to one slab of
// U per workgroup (lgbt_slab_kernel), so U is read and written once per NB reflectors (the
// two-block slab kernel above: once per 32). n_phi = 3000, 64 matrices: 248 ms (two-block
// slabs) -> 192 ms (two launches per pass: W2 = T V^H U by column tiles, then U -= V W2
// by 128 x 64 tiles; U read twice and written once per pass) -> 168 ms (one slab kernel)
// in complex float, 433 -> 315 -> 293 ms in complex double.

// T of pass p of matrix b, column-major (Tbig[k' NB + k] = T(k, k')); one workgroup per
// (matrix, pass). G_ml = V_m^H V_l (m < l) on the matrix cores over the rows j0 + 1 .. n - 1
// in chunks of 16 (V staged through LDS, the next chunk loaded during the current one's
// products; wave w the pairs w, w + 4, ..), then the block recurrence on the VALU in LDS.
template <typename T, int NB>
__global__ __launch_bounds__(256) void lgbt_tbig_kernel(const T *__restrict__ Abase, long strideA, int n,
                                                        const T *__restrict__ Tin, long strideT,
                                                        T *__restrict__ Tbig, long strideTb, int nbb, int batch) {
    using M = Mf<T>;
    using v4 = typename M::v4;
    using CT = cx<T>;
    constexpr int NS = NB / 16, NG = NS * (NS - 1) / 2, NTB = NS * (NS + 1) / 2, LP = 17;
    constexpr int NGW = (NG + 3) / 4;                  // pairs per wave
    constexpr int KC = 16, VP = NB + 1, NVL = KC * NB / 256;
    constexpr int STG = 2 * KC * VP;                   // double-buffered chunk of V
    constexpr int R0 = STG > NG * 16 * LP ? STG : NG * 16 * LP;
    __shared__ CT sm[R0 + (NTB + NS - 1) * 16 * LP];
    CT *vst = sm;                                      // phase A: V chunks
    CT *gsh = sm;                                      // then G_ml, m < l, at l (l - 1) / 2 + m
    CT *tsh = sm + R0;                                 // T_il, i <= l, at l (l + 1) / 2 + i
    CT *xsh = tsh + NTB * 16 * LP;                     // X_m = G_ml T_ll of the current l
    const long b = blockIdx.x / nbb;
    const int p = (int)(blockIdx.x % nbb);
    if (b >= batch) return;                            // whole workgroup
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(t / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const T *A = Abase + 2 * b * strideA;
    const CT *T16 = reinterpret_cast<const CT *>(Tin + 2 * b * strideT);
    CT *Tb = reinterpret_cast<CT *>(Tbig + 2 * b * strideTb) + (long)p * NB * NB;
    const int j0 = p * NB, nblk = (n - 1 + 15) / 16;
    int pm[NGW], pl[NGW];
#pragma unroll
    for (int q = 0; q < NGW; q++) {
        const int pi = w + 4 * q;
        int l = 1;
        while (l * (l + 1) / 2 <= pi) l++;
        pl[q] = pi < NG ? l : 1;                       // a spare slot computes pair (0, 1) unused
        pm[q] = pi < NG ? pi - l * (l - 1) / 2 : 0;
    }
    const v4 z4 = {0, 0, 0, 0};
    v4 gq[NGW], gr[NGW], gi[NGW];
#pragma unroll
    for (int q = 0; q < NGW; q++) gq[q] = gr[q] = gi[q] = z4;
    const int rs = j0 + 1, nch = (n - rs + KC - 1) / KC;
    CT vreg[NVL];
    auto gload = [&](int ch) {
#pragma unroll
        for (int it = 0; it < NVL; it++) {
            const int e = t + 256 * it;
            vreg[it] = lg_refl(A, n, j0, rs + ch * KC + e % KC, e / KC);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < NVL; it++) {
            const int e = t + 256 * it;
            vst[buf * KC * VP + (e % KC) * VP + e / KC] = vreg[it];
        }
    };
    if (nch > 0) {
        gload(0);
        sstore(0);
    }
    for (int e = t; e < NS * 256; e += 256) {
        const int l = e >> 8, k2 = (e >> 4) & 15, k1 = e & 15;
        const int bk = j0 / 16 + l;
        tsh[(l * (l + 1) / 2 + l) * 16 * LP + k1 * LP + k2] =
            bk < nblk ? T16[(long)bk * 256 + k2 * 16 + k1] : CT{(T)0, (T)0};
    }
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
        const int cur = ch & 1;
        if (ch + 1 < nch) gload(ch + 1);
        const CT *vs = vst + cur * KC * VP;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int rr = 4 * s + lq;
#pragma unroll
            for (int q = 0; q < NGW; q++) {
                const CT x = vs[rr * VP + 16 * pm[q] + li];    // A operand: conj V_m[row][k = li]
                const CT y = vs[rr * VP + 16 * pl[q] + li];    // B operand: V_l[row][k = li]
                gq[q] = M::mma(x.x - x.y, y.x, gq[q]);
                gr[q] = M::mma(x.y, y.x + y.y, gr[q]);
                gi[q] = M::mma(x.x, y.y - y.x, gi[q]);
            }
        }
        if (ch + 1 < nch) sstore(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NGW; q++) {
        if (w + 4 * q >= NG) continue;
        CT *g = gsh + (pl[q] * (pl[q] - 1) / 2 + pm[q]) * 16 * LP;
#pragma unroll
        for (int v = 0; v < 4; v++) g[M::row(lane, v) * LP + li] = CT{gq[q][v] + gr[q][v], gq[q][v] + gi[q][v]};
    }
    __syncthreads();
    const int er = t >> 4, ec = t & 15;      // element (er, ec) of a 16 x 16 block
    for (int l = 1; l < NS; l++) {
        const CT *tll = tsh + (l * (l + 1) / 2 + l) * 16 * LP;
        for (int m = 0; m < l; m++) {
            const CT *g = gsh + (l * (l - 1) / 2 + m) * 16 * LP;
            CT acc{(T)0, (T)0};
#pragma unroll
            for (int q = 0; q < 16; q++) fmax_(acc, g[er * LP + q], tll[q * LP + ec]);
            xsh[m * 16 * LP + er * LP + ec] = acc;
        }
        __syncthreads();
        for (int i = 0; i < l; i++) {
            CT acc{(T)0, (T)0};
            for (int m = i; m < l; m++) {
                const CT *tim = tsh + (m * (m + 1) / 2 + i) * 16 * LP;
#pragma unroll
                for (int q = 0; q < 16; q++) fmax_(acc, tim[er * LP + q], xsh[m * 16 * LP + q * LP + ec]);
            }
            tsh[(l * (l + 1) / 2 + i) * 16 * LP + er * LP + ec] = CT{-acc.x, -acc.y};
        }
        __syncthreads();
    }
    for (int e = t; e < NB * NB; e += 256) {
        const int kc = e / NB, kr = e % NB, i = kr >> 4, l = kc >> 4;
        Tb[e] = (i <= l) ? tsh[(l * (l + 1) / 2 + i) * 16 * LP + (kr & 15) * LP + (kc & 15)] : CT{(T)0, (T)0};
    }
}

// All passes for one slab of 64 eigenvector columns per workgroup (512 threads): the slab
// streams through in chunks of 32 rows; in each chunk the waves first apply the current
// pass p (U -= V_p W2_p, wave w the 16 x 16 tile (w >> 2, w & 3), K = NB), then
// accumulate the next pass's W' = V_{p-1}^H U from the updated rows (wave w: k tile
// w % (NB/16), its column tiles), so U is read and written once per NB reflectors. At
// the end of a pass W2_{p-1} = T_{p-1} W' replaces W2_p in LDS. The next chunk's U rows
// and V operands are loaded while the current chunk's W' products run (staging V_p's rows
// in LDS instead, shared by the four waves of a row tile, measured slower: 163 -> 178 ms
// at n = 3000, 64 matrices). Pass nbb is a
// virtual one that only accumulates W' of the last pass. The slabs of one matrix share
// blockIdx % 8 (one XCD, whose L2 serves V and T).
template <typename T, int NB>
__global__ __launch_bounds__(512) void lgbt_slab_kernel(const T *__restrict__ Abase, long strideA, int n,
                                                        const T *__restrict__ Tbig, long strideTb,
                                                        T *__restrict__ Uio, long strideU, int nslab, int batch) {
    using M = Mf<T>;
    using v4 = typename M::v4;
    using CT = cx<T>;
    constexpr int CW = 64, RC = 32, WP = CW + 1, UP = CW + 1;
    constexpr int NKB = NB / 16;                 // k tiles of a pass
    constexpr int NCW = 4 * NKB / 8;             // W' column tiles per wave
    constexpr int KS = NB / 4;                   // K-steps of the update
    constexpr int WS = RC / 4;                   // K-steps of the W' accumulation per chunk
    constexpr int NUL = RC * CW / 512;           // U elements per thread per chunk
    __shared__ CT wsh[NB * WP];                  // W2 of the pass being applied; W' at a pass end
    __shared__ CT ush[2][RC * UP];               // the chunk's U rows (double buffer)
    const long b = (blockIdx.x & 7) + 8L * (blockIdx.x / (8 * nslab));
    const int slab = (int)((blockIdx.x >> 3) % nslab);
    if (b >= batch) return;                      // whole workgroup
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(t / kWave);
    const int li = lane & 15, lq = lane >> 4;
    const int utr = w >> 2, utc = w & 3;         // update tile
    const int wk = w % NKB, wc0 = (w / NKB) * NCW;
    const T *A = Abase + 2 * b * strideA;
    T *U = Uio + 2 * b * strideU;
    const CT *Tm = reinterpret_cast<const CT *>(Tbig + 2 * b * strideTb);
    const int c0 = slab * CW;
    const int nbb = (n - 1 + NB - 1) / NB;
    auto rlo = [&](int pp) { return pp >= 1 ? (pp - 1) * NB + 1 : 1; };
    auto nchunks = [&](int pp) { return (n - rlo(pp) + RC - 1) / RC; };
    // step (pp, cc) applies pass pp: some of its rows are at or below the pass's first row
    auto upd = [&](int pp, int cc) { return pp < nbb && rlo(pp) + cc * RC + RC - 1 >= pp * NB + 1; };
    const v4 z4 = {0, 0, 0, 0};
    CT xp[KS], xw[WS], xwn[WS], un[NUL];
    auto load_step = [&](int pp, int cc) {       // V operands and U rows of step (pp, cc)
        const int r0 = rlo(pp) + cc * RC;
        if (upd(pp, cc)) {
#pragma unroll
            for (int s = 0; s < KS; s++) xp[s] = lg_refl(A, n, pp * NB, r0 + 16 * utr + li, 4 * s + lq);
        }
        if (pp >= 1) {
#pragma unroll
            for (int s = 0; s < WS; s++) xwn[s] = lg_refl(A, n, (pp - 1) * NB, r0 + 4 * s + lq, 16 * wk + li);
        }
#pragma unroll
        for (int it = 0; it < NUL; it++) {
            const int e = t + 512 * it, rr = e / CW, cc2 = e % CW;
            const int row = r0 + rr, col = c0 + cc2;
            un[it] = (row < n && col < n) ? ldx(U, (long)row * n + col) : CT{(T)0, (T)0};
        }
    };
    auto store_u = [&](CT *us) {
#pragma unroll
        for (int it = 0; it < NUL; it++) {
            const int e = t + 512 * it;
            us[(e / CW) * UP + e % CW] = un[it];
        }
    };
    v4 wq[NCW], wr[NCW], wi[NCW];                // W' partial sums of the wave's tiles
#pragma unroll
    for (int c = 0; c < NCW; c++) wq[c] = wr[c] = wi[c] = z4;
    load_step(nbb, 0);
    store_u(ush[0]);
#pragma unroll
    for (int s = 0; s < WS; s++) xw[s] = xwn[s];
    __syncthreads();
    int p = nbb, ch = 0, cur = 0;
    while (true) {
        const int r0 = rlo(p) + ch * RC;
        int pn = p, cn = ch + 1;
        if (cn >= nchunks(p)) {
            pn = p - 1;
            cn = 0;
        }
        const bool has_next = pn >= 0;
        CT *us = ush[cur];
        // U -= V_p W2_p on the wave's tile (rows below the pass's first row are unchanged)
        if (upd(p, ch)) {
            v4 aq = z4, ar, ai;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const CT x = us[(16 * utr + M::row(lane, v)) * UP + 16 * utc + li];
                ar[v] = x.x;
                ai[v] = x.y;
            }
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const CT x = xp[s];                                          // V_p: A[row][k]
                const CT wv = wsh[(4 * s + lq) * WP + 16 * utc + li];        // W2_p: B[k][col]
                aq = M::mma(x.x + x.y, wv.x, aq);
                ar = M::mma(x.y, wv.x + wv.y, ar);
                ai = M::mma(-x.x, wv.y - wv.x, ai);
            }
            const int col = c0 + 16 * utc + li;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int rl = 16 * utr + M::row(lane, v);
                const CT y{ar[v] - aq[v], ai[v] - aq[v]};
                us[rl * UP + 16 * utc + li] = y;
                if (r0 + rl < n && col < n) stx(U, (long)(r0 + rl) * n + col, y);
            }
        }
        __syncthreads();
        if (p >= 1) {
#pragma unroll
            for (int s = 0; s < WS; s++) xw[s] = xwn[s];
        }
        if (has_next) load_step(pn, cn);
        // W' += V_{p-1}^H U over the chunk's (updated) rows
        if (p >= 1) {
#pragma unroll
            for (int s = 0; s < WS; s++) {
                const CT x = xw[s];                                          // conj V_{p-1}: A[k][row]
                const T xd = x.x - x.y;
#pragma unroll
                for (int c = 0; c < NCW; c++) {
                    const CT u = us[(4 * s + lq) * UP + 16 * (wc0 + c) + li];   // U: B[row][col]
                    wq[c] = M::mma(xd, u.x, wq[c]);
                    wr[c] = M::mma(x.y, u.x + u.y, wr[c]);
                    wi[c] = M::mma(x.x, u.y - u.x, wi[c]);
                }
            }
        }
        if (has_next) store_u(ush[cur ^ 1]);
        if (pn != p && p >= 1) {
            // end of pass p: W2_{p-1} = T_{p-1} W' replaces W2_p
#pragma unroll
            for (int c = 0; c < NCW; c++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    wsh[(16 * wk + M::row(lane, v)) * WP + 16 * (wc0 + c) + li] =
                        CT{wq[c][v] + wr[c][v], wq[c][v] + wi[c][v]};
            __syncthreads();
            const CT *Tp = Tm + (long)(p - 1) * NB * NB;
#pragma unroll
            for (int c = 0; c < NCW; c++) wq[c] = wr[c] = wi[c] = z4;
            for (int s = 4 * wk; s < KS; s++) {                              // T upper triangular
                const CT tv = Tp[(long)(4 * s + lq) * NB + 16 * wk + li];    // T: A[k][k']
                const T ts = tv.x + tv.y;
#pragma unroll
                for (int c = 0; c < NCW; c++) {
                    const CT wv = wsh[(4 * s + lq) * WP + 16 * (wc0 + c) + li];   // W': B[k'][col]
                    wq[c] = M::mma(ts, wv.x, wq[c]);
                    wr[c] = M::mma(-tv.y, wv.x + wv.y, wr[c]);
                    wi[c] = M::mma(tv.x, wv.y - wv.x, wi[c]);
                }
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NCW; c++) {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    wsh[(16 * wk + M::row(lane, v)) * WP + 16 * (wc0 + c) + li] =
                        CT{wq[c][v] + wr[c][v], wq[c][v] + wi[c][v]};
                wq[c] = wr[c] = wi[c] = z4;
            }
        }
        __syncthreads();
        if (!has_next) break;
        p = pn;
        ch = cn;
        cur ^= 1;
    }
}

template <typename T>
constexpr int lgbt_nb() { return sizeof(T) == sizeof(float) ? 128 : 64; }

// The pass back-transform over `batch` matrices, in groups whose pass factors fit
// `cap_elems` elements of T of workspace at `work`: the factors (lgbt_tbig_kernel), then
// one launch of the slab kernel.
template <typename T>
static int launch_bt_passes(const T *A, long strideA, int n, const T *tws, long strideT, T *U, long strideU,
                            int batch, T *work, long cap_elems, hipStream_t s) {
    constexpr int NB = lgbt_nb<T>(), CW = 64;
    const int nbb = (n - 1 + NB - 1) / NB;
    const int nslab = (n + CW - 1) / CW;
    const long strideTb = (long)nbb * NB * NB;   // complex elements per matrix
    const long mg = std::min<long>(batch, cap_elems / (2 * strideTb));
    if (mg < 1) return -12;
    for (long b0 = 0; b0 < batch; b0 += mg) {
        const int nbm = (int)std::min<long>(mg, batch - b0);
        const T *Ab = A + 2 * b0 * strideA;
        hipLaunchKernelGGL((lgbt_tbig_kernel<T, NB>), dim3((unsigned)((long)nbm * nbb)), dim3(256), 0, s, Ab, strideA,
                           n, tws + 2 * b0 * strideT, strideT, work, strideTb, nbb, nbm);
        const long groups = (nbm + 7) / 8;
        hipLaunchKernelGGL((lgbt_slab_kernel<T, NB>), dim3((unsigned)(groups * 8 * nslab)), dim3(512), 0, s, Ab,
                           strideA, n, work, strideTb, U + 2 * b0 * strideU, strideU, nslab, nbm);
        RLM_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// RLM_LG_BT=pairs selects lg_backtransform2_kernel (two blocks of 16 reflectors per pass) for
// n > 512, for comparison; default: passes of lgbt_nb<T>() reflectors (tbig + slab kernels).
static bool lg_bt_pairs() {
    const char *e = getenv("RLM_LG_BT");
    return e && std::string(e) == "pairs";
}

// ----------------------------------------------------------------------------
// host
template <typename T>
static int heevd_large_t(T *A, int n, long batch, double *D, double *E, T *TAU, long strideV, T *aux, double *W,
                         T *U, double *work, int grid_dc, T *tws, hipStream_t s, hipEvent_t mid, hipEvent_t mid2) {
    if (n <= 512 || n > heevd_large_max_n() || batch < 1) return batch < 1 ? 0 : -22;
    const long strideA = (long)n * n;
    const long strideAux = heevd_large_aux_complex(n);
    const int nb = (int)batch;
    const int ns0 = (int)lg_strips(n);
    for (int k0 = 0; k0 < n - 1; k0 += kLgNB) {
        const int nbk = std::min(kLgNB, n - 1 - k0);
        for (int i = 0; i <= nbk; i++) {
            hipLaunchKernelGGL(lg_col_kernel<T>, dim3(nb), dim3(kLgColNT), 0, s, A, strideA, n, D, E, TAU,
                               strideV, aux, strideAux, k0, i, nbk);
            if (i == nbk) break;
            // strips of rows base .. n-1 (base = c + 1 rounded down to a multiple of 8) and the
            // workgroups per matrix: row block ib has floor(min(4 ib + 3, ns - 1) / 16) + 1
            const int mb = n - ((k0 + i + 1) & ~7);
            const int nsm = (mb + kLgTS - 1) / kLgTS, nib = (nsm + kLgTR - 1) / kLgTR;
            int per = 0;
            // column blocks of 16 tiles, or of 8 where 16 would leave fewer than about three
            // rounds of workgroups on the chip (the last, partly empty round then costs less)
            int tc = kLgTC;
            for (int ib = 0; ib < nib; ib++) per += std::min(kLgTR * ib + kLgTR - 1, nsm - 1) / tc + 1;
            if ((long)nb * per < 3072) {
                tc = kLgTCmin;
                per = 0;
                for (int ib = 0; ib < nib; ib++) per += std::min(kLgTR * ib + kLgTR - 1, nsm - 1) / tc + 1;
            }
            hipEvent_t te0, te1;
            kt_pair(s, te0, te1);
            if (te0) RLM_HIP_CHECK(hipEventRecord(te0, s));
            hipLaunchKernelGGL(lg_symv_lower_kernel<T>, dim3((unsigned)((long)nb * per + nb)), dim3(256), 0, s, A,
                               strideA, n, aux, strideAux, k0, i, per, ns0, nb, tc);
            if (te1) RLM_HIP_CHECK(hipEventRecord(te1, s));
            hipLaunchKernelGGL(lg_symv_reduce_kernel<T>, dim3((unsigned)((long)nb * nsm)), dim3(64), 0, s, A, strideA,
                               aux, strideAux, n, k0, i, nbk, ns0, tc);
        }
        const int s0 = k0 + nbk;
        if (s0 < n - 1) {
            const int nt = (n - s0 + 63) / 64;
            hipLaunchKernelGGL(lg_her2k_kernel<T>, dim3((unsigned)((long)nb * nt * (nt + 1) / 2)), dim3(256), 0, s, A,
                               strideA,
                               n, aux, strideAux, s0, nbk, nt);
        }
    }
    RLM_HIP_CHECK(hipGetLastError());
    if (mid) RLM_HIP_CHECK(hipEventRecord(mid, s));
    int rc;
    if ((rc = launch_stedc_global(n, D, E, strideV, W, n, U, strideA, work, nb, grid_dc, s))) return rc;
    if (mid2) RLM_HIP_CHECK(hipEventRecord(mid2, s));
    const long strideT = backtransform_tws_complex(n);
    const int nblk = (n - 1 + 15) / 16;
    if ((rc = launch_larft(A, strideA, n, n, TAU, strideV, tws, strideT, nb, (nblk + 3) / 4, s))) return rc;
    if (!lg_bt_pairs()) {
        // the divide & conquer workspace is free now: the pass factors live there
        const long cap = (long)grid_dc * stedc_global_work_doubles(n) * (long)(sizeof(double) / sizeof(T));
        return launch_bt_passes<T>(A, strideA, n, tws, strideT, U, strideA, nb, reinterpret_cast<T *>(work), cap, s);
    }
    // 32 columns per slab for complex float, 16 for complex double (the W reduction
    // buffers then stay at 32 KB of LDS)
    constexpr int NCB = sizeof(T) == sizeof(float) ? 2 : 1;
    const int nslab = (n + 16 * NCB - 1) / (16 * NCB);
    const long groups = (batch + 7) / 8;
    hipLaunchKernelGGL((lg_backtransform2_kernel<T, NCB>), dim3((unsigned)(groups * 8 * nslab)), dim3(256), 0, s, A,
                       strideA, n, tws, strideT, U, strideA, nslab, nb);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
static int launch_bt_stream_t(const T *A, long strideA, int n, const T *tws, long strideT, T *U, long strideU,
                              int batch, hipStream_t s) {
    if (batch < 1 || n < 2) return 0;
    constexpr int NCB = sizeof(T) == sizeof(float) ? 2 : 1;
    const int nslab = (n + 16 * NCB - 1) / (16 * NCB);
    const long groups = (batch + 7) / 8;
    hipLaunchKernelGGL((lg_backtransform2_kernel<T, NCB>), dim3((unsigned)(groups * 8 * nslab)), dim3(256), 0, s, A,
                       strideA, n, tws, strideT, U, strideU, nslab, batch);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}
int launch_backtransform_stream(const double *A, long strideA, int n, const double *tws, long strideT, double *U,
                                long strideU, int batch, hipStream_t s) {
    return launch_bt_stream_t<double>(A, strideA, n, tws, strideT, U, strideU, batch, s);
}
int launch_backtransform_stream(const float *A, long strideA, int n, const float *tws, long strideT, float *U,
                                long strideU, int batch, hipStream_t s) {
    return launch_bt_stream_t<float>(A, strideA, n, tws, strideT, U, strideU, batch, s);
}

int heevd_large(double *A, int n, long batch, double *D, double *E, double *TAU, long strideV, double *aux, double *W,
                double *U, double *work, int grid_dc, double *tws, hipStream_t s, hipEvent_t mid, hipEvent_t mid2) {
    return heevd_large_t<double>(A, n, batch, D, E, TAU, strideV, aux, W, U, work, grid_dc, tws, s, mid, mid2);
}
int heevd_large(float *A, int n, long batch, double *D, double *E, float *TAU, long strideV, float *aux, double *W,
                float *U, double *work, int grid_dc, float *tws, hipStream_t s, hipEvent_t mid, hipEvent_t mid2) {
    return heevd_large_t<float>(A, n, batch, D, E, TAU, strideV, aux, W, U, work, grid_dc, tws, s, mid, mid2);
}

}  // namespace rlm
