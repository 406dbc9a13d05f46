// C ABI of librlm_hip.so: diag_wrappers-level entry points, rlm model entry
// points and the batched rlm pipeline. Declarations and reference citations are
// in include/hip_diag.h.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/hip_diag.h"
#include "chern.hpp"
#include "common.hpp"
#include "eig.hpp"
#include "hlm_kernels.hpp"
#include "host_rng.hpp"
#include "refine.hpp"
#include "rlm_kernels.hpp"
#include "rng.hpp"

using namespace rlm;

namespace {

constexpr int kGridTrd = 1024;   // persistent workgroups for the tridiagonalisation
constexpr int kGridTail = 768;   // persistent workgroups for D&C (three per CU) + back-transform

// ---------------------------------------------------------------- small kernels
// Elementwise kernels over whole matrices take nch workgroups per matrix
// (blockIdx.x = b * nch + chunk; the chunks stride over the n*n elements), so a
// single n = 3000 matrix is spread over the chip; small n use nch = 1.
static int chunks_per_matrix(int n) {
    const long nn = (long)n * n;
    const long c = (nn + 256L * 64 - 1) / (256L * 64);
    return (int)(c < 1 ? 1 : (c > 256 ? 256 : c));
}

// Fill the triangle opposite to `uplo` by conjugate mirroring; real diagonal.
__global__ void hermitize_kernel(double *A, long lda, long strideA, int n, int from_lower, int nch) {
    double *M = A + 2 * (blockIdx.x / nch) * strideA;
    for_2d((int)(blockIdx.x % nch) * (int)blockDim.x + (int)threadIdx.x, nch * (int)blockDim.x, n, n, [&](int c, int r) {
        if (r == c) {
            M[2 * ((long)c * lda + r) + 1] = 0.0;
            return;
        }
        const bool dst = from_lower ? (r < c) : (r > c);
        if (!dst) return;
        const cd v = ld(M, (long)r * lda + c);  // (c, r) element
        st(M, (long)c * lda + r, cconj(v));
    });
}

// A(r, j) = U[r*n + j] (component-major -> column-major eigenvectors)
__global__ void transpose_out_kernel(const double *U, long strideU, int n, double *A, long lda, long strideA,
                                     int nch) {
    const long b = blockIdx.x / nch;
    const double *Ub = U + 2 * b * strideU;
    double *M = A + 2 * b * strideA;
    for_2d((int)(blockIdx.x % nch) * (int)blockDim.x + (int)threadIdx.x, nch * (int)blockDim.x, n, n,
           [&](int j, int r) { st(M, (long)j * lda + r, ld(Ub, (long)r * n + j)); });
}

// Full-matrix copy with lda (column-major) -> packed lda = n
__global__ void pack_kernel(const double *A, long lda, long strideA, int n, double *P, int nch) {
    const long b = blockIdx.x / nch;
    const double *M = A + 2 * b * strideA;
    double *Q = P + 2 * b * n * n;
    for_2d((int)(blockIdx.x % nch) * (int)blockDim.x + (int)threadIdx.x, nch * (int)blockDim.x, n, n,
           [&](int c, int r) { st(Q, (long)c * n + r, ld(M, (long)c * lda + r)); });
}

// FP64 -> FP32 copy (count values)
__global__ void to_f32_kernel(const double *P, long count, float *Q) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
        Q[i] = (float)P[i];
}

// Largest |real or imaginary part| of each chunk of the packed complex matrices:
// part[b * nch + chunk]. 256 threads.
__global__ __launch_bounds__(256) void maxabs_kernel(const double *P, long nn, int nch, double *part) {
    __shared__ double red[256 / kWave];
    const double *M = P + 2 * (blockIdx.x / nch) * nn;
    const long step = (long)nch * blockDim.x;
    double mx = 0.0;
    for (long i = (blockIdx.x % nch) * (long)blockDim.x + threadIdx.x; i < nn; i += step)
        mx = fmax(mx, fmax(fabs(M[2 * i]), fabs(M[2 * i + 1])));
    for (int off = kWave / 2; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) mx = fmax(mx, red[w]);
        part[blockIdx.x] = mx;
    }
}

// 1 in flag[b * nch + chunk] if that chunk of packed matrix b holds a NaN or an
// infinity (real or imaginary part), else 0. 256 threads.
__global__ __launch_bounds__(256) void nonfinite_kernel(const double *P, long nn, int nch, int *flag) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    const double *M = P + 2 * (blockIdx.x / nch) * nn;
    const long step = (long)nch * blockDim.x;
    int bad = 0;
    for (long i = (blockIdx.x % nch) * (long)blockDim.x + threadIdx.x; i < nn; i += step)
        bad |= !(isfinite(M[2 * i]) && isfinite(M[2 * i + 1]));
    if (bad) any = 1;   // same-value writes
    __syncthreads();
    if (threadIdx.x == 0) flag[blockIdx.x] = any;
}

// RAII roctx range (rocprofv3 --marker-trace shows the stage launches of a plan)
struct Range {
    explicit Range(const char *name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
};

// FP64 -> FP32 copy of packed complex matrices (mixed-precision solver input), each
// matrix scaled by a power of two to max |entry| in [1/2, 1) (from the maxabs_kernel
// partials) so that the FP32 reduction neither overflows nor underflows;
// eigenvectors do not depend on the scale, and the eigenvalues are taken from the
// FP64 matrix afterwards.
__global__ __launch_bounds__(256) void to_f32_scaled_kernel(const double *P, long nn, const double *part, int nch,
                                                            float *Q) {
    const long b = blockIdx.x / nch;
    const double *M = P + 2 * b * nn;
    float *R = Q + 2 * b * nn;
    double mx = 0.0;
    for (int k = 0; k < nch; k++) mx = fmax(mx, part[b * nch + k]);
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    const double sc = ldexp(1.0, -e);
    const long step = (long)nch * blockDim.x;
    for (long i = (blockIdx.x % nch) * (long)blockDim.x + threadIdx.x; i < 2 * nn; i += step)
        R[i] = (float)(M[i] * sc);
}

// A(r, j) = U[r*n + j] from FP32 eigenvectors (component-major -> column-major FP64)
__global__ void transpose_out_f32_kernel(const float *U, long strideU, int n, double *A, long lda, long strideA,
                                         int nch) {
    const long b = blockIdx.x / nch;
    const float *Ub = U + 2 * b * strideU;
    double *M = A + 2 * b * strideA;
    for_2d((int)(blockIdx.x % nch) * (int)blockDim.x + (int)threadIdx.x, nch * (int)blockDim.x, n, n, [&](int j, int r) {
        const cx<float> u = ldx(Ub, (long)r * n + j);
        st(M, (long)j * lda + r, cd{(double)u.x, (double)u.y});
    });
}

// ---------------------------------------------------------------- device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t b) {
        if (b <= bytes) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, b) != hipSuccess) return -12;
        bytes = b;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    double *d() { return static_cast<double *>(p); }
};

// Device copy of the jumps of a run's disorder streams (rng.hip): the jump of each
// sample of a batch relative to the batch's first sample (B, fixed at creation) and the
// jump from the seed's initial window to the batch's first sample (A, per batch).
struct StreamJumps {
    long stride = 0;   // words per sample
    MtJumpTable b;
    DevBuf dB, dBp, dA;
    long a_first = -1;
    int a_kind = 0, a_mono = 0;
    int create(long count, long stride_words) {
        if (count < 1 || stride_words < 1) return -22;
        if (!mt_jump_ready()) return -1;
        stride = stride_words;
        b.set(count);
        for (long k = 0; k < count; k++) b.fill(k, (uint64_t)k * (uint64_t)stride);
        int rc;
        if ((rc = dB.ensure(sizeof(int) * 2 * (size_t)count)) || (rc = dBp.ensure(8 * b.poly.size())) ||
            (rc = dA.ensure(8 * (size_t)mtj::kPolyWords)))
            return rc;
        std::vector<int> km(2 * count);
        std::copy(b.kind.begin(), b.kind.end(), km.begin());
        std::copy(b.mono.begin(), b.mono.end(), km.begin() + count);
        RLM_HIP_CHECK(hipMemcpy(dB.p, km.data(), sizeof(int) * km.size(), hipMemcpyHostToDevice));
        RLM_HIP_CHECK(hipMemcpy(dBp.p, b.poly.data(), 8 * b.poly.size(), hipMemcpyHostToDevice));
        return 0;
    }
    // jump A to sample `first`; a new polynomial is copied once the stream's earlier
    // work (which may still read the old one) has finished
    int select(long first, hipStream_t s) {
        if (first < 0) return -22;
        if (first == a_first) return 0;
        MtJumpTable a;
        a.set(1);
        a.fill(0, (uint64_t)first * (uint64_t)stride);
        if (a.kind[0] == 2) {
            RLM_HIP_CHECK(hipStreamSynchronize(s));
            RLM_HIP_CHECK(hipMemcpy(dA.p, a.poly.data(), 8 * (size_t)mtj::kPolyWords, hipMemcpyHostToDevice));
        }
        a_kind = a.kind[0];
        a_mono = a.mono[0];
        a_first = first;
        return 0;
    }
    int launch(uint32_t seed, int count, int mode, long words, long n_sec, double s0, double s1, double *o0,
               double *o1, double *o2, hipStream_t s) const {
        if (count > b.count) return -22;
        const int *kb = static_cast<const int *>(dB.p);
        return launch_mt_streams(seed, a_kind, a_mono, static_cast<const uint64_t *>(dA.p), kb, kb + b.count,
                                 static_cast<const uint64_t *>(dBp.p), count, words, mode, n_sec, s0, s1, o0, o1, o2, s);
    }
    void release() {
        dB.release();
        dBp.release();
        dA.release();
        a_first = -1;
    }
};

// n above this goes to the multi-workgroup solver (eig_large.hip)
constexpr int kSmallMaxN = 512;
// divide & conquer workgroups (each with its own n*n workspaces) for n > kSmallMaxN
constexpr int kGridTailLarge = 64;

// n <= 128: the divide & conquer runs two-wave workgroups, twice as many fit the CUs
static int tail_grid(int n, long batch) {
    const long g = n > kSmallMaxN ? kGridTailLarge : n <= 128 ? 2 * kGridTail : kGridTail;
    return (int)(batch < g ? batch : g);
}

// Workspace of the batched eigensolver.
struct EigWork {
    DevBuf dev, U, work, tw, aux;
    int ensure(int n, long batch, bool with_u = true) {
        int rc;
        if ((rc = dev.ensure(sizeof(double) * 4 * (size_t)n * batch))) return rc;
        if (with_u && (rc = U.ensure(sizeof(double) * 2 * (size_t)n * n * batch))) return rc;
        const long g = tail_grid(n, batch);
        const size_t per_wg = n > kSmallMaxN ? (size_t)stedc_global_work_doubles(n) : 3 * (size_t)n * n;
        if ((rc = work.ensure(sizeof(double) * per_wg * g))) return rc;
        // triangular factors of the matrix-core back-transform (complex double covers float)
        if ((rc = tw.ensure(sizeof(double) * 2 * (size_t)backtransform_tws_complex(n) * batch))) return rc;
        if (n > kSmallMaxN && (rc = aux.ensure(sizeof(double) * 2 * (size_t)heevd_large_aux_complex(n) * batch)))
            return rc;
        return 0;
    }
    void release() {
        dev.release();
        U.release();
        work.release();
        tw.release();
        aux.release();
    }
};

// Eigen-decomposition of `batch` packed (lda = n, stride n*n) full Hermitian
// matrices in A. Eigenvalues -> W (stride n), eigenvectors component-major -> Uout.
int eig_packed(double *A, int n, long batch, double *W, double *Uout, EigWork &ws, hipStream_t s,
               hipEvent_t mid = nullptr, hipEvent_t mid2 = nullptr) {
    int rc;
    const long strideV = 4L * n;   // d | e | tau(re,im) per matrix, interleaved blocks
    if ((size_t)batch * strideV * sizeof(double) > ws.dev.bytes) return -22;
    double *base = ws.dev.d();
    // layout per matrix: D at base + b*strideV, E at +n, TAU at +2n (complex, 2n doubles)
    // launch_* take separate arrays with a common stride:
    double *D = base, *E = base + n, *TAU = base + 2 * n;
    if (n > kSmallMaxN) {
        const int gl = tail_grid(n, batch);
        if ((size_t)gl * stedc_global_work_doubles(n) * sizeof(double) > ws.work.bytes ||
            (size_t)batch * heevd_large_aux_complex(n) * 2 * sizeof(double) > ws.aux.bytes ||
            (size_t)batch * backtransform_tws_complex(n) * 2 * sizeof(double) > ws.tw.bytes)
            return -22;
        return heevd_large(A, n, batch, D, E, TAU, strideV, ws.aux.d(), W, Uout, ws.work.d(), gl, ws.tw.d(), s, mid,
                           mid2);
    }
    // TAU is read with a complex stride of strideV doubles -> pass complex-stride strideV/2
    const int gt = (int)(batch < kGridTrd ? batch : kGridTrd);
    if ((rc = launch_zhetrd(A, (long)n * n, n, n, D, E, TAU, strideV, (int)batch, gt, s))) return rc;
    if (mid) RLM_HIP_CHECK(hipEventRecord(mid, s));
    const int gl = tail_grid(n, batch);
    if ((size_t)gl * 3 * n * n * sizeof(double) > ws.work.bytes) return -22;
    if ((size_t)batch * backtransform_tws_complex(n) * 2 * sizeof(double) > ws.tw.bytes) return -22;
    if ((rc = launch_tail(A, (long)n * n, n, n, D, E, TAU, strideV, W, n, Uout, (long)n * n, ws.work.d(), ws.tw.d(),
                          (int)batch, gl, s, mid2)))
        return rc;
    return 0;
}

// FP32 tridiagonalisation and back-transform (mixed precision); D&C in FP64.
// Same per-matrix layout of the d | e | tau slot as eig_packed (4n doubles); tau holds
// 2(n-1) floats inside its 2n-double region.
// Hb (optional): read the matrices from their band form of offset hb_off instead
// of A (A then only receives the reflectors).
int eig_packed_f32(float *A, int n, long batch, double *W, float *Uout, EigWork &ws, hipStream_t s,
                   hipEvent_t mid = nullptr, hipEvent_t mid2 = nullptr, const double *Hb = nullptr,
                   int hb_off = 0) {
    int rc;
    const long strideV = 4L * n;   // doubles per matrix, for D, E and TAU alike
    if ((size_t)batch * strideV * sizeof(double) > ws.dev.bytes) return -22;
    double *base = ws.dev.d();
    double *D = base, *E = base + n;
    float *TAU = reinterpret_cast<float *>(base + 2 * n);
    if (n > kSmallMaxN) {
        const int gl = tail_grid(n, batch);
        if (Hb || (size_t)gl * stedc_global_work_doubles(n) * sizeof(double) > ws.work.bytes ||
            (size_t)batch * heevd_large_aux_complex(n) * 2 * sizeof(float) > ws.aux.bytes ||
            (size_t)batch * backtransform_tws_complex(n) * 2 * sizeof(float) > ws.tw.bytes)
            return -22;
        return heevd_large(A, n, batch, D, E, TAU, strideV, static_cast<float *>(ws.aux.p), W, Uout, ws.work.d(), gl,
                           static_cast<float *>(ws.tw.p), s, mid, mid2);
    }
    const int gt = (int)(batch < kGridTrd ? batch : kGridTrd);
    if (Hb) {
        if ((rc = launch_chetrd_band(A, (long)n * n, n, n, D, E, TAU, strideV, (int)batch, s, Hb, hb_off))) return rc;
    } else if ((rc = launch_chetrd(A, (long)n * n, n, n, D, E, TAU, strideV, (int)batch, gt, s))) {
        return rc;
    }
    if (mid) RLM_HIP_CHECK(hipEventRecord(mid, s));
    const int gl = tail_grid(n, batch);
    if ((size_t)gl * 3 * n * n * sizeof(double) > ws.work.bytes) return -22;
    if ((size_t)batch * backtransform_tws_complex(n) * 2 * sizeof(float) > ws.tw.bytes) return -22;
    if ((rc = launch_tail_f32(A, (long)n * n, n, n, D, E, TAU, strideV, W, n, Uout, (long)n * n, ws.work.d(),
                              static_cast<float *>(ws.tw.p), (int)batch, gl, s, mid2)))
        return rc;
    return 0;
}

std::mutex g_single_mu;
EigWork g_single_ws;
DevBuf g_single_pack, g_single_a32, g_single_pair, g_single_flag;

}  // namespace

// ============================================================================
// diag_wrappers-level
// ============================================================================
extern "C" int hip_diag_heevd_batched(int n, double *d_A, long lda, long strideA, double *d_W, long strideW,
                                      int batch, char uplo, int precision, void *stream) {
    if (n < 1 || n > heevd_large_max_n() || lda < n || batch < 0 || strideW < n || strideA < lda * (long)n)
        return -22;
    if (precision != HIP_DIAG_FP64 && precision != HIP_DIAG_MIXED) return -22;
    if (batch == 0) return 0;
    if (uplo != 'L' && uplo != 'U' && uplo != 'l' && uplo != 'u') return -22;
    if (!d_A || !d_W) return -22;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_single_mu);
    int rc;
    if (n <= kSmallMaxN && (rc = tune_hetrd(n, precision == HIP_DIAG_MIXED, batch))) return rc;
    if ((rc = g_single_ws.ensure(n, batch))) return rc;
    if ((rc = g_single_pack.ensure(sizeof(double) * 2 * (size_t)n * n * batch + sizeof(double) * (size_t)n * batch)))
        return rc;
    double *P = g_single_pack.d();
    double *Wp = P + 2L * n * n * batch;
    const long nn = (long)n * n;
    const int nch = chunks_per_matrix(n);
    dim3 grid((unsigned)((long)batch * nch));
    hipLaunchKernelGGL(pack_kernel, grid, dim3(256), 0, s, d_A, lda, strideA, n, P, nch);
    hipLaunchKernelGGL(hermitize_kernel, grid, dim3(256), 0, s, P, (long)n, nn, n,
                       (uplo == 'L' || uplo == 'l') ? 1 : 0, nch);
    RLM_HIP_CHECK(hipGetLastError());
    {
        // a NaN or infinity in the triangle read stops the call before any solver kernel:
        // status 1 + the index of the first such matrix (LAPACK's info > 0 convention)
        const size_t nfl = (size_t)batch * nch;
        if ((rc = g_single_flag.ensure(sizeof(int) * nfl))) return rc;
        int *fl = static_cast<int *>(g_single_flag.p);
        hipLaunchKernelGGL(nonfinite_kernel, grid, dim3(256), 0, s, P, nn, nch, fl);
        RLM_HIP_CHECK(hipGetLastError());
        std::vector<int> h(nfl);
        RLM_HIP_CHECK(hipMemcpyAsync(h.data(), fl, sizeof(int) * nfl, hipMemcpyDeviceToHost, s));
        RLM_HIP_CHECK(hipStreamSynchronize(s));
        for (size_t k = 0; k < nfl; k++)
            if (h[k]) return 1 + (int)(k / nch);
    }
    if (precision == HIP_DIAG_MIXED) {
        // FP32 reduction and back-transform, FP64 divide & conquer, then FP64
        // Rayleigh quotients of the FP32 vectors against the FP64 matrix
        if ((rc = g_single_a32.ensure(sizeof(float) * 2 * (size_t)nn * batch + sizeof(double) * (size_t)batch * nch)))
            return rc;
        float *A32 = static_cast<float *>(g_single_a32.p);
        double *part = reinterpret_cast<double *>(A32 + 2 * (size_t)nn * batch);
        float *U32 = reinterpret_cast<float *>(g_single_ws.U.d());   // holds 2*nn*batch doubles >= floats
        hipLaunchKernelGGL(maxabs_kernel, grid, dim3(256), 0, s, P, nn, nch, part);
        hipLaunchKernelGGL(to_f32_scaled_kernel, grid, dim3(256), 0, s, P, nn, part, nch, A32);
        RLM_HIP_CHECK(hipGetLastError());
        if ((rc = eig_packed_f32(A32, n, batch, Wp, U32, g_single_ws, s))) return rc;
        if ((rc = launch_rayleigh(P, U32, n, n, batch, Wp, nullptr, s))) return rc;
        if (n <= kSmallMaxN) {
            // exact Rayleigh-Ritz of near-degenerate runs (n > 512 has no such step here;
            // the rlm plans refine those sizes in FP64)
            const int cap = pair_list_capacity(n, batch);
            if ((rc = g_single_pair.ensure(sizeof(int) * (3 * (size_t)cap + 4)))) return rc;
            int *pl = static_cast<int *>(g_single_pair.p);
            if ((rc = launch_pair_rr(P, nullptr, U32, n, n, batch, Wp, pl + 4, cap, pl, s))) return rc;
            int cnt[kPairCounters];
            RLM_HIP_CHECK(hipMemcpyAsync(cnt, pl, sizeof(cnt), hipMemcpyDeviceToHost, s));
            RLM_HIP_CHECK(hipStreamSynchronize(s));
            if (cnt[0] > cap || cnt[2] > 0)
                fprintf(stderr,
                        "[rlm_hip] mixed solver: %d near-degenerate chunks beyond the list capacity, %d without a "
                        "Cholesky factor (their FP64 Rayleigh quotients stand)\n",
                        cnt[0] > cap ? cnt[0] - cap : 0, cnt[2]);
        }
        hipLaunchKernelGGL(transpose_out_f32_kernel, grid, dim3(256), 0, s, U32, nn, n, d_A, lda, strideA, nch);
    } else {
        if ((rc = eig_packed(P, n, batch, Wp, g_single_ws.U.d(), g_single_ws, s))) return rc;
        hipLaunchKernelGGL(transpose_out_kernel, grid, dim3(256), 0, s, g_single_ws.U.d(), nn, n, d_A, lda, strideA,
                           nch);
    }
    RLM_HIP_CHECK(hipGetLastError());
    RLM_HIP_CHECK(hipMemcpy2DAsync(d_W, sizeof(double) * strideW, Wp, sizeof(double) * n, sizeof(double) * n, batch,
                                   hipMemcpyDeviceToDevice, s));
    RLM_HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int hip_diag_hamil(double *d_A, double *d_w, int n) {
    return hip_diag_heevd_batched(n, d_A, n, (long)n * n, d_w, n, 1, 'L', HIP_DIAG_FP64, nullptr);
}

// ---------------------------------------------------------------- allocation-free entry
// hip_diag_hamil_noalloc keeps every intermediate (packed matrix, eigenvalues, d / e /
// tau, eigenvectors, divide & conquer workspace, triangular factors, status flags) in
// one file-scope device array: kNoAllocMaxN = 256 needs 468 224 doubles (3.6 MiB); the
// array holds 4 MiB. Each GPU holds its own copy; its address on the current device
// comes from hipGetSymbolAddress once per device (the array's name is a host-side
// stand-in, never a device address).
constexpr int kNoAllocMaxN = 256;
constexpr size_t kScratchBytes = 4u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                        // serialises the calls of the process
void *g_scratch_ptr[kMaxDevices] = {};

// regions of the array, each starting on a 256-byte boundary (32 doubles)
static size_t pad32(size_t d) { return (d + 31) & ~(size_t)31; }
static size_t noalloc_doubles(int n) {
    const size_t nn = (size_t)n * n;
    return pad32(2 * nn) + pad32(n) + pad32(4 * (size_t)n) + pad32(2 * nn) + pad32(3 * nn) +
           pad32(2 * (size_t)backtransform_tws_complex(n)) + pad32((size_t)chunks_per_matrix(n));   // + flags
}

extern "C" int hip_diag_hamil_noalloc(double *d_A, double *d_w, int n) {
    if (!d_A || !d_w || n < 1 || n > kNoAllocMaxN) return -22;
    if (noalloc_doubles(n) * sizeof(double) > kScratchBytes) return -22;
    int dev = 0;
    RLM_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return -22;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        RLM_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
    }
    double *base = static_cast<double *>(g_scratch_ptr[dev]);
    const long nn = (long)n * n;
    const size_t tws = 2 * (size_t)backtransform_tws_complex(n);
    double *P = base;                                  // packed full Hermitian matrix (2 nn)
    double *Wp = P + pad32(2 * nn);                    // eigenvalues (n)
    EigWork view;                                      // non-owning views (never released)
    view.dev.p = Wp + pad32(n);                        // d | e | tau (4 n)
    view.dev.bytes = sizeof(double) * 4 * (size_t)n;
    view.U.p = static_cast<double *>(view.dev.p) + pad32(4 * (size_t)n);   // eigenvectors (2 nn)
    view.U.bytes = sizeof(double) * 2 * (size_t)nn;
    view.work.p = static_cast<double *>(view.U.p) + pad32(2 * nn);          // divide & conquer (3 nn)
    view.work.bytes = sizeof(double) * 3 * (size_t)nn;
    view.tw.p = static_cast<double *>(view.work.p) + pad32(3 * nn);         // triangular factors
    view.tw.bytes = sizeof(double) * tws;
    int *flags = reinterpret_cast<int *>(static_cast<double *>(view.tw.p) + pad32(tws));
    const int nch = chunks_per_matrix(n);
    const dim3 grid((unsigned)nch);
    hipStream_t s = nullptr;
    hipLaunchKernelGGL(pack_kernel, grid, dim3(256), 0, s, d_A, (long)n, nn, n, P, nch);
    hipLaunchKernelGGL(hermitize_kernel, grid, dim3(256), 0, s, P, (long)n, nn, n, 1, nch);
    hipLaunchKernelGGL(nonfinite_kernel, grid, dim3(256), 0, s, P, nn, nch, flags);
    RLM_HIP_CHECK(hipGetLastError());
    int hflag[256];
    RLM_HIP_CHECK(hipMemcpyAsync(hflag, flags, sizeof(int) * nch, hipMemcpyDeviceToHost, s));
    RLM_HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < nch; k++)
        if (hflag[k]) return 1;
    int rc;
    if ((rc = eig_packed(P, n, 1, Wp, view.U.d(), view, s))) return rc;
    hipLaunchKernelGGL(transpose_out_kernel, grid, dim3(256), 0, s, view.U.d(), nn, n, d_A, (long)n, nn, nch);
    RLM_HIP_CHECK(hipGetLastError());
    RLM_HIP_CHECK(hipMemcpyAsync(d_w, Wp, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    RLM_HIP_CHECK(hipStreamSynchronize(s));
    return 0;
}

extern "C" void hip_diag_hamil_v(double *d_A, double *d_w, int n) {
    int rc = hip_diag_hamil(d_A, d_w, n);
    if (rc) fprintf(stderr, "[rlm_hip] hip_diag_hamil failed with status %d\n", rc);
}

extern "C" int hip_heevd(double *A, double *w, int n) {
    if (n < 1 || !A || !w) return -22;
    double *dA = nullptr, *dw = nullptr;
    const size_t bA = sizeof(double) * 2 * (size_t)n * n;
    if (hipMalloc(&dA, bA) != hipSuccess) return -12;
    if (hipMalloc(&dw, sizeof(double) * n) != hipSuccess) {
        (void)hipFree(dA);
        return -12;
    }
    int rc = 0;
    if (hipMemcpy(dA, A, bA, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    if (!rc) rc = hip_diag_heevd_batched(n, dA, n, (long)n * n, dw, n, 1, 'U', HIP_DIAG_FP64, nullptr);
    if (!rc && hipMemcpy(A, dA, bA, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (!rc && hipMemcpy(w, dw, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    (void)hipFree(dA);
    (void)hipFree(dw);
    return rc;
}

// Device copy of a 2-slot ring (cal_Chern and the wfs file variant): grown as needed,
// shared by the calls of the process (serialised by g_ring_mu).
std::mutex g_ring_mu;
DevBuf g_ring_buf, g_ring_out;

static int ring_chern_device(const double *host_ring, double *chern_theta, int n_phi, int dim_vec, int n_mesh,
                             int theta_1, int io_order) {
    const size_t ring_doubles = 2 * 2 * (size_t)(n_mesh + 1) * n_phi * dim_vec;
    const size_t out_doubles = (size_t)n_mesh * n_phi;
    std::lock_guard<std::mutex> lk(g_ring_mu);
    int rc;
    if ((rc = g_ring_buf.ensure(sizeof(double) * ring_doubles)) || (rc = g_ring_out.ensure(sizeof(double) * out_doubles)))
        return rc;
    RLM_HIP_CHECK(hipMemcpy(g_ring_buf.p, host_ring, sizeof(double) * ring_doubles, hipMemcpyHostToDevice));
    if ((rc = launch_ring_plaquettes(g_ring_buf.d(), n_phi, dim_vec, n_mesh, theta_1, io_order, g_ring_out.d(),
                                     nullptr)))
        return rc;
    RLM_HIP_CHECK(hipMemcpy(chern_theta, g_ring_out.p, sizeof(double) * out_doubles, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hip_cal_chern(const double *wfs_full, double *chern_theta, int n_phi, int dim_vec, int n_mesh,
                             int theta_1) {
    if (!wfs_full || !chern_theta || n_phi < 1 || dim_vec < 1 || n_mesh < 1 || theta_1 < 1) return -22;
    return ring_chern_device(wfs_full, chern_theta, n_phi, dim_vec, n_mesh, theta_1, 0);
}

static std::string wfs_path(const char *dir, int theta_1, int theta_2) {
    char name[64];
    snprintf(name, sizeof(name), "wfs_%d_%d.bin", theta_1, theta_2);   // diag_wrappers/wfs_file.cpp:3-4
    if (!dir || !*dir) return std::string(name);
    std::string p(dir);
    if (p.back() != '/') p += '/';
    return p + name;
}

// raw complex values, no header; complex double input written as elem_bytes = 16
// (complex double) or 8 (complex float)
static int write_raw(const std::string &path, const double *v, size_t count, int elem_bytes) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return -5;
    size_t ok;
    if (elem_bytes == 16) {
        ok = fwrite(v, 16, count, f);
    } else {
        std::vector<float> tmp(2 * count);
        for (size_t k = 0; k < 2 * count; k++) tmp[k] = (float)v[k];
        ok = fwrite(tmp.data(), 8, count, f);
    }
    const int rc = fclose(f);
    return (ok == count && rc == 0) ? 0 : -5;
}

extern "C" int hip_write_wfs(const char *dir, int theta_1, int theta_2, const double *wfs, unsigned long size,
                             int elem_bytes) {
    if (!wfs || (elem_bytes != 8 && elem_bytes != 16)) return -22;
    return write_raw(wfs_path(dir, theta_1, theta_2), wfs, size, elem_bytes);
}

extern "C" int hip_cal_chern_wfs_io(double *chern_theta, int n_phi, int dim_vec, int n_mesh, int theta_1,
                                    const char *dir, int elem_bytes) {
    if (!chern_theta || n_phi < 1 || dim_vec < 1 || n_mesh < 1 || theta_1 < 1 || (elem_bytes != 8 && elem_bytes != 16))
        return -22;
    const size_t dim_wfs = (size_t)n_phi * dim_vec;
    std::vector<double> ring(2 * 2 * (size_t)(n_mesh + 1) * dim_wfs);
    std::vector<float> tmp(elem_bytes == 8 ? 2 * dim_wfs : 0);
    for (int slot = 0; slot < 2; slot++)
        for (int t2 = 0; t2 <= n_mesh; t2++) {
            FILE *f = fopen(wfs_path(dir, slot, t2).c_str(), "rb");
            if (!f) return -2;
            double *dst = ring.data() + 2 * ((size_t)slot * (n_mesh + 1) + t2) * dim_wfs;
            size_t got = elem_bytes == 16 ? fread(dst, 16, dim_wfs, f) : fread(tmp.data(), 8, dim_wfs, f);
            fclose(f);
            if (got != dim_wfs) return -5;
            if (elem_bytes == 8)
                for (size_t k = 0; k < 2 * dim_wfs; k++) dst[k] = (double)tmp[k];
        }
    return ring_chern_device(ring.data(), chern_theta, n_phi, dim_vec, n_mesh, theta_1, 1);
}

// ============================================================================
// rlm model entry points
// ============================================================================
extern "C" int rlm_prepare_coeff(double *coeff_mn, double *coeff_m_theta, double *coeff_jm, int n_phi, int off_head,
                                 int dim_m, int dim_n, int n_mesh, double L1, double L2) {
    if (!coeff_mn || !coeff_m_theta || !coeff_jm || n_phi < 1 || n_mesh < 1 || dim_m != dim_n) return -22;
    host_prepare_coeff(coeff_mn, coeff_m_theta, coeff_jm, n_phi, off_head, dim_m, n_mesh, L1, L2);
    return 0;
}

extern "C" int rlm_generate_disorder_potential(double *x, double *y, double *intensity, long impurity_num, double L1,
                                               double L2, unsigned long seed, long sample_offset) {
    if (!x || !y || !intensity || impurity_num < 0 || sample_offset < 0) return -22;
    host_rlm_disorder(x, y, intensity, impurity_num, L1, L2, (uint32_t)seed, sample_offset);
    return 0;
}

extern "C" int hip_generate_disorder_potential(double *d_x, double *d_y, double *d_intensity, long impurity_num,
                                               double L1, double L2, unsigned long seed, long sample_offset,
                                               int n_samples) {
    if (!d_x || !d_y || !d_intensity || impurity_num < 0 || sample_offset < 0 || n_samples < 0) return -22;
    if (impurity_num == 0 || n_samples == 0) return 0;
    StreamJumps j;
    int rc = j.create(n_samples, 3 * impurity_num);
    if (!rc) rc = j.select(sample_offset, nullptr);
    if (!rc) rc = j.launch((uint32_t)seed, n_samples, 0, 3 * impurity_num, impurity_num, L1, L2, d_x, d_y, d_intensity,
                           nullptr);
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) rc = -1;
    j.release();
    return rc;
}

extern "C" int hip_hlm_generate_disorder(double *d_V, long n_sites, long n_samples, unsigned long seed,
                                         long first_sample, double W) {
    if (!d_V || n_sites < 1 || n_samples < 0 || first_sample < 0 || n_samples > (1L << 30)) return -22;
    if (n_samples == 0) return 0;
    StreamJumps j;
    int rc = j.create(n_samples, n_sites);
    if (!rc) rc = j.select(first_sample, nullptr);
    if (!rc) rc = j.launch((uint32_t)seed, (int)n_samples, 1, n_sites, 0, 0.5 * W, 0.0, d_V, nullptr, nullptr, nullptr);
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess) rc = -1;
    j.release();
    return rc;
}

extern "C" int hip_prepare_potential_coeff(const double *d_x, const double *d_y, const double *d_intensity,
                                           double *d_Vmn, const double *d_coeff_mn, int dim_m, int dim_n,
                                           int off_head, long impurity_num) {
    if (dim_m != dim_n || dim_m < 1 || impurity_num < 0) return -22;
    int rc = launch_potential(d_x, d_y, d_intensity, impurity_num, d_coeff_mn, dim_m, off_head, d_Vmn, 1, nullptr);
    if (rc) return rc;
    RLM_HIP_CHECK(hipDeviceSynchronize());
    return 0;
}

extern "C" int hip_set_hamil_matrix(const double *d_coeff_m_theta_1, const double *d_coeff_m_theta_2,
                                    const double *d_coeff_jm, const double *d_Vmn, double *d_wfs, int n_phi,
                                    int dim_n, int off_head) {
    if (n_phi < 1 || dim_n != 2 * off_head + 1) return -22;
    // Build a 2-row coeff_m_theta table (rows: theta_1, theta_2) and reuse the batched kernels
    // with n_mesh = 1, theta_1 = 0 row, theta_2 = 1 row.
    double *cmt = nullptr, *E = nullptr, *Wt = nullptr;
    const int d = dim_n;
    int rc = 0;
    if (hipMalloc(&cmt, sizeof(double) * 4 * d) != hipSuccess) return -12;
    if (hipMalloc(&E, sizeof(double) * 2 * 2L * d * n_phi) != hipSuccess) rc = -12;
    if (!rc && hipMalloc(&Wt, sizeof(double) * 2 * 2L * d * n_phi) != hipSuccess) rc = -12;
    if (!rc && hipMemcpy(cmt, d_coeff_m_theta_1, sizeof(double) * 2 * d, hipMemcpyDeviceToDevice) != hipSuccess)
        rc = -1;
    if (!rc && hipMemcpy(cmt + 2 * d, d_coeff_m_theta_2, sizeof(double) * 2 * d, hipMemcpyDeviceToDevice) !=
                   hipSuccess)
        rc = -1;
    if (!rc) rc = launch_etable(cmt, d_coeff_jm, n_phi, d, 1, E, nullptr);
    if (!rc) rc = launch_wtable(d_Vmn, cmt, n_phi, d, off_head, 1, 1, Wt, nullptr);
    // k-point (t1 = 0, t2 = 1) of a 2x2 mesh is kp = 0*2 + 1
    if (!rc) rc = launch_assemble(E, Wt, n_phi, d, off_head, 1, 1, 1, d_wfs, (long)n_phi * n_phi, nullptr);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = -1;
    (void)hipFree(cmt);
    (void)hipFree(E);
    (void)hipFree(Wt);
    return rc;
}

// ============================================================================
// rlm batched pipeline
// ============================================================================
// energies [SN] then Chern values [SN] of a fetched batch: 1 + the first sample holding a
// non-finite value (every output is still copied), 0 if all are finite
static int first_nonfinite_sample(const double *h, long SN, long n_phi) {
    for (long k = 0; k < SN; k++)
        if (!std::isfinite(h[k]) || !std::isfinite(h[SN + k])) return 1 + (int)(k / n_phi);
    return 0;
}

struct rlm_plan_s {
    long n_phi, n_imp, off, d, g, spb;
    double L1, L2;
    bool gpu_conv = false;   // HIP_DIAG_CONV_GPU_FP32: the single-precision GPU builds' conventions
    uint32_t seed;
    int precision;
    hipStream_t stream = nullptr;
    // host
    std::vector<double> cmn, cmt, cjm;
    // disorder streams, generated on the device at every execute
    StreamJumps jumps;
    // device
    DevBuf dcmn, dcmt, dE, dX, dY, dS, dV, dW, dH, dA, dU, dEig, dL, dOut, dA32, dResid;
    // FP64 refinement of the mixed-precision eigenvectors (refine.hip): on by default for
    // n_phi > 512, RLM_REFINE=0 / 1 at rlm_plan_create turns it off / on
    bool refine = false;
    DevBuf dXd, dXr, dSw, dGw, dQg, dCl, dFlag;
    DevBuf dPair;   // near-degenerate runs of the unrefined mixed path: list + 2 counters
    DevBuf dWfs;    // staging of rlm_plan_write_wfs
    EigWork ws;
    int nb = 0;              // band width of the band form (dW: S*(g+1) x d x nb, dH: nkp x N x nb)
    bool band_trd = false;   // mixed plans: the FP32 reduction reads the band form (no dense matrices)
    // Solve only the twist points theta_1, theta_2 < g (pl = g points per line) and
    // derive the theta = 2 pi edges in the link kernel; RLM_MESH=full solves all
    // (g+1)^2 points (pl = g+1).
    int pl = 0;
    int loaded = 0;
    long first = 0;
    static constexpr int kStages = 8;
    hipEvent_t ev[kStages + 1] = {};
    // graph mode (rlm_plan_set_graph, or RLM_GRAPH=1 at creation)
    bool graph = false;
    bool staged = false;   // the last execute recorded every stage event (eager mode)
    hipGraphExec_t gexec = nullptr;
    long gkey[3] = {0, 0, 0};
};

extern "C" int rlm_plan_create(rlm_plan_t *plan, long n_phi, double q, double imp, long n_mesh, long off_override,
                               unsigned long seed, int samples_per_batch, int precision) {
    return rlm_plan_create_ex(plan, n_phi, q, imp, n_mesh, off_override, seed, samples_per_batch, precision, 0);
}

extern "C" int rlm_plan_create_ex(rlm_plan_t *plan, long n_phi, double q, double imp, long n_mesh, long off_override,
                                  unsigned long seed, int samples_per_batch, int precision, int flags) {
    if (!plan || n_phi < 1 || n_phi > heevd_large_max_n() || n_mesh < 1 || samples_per_batch < 1 || q <= 0.0)
        return -22;
    if (precision != HIP_DIAG_FP64 && precision != HIP_DIAG_MIXED) return -22;
    if (flags & ~HIP_DIAG_CONV_GPU_FP32) return -22;
    auto *p = new rlm_plan_s();
    p->n_phi = n_phi;
    p->gpu_conv = (flags & HIP_DIAG_CONV_GPU_FP32) != 0;
    if (p->gpu_conv) {
        // the single-precision GPU builds' sizes (rlm/main_rlm_gpu.cu:29-33: float L1, L2 and
        // a float impurity count)
        p->L1 = p->L2 = (double)std::sqrt((float)n_phi * (float)q);
        p->n_imp = (long)((float)imp * (float)n_phi / (float)q);
    } else {
        p->L1 = p->L2 = std::sqrt((double)n_phi * q);                 // rlm/main_rlm_cpu.cpp:29-30
        p->n_imp = (long)(imp * (double)n_phi / q);                   // :31
    }
    p->off = off_override >= 0 ? off_override : (long)(std::sqrt((double)p->n_imp) / 2 + 1);  // :32
    p->d = 2 * p->off + 1;                                            // :33
    p->g = n_mesh;
    p->spb = samples_per_batch;
    p->seed = (uint32_t)seed;
    p->precision = precision;
    {
        const char *gm = getenv("RLM_GRAPH");
        p->graph = gm && strcmp(gm, "1") == 0;
    }
    const long N = n_phi, d = p->d, g1 = n_mesh + 1, S = samples_per_batch;
    {
        // the GPU builds' conventions solve every point (the derived-edge relations hold
        // for the CPU build's phase conventions)
        const char *mesh = getenv("RLM_MESH");
        p->pl = ((mesh && !strcmp(mesh, "full")) || p->gpu_conv) ? (int)g1 : (int)n_mesh;
    }
    const long nkp = S * p->pl * p->pl;   // solved twist points
    const long nlk = S * g1 * g1;         // twist points with links
    p->nb = band_width((int)N, (int)p->off);
    int rc0 = 0;
    if (precision == HIP_DIAG_MIXED) {
        const char *rf = getenv("RLM_REFINE");
        p->refine = rf ? strcmp(rf, "0") != 0 : N > kSmallMaxN;
    }
    // autotuned tridiagonalisation kernel for this size and precision (decides whether the
    // mixed reduction can read the band form)
    if (N <= kSmallMaxN && (rc0 = tune_hetrd((int)N, precision == HIP_DIAG_MIXED, nkp))) {
        rlm_plan_destroy(p);
        return rc0;
    }
    p->band_trd = precision == HIP_DIAG_MIXED && chetrd_band_supported((int)N);
    const long nb = p->nb;
    p->cmn.resize(2 * d * d);
    p->cmt.resize(2 * g1 * d);
    p->cjm.resize(2 * N * d);
    if (p->gpu_conv) {
        host_prepare_coeff_f32(p->cmn.data(), p->cmt.data(), p->cjm.data(), (int)N, (int)p->off, (int)d, (int)n_mesh,
                               (float)p->L1, (float)p->L2);
        // h_jk = sum_m coeff_jm[j][m] c2[m] sum_n V[m][n] conj(c1[n]) in the GPU builds
        // (rlm/hamiltonian_rlm_gpu.cu:23,36) against conj(c2[m]) coeff_jm[j][m] and c1[n] in
        // the CPU build: the tables take conj(coeff_m_theta)
        for (size_t k = 1; k < p->cmt.size(); k += 2) p->cmt[k] = -p->cmt[k];
    } else {
        host_prepare_coeff(p->cmn.data(), p->cmt.data(), p->cjm.data(), (int)N, (int)p->off, (int)d, (int)n_mesh,
                           p->L1, p->L2);
    }
    int rc = 0;
    auto fail = [&](int r) {
        rlm_plan_destroy(p);
        return r;
    };
    if (hipStreamCreate(&p->stream) != hipSuccess) return fail(-1);
    for (int k = 0; k <= rlm_plan_s::kStages; k++)
        if (hipEventCreate(&p->ev[k]) != hipSuccess) return fail(-1);
    if ((rc = p->dcmn.ensure(8 * p->cmn.size())) || (rc = p->dcmt.ensure(8 * p->cmt.size())) ||
        (rc = p->dE.ensure(16 * g1 * d * N)) || (rc = p->dX.ensure(8 * S * p->n_imp + 8)) ||
        (rc = p->dY.ensure(8 * S * p->n_imp + 8)) || (rc = p->dS.ensure(8 * S * p->n_imp + 8)) ||
        (rc = p->dV.ensure(16 * S * d * d)) || (rc = p->dW.ensure(16 * S * g1 * d * nb)) ||
        (rc = p->dH.ensure(16 * nkp * N * nb)) || (!p->band_trd && (rc = p->dA.ensure(16 * nkp * N * N))) ||
        (rc = p->dU.ensure((precision == HIP_DIAG_MIXED ? 8 : 16) * nkp * N * N)) ||
        (precision == HIP_DIAG_MIXED && (rc = p->dA32.ensure(8 * nkp * N * N))) ||
        (rc = p->dResid.ensure(8 * nkp * rayleigh_resid_slots((int)N))) ||
        (rc = p->dEig.ensure(8 * nkp * N)) || (rc = p->dL.ensure(32 * nlk * N)) ||
        (rc = p->dOut.ensure(8 * 3 * S * N)) || (rc = p->ws.ensure((int)N, nkp, false)))
        return fail(rc);
    if (precision == HIP_DIAG_MIXED && !p->refine && N <= kSmallMaxN &&
        (rc = p->dPair.ensure(sizeof(int) * (3 * (size_t)pair_list_capacity((int)N, nkp) + 4))))
        return fail(rc);
    if (p->refine &&
        ((rc = p->dXd.ensure(16 * nkp * N * N)) || (rc = p->dXr.ensure(16 * nkp * N * N)) ||
         (rc = p->dSw.ensure(16 * nkp * N * refine_window())) || (rc = p->dGw.ensure(16 * nkp * N * refine_window())) ||
         (rc = p->dQg.ensure(16 * nkp * N * refine_max_cluster())) || (rc = p->dCl.ensure(4 * nkp * N)) ||
         (rc = p->dFlag.ensure(4 * nkp))))
        return fail(rc);
    if ((rc = p->jumps.create(S, 3 * p->n_imp))) return fail(rc);
    if (hipMemcpy(p->dcmn.p, p->cmn.data(), 8 * p->cmn.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->dcmt.p, p->cmt.data(), 8 * p->cmt.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(-1);
    {
        DevBuf dcjm;
        if ((rc = dcjm.ensure(8 * p->cjm.size()))) return fail(rc);
        if (hipMemcpy(dcjm.p, p->cjm.data(), 8 * p->cjm.size(), hipMemcpyHostToDevice) != hipSuccess) return fail(-1);
        if ((rc = launch_etable(p->dcmt.d(), dcjm.d(), (int)N, (int)d, (int)n_mesh, p->dE.d(), p->stream)))
            return fail(rc);
        if (hipStreamSynchronize(p->stream) != hipSuccess) return fail(-1);
        dcjm.release();
    }
    *plan = p;
    return 0;
}

extern "C" int rlm_plan_load(rlm_plan_t p, long first_sample, int count) {
    if (!p || count < 1 || count > p->spb || first_sample < 0) return -22;
    int rc;
    if ((rc = p->jumps.select(first_sample, p->stream))) return rc;
    p->loaded = count;
    p->first = first_sample;
    return 0;
}

// Every launch of one execute on the plan's stream (stage events when ev).
static int rlm_launches(rlm_plan_t p, bool ev) {
    const int S = p->loaded;
    const int N = (int)p->n_phi, d = (int)p->d, off = (int)p->off, g = (int)p->g;
    const long g1 = g + 1, nkp = (long)S * p->pl * p->pl, nlk = (long)S * g1 * g1;
    const bool derive = p->pl == g;
    hipStream_t s = p->stream;
    int rc;
    Range range("rlm_plan_execute");
    roctxMarkA("rlm: disorder streams + V_mn + band tables");
    if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[0], s));
    // disorder of the loaded samples from the seed (rlm/disorder_potential.cpp:3-21)
    if ((rc = p->jumps.launch(p->seed, S, p->gpu_conv ? 2 : 0, 3 * p->n_imp, p->n_imp, p->L1, p->L2, p->dX.d(),
                              p->dY.d(), p->dS.d(), s)))
        return rc;
    if ((rc = launch_potential(p->dX.d(), p->dY.d(), p->dS.d(), p->n_imp, p->dcmn.d(), d, off, p->dV.d(), S, s)))
        return rc;
    if ((rc = launch_wband(p->dV.d(), p->dcmt.d(), N, d, off, g, S, p->dW.d(), s))) return rc;
    if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[1], s));
    roctxMarkA("rlm: assembly");
    // band form of every k-point's Hamiltonian; dense FP64 matrices where a kernel needs them
    const bool mixed = p->precision == HIP_DIAG_MIXED;
    float *A32 = mixed ? static_cast<float *>(p->dA32.p) : nullptr;
    if ((rc = launch_hband(p->dE.d(), p->dW.d(), N, d, off, g, p->pl, 0, nkp, p->dH.d(), s))) return rc;
    if (!p->band_trd) {
        if ((rc = launch_expand_band(p->dH.d(), N, off, nkp, p->dA.d(), (long)N * N, s))) return rc;
        if (mixed) {
            hipLaunchKernelGGL(to_f32_kernel, dim3(2048), dim3(256), 0, s, p->dA.d(), 2 * nkp * N * (long)N, A32);
            RLM_HIP_CHECK(hipGetLastError());
        }
    }
    if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[2], s));
    roctxMarkA("rlm: eigensolve + refinement + links");
    double *L1 = p->dL.d(), *L2 = L1 + 2 * nlk * N;
    if (!mixed) {
        if ((rc = eig_packed(p->dA.d(), N, nkp, p->dEig.d(), p->dU.d(), p->ws, s, ev ? p->ev[3] : nullptr, ev ? p->ev[4] : nullptr))) return rc;
        if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[5], s));
        if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[6], s));
        if (derive)
            rc = launch_links_derived(p->dU.p, false, (long)N * N, N, g, S, L1, L2, s);
        else
            rc = launch_links(p->dU.d(), (long)N * N, N, N, (int)g1, (int)g1, g, 0, nkp, L1, L2, s);
        if (rc) return rc;
    } else {
        float *U32 = static_cast<float *>(p->dU.p);
        if ((rc = eig_packed_f32(A32, N, nkp, p->dEig.d(), U32, p->ws, s, ev ? p->ev[3] : nullptr, ev ? p->ev[4] : nullptr,
                                 p->band_trd ? p->dH.d() : nullptr, off)))
            return rc;
        if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[5], s));
        if (p->refine) {
            // FP64 eigenvector refinement, then Rayleigh quotients and links of the
            // refined FP64 vectors
            double *Xr = p->dXr.d();
            if ((rc = launch_refine(p->dH.d(), U32, N, off, nkp, p->dXd.d(), Xr, p->dSw.d(), p->dGw.d(), p->dQg.d(),
                                    static_cast<int *>(p->dCl.p), static_cast<int *>(p->dFlag.p), s)))
                return rc;
            if ((rc = launch_rayleigh(nullptr, static_cast<const double *>(Xr), N, off, nkp, p->dEig.d(), nullptr, s,
                                      p->dH.d())))
                return rc;
            if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[6], s));
            if (derive)
                rc = launch_links_derived(Xr, false, (long)N * N, N, g, S, L1, L2, s);
            else
                rc = launch_links(Xr, (long)N * N, N, N, (int)g1, (int)g1, g, 0, nkp, L1, L2, s);
            if (rc) return rc;
        } else {
            // FP64 Rayleigh quotients against the FP64 matrix (band |c-r| <= off_head), then
            // exact Rayleigh-Ritz of the near-degenerate runs
            if ((rc = launch_rayleigh(p->band_trd ? nullptr : p->dA.d(), U32, N, off, nkp, p->dEig.d(), nullptr, s,
                                      p->band_trd ? p->dH.d() : nullptr)))
                return rc;
            if (N <= kSmallMaxN) {
                const int cap = pair_list_capacity(N, nkp);
                int *pl = static_cast<int *>(p->dPair.p);
                if ((rc = launch_pair_rr(nullptr, p->dH.d(), U32, N, off, nkp, p->dEig.d(), pl + 4, cap, pl, s)))
                    return rc;
            }
            if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[6], s));
            if (derive)
                rc = launch_links_derived(U32, true, (long)N * N, N, g, S, L1, L2, s);
            else
                rc = launch_links_f32(U32, (long)N * N, N, N, (int)g1, (int)g1, g, 0, nkp, L1, L2, s);
            if (rc) return rc;
        }
    }
    if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[7], s));
    roctxMarkA("rlm: plaquettes + energies");
    double *out = p->dOut.d();
    if ((rc = launch_reduce(L1, L2, p->dEig.d(), N, g, (int)g1, derive ? 2 : 0, (double)N, S, out, out + (long)S * N,
                            out + 2L * S * N, s)))
        return rc;
    if (ev) RLM_HIP_CHECK(hipEventRecord(p->ev[8], s));
    return 0;
}

// Graph mode: the launches of one execute captured once into a hipGraph (per loaded
// count and first-sample jump, the kernel arguments that change between batches) and
// replayed with one hipGraphLaunch. The stage events are captured with it.
static int rlm_graph_execute(rlm_plan_t p) {
    const long key[3] = {p->loaded, p->jumps.a_kind, p->jumps.a_mono};
    if (!p->gexec || memcmp(key, p->gkey, sizeof(key)) != 0) {
        if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
        p->gexec = nullptr;
        RLM_HIP_CHECK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        const int rc = rlm_launches(p, false);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(p->stream, &gr);
        if (rc || ec != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc ? rc : -(int)ec;
        }
        const hipError_t ei = hipGraphInstantiate(&p->gexec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) {
            p->gexec = nullptr;
            return -(int)ei;
        }
        memcpy(p->gkey, key, sizeof(key));
    }
    // events recorded inside a graph carry no timing: time the whole launch
    RLM_HIP_CHECK(hipEventRecord(p->ev[0], p->stream));
    RLM_HIP_CHECK(hipGraphLaunch(p->gexec, p->stream));
    RLM_HIP_CHECK(hipEventRecord(p->ev[rlm_plan_s::kStages], p->stream));
    p->staged = false;
    return 0;
}

extern "C" int rlm_plan_execute(rlm_plan_t p) {
    if (!p || p->loaded < 1) return -22;
    if (p->graph) return rlm_graph_execute(p);
    p->staged = true;
    return rlm_launches(p, true);
}

extern "C" int rlm_plan_elapsed_ms(rlm_plan_t p, double *ms) {
    if (!p || !ms) return -22;
    RLM_HIP_CHECK(hipEventSynchronize(p->ev[rlm_plan_s::kStages]));
    float t = 0.f;
    RLM_HIP_CHECK(hipEventElapsedTime(&t, p->ev[0], p->ev[rlm_plan_s::kStages]));
    *ms = t;
    return 0;
}

extern "C" int rlm_plan_set_graph(rlm_plan_t p, int enable) {
    if (!p) return -22;
    p->graph = enable != 0;
    return 0;
}

extern "C" int rlm_plan_fetch(rlm_plan_t p, double *energy, double *chern, double *margin) {
    if (!p || p->loaded < 1) return -22;
    const long SN = (long)p->loaded * p->n_phi;
    RLM_HIP_CHECK(hipStreamSynchronize(p->stream));
    std::vector<double> h(3 * SN);
    RLM_HIP_CHECK(hipMemcpy(h.data(), p->dOut.p, 8 * 3 * SN, hipMemcpyDeviceToHost));
    if (energy) memcpy(energy, h.data(), 8 * SN);
    if (chern) memcpy(chern, h.data() + SN, 8 * SN);
    if (margin) memcpy(margin, h.data() + 2 * SN, 8 * SN);
    return first_nonfinite_sample(h.data(), SN, p->n_phi);
}

extern "C" int rlm_plan_run(rlm_plan_t p, long first_sample, long n_samples, double *energy, double *chern,
                            double *margin) {
    if (!p || n_samples < 0) return -22;
    for (long s0 = 0; s0 < n_samples; s0 += p->spb) {
        const int cnt = (int)((n_samples - s0) < p->spb ? (n_samples - s0) : p->spb);
        int rc;
        if ((rc = rlm_plan_load(p, first_sample + s0, cnt))) return rc;
        if ((rc = rlm_plan_execute(p))) return rc;
        const long off = s0 * p->n_phi;
        if ((rc = rlm_plan_fetch(p, energy ? energy + off : nullptr, chern ? chern + off : nullptr,
                                 margin ? margin + off : nullptr)))
            return rc > 0 ? rc + (int)s0 : rc;   // > 0: 1 + the sample (of the run) with a non-finite value
    }
    return 0;
}

extern "C" void *rlm_plan_stream(rlm_plan_t p) { return p ? (void *)p->stream : nullptr; }

extern "C" int rlm_plan_loaded(rlm_plan_t p) { return p ? p->loaded : -22; }

extern "C" int rlm_plan_info(rlm_plan_t p, long *o) {
    if (!p || !o) return -22;
    o[0] = p->n_phi;
    o[1] = p->n_imp;
    o[2] = p->off;
    o[3] = p->d;
    o[4] = p->g;
    o[5] = p->spb;
    return 0;
}

extern "C" int rlm_plan_stage_times(rlm_plan_t p, double *ms) {
    if (!p || !ms) return -22;
    if (!p->staged) return 0;   // graph mode: rlm_plan_elapsed_ms has the total
    RLM_HIP_CHECK(hipEventSynchronize(p->ev[rlm_plan_s::kStages]));
    for (int k = 0; k < rlm_plan_s::kStages; k++) {
        float t = 0.f;
        RLM_HIP_CHECK(hipEventElapsedTime(&t, p->ev[k], p->ev[k + 1]));
        ms[k] = t;
    }
    return rlm_plan_s::kStages;
}

extern "C" int rlm_plan_max_residual(rlm_plan_t p, double *out) {
    if (!p || !out || p->loaded < 1) return -22;
    if (p->precision != HIP_DIAG_MIXED) {
        *out = 0.0;
        return 0;
    }
    // residuals of the last executed batch, from the full FP64 product (the plan's
    // refinement itself uses the half-band form and computes no residual)
    const long nkp = (long)p->loaded * p->pl * p->pl;
    const int N = (int)p->n_phi;
    int rc;
    if (p->refine)
        rc = launch_rayleigh(nullptr, static_cast<const double *>(p->dXr.p), N, (int)p->off, nkp, nullptr,
                             p->dResid.d(), p->stream, p->dH.d());
    else
        rc = launch_rayleigh(p->band_trd ? nullptr : p->dA.d(), static_cast<const float *>(p->dU.p), N, (int)p->off,
                             nkp, nullptr, p->dResid.d(), p->stream, p->band_trd ? p->dH.d() : nullptr);
    if (rc) return rc;
    const long nres = nkp * rayleigh_resid_slots(N);
    std::vector<double> h(nres);
    RLM_HIP_CHECK(hipStreamSynchronize(p->stream));
    RLM_HIP_CHECK(hipMemcpy(h.data(), p->dResid.p, 8 * nres, hipMemcpyDeviceToHost));
    double m = 0.0;
    for (double v : h) m = v > m ? v : m;
    *out = m;
    return 0;
}

extern "C" const double *rlm_plan_eigenvalues(rlm_plan_t p) { return p ? p->dEig.d() : nullptr; }
extern "C" const double *rlm_plan_eigenvectors(rlm_plan_t p) {
    return p ? (p->refine ? p->dXr.d() : p->dU.d()) : nullptr;
}

extern "C" int rlm_plan_write_wfs(rlm_plan_t p, int sample, int theta_1, const char *dir, int elem_bytes) {
    if (!p || p->loaded < 1 || sample < 0 || sample >= p->loaded || theta_1 < 0 || theta_1 > p->g ||
        (elem_bytes != 8 && elem_bytes != 16))
        return -22;
    const int N = (int)p->n_phi, g = (int)p->g;
    const size_t per = (size_t)N * N;
    const size_t count = per * (g + 1);
    int rc;
    if ((rc = p->dWfs.ensure((size_t)elem_bytes * count))) return rc;
    const bool u32 = p->precision == HIP_DIAG_MIXED && !p->refine;
    const void *U = rlm_plan_eigenvectors(p);
    if ((rc = launch_gather_line(U, u32, (long)N * N, N, g, p->pl, sample, theta_1, p->dWfs.p, elem_bytes == 8,
                                 p->stream)))
        return rc;
    std::vector<unsigned char> h((size_t)elem_bytes * count);
    RLM_HIP_CHECK(hipStreamSynchronize(p->stream));
    RLM_HIP_CHECK(hipMemcpy(h.data(), p->dWfs.p, h.size(), hipMemcpyDeviceToHost));
    for (int t2 = 0; t2 <= g; t2++) {
        FILE *f = fopen(wfs_path(dir, theta_1 % 2, t2).c_str(), "wb");
        if (!f) return -5;
        const size_t ok = fwrite(h.data() + (size_t)elem_bytes * per * t2, elem_bytes, per, f);
        if (fclose(f) != 0 || ok != per) return -5;
    }
    return 0;
}

extern "C" int rlm_plan_pair_info(rlm_plan_t p, long *out) {
    if (!p || !out) return -22;
    for (int k = 0; k < 4; k++) out[k] = 0;
    if (!p->dPair.p || p->loaded < 1) return 0;
    int cnt[kPairCounters];
    RLM_HIP_CHECK(hipStreamSynchronize(p->stream));
    RLM_HIP_CHECK(hipMemcpy(cnt, p->dPair.p, sizeof(cnt), hipMemcpyDeviceToHost));
    const long cap = pair_list_capacity((int)p->n_phi, (long)p->loaded * p->pl * p->pl);
    out[0] = cnt[0] < cap ? cnt[0] : cap;
    out[1] = cnt[1];
    out[2] = cnt[0] > cap ? cnt[0] - cap : 0;
    out[3] = cnt[2];
    return 0;
}

extern "C" int rlm_plan_refine_info(rlm_plan_t p, long *out) {
    if (!p || !out) return -22;
    out[0] = p->refine ? 1 : 0;
    out[1] = 0;
    if (p->refine && p->loaded > 0) {
        const long nkp = (long)p->loaded * p->pl * p->pl;
        std::vector<int> h(nkp);
        RLM_HIP_CHECK(hipStreamSynchronize(p->stream));
        RLM_HIP_CHECK(hipMemcpy(h.data(), p->dFlag.p, sizeof(int) * nkp, hipMemcpyDeviceToHost));
        for (int v : h) out[1] += v;
    }
    return 0;
}

extern "C" int rlm_plan_destroy(rlm_plan_t p) {
    if (!p) return 0;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    DevBuf *bufs[] = {&p->dcmn, &p->dcmt, &p->dE,   &p->dX,   &p->dY,   &p->dS,   &p->dV,
                      &p->dW,   &p->dH,   &p->dA,   &p->dU,   &p->dEig, &p->dL,   &p->dOut,
                      &p->dA32, &p->dResid, &p->dXd, &p->dXr, &p->dSw,  &p->dGw,  &p->dQg,
                      &p->dCl,  &p->dFlag,  &p->dPair, &p->dWfs};
    for (DevBuf *b : bufs) b->release();
    p->ws.release();
    p->jumps.release();
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    for (int k = 0; k <= rlm_plan_s::kStages; k++)
        if (p->ev[k]) (void)hipEventDestroy(p->ev[k]);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

// ============================================================================
// rlm plans of several sizes in one graph launch
// ============================================================================
struct rlm_multi_s {
    std::vector<rlm_plan_t> plans;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr, e0 = nullptr, e1 = nullptr;
    std::vector<hipEvent_t> joins;
    hipGraphExec_t gexec = nullptr;
    std::vector<long> gkey;
};

extern "C" int rlm_multi_destroy(rlm_multi_t mp) {
    if (!mp) return 0;
    if (mp->origin) (void)hipStreamSynchronize(mp->origin);
    if (mp->gexec) (void)hipGraphExecDestroy(mp->gexec);
    for (rlm_plan_t p : mp->plans) rlm_plan_destroy(p);
    for (hipEvent_t e : mp->joins)
        if (e) (void)hipEventDestroy(e);
    hipEvent_t evs[] = {mp->fork, mp->e0, mp->e1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (mp->origin) (void)hipStreamDestroy(mp->origin);
    delete mp;
    return 0;
}

extern "C" int rlm_multi_create(rlm_multi_t *out, int nsizes, const long *n_phi, double q, double imp, long n_mesh,
                                unsigned long seed, int samples_per_batch, int precision) {
    if (!out || nsizes < 1 || !n_phi) return -22;
    auto *mp = new rlm_multi_s();
    auto fail = [&](int r) {
        rlm_multi_destroy(mp);
        return r;
    };
    for (int k = 0; k < nsizes; k++) {
        rlm_plan_t p = nullptr;
        const int rc = rlm_plan_create(&p, n_phi[k], q, imp, n_mesh, -1, seed, samples_per_batch, precision);
        if (rc) return fail(rc);
        mp->plans.push_back(p);
    }
    if (hipStreamCreate(&mp->origin) != hipSuccess || hipEventCreate(&mp->e0) != hipSuccess ||
        hipEventCreate(&mp->e1) != hipSuccess ||
        hipEventCreateWithFlags(&mp->fork, hipEventDisableTiming) != hipSuccess)
        return fail(-1);
    mp->joins.assign(nsizes, nullptr);
    for (int k = 0; k < nsizes; k++)
        if (hipEventCreateWithFlags(&mp->joins[k], hipEventDisableTiming) != hipSuccess) return fail(-1);
    *out = mp;
    return 0;
}

extern "C" int rlm_multi_load(rlm_multi_t mp, long first_sample, int count) {
    if (!mp) return -22;
    for (rlm_plan_t p : mp->plans) {
        const int rc = rlm_plan_load(p, first_sample, count);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int rlm_multi_execute(rlm_multi_t mp) {
    if (!mp) return -22;
    std::vector<long> key;
    for (rlm_plan_t p : mp->plans) {
        if (p->loaded < 1) return -22;
        key.push_back(p->loaded);
        key.push_back(p->jumps.a_kind);
        key.push_back(p->jumps.a_mono);
    }
    if (!mp->gexec || key != mp->gkey) {
        if (mp->gexec) (void)hipGraphExecDestroy(mp->gexec);
        mp->gexec = nullptr;
        // one branch per size: each plan's launches on its own stream, forked from and
        // joined back into the capture of the origin stream
        RLM_HIP_CHECK(hipStreamBeginCapture(mp->origin, hipStreamCaptureModeThreadLocal));
        int rc = hipEventRecord(mp->fork, mp->origin) == hipSuccess ? 0 : -1;
        for (size_t k = 0; k < mp->plans.size() && !rc; k++) {
            rlm_plan_t p = mp->plans[k];
            if (hipStreamWaitEvent(p->stream, mp->fork, 0) != hipSuccess) rc = -1;
            if (!rc) rc = rlm_launches(p, false);
            if (!rc && hipEventRecord(mp->joins[k], p->stream) != hipSuccess) rc = -1;
            if (!rc && hipStreamWaitEvent(mp->origin, mp->joins[k], 0) != hipSuccess) rc = -1;
        }
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(mp->origin, &gr);
        if (rc || ec != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc ? rc : -(int)ec;
        }
        const hipError_t ei = hipGraphInstantiate(&mp->gexec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) {
            mp->gexec = nullptr;
            return -(int)ei;
        }
        mp->gkey = key;
    }
    RLM_HIP_CHECK(hipEventRecord(mp->e0, mp->origin));
    RLM_HIP_CHECK(hipGraphLaunch(mp->gexec, mp->origin));
    RLM_HIP_CHECK(hipEventRecord(mp->e1, mp->origin));
    return 0;
}

extern "C" int rlm_multi_fetch(rlm_multi_t mp, int size_index, double *energy, double *chern, double *margin) {
    if (!mp || size_index < 0 || size_index >= (int)mp->plans.size()) return -22;
    RLM_HIP_CHECK(hipStreamSynchronize(mp->origin));
    return rlm_plan_fetch(mp->plans[size_index], energy, chern, margin);
}

extern "C" int rlm_multi_elapsed_ms(rlm_multi_t mp, double *ms) {
    if (!mp || !ms) return -22;
    RLM_HIP_CHECK(hipEventSynchronize(mp->e1));
    float t = 0.f;
    RLM_HIP_CHECK(hipEventElapsedTime(&t, mp->e0, mp->e1));
    *ms = t;
    return 0;
}

// ============================================================================
// hlm batched pipeline
// ============================================================================
extern "C" int hlm_generate_disorder(double *V, long n_sites, long n_samples, unsigned long seed, long first_sample,
                                     double W) {
    if (!V || n_sites < 1 || n_samples < 0 || first_sample < 0) return -22;
    uint32_t win[mtj::kN];
    mtj::seed_window((uint32_t)seed, (uint64_t)n_sites * (uint64_t)first_sample, win);
    mtj::WindowGen gen(win);
    const double half = 0.5 * W;
    for (long i = 0; i < n_sites * n_samples; i++) {
        const double u = mt_uniform(gen, -1.0, 1.0);   // hlm/main_hlm_cpu.cpp:79
        V[i] = (W == 2.0) ? u : half * u;
    }
    return 0;
}

struct hlm_plan_s {
    int lx, ly, Q, band, g, n_sites, n_phi, spb;
    uint32_t seed;
    double W;
    int precision;
    long ntw;                       // (g+1)*g twist points per sample
    hipStream_t stream = nullptr;
    StreamJumps jumps;   // the run's single disorder stream, n_sites words per sample
    DevBuf dPhi, dM1, dM2, dV, dH, dU, dEig, dL, dOut;
    EigWork ws;
    int loaded = 0;
    static constexpr int kStages = 7;
    hipEvent_t ev[kStages + 1] = {};
};

extern "C" int hlm_plan_destroy(hlm_plan_t p);

extern "C" int hlm_plan_create(hlm_plan_t *plan, int lx, int ly, int Q, int band, int n_mesh, unsigned long seed,
                               double W, int samples_per_batch, int precision) {
    if (!plan || lx < 1 || ly < 1 || Q < 1 || lx % Q != 0 || band < 0 || band >= Q || n_mesh < 1 ||
        samples_per_batch < 1 || precision != HIP_DIAG_FP64 || !(W > 0.0))
        return -22;
    const int n_sites = lx * ly, n_phi = n_sites / Q;
    if (n_phi > 512 || Q > 512) return -22;
    auto *p = new hlm_plan_s();
    p->lx = lx; p->ly = ly; p->Q = Q; p->band = band; p->g = n_mesh;
    p->n_sites = n_sites; p->n_phi = n_phi; p->spb = samples_per_batch;
    p->seed = (uint32_t)seed; p->W = W; p->precision = precision;
    p->ntw = (long)(n_mesh + 1) * n_mesh;
    auto fail = [&](int r) {
        hlm_plan_destroy(p);
        return r;
    };
    if (hipStreamCreate(&p->stream) != hipSuccess) return fail(-1);
    for (int k = 0; k <= hlm_plan_s::kStages; k++)
        if (hipEventCreate(&p->ev[k]) != hipSuccess) return fail(-1);
    const long N = n_phi, S = samples_per_batch, ntw = p->ntw, nkp = S * ntw;
    int rc;
    if ((rc = tune_hetrd((int)N, false, nkp))) return fail(rc);
    if ((rc = p->dPhi.ensure(16 * ntw * n_sites * N)) || (rc = p->dM1.ensure(16 * ntw * N * N)) ||
        (rc = p->dM2.ensure(16 * ntw * N * N)) || (rc = p->dV.ensure(8 * S * n_sites)) ||
        (rc = p->dH.ensure(16 * nkp * N * N)) || (rc = p->dU.ensure(16 * nkp * N * N)) ||
        (rc = p->dEig.ensure(8 * nkp * N)) || (rc = p->dL.ensure(32 * nkp * N)) ||
        (rc = p->dOut.ensure(8 * 3 * S * N)) || (rc = p->ws.ensure((int)N, nkp, false)) ||
        (rc = p->jumps.create(S, n_sites)))
        return fail(rc);
    // Bloch basis for every twist (disorder independent)
    {
        const long nmat = ntw * N;
        DevBuf bH, bW, bU;
        EigWork bws;
        if ((rc = bH.ensure(16 * nmat * Q * Q)) || (rc = bW.ensure(8 * nmat * Q)) ||
            (rc = bU.ensure(16 * nmat * Q * Q)) || (rc = bws.ensure(Q, nmat, false)))
            return fail(rc);
        if ((rc = launch_bloch_build(lx, ly, Q, n_mesh, nmat, bH.d(), p->stream))) return fail(rc);
        if ((rc = eig_packed(bH.d(), Q, nmat, bW.d(), bU.d(), bws, p->stream))) return fail(rc);
        if ((rc = launch_phi_build(lx, ly, Q, band, n_mesh, nmat, bU.d(), p->dPhi.d(), p->stream))) return fail(rc);
        if (hipStreamSynchronize(p->stream) != hipSuccess) return fail(-1);
    }
    // link matrices M1(t1,t2) = Phi(t1,t2)^H Phi(t1+1,t2), M2(t1,t2) = Phi(t1,t2)^H Phi(t1,(t2+1)%g)
    {
        const long sz = (long)n_sites * N;        // complex elements per Phi
        const long nn = N * N;
        const double *Phi = p->dPhi.d();
        const int g = n_mesh;
        // M1: batch over tw in [0, g*g)
        if ((rc = launch_zgemm((int)N, (int)N, n_sites, Phi, 1, n_sites, sz, 0, 1, Phi + 2 * g * sz, 1, n_sites, sz,
                               0, nullptr, 0, 0, p->dM1.d(), 1, N, nn, 0, g * g, 1, p->stream)))
            return fail(rc);
        // M2 for t2 < g-1: b1 = t1 (stride g), b2 = t2
        if (g > 1 &&
            (rc = launch_zgemm((int)N, (int)N, n_sites, Phi, 1, n_sites, g * sz, sz, 1, Phi + 2 * sz, 1, n_sites,
                               g * sz, sz, nullptr, 0, 0, p->dM2.d(), 1, N, g * nn, nn, g + 1, g - 1, p->stream)))
            return fail(rc);
        // M2 for t2 = g-1 -> t2 = 0 of the same line
        if ((rc = launch_zgemm((int)N, (int)N, n_sites, Phi + 2 * (g - 1) * sz, 1, n_sites, g * sz, 0, 1, Phi, 1,
                               n_sites, g * sz, 0, nullptr, 0, 0, p->dM2.d() + 2 * (g - 1) * nn, 1, N, g * nn, 0,
                               g + 1, 1, p->stream)))
            return fail(rc);
        if (hipStreamSynchronize(p->stream) != hipSuccess) return fail(-1);
    }
    *plan = p;
    return 0;
}

extern "C" int hlm_plan_load(hlm_plan_t p, long first_sample, int count) {
    if (!p || count < 1 || count > p->spb || first_sample < 0) return -22;
    int rc;
    if ((rc = p->jumps.select(first_sample, p->stream))) return rc;
    p->loaded = count;
    return 0;
}

extern "C" int hlm_plan_execute(hlm_plan_t p) {
    if (!p || p->loaded < 1) return -22;
    const int S = p->loaded, N = p->n_phi, ns = p->n_sites, g = p->g;
    const long ntw = p->ntw, nkp = S * ntw, nn = (long)N * N;
    hipStream_t s = p->stream;
    int rc;
    Range range("hlm_plan_execute");
    RLM_HIP_CHECK(hipEventRecord(p->ev[0], s));
    // on-site disorder of the loaded samples (hlm/main_hlm_cpu.cpp:79, width W)
    if ((rc = p->jumps.launch(p->seed, S, 1, ns, 0, 0.5 * p->W, 0.0, p->dV.d(), nullptr, nullptr, s))) return rc;
    // Htr[(s*ntw + tw)] = Phi(tw)^H diag(V_s) Phi(tw), upper triangle
    if ((rc = launch_hlm_project(p->dPhi.d(), p->dV.d(), N, ns, S, ntw, p->dH.d(), s))) return rc;
    // the reference's zheevd reads the upper triangle of Htr (mkl_diag.cpp:6)
    {
        hipLaunchKernelGGL(hermitize_kernel, dim3((unsigned)nkp), dim3(256), 0, s, p->dH.d(), (long)N, nn, N, 0, 1);
        RLM_HIP_CHECK(hipGetLastError());
    }
    RLM_HIP_CHECK(hipEventRecord(p->ev[1], s));
    if ((rc = eig_packed(p->dH.d(), N, nkp, p->dEig.d(), p->dU.d(), p->ws, s, p->ev[2], p->ev[3]))) return rc;
    RLM_HIP_CHECK(hipEventRecord(p->ev[4], s));
    double *L1 = p->dL.d(), *L2 = L1 + 2 * nkp * N;
    if ((rc = launch_hlm_links(p->dU.d(), N, g, nkp, p->dM1.d(), p->dM2.d(), L1, L2, s))) return rc;
    RLM_HIP_CHECK(hipEventRecord(p->ev[5], s));
    double *out = p->dOut.d();
    if ((rc = launch_reduce(L1, L2, p->dEig.d(), N, g, g, 1, (double)N, S, out, out + (long)S * N, out + 2L * S * N,
                            s)))
        return rc;
    RLM_HIP_CHECK(hipEventRecord(p->ev[6], s));
    RLM_HIP_CHECK(hipEventRecord(p->ev[7], s));
    return 0;
}

extern "C" int hlm_plan_fetch(hlm_plan_t p, double *energy, double *chern, double *margin) {
    if (!p || p->loaded < 1) return -22;
    const long SN = (long)p->loaded * p->n_phi;
    RLM_HIP_CHECK(hipStreamSynchronize(p->stream));
    std::vector<double> h(3 * SN);
    RLM_HIP_CHECK(hipMemcpy(h.data(), p->dOut.p, 8 * 3 * SN, hipMemcpyDeviceToHost));
    if (energy) memcpy(energy, h.data(), 8 * SN);
    if (chern) memcpy(chern, h.data() + SN, 8 * SN);
    if (margin) memcpy(margin, h.data() + 2 * SN, 8 * SN);
    return first_nonfinite_sample(h.data(), SN, p->n_phi);
}

extern "C" int hlm_plan_run(hlm_plan_t p, long first_sample, long n_samples, double *energy, double *chern,
                            double *margin) {
    if (!p || n_samples < 0) return -22;
    for (long s0 = 0; s0 < n_samples; s0 += p->spb) {
        const int cnt = (int)((n_samples - s0) < p->spb ? (n_samples - s0) : p->spb);
        int rc;
        if ((rc = hlm_plan_load(p, first_sample + s0, cnt))) return rc;
        if ((rc = hlm_plan_execute(p))) return rc;
        const long off = s0 * p->n_phi;
        if ((rc = hlm_plan_fetch(p, energy ? energy + off : nullptr, chern ? chern + off : nullptr,
                                 margin ? margin + off : nullptr)))
            return rc > 0 ? rc + (int)s0 : rc;   // > 0: 1 + the sample (of the run) with a non-finite value
    }
    return 0;
}

extern "C" int hlm_plan_stage_times(hlm_plan_t p, double *ms) {
    if (!p || !ms) return -22;
    RLM_HIP_CHECK(hipEventSynchronize(p->ev[hlm_plan_s::kStages]));
    for (int k = 0; k < hlm_plan_s::kStages; k++) {
        float t = 0.f;
        RLM_HIP_CHECK(hipEventElapsedTime(&t, p->ev[k], p->ev[k + 1]));
        ms[k] = t;
    }
    return hlm_plan_s::kStages;
}

extern "C" void *hlm_plan_stream(hlm_plan_t p) { return p ? (void *)p->stream : nullptr; }

extern "C" int hlm_plan_loaded(hlm_plan_t p) { return p ? p->loaded : -22; }

extern "C" const char *hip_diag_tuned_kernel(int n, int precision) {
    return hetrd_choice_name(n, precision == HIP_DIAG_MIXED);
}

extern "C" int hip_diag_symv_timing(int enable, double *total_ms, long *launches) {
    return symv_timing(enable, total_ms, launches);
}

extern "C" int hlm_plan_destroy(hlm_plan_t p) {
    if (!p) return 0;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    DevBuf *bufs[] = {&p->dPhi, &p->dM1, &p->dM2, &p->dV, &p->dH, &p->dU, &p->dEig, &p->dL, &p->dOut};
    for (DevBuf *b : bufs) b->release();
    p->ws.release();
    p->jumps.release();
    for (int k = 0; k <= hlm_plan_s::kStages; k++)
        if (p->ev[k]) (void)hipEventDestroy(p->ev[k]);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}
