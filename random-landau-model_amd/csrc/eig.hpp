// Host launchers of the batched FP64 Hermitian eigensolver (eig.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace rlm {

// Householder tridiagonalisation of `batch` n x n complex matrices (column-major,
// lda, full Hermitian storage, real diagonal). Writes d (n), e (n-1) and the
// complex tau (n-1) at stride strideV per matrix; reflectors stay in A.
int launch_zhetrd(double *A, long strideA, int lda, int n, double *D, double *E, double *TAU, long strideV, int batch,
                  int grid, hipStream_t stream);

// Divide & conquer on (d, e) + back-transform with the reflectors in A (the
// back-transform runs on the matrix cores for n <= 256 when tws is given; `mid`,
// if given, is recorded between the divide & conquer and the back-transform).
// W: eigenvalues ascending (strideW per matrix). U: eigenvectors, component-major
// complex (U[r*n + j] = component r of state j), strideU complex elements per matrix.
// work: grid * 3*n*n doubles. tws: batch * backtransform_tws_complex(n) complex
// elements of the precision of A (or null).
int launch_tail(const double *A, long strideA, int lda, int n, const double *D, const double *E, const double *TAU,
                long strideV, double *W, long strideW, double *U, long strideU, double *work, double *tws, int batch,
                int grid, hipStream_t stream, hipEvent_t mid = nullptr);
long backtransform_tws_complex(int n);

// FP32 variants (mixed precision): A, TAU, U single-precision complex; D, E, W and the
// divide & conquer stay FP64.
int launch_chetrd(float *A, long strideA, int lda, int n, double *D, double *E, float *TAU, long strideV, int batch,
                  int grid, hipStream_t stream);
// FP32 tridiagonalisation reading the matrices from their band form (complex
// double, n x band_nb(n, hb_off) each); A receives the reflectors only. Supported
// where chetrd_band_supported(n) (register-resident kernel, n <= 256).
bool chetrd_band_supported(int n);
int launch_chetrd_band(float *A, long strideA, int lda, int n, double *D, double *E, float *TAU, long strideV,
                       int batch, hipStream_t stream, const double *Hb, int hb_off);
int launch_tail_f32(const float *A, long strideA, int lda, int n, const double *D, const double *E, const float *TAU,
                    long strideV, double *W, long strideW, float *U, long strideU, double *work, float *tws, int batch,
                    int grid, hipStream_t stream, hipEvent_t mid = nullptr);

size_t heevd_dc_lds_bytes(int n);

// Autotuner of the tridiagonalisation kernel (n <= 512): times every valid variant on
// min(batch_hint, 256..1024) random Hermitian matrices and keeps the fastest for
// (n, precision) in this process (and in the tune file). RLM_TUNE=0 or RLM_HETRD set:
// no tuning. hetrd_choice_name: the variant launches of (n, precision) use now.
int tune_hetrd(int n, bool fp32, long batch_hint);
const char *hetrd_choice_name(int n, bool fp32);

// ---------------------------------------------------------------- 512 < n <= 4096
// Doubles of global workspace per workgroup of the divide & conquer with its node
// arrays in global memory: Qa, Qb, S (n*n each), 13 double and 16 int arrays of n,
// 8 doubles of reduction scratch (+ 8 of padding).
__host__ __device__ inline long stedc_global_work_doubles(int n) { return 3L * n * n + 13L * n + 8 + 8L * n + 8; }
int launch_stedc_global(int n, const double *D, const double *E, long strideV, double *W, long strideW, double *U,
                        long strideU, double *work, int batch, int grid, hipStream_t stream);
int launch_stedc_global(int n, const double *D, const double *E, long strideV, double *W, long strideW, float *U,
                        long strideU, double *work, int batch, int grid, hipStream_t stream);
// zlarft factors of every 16-reflector block (tws layout of backtransform_tws_complex),
// wpm workgroups per matrix.
int launch_larft(const double *A, long strideA, int lda, int n, const double *TAU, long strideV, double *tws,
                 long strideT, int batch, int wpm, hipStream_t stream);
int launch_larft(const float *A, long strideA, int lda, int n, const float *TAU, long strideV, float *tws,
                 long strideT, int batch, int wpm, hipStream_t stream);

// Whole eigensolver for 512 < n <= heevd_large_max_n() (eig_large.hip): A packed
// (lda = n, stride n*n, full Hermitian storage, overwritten by the reflectors);
// d | e | tau slots as launch_zhetrd (strideV doubles per matrix); aux: batch *
// heevd_large_aux_complex(n) complex elements of the precision of A; work: grid_dc *
// stedc_global_work_doubles(n) doubles; tws: batch * backtransform_tws_complex(n)
// complex. W: eigenvalues (stride n), U: eigenvectors component-major (stride n*n).
// Streaming matrix-core back-transform U = Q U (lda = n, reflectors and T factors as
// larft_kernel leaves them; any n >= 2), a slab of eigenvector columns per workgroup.
int launch_backtransform_stream(const double *A, long strideA, int n, const double *tws, long strideT, double *U,
                                long strideU, int batch, hipStream_t s);
int launch_backtransform_stream(const float *A, long strideA, int n, const float *tws, long strideT, float *U,
                                long strideU, int batch, hipStream_t s);
int heevd_large_max_n();
// Launch timing of lg_symv_lower_kernel (measurement): enable = 1 starts recording events
// around its eager launches; enable = 0 stops, waits, and returns the summed time and count.
int symv_timing(int enable, double *total_ms, long *launches);
long heevd_large_aux_complex(int n);
int heevd_large(double *A, int n, long batch, double *D, double *E, double *TAU, long strideV, double *aux, double *W,
                double *U, double *work, int grid_dc, double *tws, hipStream_t s, hipEvent_t mid = nullptr,
                hipEvent_t mid2 = nullptr);
int heevd_large(float *A, int n, long batch, double *D, double *E, float *TAU, long strideV, float *aux, double *W,
                float *U, double *work, int grid_dc, float *tws, hipStream_t s, hipEvent_t mid = nullptr,
                hipEvent_t mid2 = nullptr);

}  // namespace rlm
