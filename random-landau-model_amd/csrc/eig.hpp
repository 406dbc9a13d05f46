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

}  // namespace rlm
