// Host launchers of the rlm model kernels (rlm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rlm {
int launch_potential(const double *X, const double *Y, const double *S, long n_imp, const double *cmn, int d, int off,
                     double *V, int n_samples, hipStream_t stream);
int launch_etable(const double *cmt, const double *cjm, int n_phi, int d, int n_mesh, double *E, hipStream_t stream);
int launch_wtable(const double *V, const double *cmt, int n_phi, int d, int off, int n_mesh, int n_samples, double *Wt,
                  hipStream_t stream);
int launch_assemble(const double *E, const double *Wt, int n_phi, int d, int off, int n_mesh, long kp0, long nkp,
                    double *A, long strideA, hipStream_t stream, float *A32 = nullptr);
// FP64 Rayleigh quotients of FP32 eigenvectors against the FP64 matrices (circular
// band |c - r| <= bw). resid (optional): nkp * rayleigh_resid_slots(n) values.
int launch_rayleigh(const double *A, const float *U, int n, int bw, long nkp, double *W, double *resid,
                    hipStream_t stream);
int rayleigh_resid_slots(int n);
}  // namespace rlm
