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
// Band form of the projected Hamiltonians (see rlm_kernels.hip): nb = band_width
// columns per row; Wb: S*(g+1) x d x nb; H: nkp x N x nb (complex).
int band_width(int N, int off);
int launch_wband(const double *V, const double *cmt, int n_phi, int d, int off, int n_mesh, int n_samples, double *Wb,
                 hipStream_t stream);
// pl: points per line of the k-point mesh enumerated (g: solved points, g+1: all)
int launch_hband(const double *E, const double *Wb, int n_phi, int d, int off, int n_mesh, int pl, long kp0,
                 long nkp, double *H, hipStream_t stream);
int launch_expand_band(const double *H, int n_phi, int off, long nkp, double *A, long strideA, hipStream_t stream);
// FP64 Rayleigh quotients of FP32 eigenvectors against the FP64 matrices (circular
// band |c - r| <= bw), dense A or, with Hb, the band form of offset bw, into W.
// Without resid: from the upper half of the band (u^H A u = 2 Re u^H B u). With
// resid (nkp * rayleigh_resid_slots(n) values): from the full product, writing the
// residual norms ||A u - lambda u|| / ||u|| as well (W may then be null).
int launch_rayleigh(const double *A, const float *U, int n, int bw, long nkp, double *W, double *resid,
                    hipStream_t stream, const double *Hb = nullptr);
// the same for FP64 eigenvectors (the refined vectors of large mixed plans)
int launch_rayleigh(const double *A, const double *U, int n, int bw, long nkp, double *W, double *resid,
                    hipStream_t stream, const double *Hb = nullptr);
int rayleigh_resid_slots(int n);
}  // namespace rlm
