// Host launchers of the hlm kernels (hlm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rlm {
int launch_bloch_build(int lx, int ly, int Q, int n_mesh, long nmat, double *H, hipStream_t s);
int launch_phi_build(int lx, int ly, int Q, int band, int n_mesh, long nmat, const double *Ub, double *Phi,
                     hipStream_t s);
int launch_zgemm(int M, int N, int K, const double *A, long rsA, long csA, long sA1, long sA2, int conjA,
                 const double *B, long rsB, long csB, long sB1, long sB2, const double *w, long sw1, long sw2,
                 double *C, long rsC, long csC, long sC1, long sC2, int nb1, int nb2, hipStream_t s);
int launch_hlm_links(const double *U, int ns, int n_mesh, long nkp, const double *M1, const double *M2, double *Lk1,
                     double *Lk2, hipStream_t s);
int launch_hlm_project(const double *Phi, const double *V, int N, int ns, int S, long ntw, double *H, hipStream_t s);
}  // namespace rlm
