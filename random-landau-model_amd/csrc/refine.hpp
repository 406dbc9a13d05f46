// Host launcher of the FP64 eigenvector refinement (refine.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rlm {
// window entries per state (2 W + 1) and the largest Rayleigh-Ritz chunk
int refine_window();
int refine_max_cluster();
// Refines the FP32 eigenvectors U32 (component-major, eigenvalues ascending) of nkp
// matrices given in band form Hb (offset bw): the refined FP64 vectors are written to
// Y. Workspace: Xd (nkp*n*n complex), Sw, Gw (nkp*n*refine_window() complex), Qg
// (nkp*n*refine_max_cluster() complex), Cl (nkp*n ints), nflag (nkp ints: chunks cut
// at the size limit or without a Cholesky factor).
int launch_refine(const double *Hb, const float *U32, int n, int bw, long nkp, double *Xd, double *Y, double *Sw,
                  double *Gw, double *Qg, int *Cl, int *nflag, hipStream_t s);
// Exact FP64 Rayleigh-Ritz of the runs of near-degenerate neighbouring states (gaps
// below 1e-4 of the spectrum's extent; runs longer than 8 states split into balanced
// chunks of at most 8) after the Rayleigh quotients: eigenvalues W updated, FP32 vectors
// U32 rotated. Matrices from the band form Hb (offset bw) or, with Hb null, dense
// column-major A. list: pair_list_capacity(n, nkp) x 3 ints; count: kPairCounters ints
// (chunks found (those beyond the list capacity are not solved), runs split, chunks
// without a Cholesky factor of their Gram matrix).
constexpr int kPairCounters = 3;
int pair_list_capacity(int n, long nkp);
int launch_pair_rr(const double *A, const double *Hb, float *U32, int n, int bw, long nkp, double *W, int *list,
                   int cap, int *count, hipStream_t s);
}  // namespace rlm
