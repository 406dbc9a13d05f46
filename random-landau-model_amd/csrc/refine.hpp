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
}  // namespace rlm
