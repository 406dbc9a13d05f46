// Host launchers of the Chern / energy kernels (chern.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rlm {
int launch_links(const double *U, long strideU, int dim_vec, int ns, int lines, int L2n, int nlink2, long kp0,
                 long nkp, double *Lk1, double *Lk2, hipStream_t stream);
int launch_links_f32(const float *U, long strideU, int dim_vec, int ns, int lines, int L2n, int nlink2, long kp0,
                     long nkp, double *Lk1, double *Lk2, hipStream_t stream);
int launch_reduce(const double *Lk1, const double *Lk2, const double *Wv, int ns, int n_mesh, int L2n, int mode,
                  double n_phi_d, int n_samples, double *energy, double *chern, double *margin, hipStream_t stream);
}  // namespace rlm
