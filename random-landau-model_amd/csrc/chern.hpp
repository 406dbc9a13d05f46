// Host launchers of the Chern / energy kernels (chern.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace rlm {
int launch_links(const double *U, long strideU, int dim_vec, int ns, int lines, int L2n, int nlink2, long kp0,
                 long nkp, double *Lk1, double *Lk2, hipStream_t stream);
int launch_links_f32(const float *U, long strideU, int dim_vec, int ns, int lines, int L2n, int nlink2, long kp0,
                     long nkp, double *Lk1, double *Lk2, hipStream_t stream);
// rlm links over the (g+1)^2 mesh from eigenvectors of the solved points only
// (theta_1, theta_2 < g; see links_derived_kernel); U complex double or float (fp32).
int launch_links_derived(const void *U, bool fp32, long strideU, int ns, int g, int n_samples, double *Lk1,
                         double *Lk2, hipStream_t stream);
// mode 0 = rlm, 1 = hlm, 2 = rlm with eigenvalues of the solved points only
int launch_reduce(const double *Lk1, const double *Lk2, const double *Wv, int ns, int n_mesh, int L2n, int mode,
                  double n_phi_d, int n_samples, double *energy, double *chern, double *margin, hipStream_t stream);
// One theta_1 line of the reference's 2-slot ring buffer (device copy, complex double,
// state-major): chern_theta[m + n*n_mesh] (io_order: cal_Chern_wfs_IO's factor order).
int launch_ring_plaquettes(const double *ring, int n_phi, int dim_vec, int n_mesh, int theta_1, int io_order,
                           double *chern_theta, hipStream_t stream);
// Line t1 of sample s of a plan's eigenvectors in the ring layout (g + 1 points,
// state-major), complex double or float in and out; pl = g (derived edges) or g + 1.
int launch_gather_line(const void *U, bool fp32_in, long strideU, int ns, int g, int pl, long s, int t1, void *out,
                       bool fp32_out, hipStream_t stream);
}  // namespace rlm
