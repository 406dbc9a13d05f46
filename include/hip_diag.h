/*
 * hip_diag.h -- C ABI of the MI355X-native rlm/hlm solver (librlm_hip.so).
 *
 * Drop-in boundary for qiongz/Random-Landau-Model's diag_wrappers plugin layer
 * and the rlm/hlm hot path. Every entry point takes plain pointers and sizes.
 * Complex arrays are interleaved (re, im) doubles, i.e. std::complex<double> /
 * hipDoubleComplex layout. "d_" pointers are device (HBM) memory; the others
 * are host memory.
 *
 * Status: 0 = OK; > 0 = solver info (as LAPACK's info: 1 + the index of the
 * first matrix, or sample, whose input or result is not finite); < 0 = error
 * (-22 invalid argument, -12 out of memory, -2 / -5 file missing / I/O error,
 * other negatives = -hipError_t).
 * Every entry point is synchronous unless it names a stream.
 */
#ifndef HIP_DIAG_H
#define HIP_DIAG_H

#ifdef __cplusplus
extern "C" {
#endif

/* Precision modes of the batched solver.
 * HIP_DIAG_FP64: every stage in FP64.
 * HIP_DIAG_MIXED (rlm plans): FP32 tridiagonalisation and back-transform with FP32
 * accumulation, FP64 divide & conquer, FP64 Rayleigh quotients of the FP32
 * eigenvectors against the FP64 matrix (the reported energies), and link overlaps
 * of the FP32 vectors accumulated in FP64 (each product of two floats is exact in
 * FP64). */
#define HIP_DIAG_FP64 0
#define HIP_DIAG_MIXED 1

/* Flags of rlm_plan_create_ex.
 * HIP_DIAG_CONV_GPU_FP32: the conventions of the reference's single-precision GPU
 * builds (rlm_gpu_magma / rlm_gpu_cusolver, which write energy_Chern_No.dat):
 * float L1, L2 and impurity count (rlm/main_rlm_gpu.cu:29-33), float tables
 * (matrix_coefficients.cpp with dtype float), float impurity positions
 * (vsRngUniform: L * float(word) * 2^-32 in float), V_mn = coeff_mn sum_i s_i
 * e^{+i phase} (rlm/potential_coeff.cu:14-15) and h_jk = sum_m coeff_jm[j][m]
 * c2[m] sum_n V[m][n] conj(c1[n]) (rlm/hamiltonian_rlm_gpu.cu:23,36), every twist
 * point solved. The arithmetic after those float inputs is this library's (FP64,
 * or the mixed solver): results agree with the FP32 builds to their FP32
 * accuracy and with an FP64 evaluation of their formulas to 1e-10. */
#define HIP_DIAG_CONV_GPU_FP32 1

/* ------------------------------------------------------------------------ */
/* diag_wrappers-level entry points                                          */
/* ------------------------------------------------------------------------ */

/* Replaces cusolver_diag_hamil (diag_wrappers/cusolver_diag.h:6,
 * diag_wrappers/cusolver_diag.cu:3-22) and magma_diag_hamil
 * (diag_wrappers/magma_diag.h:9): Hermitian eigen-decomposition in place of a
 * device n x n complex128 column-major matrix (lda = n), lower triangle read,
 * eigenvalues ascending into device d_w, eigenvectors overwrite d_A (column j
 * = state j). FP64 (the reference wrappers are FP32). */
int hip_diag_hamil(double *d_A, double *d_w, int n);

/* The same with the reference's void signature, for call sites that expect it
 * (rlm/main_rlm_gpu.cu:176,216). Errors are printed to stderr. */
void hip_diag_hamil_v(double *d_A, double *d_w, int n);

/* No counterpart in the reference: hip_diag_hamil for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same contract as hip_diag_hamil (FP64, lower
 * triangle read, in place, synchronous) for n <= 256; n > 256 returns -22. Every
 * intermediate lives in one 4 MiB device array of the library (per GPU, held
 * until the process ends, located with hipGetSymbolAddress at the first call on
 * each device); no hipMalloc, hipFree or allocating library call. Calls are
 * serialised within the process and share the array, which keeps the last call's
 * contents; calls from several processes use their own copies. */
int hip_diag_hamil_noalloc(double *d_A, double *d_w, int n);

/* Replaces mkl_heevd (diag_wrappers/mkl_diag.h:13, mkl_diag.cpp:2-9):
 * LAPACKE_zheevd(COL_MAJOR, 'V', 'U', n, A, n, w) on host memory, computed on
 * the GPU. */
int hip_heevd(double *A, double *w, int n);

/* Batched form of the two above (no reference counterpart: the reference calls
 * its wrappers once per k-point, rlm/main_rlm_gpu.cu:198-229). `batch`
 * matrices at d_A + 2*b*strideA (lda >= n), uplo 'L' or 'U' names the triangle
 * read; eigenvectors overwrite the matrices, eigenvalues go to
 * d_W + b*strideW. stream: a hipStream_t or NULL for the default stream.
 * precision HIP_DIAG_FP64: FP64 throughout. HIP_DIAG_MIXED: FP32 Householder
 * reduction and back-transform, FP64 divide & conquer, eigenvalues = FP64
 * Rayleigh quotients of the FP32 eigenvectors against the FP64 matrix
 * (eigenvectors carry FP32 accuracy, ~1e-6 relative residual).
 * A NaN or infinity in the triangle read returns 1 + the index of the first such
 * matrix before any solver kernel runs (the matrices are left as packed copies
 * never written back; d_A and d_W are untouched). */
int hip_diag_heevd_batched(int n, double *d_A, long lda, long strideA, double *d_W, long strideW, int batch,
                           char uplo, int precision, void *stream);

/* Replaces cal_Chern (diag_wrappers/chern.h:29, chern.cpp:3-35): for one
 * theta_1 line, chern_theta[m + n*n_mesh] = -arg(prod of the 4 link overlaps
 * of plaquette m for state n)/(2 pi), reading the 2-slot host ring buffer
 * wfs_full[(slot*(n_mesh+1) + theta_2)*n_phi*dim_vec + state*dim_vec + i]
 * (complex128; slot theta_1 % 2 holds line theta_1, the other slot line
 * theta_1 - 1). The ring is copied to the device once; overlaps, products and
 * phases run there (one wave per plaquette and state). Host chern_theta. */
int hip_cal_chern(const double *wfs_full, double *chern_theta, int n_phi, int dim_vec, int n_mesh, int theta_1);

/* Replaces write_wfs (diag_wrappers/wfs_file.h:11, wfs_file.cpp:2-12) of the
 * -DwfsIO builds: writes `size` complex values of host wfs (complex128) raw, no
 * header, to <dir>/wfs_<theta_1>_<theta_2>.bin (dir NULL or "": the working
 * directory, as the reference), as complex128 (elem_bytes 16) or complex64
 * (elem_bytes 8, the reference's FP32 builds). Callers pass theta_1 % 2, as
 * rlm/main_rlm_gpu.cu:178,215 do. Returns -5 on an I/O error. */
int hip_write_wfs(const char *dir, int theta_1, int theta_2, const double *wfs, unsigned long size, int elem_bytes);

/* Replaces cal_Chern_wfs_IO (diag_wrappers/chern.h:30, chern.cpp:77-123): the
 * theta_1 line's chern_theta[m + n*n_mesh] from the files
 * <dir>/wfs_<slot>_<theta_2>.bin of both ring slots (theta_2 = 0..n_mesh, each
 * n_phi state vectors of dim_vec components, complex128 or complex64 by
 * elem_bytes), with the reference's factor order <p_{l+1}|p_l> and sign
 * (+arg/2 pi), on the device. Returns -2 if a file is missing, -5 if short. */
int hip_cal_chern_wfs_io(double *chern_theta, int n_phi, int dim_vec, int n_mesh, int theta_1, const char *dir,
                         int elem_bytes);

/* ------------------------------------------------------------------------ */
/* rlm model entry points                                                    */
/* ------------------------------------------------------------------------ */

/* Replaces prepare_coeff (rlm/matrix_coefficients.h:139-140,
 * matrix_coefficients.cpp:3-25), double-precision build: host tables
 * coeff_mn[dim_m*dim_n], coeff_m_theta[(n_mesh+1)*dim_m], coeff_jm[n_phi*dim_m]
 * (complex), including the single-precision expf and 1.0f/n_phi of :11. */
int rlm_prepare_coeff(double *coeff_mn, double *coeff_m_theta, double *coeff_jm, int n_phi, int off_head, int dim_m,
                      int dim_n, int n_mesh, double L1, double L2);

/* Replaces generate_disorder_potential (rlm/disorder_potential.h:9,
 * disorder_potential.cpp:3-21): MKL-compatible MT19937 stream of `seed`
 * (init_by_array), skip 3*impurity_num*sample_offset words, x and y uniform
 * on [0,L1) and [0,L2), intensity +1 for an even word and -1 for an odd one. */
int rlm_generate_disorder_potential(double *x, double *y, double *intensity, long impurity_num, double L1, double L2,
                                    unsigned long seed, long sample_offset);

/* The same stream chain generated on the device for samples sample_offset ..
 * sample_offset + n_samples - 1 (sample k's values at d_x + k*impurity_num,
 * d_y + k*impurity_num, d_intensity + k*impurity_num): one workgroup per sample
 * seeds the MT19937 window, jumps it by 3*impurity_num*sample words (F2
 * jump-ahead polynomial t^J mod phi, cost O(log J) on the host and O(19937)
 * on the device, independent of the sample index) and generates its words.
 * Bit-identical to rlm_generate_disorder_potential. Synchronous. */
int hip_generate_disorder_potential(double *d_x, double *d_y, double *d_intensity, long impurity_num, double L1,
                                    double L2, unsigned long seed, long sample_offset, int n_samples);

/* Replaces prepare_potential_coeff (rlm/potential_coeff.h,
 * potential_coeff.cu:21-26) in FP64 with the CPU build's sign convention
 * (get_potential_coeff, matrix_coefficients.cpp:27-44):
 * d_Vmn[m*dim_n+n] = coeff_mn * sum_i s_i e^{-i 2pi((m-o)x_i + (n-o)y_i)}. */
int hip_prepare_potential_coeff(const double *d_x, const double *d_y, const double *d_intensity, double *d_Vmn,
                                const double *d_coeff_mn, int dim_m, int dim_n, int off_head, long impurity_num);

/* Replaces set_hamil_matrix (rlm/hamiltonian_rlm_gpu.h,
 * hamiltonian_rlm_gpu.cu:43-50) in FP64 with the CPU build's conventions
 * (hamiltonian_rlm_cpu.cpp:14-48): d_wfs receives the full Hermitian n_phi x
 * n_phi matrix, column-major, exactly the matrix LAPACK sees in the CPU build
 * (A(r,c) = conj(h_rc) for c >= r). Rows of coeff_m_theta are the theta_1 and
 * theta_2 rows; d_coeff_jm is the n_phi x dim_n table. Swapped into
 * rlm/main_rlm_gpu.cu as is, it builds the CPU build's matrix, not the GPU kernel's
 * (conj(c1[n]) c2[m] with e^{+i phase} potentials, hamiltonian_rlm_gpu.cu:29-34,43 and
 * potential_coeff.cu:14-15); conjugated theta rows and negated impurity positions give
 * the GPU build's matrix exactly (the equivalence HIP_DIAG_CONV_GPU_FP32 uses). */
int hip_set_hamil_matrix(const double *d_coeff_m_theta_1, const double *d_coeff_m_theta_2, const double *d_coeff_jm,
                         const double *d_Vmn, double *d_wfs, int n_phi, int dim_n, int off_head);

/* ------------------------------------------------------------------------ */
/* Batched pipelines                                                         */
/* ------------------------------------------------------------------------ */

/* Replaces the sample / theta_1 / theta_2 loops of rlm/main_rlm_cpu.cpp:109-243
 * (and rlm/main_rlm_gpu.cu:138-300): for each disorder sample, assemble,
 * diagonalise and Chern-reduce the whole (n_mesh+1)^2 twist mesh on the GPU
 * (n_mesh^2 points diagonalised, the theta = 2 pi edges derived from them).
 * off_override < 0 keeps the reference's off_head = (long)(sqrt(N_imp)/2+1). */
typedef struct rlm_plan_s *rlm_plan_t;
int rlm_plan_create(rlm_plan_t *plan, long n_phi, double quanta_concentration, double impurity_concentration,
                    long n_mesh, long off_override, unsigned long seed, int samples_per_batch, int precision);
/* rlm_plan_create with flags (0 or HIP_DIAG_CONV_GPU_FP32). */
int rlm_plan_create_ex(rlm_plan_t *plan, long n_phi, double quanta_concentration, double impurity_concentration,
                       long n_mesh, long off_override, unsigned long seed, int samples_per_batch, int precision,
                       int flags);
/* Selects samples first..first+count-1 (count <= samples_per_batch) for the
 * next execute. Their disorder is generated on the device by rlm_plan_execute
 * from the seed (the jump to the first sample is computed here on the host,
 * O(log first) polynomial work that does not depend on the seed). */
int rlm_plan_load(rlm_plan_t plan, long first_sample, int count);
/* Runs every GPU stage for the loaded samples on the plan's stream: disorder
 * streams, V_mn, assembly, eigensolve, links, plaquettes and energies. */
int rlm_plan_execute(rlm_plan_t plan);
/* Copies energies / Chern values / plaquette margins ([count][n_phi], host)
 * after synchronising the stream. margin may be NULL. Returns 1 + the batch
 * sample holding a non-finite energy or Chern value (every value is copied
 * anyway), 0 when all are finite; rlm_plan_run returns 1 + the run's sample. */
int rlm_plan_fetch(rlm_plan_t plan, double *energy, double *chern, double *margin);
/* load + execute + fetch over any number of samples, batch by batch. */
int rlm_plan_run(rlm_plan_t plan, long first_sample, long n_samples, double *energy, double *chern, double *margin);
/* GPU time of each stage of the last rlm_plan_execute, in ms, from hipEvents
 * recorded on the plan's stream: [0] V_mn + band tables, [1] assembly,
 * [2] tridiagonalisation, [3] divide & conquer, [4] back-transform, [5] FP64
 * refinement (mixed plans; 0 otherwise), [6] link overlaps, [7] plaquettes +
 * energies. Returns 8 (the number of stages), or 0 after an execute in graph mode
 * (whose events carry no stage timing). */
int rlm_plan_stage_times(rlm_plan_t plan, double *ms);
/* GPU time of the whole last rlm_plan_execute in ms (eager and graph mode). */
int rlm_plan_elapsed_ms(rlm_plan_t plan, double *ms);
/* The plan's hipStream_t. */
void *rlm_plan_stream(rlm_plan_t plan);
/* Samples of the current batch (the count of the last rlm_plan_load, also inside
 * rlm_plan_run): rlm_plan_fetch writes loaded * n_phi values per output. */
int rlm_plan_loaded(rlm_plan_t plan);
/* Sizes derived as rlm/main_rlm_cpu.cpp:29-34: out[0]=n_phi, [1]=impurity_num,
 * [2]=off_head, [3]=dim_m, [4]=n_mesh, [5]=samples_per_batch. */
int rlm_plan_info(rlm_plan_t plan, long *out6);
/* Mixed precision: the largest ||A u - lambda u|| / ||u|| (FP64, against the FP64
 * matrix) over every state and twist point of the last execute, computed by this
 * call (a full FP64 product on the plan's stream); 0 for FP64 plans. */
int rlm_plan_max_residual(rlm_plan_t plan, double *out);
/* Device pointer of the eigenvalues [kp][n_phi] / eigenvectors
 * [kp][component][state] of the last execute, for the solved twist points
 * kp = (s*g+t1)*g+t2, t1, t2 < g: the theta = 2 pi points are not solved, their
 * matrices are unitarily equivalent to the theta = 0 ones (same eigenvalues;
 * u(t1, g)[i] = u(t1, 0)[(i+1) % n_phi] and u(g, t2)[i] = e^{i 2 pi i/n_phi}
 * u(0, t2)[i] up to a phase per state). With RLM_MESH=full in the environment at
 * rlm_plan_create, every point is solved and kp = (s*(g+1)+t1)*(g+1)+t2.
 * Eigenvectors are complex128 (FP64 plans and refined mixed plans) or complex64
 * (mixed plans without refinement). */
const double *rlm_plan_eigenvalues(rlm_plan_t plan);
const double *rlm_plan_eigenvectors(rlm_plan_t plan);
/* -DwfsIO dump of the plan (rlm/main_rlm_gpu.cu:177-179,214-216 with
 * diag_wrappers/wfs_file.cpp:2-12): after rlm_plan_execute, writes the
 * eigenvectors of line theta_1 (0..n_mesh) of batch sample `sample` (0..loaded-1)
 * as <dir>/wfs_<theta_1 % 2>_<theta_2>.bin for theta_2 = 0..n_mesh, each
 * n_phi x n_phi complex values state-major (state n's components contiguous, the
 * column-major eigenvector matrix the reference's solvers leave), raw, as
 * complex128 (elem_bytes 16) or complex64 (elem_bytes 8). The theta = 2 pi points
 * the plan derives are written from the exact relations rlm_plan_eigenvectors
 * states. Writing lines theta_1 - 1 and theta_1 and calling hip_cal_chern_wfs_io
 * reproduces the plan's plaquettes of that line. */
int rlm_plan_write_wfs(rlm_plan_t plan, int sample, int theta_1, const char *dir, int elem_bytes);
/* Mixed plans refine their FP32 eigenvectors in FP64 (one windowed Ogita-Aishima
 * step after exact Rayleigh-Ritz rotations of near-degenerate chunks) when n_phi > 512,
 * or as RLM_REFINE=0 / 1 in the environment at rlm_plan_create says. out2[0] = 1 if
 * the plan refines, out2[1] = chunks of the last execute that were cut at the chunk
 * size limit or had no Cholesky factor (0 when every coupled run was solved exactly).
 * No counterpart in the reference (whose FP32 builds have no refinement). */
int rlm_plan_refine_info(rlm_plan_t plan, long *out2);
/* Mixed plans up to n_phi = 512 without the refinement give every run of
 * near-degenerate neighbouring states (gaps below 1e-4 of the spectrum's extent)
 * an exact FP64 Rayleigh-Ritz; runs longer than 8 states are split into balanced
 * chunks of at most 8. For the last execute: out4[0] = chunks solved, [1] = runs
 * split, [2] = chunks not solved (beyond the list capacity), [3] = chunks left
 * unrotated (Gram matrix without a Cholesky factor). Nonzero [2] or [3] means some
 * energies may carry the FP32 vectors' coupling error; all zero for FP64 plans. No
 * counterpart in the reference. */
int rlm_plan_pair_info(rlm_plan_t plan, long *out4);
int rlm_plan_destroy(rlm_plan_t plan);

/* Graph mode of an rlm plan (no reference counterpart; the reference launches
 * every kernel from the host, rlm/main_rlm_gpu.cu:198-229): with enable != 0 the
 * launches of rlm_plan_execute are captured once into a hipGraph (again when the
 * loaded count or the first sample's jump changes) and replayed with one
 * hipGraphLaunch; rlm_plan_elapsed_ms times it. RLM_GRAPH=1 in the environment at
 * rlm_plan_create enables it. */
int rlm_plan_set_graph(rlm_plan_t plan, int enable);

/* Several rlm sizes in one launch (finite-size scaling batches of different n_phi;
 * no reference counterpart): one plan per n_phi[k] (same q, impurity
 * concentration, mesh, seed, samples per batch and precision), executed as ONE
 * hipGraph launch whose branches, one per size, run concurrently on the GPU.
 * rlm_multi_fetch(k) returns size k's results as rlm_plan_fetch does;
 * rlm_multi_elapsed_ms the GPU time of the last launch. */
typedef struct rlm_multi_s *rlm_multi_t;
int rlm_multi_create(rlm_multi_t *multi, int nsizes, const long *n_phi, double quanta_concentration,
                     double impurity_concentration, long n_mesh, unsigned long seed, int samples_per_batch,
                     int precision);
int rlm_multi_load(rlm_multi_t multi, long first_sample, int count);
int rlm_multi_execute(rlm_multi_t multi);
int rlm_multi_fetch(rlm_multi_t multi, int size_index, double *energy, double *chern, double *margin);
int rlm_multi_elapsed_ms(rlm_multi_t multi, double *ms);
int rlm_multi_destroy(rlm_multi_t multi);

/* Replaces the hlm sample / theta loops of hlm/main_hlm_cpu.cpp:75-196 (with
 * hlm/hamiltonian_hlm_cpu.cpp:2-56,149-184): lx x ly lattice, flux 1/Q, the
 * Hamiltonian projected on Bloch band `band` (n_phi = lx*ly/Q states), twist
 * mesh theta_1 = 0..g, theta_2 = 0..g-1 (theta_2 = g is theta_2 = 0). The
 * on-site disorder of sample s is the reference's single MT19937 stream,
 * V = (W/2) * U[-1,1) (W = 2 is the reference, hlm/main_hlm_cpu.cpp:79).
 * Requires lx % Q == 0 (the reference's k-space basis assumes it). */
typedef struct hlm_plan_s *hlm_plan_t;
int hlm_plan_create(hlm_plan_t *plan, int lx, int ly, int Q, int band, int n_mesh, unsigned long seed, double W,
                    int samples_per_batch, int precision);
int hlm_plan_load(hlm_plan_t plan, long first_sample, int count);
int hlm_plan_execute(hlm_plan_t plan);
int hlm_plan_fetch(hlm_plan_t plan, double *energy, double *chern, double *margin);
int hlm_plan_run(hlm_plan_t plan, long first_sample, long n_samples, double *energy, double *chern, double *margin);
/* [0] projection Phi^H diag(V) Phi, [1] tridiagonalisation, [2] divide & conquer,
 * [3] back-transform, [4] link overlaps, [5] plaquettes + energies, [6] unused.
 * Returns 7. */
int hlm_plan_stage_times(hlm_plan_t plan, double *ms);
void *hlm_plan_stream(hlm_plan_t plan);
/* Samples of the current batch, as rlm_plan_loaded. */
int hlm_plan_loaded(hlm_plan_t plan);
int hlm_plan_destroy(hlm_plan_t plan);

/* The tridiagonalisation kernel the library's autotuner chose for matrices of size
 * n (1..512) in precision HIP_DIAG_FP64 / HIP_DIAG_MIXED (timed at the first plan or
 * batched call of that size and kept per process and in the tune file
 * RLM_TUNE_CACHE, default $HOME/.cache/rlm_hip_tune.txt): "resident", "lower",
 * "lower8", "hybrid", "fused", "blocked", "small", or "auto" (not tuned: fixed
 * defaults). RLM_TUNE=0 disables tuning, RLM_HETRD=<name> forces a kernel. */
const char *hip_diag_tuned_kernel(int n, int precision);
/* Measurement aid, no counterpart in the reference: enable = 1 starts recording HIP
 * events around every eager launch of the n > 512 solver's Hermitian product kernel
 * (lg_symv_lower_kernel, one launch per column; launches under graph capture are not
 * timed); enable = 0 stops, waits for the recorded launches and returns their summed
 * time in ms and their count. */
int hip_diag_symv_timing(int enable, double *total_ms, long *launches);
/* hlm/main_hlm_cpu.cpp:66-67,79: samples first..first+n_samples-1 of the run's
 * single stream, n_sites values each, scaled to [-W/2, W/2). Host memory. */
int hlm_generate_disorder(double *V, long n_sites, long n_samples, unsigned long seed, long first_sample, double W);
/* The same values generated on the device into d_V (n_samples x n_sites), one
 * workgroup per sample jumping to word n_sites*(first_sample + k) of the stream
 * (as hip_generate_disorder_potential). hlm plans generate their samples this
 * way at every hlm_plan_execute. Synchronous. */
int hip_hlm_generate_disorder(double *d_V, long n_sites, long n_samples, unsigned long seed, long first_sample,
                              double W);

#ifdef __cplusplus
}
#endif
#endif /* HIP_DIAG_H */
