/*
 * binder_c.h -- C ABI of the MI355X (gfx950) backend for BINDER's
 * registration-and-sampling path.
 *
 * Each entry point replaces one method of the reference's pybind11 module
 * `BINDER._core` (GPU bindings: BINDER/src/gpu/bindings.cpp, CPU bindings:
 * BINDER/src/cpu/bindings.cpp); the citation next to each declaration names
 * the binding it stands in for.  A ctypes shim (binder_amd/_core.py) turns
 * these calls back into the reference's Python classes with unchanged names
 * and keyword arguments.
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero code on failure:
 *      BINDER_EINVAL  (1)  bad argument (the shim raises ValueError)
 *      BINDER_EDEVICE (2)  HIP runtime/launch error (RuntimeError)
 *      BINDER_ESTATE  (3)  call out of order, e.g. EM before setMILikelihood
 *    binder_last_error() returns a thread-local message for the last failure.
 *    No C++ exception crosses this boundary.
 *  - Arrays are host pointers in C order with the reference layout:
 *    fields are [Nx, Ny, Nz, N] with interleaved components (N = 2 when
 *    Nz == 1, else 3); images/bins are [Nx, Ny, Nz].  Inputs are copied to
 *    device memory at set time; outputs are written to caller buffers.
 *  - A handle owns its device memory and one HIP stream (default priority).
 *    Handles are not thread-safe; distinct handles may be used concurrently.
 *  - `threads` arguments are accepted for signature parity and ignored.
 */
#ifndef BINDER_C_H
#define BINDER_C_H

#ifdef __cplusplus
extern "C" {
#endif

#define BINDER_OK 0
#define BINDER_EINVAL 1
#define BINDER_EDEVICE 2
#define BINDER_ESTATE 3

/* Message of the last failing call on this thread ("" if none). */
const char *binder_last_error(void);

/* Backend name, "hip" (reference attribute `backend`,
 * BINDER/src/gpu/bindings.cpp:16). */
const char *binder_backend(void);

/* Philox4x32-10 block function (host-side, no device work): the generator
 * behind every random draw of the sampler, exported so its known-answer
 * vectors can be checked.  No counterpart in the reference (which uses
 * std::mt19937 / cuRAND). */
int binder_philox4x32_10(const unsigned int *ctr4, const unsigned int *key2,
                         unsigned int *out4);

/* ------------------------------------------------------------------ */
/* Registration (BINDER/src/gpu/bindings.cpp:17-157)                   */
/* ------------------------------------------------------------------ */
typedef struct binder_registration binder_registration;

/* Registration(final_locations, threads, voxel_pos, node_pos, Nx_v, Ny_v,
 * Nz_v, Nx_n, Ny_n, Nz_n, spline_order, non_zero_voxels)
 * -- BINDER/src/gpu/bindings.cpp:18-47, BINDER/src/cpu/Registration.cpp:25-70.
 * final_locations and voxel_pos are [Nx_v,Ny_v,Nz_v,N]; node_pos is unused
 * after construction in the reference and may be NULL here.
 * spline_order is 1, 2 or 3 (linear, quadratic or cubic B-spline node
 * weights, BINDER/src/cpu/Spline.cpp:8-110; others -> BINDER_EINVAL). */
int binder_registration_create(binder_registration **out,
                               const float *final_locations, int threads,
                               const float *voxel_pos, const float *node_pos,
                               int Nx_v, int Ny_v, int Nz_v,
                               int Nx_n, int Ny_n, int Nz_n,
                               int spline_order, double non_zero_voxels);

/* A new Registration on a new grid whose current field is `src`'s current
 * field resampled on the device: Gaussian anti-aliasing where an axis
 * shrinks, linear zoom, clip to the source range, times `scale` -- the
 * resize that BINDER/FastRegistration.py:566-568 applies between pyramid
 * levels followed by the per-level constructor (:443-450), without the
 * field's round trip through host memory.  No reference counterpart; other
 * arguments as binder_registration_create, except that voxel_pos may be NULL
 * for voxel positions equal to the grid indices (i, j[, k]), which are then
 * generated on the device.  `src` stays valid. */
int binder_registration_create_resampled(binder_registration **out,
                                         binder_registration *src, double scale,
                                         const float *voxel_pos, int Nx_v, int Ny_v,
                                         int Nz_v, int Nx_n, int Ny_n, int Nz_n,
                                         int spline_order, double non_zero_voxels);
void binder_registration_destroy(binder_registration *reg);

/* setMILikelihood(alpha, theta, K, L, threads, binned_nodes, binned_voxels)
 * -- BINDER/src/gpu/bindings.cpp:49-70, BINDER/src/cpu/Registration.cpp:455-464.
 * theta is K x L row-major (k = moving bin, l = fixed bin).  binned_nodes
 * are [Nx_n,Ny_n,Nz_n] in [0,K), binned_voxels [Nx_v,Ny_v,Nz_v] in [0,L).
 * K and L must be in [1,256] with K*L <= 4096 (the likelihood table and
 * the block histogram are held in LDS). */
int binder_registration_set_mi_likelihood(binder_registration *reg,
                                          double alpha, const double *theta,
                                          int K, int L, int threads,
                                          const int *binned_nodes,
                                          const int *binned_voxels);

/* setSSDLikelihood(bins, threads, sigma_sq, scale_shift, scale, shift,
 *                  binned_nodes, binned_voxels, alpha_0, beta_0)
 * -- BINDER/src/cpu/bindings.cpp:120-152, BINDER/src/cpu/Registration.cpp:472-479,
 * BINDER/src/cpu/SSDLikelihood.cpp.  Both images binned to [0, bins), bins in
 * [2, 64]; scale_shift 0 compares the bins directly, 1 fits shift and scale.
 * The E-step, EM loop and sampler run as with the MI likelihood;
 * get_likelihood_parameters then returns [scale, shift]. */
int binder_registration_set_ssd_likelihood(binder_registration *reg, int bins,
                                           int threads, double sigma_sq, int scale_shift,
                                           double scale, double shift,
                                           const int *binned_nodes,
                                           const int *binned_voxels, double alpha_0,
                                           double beta_0);
/* No counterpart in the reference (its SSD parameters are printed only):
 * (scale, shift, sigma_sq) of the SSD likelihood as they stand. */
int binder_registration_get_ssd_parameters(binder_registration *reg, double *out3);

/* setAffineTransformation(A, t) -- BINDER/src/cpu/bindings.cpp:223-235,
 * BINDER/src/cpu/Registration.cpp:507-510, AffineTransformation.cpp:25-130.
 * A is (N+1) x (N+1) row-major (homogeneous), t has N entries; both are the
 * starting values, which the least-squares fit of each EM iteration
 * replaces.  Replaces the DCT prior as the transformation model. */
int binder_registration_set_affine_transformation(binder_registration *reg,
                                                  const double *A, const double *t);
/* setRigidTransformation(R, t) -- BINDER/src/cpu/bindings.cpp:236-247,
 * BINDER/src/cpu/Registration.cpp:512-516, RigidTransformation.cpp:57-154.
 * R is N x N row-major, t has N entries. */
int binder_registration_set_rigid_transformation(binder_registration *reg,
                                                 const double *R, const double *t);
/* No counterpart in the reference (it prints A / R and t): the fitted
 * matrix (affine (N+1)^2, rigid N^2, row-major) followed by t; *count gets
 * the number of values written (at most 19). */
int binder_registration_get_linear_parameters(binder_registration *reg, double *out,
                                              int *count);

/* setNonLinearTransformation(gamma_x..z, voxels_size_x..z, nodes_size_x..z, D)
 * -- BINDER/src/gpu/bindings.cpp:72-95, BINDER/src/cpu/Registration.cpp:492-502.
 * D is the [Nx_v,Ny_v,Nz_v] DCT-domain eigenvalue array. */
int binder_registration_set_nonlinear_transformation(
    binder_registration *reg, double gamma_x, double gamma_y, double gamma_z,
    double voxels_size_x, double voxels_size_y, double voxels_size_z,
    double nodes_size_x, double nodes_size_y, double nodes_size_z,
    const float *D);

/* setNonLinearTransformation with the prior spectrum D of
 * binder_dct_prior_eigenvalues(Nx, Ny, Nz, N, order, voxel sizes) made on the
 * device from its separable terms (bit-identical float32 values, order 1 or
 * 2), so no D array crosses PCIe.  No reference counterpart: the reference
 * takes D from BINDER/utils.py:331-349 through setNonLinearTransformation
 * (BINDER/src/gpu/bindings.cpp:72-95). */
int binder_registration_set_nonlinear_transformation_dct_prior(
    binder_registration *reg, double gamma_x, double gamma_y, double gamma_z,
    double voxels_size_x, double voxels_size_y, double voxels_size_z,
    double nodes_size_x, double nodes_size_y, double nodes_size_z, int order);

/* EM(max_EM_iterations, convergence_th, update_likelihood_parameters, debug)
 * -- BINDER/src/gpu/bindings.cpp:99-103, BINDER/src/cpu/Registration.cpp:122-217.
 * Iterates E-step -> theta M-step -> deformation M-step until
 * (mlp_old - mlp)/non_zero_voxels < convergence_th or the cap is reached.
 * debug != 0 prints the per-iteration objective like the reference. */
int binder_registration_em(binder_registration *reg, int max_EM_iterations,
                           double convergence_th,
                           int update_likelihood_parameters, int debug);

/* sampler(seed, N_b, N_s, sample_gamma, sample_gamma_every,
 *         sample_posteriors, sample_likelihood_parameters,
 *         sample_transformation_parameters) -> (mean, cov)
 * -- BINDER/src/gpu/bindings.cpp:105-131, BINDER/src/cpu/Registration.cpp:515-642.
 * mean_out: [Nx,Ny,Nz,N] doubles; cov_out: [Nx,Ny,Nz,N,N] doubles
 * (= M2/(N_s-1)).  Random numbers come from stateless Philox4x32-10 streams
 * keyed by (seed, sweep, site); accumulators are FP64. */
int binder_registration_sampler(binder_registration *reg, int seed, int N_b,
                                int N_s, int sample_gamma,
                                int sample_gamma_every, int sample_posteriors,
                                int sample_likelihood_parameters,
                                int sample_transformation_parameters,
                                double *mean_out, double *cov_out);

/* The two halves of binder_registration_sampler: run_ executes the sweeps
 * and keeps mean/M2 on the device; result_ copies them out (mean, and
 * M2/(N_s-1) expanded to the full N x N matrix).  Splitting them lets a
 * caller time the sweeps with all state resident in HBM. */
int binder_registration_sampler_run(binder_registration *reg, int seed, int N_b,
                                    int N_s, int sample_gamma,
                                    int sample_gamma_every,
                                    int sample_posteriors,
                                    int sample_likelihood_parameters,
                                    int sample_transformation_parameters);
int binder_registration_sampler_result(binder_registration *reg,
                                       double *mean_out, double *cov_out);

/* Independent Gibbs chains from the handle's current state: chain c runs
 * N_b burn-in and N_s sample sweeps exactly as
 * binder_registration_sampler_run(seeds[c], N_b, N_s, 0, 1, sp, sl, st)
 * would (same draws), each from the same start field and theta; the FP64
 * Welford accumulators fold all chains' samples, chain after chain, read
 * with binder_registration_sampler_result (covariance = M2 / (nchains N_s
 * - 1)).  The handle's field and theta are restored afterwards.
 * use_graph != 0 captures one chain's sweeps into a HIP graph and replays
 * it for chains 1..nchains-1 (the Philox key and the Welford sample offset
 * are read from device memory, so the graph is chain-independent).  No nu
 * sampling (sample_gamma = 0).  No reference counterpart: the reference
 * runs one chain per sampler() call (BINDER/src/cpu/Registration.cpp:515-642). */
int binder_registration_sampler_chains(binder_registration *reg, int nchains,
                                       const int *seeds, int N_b, int N_s,
                                       int sample_posteriors,
                                       int sample_likelihood_parameters,
                                       int sample_transformation_parameters,
                                       int use_graph);

/* Several registrations iterated in lockstep (batched EM).  No reference
 * counterpart: the reference registers one pair per Registration object
 * (BINDER/FastRegistration.py:443-450, BINDER/src/gpu/Registration.cu:529-530).
 * A batch borrows npairs handles that share K and L (K <= 32) and
 * dimensionality, each with setMILikelihood and setNonLinearTransformation
 * done; the handles must outlive the batch.  binder_batch_em runs
 * binder_registration_em(regs[p], max_iterations[p], convergence_th,
 * update, 0) for every p at once: each iteration runs all active
 * registrations' E-steps in one launch (a work list of (pair, logical block)
 * items), their histogram reductions and theta steps in one launch each,
 * their DCT smoothings on a pool of streams, and copies all objective terms
 * to the host once.  Each registration's kernels compute what its own EM
 * call computes, so its field, theta, objective trace and iteration count
 * (iterations_out[p]) are those of the single call. */
typedef struct binder_batch binder_batch;
/* npairs slots; regs may be NULL (all slots empty) or hold NULL entries. */
int binder_batch_create(binder_batch **out, int npairs, binder_registration *const *regs);
void binder_batch_destroy(binder_batch *b);
int binder_batch_em(binder_batch *b, const int *max_iterations, double convergence_th,
                    int update_likelihood_parameters, int *iterations_out);
/* Slot-wise use, so that pairs at different pyramid levels share launches:
 * binder_batch_set puts reg (NULL: nothing) into `slot` with an EM call of
 * max_iterations ahead of it (running slots share the update flag);
 * binder_batch_run iterates every running slot until at least min_finished
 * of them finish their EM call (or all have) and writes, per slot, the
 * iteration count of a call that finished during this run, else -1.  A slot
 * left running resumes exactly where it stands at the next run; a finished
 * slot's handle holds the state its own EM call would leave. */
int binder_batch_set(binder_batch *b, int slot, binder_registration *reg, int max_iterations,
                     int update_likelihood_parameters);
int binder_batch_run(binder_batch *b, double convergence_th, int min_finished, int *finished);

/* get_final_locations_as_numpy_array() -- BINDER/src/gpu/bindings.cpp:142-157. */
int binder_registration_get_final_locations(binder_registration *reg,
                                            float *out);
/* get_dream_locations() -- BINDER/src/gpu/bindings.cpp:133. */
int binder_registration_get_dream_locations(binder_registration *reg,
                                            float *out);
/* set_dream_locations(dream) -- BINDER/src/gpu/bindings.cpp:135-140. */
int binder_registration_set_dream_locations(binder_registration *reg,
                                            const float *dream);
/* get_likelihood_parameters() -> theta[K*L] -- BINDER/src/gpu/bindings.cpp:97. */
int binder_registration_get_likelihood_parameters(binder_registration *reg,
                                                  double *theta_out);
/* The deformation prior's current weights gamma_x, gamma_y, gamma_z: the
 * values set by setNonLinearTransformation, or, after a sampler call with
 * sample_gamma != 0, those of the last nu draw, gamma_n = nu rho_x rho_y rho_z
 * eta_n^2 Nv (BINDER/src/cpu/NonLinearTransformation.cpp:116-124, :140-148).
 * The reference keeps gamma inside its NonLinearTransformation object and
 * only prints it (:145-146); this getter has no binding there. */
int binder_registration_get_gamma(binder_registration *r, double *gamma3);
/* get_shape() -> {Nx, Ny, Nz, N} -- BINDER/src/cpu/bindings.cpp:294. */
int binder_registration_get_shape(binder_registration *reg, int *shape4);

/* Objective trace of the last EM call: rows of (mll, prior_theta, prior_d,
 * mlp).  The reference only prints these (Registration.cpp:196-202).
 * Writes min(*rows, available) rows and sets *rows to the number written. */
int binder_registration_get_objective_trace(binder_registration *reg,
                                            double *out, int *rows);

/* One EM E-step with the current theta and deformation (the voxel loop of
 * Registration.cpp:153-165 over estimate_posterior_and_dream_location,
 * :219-330).  Writes the dream locations on the device, returns
 * mll = -sum_i log Z_i and, if hist_out != NULL, the soft joint histogram
 * H[K*L] it accumulated (MILikelihood.cpp:89-124).  Used by tests/bench. */
int binder_registration_e_step(binder_registration *reg, double *mll,
                               double *hist_out);

/* Kernel timing on the handle's stream (HIP events around each launch).
 * enable != 0 starts recording; get_kernel_times returns, for up to
 * *count kernel names, the total milliseconds and launch counts since the
 * last enable.  names receives a '\n'-separated list (capacity names_cap). */
int binder_registration_set_profiling(binder_registration *reg, int enable);
int binder_registration_get_kernel_times(binder_registration *reg,
                                         char *names, int names_cap,
                                         double *total_ms, long long *launches,
                                         int *count);

/* ------------------------------------------------------------------ */
/* NonLinearTransformation (BINDER/src/gpu/bindings.cpp:159-255)       */
/* ------------------------------------------------------------------ */
typedef struct binder_nlt binder_nlt;

/* NonLinearTransformation(gamma_x..z, voxels_size_x..z, nodes_size_x..z, D,
 *   Nx_v, Ny_v, Nz_v, threads, sigma_spline, spline_offset)
 * -- BINDER/src/gpu/bindings.cpp:161-198,
 *    BINDER/src/cpu/NonLinearTransformation.cpp:15-56. */
int binder_nlt_create(binder_nlt **out, double gamma_x, double gamma_y,
                      double gamma_z, double voxels_size_x,
                      double voxels_size_y, double voxels_size_z,
                      double nodes_size_x, double nodes_size_y,
                      double nodes_size_z, const float *D, int Nx_v, int Ny_v,
                      int Nz_v, int threads, double sigma_spline,
                      double spline_offset);
void binder_nlt_destroy(binder_nlt *nlt);
/* smooth_deformation(field) -> (cost, field)
 * -- BINDER/src/gpu/bindings.cpp:201-224,
 *    BINDER/src/cpu/NonLinearTransformation.cpp:226-278. */
int binder_nlt_smooth_deformation(binder_nlt *nlt, const float *deformation,
                                  float *out, double *cost);
/* sampleDeformationField(field, S) -> [S,Nx,Ny,Nz,N]
 * -- BINDER/src/gpu/bindings.cpp:226-255,
 *    BINDER/src/cpu/NonLinearTransformation.cpp:280-341.
 * seed selects the Philox stream (the reference's GPU build fixes 1234). */
int binder_nlt_sample_deformation_field(binder_nlt *nlt,
                                        const float *deformation,
                                        int number_of_samples,
                                        unsigned long long seed, float *out);
/* up_sample(field, N_x, N_y, N_z, N_x_o, N_y_o, N_z_o, threads)
 * -- BINDER/src/cpu/bindings.cpp:408-433,
 *    BINDER/src/cpu/NonLinearTransformation.cpp:366-430.
 * Ortho DCT-II at the input size, zero-padded, ortho DCT-III at the output
 * size; field [N_x,N_y,N_z,N] -> out [N_x_o,N_y_o,N_z_o,N]. */
int binder_nlt_up_sample(const double *field, int N_x, int N_y, int N_z,
                         int N_x_o, int N_y_o, int N_z_o, int threads,
                         double *out);

/* ------------------------------------------------------------------ */
/* Jacobian determinant and topology correction                       */
/* ------------------------------------------------------------------ */

/* compute_jacobian_determinant_3D(field, Nx, Ny, Nz, vx, vy, vz)
 * -- BINDER/src/cpu/bindings.cpp:564-593, BINDER/src/cpu/utils.cpp:100-345.
 * FP64, same corner order and expression tree as the reference, no FMA
 * contraction: bit-exact with the reference CPU result.
 * Also Nz == 1 with a [Nx,Ny,2] field gives the 2D version
 * (BINDER/src/cpu/utils.cpp:10-98) when called through
 * binder_jacobian_determinant_2d. */
int binder_jacobian_determinant_3d(const double *field, int Nx, int Ny,
                                   int Nz, double voxel_size_x,
                                   double voxel_size_y, double voxel_size_z,
                                   double *out);
/* compute_jacobian_determinant_2D(field, Nx, Ny, vx, vy)
 * -- BINDER/src/cpu/bindings.cpp:530-562, BINDER/src/cpu/utils.cpp:10-98. */
int binder_jacobian_determinant_2d(const double *field, int Nx, int Ny,
                                   double voxel_size_x, double voxel_size_y,
                                   double *out);
/* Fold mask and count: mask[i] = (J[i] < e1), *count = number of set voxels.
 * Same J as binder_jacobian_determinant_3d (reference fold test:
 * BINDER/src/cpu/topology_correction_3D.cpp:62-64). mask may be NULL. */
int binder_fold_mask_3d(const double *field, int Nx, int Ny, int Nz,
                        double voxel_size_x, double voxel_size_y,
                        double voxel_size_z, double e1, unsigned char *mask,
                        long long *count);

/* binder_fold_mask_3d on device memory, allocating nothing: no counterpart
 * in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  field_planes: device pointer to the field as three FP64
 * planes [3][Nx*Ny*Nz] (x, y, z components, each C order over (x, y, z));
 * mask: device pointer to Nx*Ny*Nz bytes (mask[i] = J[i] < e1) or NULL;
 * count: device pointer to one long long, written with the number of voxels
 * with J < e1.  stream: a hipStream_t (NULL: the null stream); the call only
 * enqueues work on it and returns without synchronizing, so the results are
 * ready once the stream reaches them.  The Jacobian is the one of
 * binder_jacobian_determinant_3d (bit-exact with the reference).
 * Limit: Nx*Ny*Nz <= 2^31 - 1 voxels (larger -> BINDER_EINVAL).  The only
 * scratch is a 128 KiB per-block partial array in a file-scope device array,
 * held by each GPU the library runs on until the process ends; every call on
 * one GPU shares it, so calls that overlap on one GPU (concurrent streams or
 * threads) race. */
int binder_fold_mask_3d_noalloc(const double *field_planes, int Nx, int Ny, int Nz,
                                double voxel_size_x, double voxel_size_y,
                                double voxel_size_z, double e1, unsigned char *mask,
                                long long *count, void *stream);

/* topology_correction_3D().correct_topology(h_init, e1, e2, max_outer_it,
 *   max_it, threads, voxel_sizes, Nx_v, Ny_v, Nz_v)
 * -- BINDER/src/gpu/bindings.cpp:257-287,
 *    BINDER/src/cpu/topology_correction_3D.cpp:17-143.
 * h_init/out: [Nx,Ny,Nz,3] grid-unit positions; FP64 throughout.
 * stats (may be NULL) receives {outer iterations, inner iterations total,
 * final J_min, final J_max, final count below e1, final count above e2}. */
int binder_topology_correction_3d(const double *h_init, double e1, double e2,
                                  int max_outer_it, int max_it, int threads,
                                  const double *voxel_sizes, int Nx_v,
                                  int Ny_v, int Nz_v, double *out,
                                  double *stats);

/* Topology correction of the handle's current field in place, on the
 * device: the same computation as binder_topology_correction_3d (N = 3) or
 * _2d (N = 2) applied to the field get_final_locations would return, written
 * back rounded to float32.  No reference counterpart: it lets the
 * FastRegistration pyramid (BINDER/FastRegistration.py:521-544) skip the
 * field's round trip through host memory.  voxel_sizes: N values. */
int binder_registration_correct_topology(binder_registration *reg, double e1,
                                         double e2, int max_outer_it, int max_it,
                                         const double *voxel_sizes, double *stats);

/* initialize_fourier_domain_smoothing(size, DCT=True, order, nd, voxel_size)
 * rounded to float32 -- BINDER/utils.py:331-349.  Writes the nx[*ny[*nz]]
 * values (the first nd of the three sizes), C order; host computation. */
int binder_dct_prior_eigenvalues(int nx, int ny, int nz, int nd, int order,
                                 const double *voxel_size, float *out);

/* Spline(order, is2D).spline(d) / .log_spline(d)
 * -- BINDER/src/cpu/bindings.cpp:436-445, BINDER/src/cpu/Spline.cpp:44-145.
 * Writes the 4 weights (log_space = 0) or their logs (log_space = 1) of the
 * order-1/2/3 B-spline at fractional distance d; host computation. */
int binder_spline_weights(int order, double d, int log_space, double *out4);

/* topology_correction_2D().correct_topology(h_init, e1, e2, max_outer_it,
 *   max_it, threads, voxel_sizes, Nx_v, Ny_v)
 * -- BINDER/src/cpu/bindings.cpp:449-478,
 *    BINDER/src/cpu/topology_correction_2D.cpp:17-430.
 * h_init/out: [Nx,Ny,2] grid-unit positions; voxel_sizes: 2 values; FP64.
 * stats as for binder_topology_correction_3d. */
int binder_topology_correction_2d(const double *h_init, double e1, double e2,
                                  int max_outer_it, int max_it, int threads,
                                  const double *voxel_sizes, int Nx_v, int Ny_v,
                                  double *out, double *stats);

/* Launch-shape autotuner (no reference counterpart: the reference fixes its
 * launch shapes, BINDER/src/gpu/Registration.cu:756-761,
 * BINDER/src/gpu/NonLinearTransformation.cu:341-345).  While enabled (also
 * by BINDER_AUTOTUNE=1 in the environment), the first run of a tunable
 * kernel on a problem size times its candidate launch shapes on the
 * caller's stream and keeps the fastest for the process: the EM and sampling
 * E-steps' block counts (grids of at least 2^18 voxels) and the persistent
 * blocks per CU of the y/z DCT passes.  Every candidate computes the same
 * results bit for bit (exact integer partial sums, tiles independent of the
 * block that runs them).  binder_autotune_report writes one line per tuned
 * problem, "op key choice default ms_choice ms_default" (needed: bytes
 * including the terminating zero).  binder_autotune_set(mode): 0 off, 1 tune
 * by timing, 2 take each tunable's last non-default shape untimed (for
 * checking that non-default shapes give the default's results). */
int binder_autotune_set(int mode);
int binder_autotune_clear(void);
int binder_autotune_report(char *buf, int cap, int *needed);

/* Host helpers of bin_signal_robust (BINDER/utils.py:20-82), run over all
 * host cores with numpy's arithmetic: binder_min_max gives the range (NaN if
 * any value is NaN); binder_histogram_uniform counts a[0..n) into the nbins
 * bins of edges[0..nbins] (numpy.histogram's uniform-bin rule, values
 * outside [edges[0], edges[nbins]] dropped); binder_bin_floor writes
 * floor((a - lo) / width), 0 below lo and nbins - 1 above hi. */
int binder_min_max(const double *a, long long n, double *mn, double *mx);
int binder_histogram_uniform(const double *a, long long n, int nbins, const double *edges,
                             long long *counts);
int binder_bin_floor(const double *a, long long n, double lo, double hi, double width,
                     int nbins, int *out);

/* Pyramid resampling of a field: skimage.transform.resize(field, shape,
 * order=1) as FastRegistration.move_resolution calls it
 * (BINDER/FastRegistration.py:566-568), times `scale`.  in: [nx,ny,nz,comps],
 * out: [nxo,nyo,nzo,comps], FP64, host buffers.  Gaussian anti-aliasing on
 * shrinking axes, linear interpolation on the pixel extent, mirrored edges,
 * clipped to the input's range. */
int binder_resize_field(const double *in, int nx, int ny, int nz, int comps,
                        int nxo, int nyo, int nzo, double scale, double *out);

/* binder_resize_field of the index grid ([nx,ny,nz,comps], component c of a
 * voxel = its index along axis c; comps 2 needs nz == 1), with the grid made
 * on the device instead of uploaded: the first level of a pyramid that
 * starts from the identity (BINDER/FastRegistration.py:566-568 applied to
 * the initial positions).  out: [nxo,nyo,nzo,comps], FP64, host buffer. */
int binder_resize_index_grid(int nx, int ny, int nz, int comps, int nxo, int nyo, int nzo,
                             double scale, double *out);

/* ------------------------------------------------------------------ */
/* Learned warm start (no counterpart in the reference, whose pyramid  */
/* always starts from the identity, BINDER/FastRegistration.py:206-225) */
/* ------------------------------------------------------------------ */
/* A 3D convolutional encoder-decoder (channels 16/32/48/64, 3x3x3 convs,
 * BF16 weights and activations, FP32 accumulation, FP32 master weights and
 * Adam) mapping a binned image pair on an nx x ny x nz grid (each side a
 * multiple of 8 in [32, 128]) to coef[3][C][C][C]: the C = ncoef (<= 16)
 * lowest orthonormal DCT-II coefficients per axis of each displacement
 * component of the full-resolution field, divided by sqrt(full voxel
 * count).  Parameters are
 * one FP32 vector: per conv layer W[cout][27][cin] then bias[cout] (layers
 * 8->16, 16->16, 16->32, 32->32, 32->48, 48->48, 48->64, 3 x 64->64,
 * 112->64, 64->64), then the head W[3][64], bias[3]. */
typedef struct binder_warmstart binder_warmstart;

/* New network predicting ncoef (1..16) coefficients per axis, with
 * He-normal weights drawn from `seed`; buffers for up to max_batch (<= 64)
 * pairs.  The parameter vector does not depend on ncoef. */
int binder_warmstart_create(binder_warmstart **out, int nx, int ny, int nz,
                            int max_batch, int ncoef, unsigned long long seed);
void binder_warmstart_destroy(binder_warmstart *ws);
int binder_warmstart_num_params(binder_warmstart *ws, long long *n);
int binder_warmstart_get_params(binder_warmstart *ws, float *out);
/* Replaces the parameters and resets the Adam state. */
int binder_warmstart_set_params(binder_warmstart *ws, const float *in);

/* Device-resident training set of n pairs: fixed_bins / moving_bins are
 * [n][nx][ny][nz] bins in [0, L) / [0, K) (network inputs bin/(L-1),
 * bin/(K-1)); targets [n][3][C][C][C] or NULL (prediction only). */
int binder_warmstart_set_dataset(binder_warmstart *ws, int n,
                                 const unsigned char *fixed_bins,
                                 const unsigned char *moving_bins,
                                 const float *targets, int L, int K);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8) on the mean squared error,
 * step s using samples idx[s*batch .. s*batch+batch); losses[steps] (or
 * NULL) receives each step's loss before its update. */
int binder_warmstart_train(binder_warmstart *ws, const int *idx, int steps,
                           int batch, double lr, double *losses);
/* Loss and parameter gradient (num_params floats) of one batch; no update. */
int binder_warmstart_gradients(binder_warmstart *ws, const int *idx, int batch,
                               double *loss, float *grad_out);
/* coef_out[batch][3][C][C][C] for dataset samples idx[0..batch). */
int binder_warmstart_predict_dataset(binder_warmstart *ws, const int *idx,
                                     int batch, float *coef_out);
/* coef_out for pairs given as host bins ([batch][nx][ny][nz] each). */
int binder_warmstart_predict(binder_warmstart *ws, int batch,
                             const unsigned char *fixed_bins,
                             const unsigned char *moving_bins, int L, int K,
                             float *coef_out);

/* Sets the handle's current field to the warm-start field of a pyramid
 * level: d(p) = x(p) + u(res * p) / res, with x the handle's voxel
 * positions (the grid indices p for an index grid), u the full-resolution
 * displacement (grid fx x fy x fz) whose C x C x C lowest orthonormal DCT
 * coefficients per component (C = ncoef, 1..16) are
 * sqrt(fx fy fz) * coef[3][C][C][C] and all others zero.  The handle's grid must be ceil(f / res) per axis; 3D only.
 * No reference counterpart (it replaces the identity start of
 * BINDER/FastRegistration.py:206-225 for the first level run). */
int binder_registration_set_field_from_dct(binder_registration *reg,
                                           const float *coef, int ncoef, int fx,
                                           int fy, int fz, int res);

/* binder_registration_create (voxel_pos NULL: the grid indices, generated on
 * the device) followed by binder_registration_set_field_from_dct: a
 * registration whose field starts at the warm start, with no field upload.
 * No reference counterpart (replaces the identity-start constructor call of
 * BINDER/FastRegistration.py:443-450 for the first level run). */
int binder_registration_create_from_dct(binder_registration **out,
                                        const float *coef, int ncoef, int fx,
                                        int fy, int fz, int res,
                                        const float *voxel_pos,
                                        int Nx_v, int Ny_v, int Nz_v, int Nx_n,
                                        int Ny_n, int Nz_n, int spline_order,
                                        double non_zero_voxels);

#ifdef __cplusplus
}
#endif
#endif /* BINDER_C_H */
