// ORACLE -- test infrastructure, not product code.
// C entry points of the FP64 CPU oracle (loaded by tests/ and bench.py via
// ctypes).  Arrays follow the reference layout: fields [Nx,Ny,Nz,N]
// interleaved, volumes [Nx,Ny,Nz], theta [K,L] row-major.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif
void orc_spline(int order, double d, double *out4);
void orc_dct3(double *data, int nx, int ny, int nz, int inverse, int threads);

void *orc_reg_create(const double *final_locations, int threads,
                     const double *voxel_pos, int Nx_v, int Ny_v, int Nz_v,
                     int Nx_n, int Ny_n, int Nz_n, int spline_order,
                     double non_zero_voxels);
void orc_reg_destroy(void *h);
void orc_reg_set_mi(void *h, double alpha, const double *theta, int K, int L,
                    const int *binned_nodes, const int *binned_voxels);
void orc_reg_set_ssd(void *h, int bins, double sigma_sq, int scale_shift, double scale,
                     double shift, const int *binned_nodes, const int *binned_voxels,
                     double alpha_0, double beta_0);
void orc_reg_get_ssd(void *h, double *out3);
void orc_reg_set_affine(void *h, const double *A, const double *t);
void orc_reg_set_rigid(void *h, const double *R, const double *t);
int orc_reg_get_linear(void *h, double *out);
double orc_reg_constrain(void *h);
void orc_reg_ssd_update(void *h, int draw);
double orc_reg_ssd_cost(void *h);
void orc_reg_ssd_seed(void *h, int seed);
void orc_reg_set_nlt(void *h, double gx, double gy, double gz, double rx,
                     double ry, double rz, double ex, double ey, double ez,
                     const double *D);
int orc_reg_em(void *h, int max_it, double th, int update, int debug);
double orc_reg_estep(void *h, double *H_out);
void orc_reg_theta_update(void *h);
double orc_reg_theta_cost(void *h);
void orc_reg_sampler(void *h, int seed, int Nb, int Ns, int sg, int every,
                     int sp, int sl, int st, double *mean, double *cov);
void orc_reg_get_final(void *h, double *out);
void orc_reg_get_dream(void *h, double *out);
void orc_reg_set_dream(void *h, const double *in);
void orc_reg_get_theta(void *h, double *out);
int orc_reg_get_trace(void *h, double *out, int max_rows);
void orc_reg_get_gamma(void *h, double *out3);

void *orc_nlt_create(double gx, double gy, double gz, double rx, double ry,
                     double rz, double ex, double ey, double ez,
                     const double *D, int nx, int ny, int nz, int threads,
                     double sigma, double offset);
void orc_nlt_destroy(void *h);
double orc_nlt_smooth(void *h, const double *in, double *out);
void orc_nlt_sample(void *h, const double *in, int S, long long seed,
                    double *out);
void orc_up_sample(const double *field, int nx, int ny, int nz, int nxo,
                   int nyo, int nzo, int threads, double *out);

void orc_jacobian_3d(const double *field, int Nx, int Ny, int Nz, double vx,
                     double vy, double vz, double *out);
void orc_jacobian_2d(const double *field, int Nx, int Ny, double vx, double vy,
                     double *out);
void orc_topology_3d(const double *h_init, double e1, double e2, int max_outer,
                     int max_inner, const double *vs, int Nx, int Ny, int Nz,
                     double *out, double *stats);
void orc_corner_jacobians_3d(const double *field, int Nx, int Ny, int Nz,
                             const double *vs, double e1, double e2,
                             double *Jc_out, long long *counts);
// One inner topology iteration's steps (physical-unit fields).
void orc_topo_gradients(const double *w_phys, int Nx, int Ny, int Nz,
                        const double *vs, double *g_out);
void orc_topo_limit(double *g, const double *Jc, int Nx, int Ny, int Nz,
                    double e1, double e2);
void orc_topo_integrate(const double *g, const double *mean3, int Nx, int Ny,
                        int Nz, const double *vs, double *w_out);
// 2D topology correction and its steps (topology_correction_2D.cpp).
void orc_topology_2d(const double *h_init, double e1, double e2, int max_outer,
                     int max_inner, const double *vs, int Nx, int Ny,
                     double *out, double *stats);
void orc_corner_jacobians_2d(const double *field, int Nx, int Ny, const double *vs,
                             double e1, double e2, double *Jc_out, long long *counts);
void orc_topo2_gradients(const double *w_phys, int Nx, int Ny, const double *vs,
                         double *g_out);
void orc_topo2_limit(double *g, const double *Jc, int Nx, int Ny, double e1, double e2);
void orc_topo2_integrate(const double *g, const double *mean2, int Nx, int Ny,
                         const double *vs, double *w_out);
#ifdef __cplusplus
}
#endif
