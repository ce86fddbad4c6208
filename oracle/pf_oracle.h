/*
 * CPU FP64 oracle for the var_diff.f90 semi-implicit spectral timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker (or the timed CPU
 * baseline), never as the thing measured or shipped.
 *
 * Layouts (identical to the reference's single-process arrays):
 *   real field      (nx, ny, nz)         x fastest            src/var_diff.f90:143
 *   phi             (var, nx, ny, nz)    var fastest (AoS)    src/var_diff.f90:154
 *   half spectrum   (nx/2+1, ny, nz)     kx fastest, complex interleaved (re, im).
 *     The reference stores the same numbers as P3DFFT's STRIDE1 Z-pencil
 *     (kz, ky, kx); only the index order differs.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int idum2, iy, iv[32];
} ora_ran2_state;

typedef struct {
  double t_step, grad_coeff_phi, grad_coeff_c, D_S, D_L, D_gb, mob_phi;
  double sigma_1, sigma_2, am, ap, con_0_mat, con_0_ppt;
} ora_params;

/* Diagnostics of the state at the START of a step: the energies the reference
 * computes in loops 1-6 of that step and the extrema it computed at the end of
 * the previous step (src/var_diff.f90:1058-1118). */
typedef struct {
  double f_int, f_ch, f_eta, e_tot, c_max, c_min, phi_min, phi_3_max;
} ora_diag;

void   ora_ran2_init(ora_ran2_state* st);
double ora_ran2(ora_ran2_state* st, int* idum);

int  ora_fft_r2c(int nx, int ny, int nz, const double* in, double* out, int scale);
int  ora_fft_c2r(int nx, int ny, int nz, const double* in, double* out);
int  ora_fft_c2c_line(int n, const double* in, double* out, int inverse);
void ora_kspace_table(int nx, int ny, int nz, double* kf, double* kf_sq, double* kf_4);
void ora_twiddles(int n, double* w);

void ora_gamma(int var, double sigma_1, double sigma_2, double* gamma);
int  ora_run(const ora_params* p, int nx, int ny, int nz, int var,
             double* phi, double* con, int nsteps, ora_diag* diag);
int  ora_pointwise(const ora_params* p, int npts, int var, const double* phi,
                   const double* con, double* df_dc, double* df_dphi,
                   double* mob_c, double* f_ch, double* f_eta);
void ora_extrema(int npts, int var, const double* phi, const double* con,
                 double* out4);

int  ora_ic_random(int nx, int ny, int nz, int var, int iseed, double c0,
                   double c_amp, double eta_amp, double* phi, double* con);
int  ora_ic_voronoi(int nx, int ny, int nz, int var, int nseeds, int iseed,
                    double c0, double c_amp, double* phi, double* con);

void ora_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
