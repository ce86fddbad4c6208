/*
 * pf.h -- C-ABI of the MI355X phase-field timestep library (libpf.so).
 *
 * Drop-in boundary for the semi-implicit Fourier-spectral step of
 * smjim/phase-field-microstructure-evolution (src/var_diff.f90).  A Fortran host
 * binds these entry points with iso_c_binding (module pf_capi,
 * phase-field-microstructure-evolution_amd/fortran/pf_capi.f90).
 *
 * Conventions
 *   - Return value 0 on success, a negative PF_E* code otherwise; the message is
 *     in pf_last_error() (thread-local).  The reference has no error returns (it
 *     aborts inside P3DFFT/MPI); its Fortran host aborts on any non-zero code.
 *   - Host arrays use the reference's Fortran layouts:
 *       phi (var, nx, ny, nz), var fastest        src/var_diff.f90:154
 *       con (nx, ny, nz), x fastest               src/var_diff.f90:143
 *       half spectrum (nz, ny, nx/2+1), kz fastest (P3DFFT STRIDE1 Z-pencil,
 *       src/var_diff.f90:135-137), complex as interleaved (re, im) doubles.
 *   - Forward transforms use e^{-i} and divide by nx*ny*nz (src/f_trans.f90:11-12);
 *     inverse transforms use e^{+i} and do not scale (src/inv_trans.f90:10).
 *   - Each axis length must be one of 8 12 16 24 32 48 64 96 128 160 192 256 512.
 *   - One host thread per handle; handles are independent.
 */
#ifndef PF_H
#define PF_H
#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_EINVAL (-1)
#define PF_ENOMEM (-2)
#define PF_EHIP (-3)
#define PF_ENOTIMPL (-4)
#define PF_ENONFINITE (-5)

#define PF_PREC_FP64 0
#define PF_PREC_FP32 1

#define PF_LAYOUT_XFAST 0 /* (nx/2+1, ny, nz), kx fastest (numpy rfftn order) */
#define PF_LAYOUT_ZFAST 1 /* (nz, ny, nx/2+1), kz fastest (reference / P3DFFT order) */

/* Step forms. */
#define PF_MODE_FP64_FAITHFUL 0 /* the reference's loop sequence, FP64, real-space state */
#define PF_MODE_FP64_REDUCED 1  /* same mathematics, spectral state, 2F transforms/step, FP64 */
#define PF_MODE_FP32_MIXED 2    /* reduced form in FP32; FP64 k=0 modes and reductions */

/* Parameters, named as in input.txt (src/var_diff.f90:106-113).  D_S is read by
 * the reference but never used; D_mean = D_L (src/var_diff.f90:916).  gamma is
 * built from sigma_1/sigma_2 exactly as src/var_diff.f90:206-218. */
typedef struct pf_params {
  double t_step, grad_coeff_phi, grad_coeff_c, D_S, D_L, D_gb, mob_phi;
  double sigma_1, sigma_2, am, ap, con_0_mat, con_0_ppt;
} pf_params;

/* Diagnostics of the state at the START of step `step` (1-based, counted over the
 * handle's lifetime): the energies the reference sums in loops 1-6 of that step
 * and the extrema it computed at the end of the previous step -- i.e. the values
 * of one summary.dat row (src/var_diff.f90:1058-1118).  c_mean is the mean of c
 * (mass), reported in addition.  valid = 0 when the step was not diagnosed. */
typedef struct pf_diag {
  int step, valid;
  double f_int, f_ch, f_eta, e_tot, c_max, c_min, phi_min, phi_3_max, c_mean;
} pf_diag;

typedef struct pf_handle pf_handle;

/* ---- timestep ------------------------------------------------------------
 * Replaces the body of the time loop (src/var_diff.f90:556-1118: loops 1-14,
 * energies, extrema) and the P3DFFT setup/teardown (:128-138, :1162).
 * pf_create returns PF_ENOTIMPL when one x-line block cannot hold every field of
 * a line (512-point lines: up to 64 fields in FP32-mixed, 37 in FP64 reduced). */
int pf_create(const pf_params* p, int nx, int ny, int nz, int var, int mode, pf_handle** out);
int pf_destroy(pf_handle* h);
/* Upload / download the state (src/var_diff.f90:154,143 layouts). */
int pf_put_state(pf_handle* h, const double* phi, const double* con);
int pf_get_state(pf_handle* h, double* phi, double* con);
/* The Voronoi polycrystal of pf_ic_voronoi (same draws, same owners), built on
 * the device straight into the handle: equivalent to pf_ic_voronoi followed by
 * pf_put_state, without the host phi array (no counterpart in the reference,
 * whose polycrystal start reads a grain-map file, src/var_diff.f90:232-252;
 * this is the synthetic start of the benchmark workloads at sizes where the
 * host array would be tens of GB).  nseeds in [1, 65536]. */
int pf_put_voronoi(pf_handle* h, int nseeds, int iseed, double c0, double c_amp);
/* Advance nsteps steps.  out (nullable) receives nsteps records; a record is
 * filled when its step number is a multiple of diag_every (diag_every <= 0: none).
 * Returns PF_ENONFINITE if a diagnosed energy or mean is NaN/Inf. */
int pf_step(pf_handle* h, int nsteps, int diag_every, pf_diag* out);
/* Diagnostics of the current state (the start of the next step). */
int pf_diagnose(pf_handle* h, pf_diag* out);
/* Ensemble of independent simulations (the reference runs one var_diff.x job
 * per parameter set, tests/scaling_tests.py:53-69 and tests/run_phase_sim.py:
 * 112-150): advances each of the n handles by nsteps steps, no diagnostics.
 * Steps are enqueued round-robin on the handles' own streams, so the GPU runs
 * the simulations concurrently; returns once every stream has finished.  The
 * handles must be distinct. */
int pf_ensemble_step(pf_handle* const* hs, int n, int nsteps);
/* Graph replay of the step (no counterpart in the reference, whose loop body is
 * host code): when on (the default), pf_step and pf_ensemble_step capture one
 * period of steps -- diag_every steps, or one step without diagnostics -- into
 * a hipGraph on first use and replay it for every whole period of a call, one
 * launch per period instead of one per kernel.  Results are bit-identical to
 * stepping without graphs.  Off while stage timing is on. */
int pf_set_graphs(pf_handle* h, int on);
/* Launch shapes of the fused FP32/FP64-reduced kernels (no counterpart in the
 * reference, which plans its transforms inside P3DFFT/FFTW at setup,
 * src/var_diff.f90:128).  pf_create takes them from the library's tuned table
 * for the handle's grid, var and precision (written by the tuner,
 * phase-field-microstructure-evolution_amd/autotune.py, from timings on an
 * MI355X; PF_TUNING=default in the environment skips the table), else the
 * defaults.  The first four decide which blocks do what, not what is computed
 * per point or per column: the state is unchanged, and the summary energies,
 * which are sums of per-block partials, can change in their last bits.
 *   xline_lines   : x lines per x-line block, 0 = automatic, else 1, 2, 4 or 8
 *   xline_grid    : x-line blocks, 0 = 8192, else 256 .. 16384
 *   zline_grid    : z-line blocks, 0 = 32768, else 1024 .. 65536
 *   zline_run_rows: main-block rows per XCD run of z-line tiles, -1 = 1, 0 = none, 1 .. 8
 *   zline_plan    : radix sequence of the FP32 z transforms, 0 = default, 1 or 2
 *                   = the alternatives for nz = 64, 128, 160, 192, 256, 512; it
 *                   changes the transform's rounding, so the state differs from
 *                   the default's at the level of FP32 rounding
 * pf_set_tuning refuses a shape this grid cannot run (PF_EINVAL); pf_get_tuning
 * reports the values in force. */
typedef struct pf_tuning {
  int xline_lines, xline_grid, zline_grid, zline_run_rows, zline_plan;
} pf_tuning;
int pf_set_tuning(pf_handle* h, const pf_tuning* t);
int pf_get_tuning(const pf_handle* h, pf_tuning* t);
/* Per-stage timing with HIP events on the handle's stream: enable/disable and
 * read the accumulated milliseconds and launch counts of up to n stages. */
int pf_set_timing(pf_handle* h, int on);
/* Timing on for the stages in mask only (bit i = stage i, as pf_stage_name
 * numbers them): two events per launch of those stages, none for the others. */
int pf_set_timing_stages(pf_handle* h, unsigned mask);
int pf_stage_times(pf_handle* h, double* ms, long* count, int n);
int pf_stage_name(int stage, char* buf, int len);
/* Algorithmic HBM bytes of one launch of `stage` for this handle's grid. */
double pf_stage_bytes(pf_handle* h, int stage);
/* The HIP stream the handle launches on (hipStream_t). */
void* pf_stream(pf_handle* h);

/* ---- transforms ------------------------------------------------------------
 * pf_fft_r2c: forward R2C of `batch` real fields (device buffers), output in
 * `layout`, multiplied by `scale` (1/(nx*ny*nz) reproduces f_trans).
 * pf_fft_c2r: inverse C2R, unscaled (inv_trans); imaginary parts of the kx=0
 * and kx=nx/2 entries are ignored.  prec = PF_PREC_FP64 / PF_PREC_FP32.
 * stream: hipStream_t or NULL.  Both share one scratch buffer per device: calls
 * from several threads or streams take turns on it in stream order (each call's
 * work waits for the previous user's), so they are safe but do not overlap. */
int pf_fft_r2c(int nx, int ny, int nz, int batch, int prec, const void* d_in, void* d_out,
               int layout, double scale, void* stream);
int pf_fft_c2r(int nx, int ny, int nz, int batch, int prec, const void* d_in, void* d_out,
               int layout, void* stream);
/* pf_fft_r2c without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  Its scratch is a static device array of 64 MiB per
 * device (the split-layout copy of the batch: 8*P*batch + 16*ny*nz*batch bytes
 * in FP64, half that in FP32); larger requests return PF_EINVAL.  Every call on
 * a device shares that device's array, so calls that overlap on one device race.
 * The first call on a device sets up (device addresses of the static scratch and
 * twiddle tables, and one copy of the tables); make that call before capturing. */
int pf_fft_r2c_noalloc(int nx, int ny, int nz, int batch, int prec, const void* d_in, void* d_out,
                       int layout, double scale, void* stream);
/* Host-buffer FP64 drop-ins for f_trans(in,out,Nx,Ny,Nz,...) (src/f_trans.f90:1)
 * and inv_trans(in,out,...) (src/inv_trans.f90:1); spectrum in PF_LAYOUT_ZFAST. */
int pf_f_trans(const double* in, double* out, int nx, int ny, int nz);
int pf_inv_trans(const double* in, double* out, int nx, int ny, int nz);

/* ---- host helpers ----------------------------------------------------------- */
/* ran_2 (src/ran_2.f90:1-40) with explicit, re-entrant state. */
typedef struct pf_ran2_state {
  int idum2, iy, iv[32];
} pf_ran2_state;
void pf_ran2_init(pf_ran2_state* st);
double pf_ran2(pf_ran2_state* st, int* idum);
/* k_space (src/k_space.f90:4-67): kf, kf_sq, kf_4 as (3, nz, ny, nx/2+1) complex. */
int pf_k_space(int nx, int ny, int nz, double* kf, double* kf_sq, double* kf_4);
/* gamma(var,var) column-major (src/var_diff.f90:198-218). */
int pf_gamma(int var, double sigma_1, double sigma_2, double* gamma);
/* Synthetic initial conditions (layouts as pf_put_state).  The draw orders are
 * defined by this library (the reference has no such modes):
 *   random : per point (z, y, x loops, x fastest) c first, then eta_1..eta_var;
 *            c = c0 + c_amp*(2u-1), eta = eta_amp*(2u-1)
 *   voronoi: nseeds seeds (x,y,z each u*N) first, nearest seed by periodic
 *            minimum-image distance (lowest index on ties), eta_{(s mod var)+1}=1;
 *            then c = c0 + c_amp*(2u-1) per point in the same order. */
int pf_ic_random(int nx, int ny, int nz, int var, int iseed, double c0, double c_amp,
                 double eta_amp, double* phi, double* con);
int pf_ic_voronoi(int nx, int ny, int nz, int var, int nseeds, int iseed, double c0,
                  double c_amp, double* phi, double* con);

/* Post-processing of one state as src/paraview.f90:188-313 does per voxel
 * (phi in the reference's (var, nx, ny, nz) layout; outputs nx*ny*nz values, x
 * fastest): phi_sq = sum_i phi_i^2; var_num = the last i with phi_i > 0.8 (0 if
 * none); phi_3 = phi_3 (0 when var < 3); boundary = 1 where more than one phi_i
 * exceeds threshold; grad_phi_3 = |central-difference gradient of phi_3| on
 * interior voxels (0 on the faces, which the reference leaves unset).
 * counts[0] = boundary voxels with phi_var > threshold, counts[1] = boundary
 * voxels.  Runs on the GPU; buffers are host memory. */
int pf_analyze_state(const double* phi, int nx, int ny, int nz, int var, double threshold,
                     double* phi_sq, double* phi_3, double* grad_phi_3, double* boundary,
                     int* var_num, long* counts);

const char* pf_last_error(void);
int pf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PF_H */
