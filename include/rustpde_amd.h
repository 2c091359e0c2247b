/*
 * rustpde_amd: MI355X (gfx950) implementation of rustpde's 2-D Rayleigh-Benard time step.
 *
 * C ABI of librustpde_amd.so.  Plain pointers and sizes only.  Every entry point returns
 * 0 on success and a negative status on failure (the message is kept per thread and read
 * with rbc_last_error()); nothing throws or aborts across the ABI.  One handle owns its
 * device buffers and one HIP stream; a handle is not thread-safe.  All arithmetic is FP64.
 *
 * Array layouts are the reference's (ndarray, C order): a field of shape [n0, n1] is
 * row-major with axis 1 (y, Chebyshev) contiguous.  Spectral arrays of periodic
 * (Fourier-x) fields are complex, stored interleaved (re, im) exactly like
 * num_complex::Complex<f64> (#[repr(C)]), i.e. 2*n0*n1 doubles.
 *
 * Citations (file:line) are relative to the reference repository (preiter93/rustpde).
 */
#ifndef RUSTPDE_AMD_H
#define RUSTPDE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status / errors */
#define RBC_OK 0
#define RBC_ERR_ARG (-1)     /* bad argument (shape, index, size)            */
#define RBC_ERR_DEVICE (-2)  /* HIP runtime error                              */
#define RBC_ERR_NUMERIC (-3) /* set-up failure (e.g. complex eigenvalues)      */

/* Last error message of the calling thread ("" if none). */
const char* rbc_last_error(void);
/* Library version string. */
const char* rbc_version(void);

/* ------------------------------------------------------------------- Navier2D */
typedef struct rbc_navier rbc_navier;

/* Field indices (members of struct Navier2D, src/navier/navier.rs:153-195). */
#define RBC_TEMP 0    /* temp                         */
#define RBC_UX 1      /* ux                           */
#define RBC_UY 2      /* uy                           */
#define RBC_PRES 3    /* pres[0]  (pressure)          */
#define RBC_PSEU 4    /* pres[1]  (pseudo pressure)   */
#define RBC_FIELDBC 6 /* fieldbc  (temperature BC)    */
/* Spaces */
#define RBC_PHYS 0 /* Field2::v     (physical, real)        */
#define RBC_SPEC 1 /* Field2::vhat  (spectral, real/complex) */

/* Navier2D::new(nx, ny, ra, pr, dt, aspect, adiabatic)  navier.rs:219-307.
 * Fields start at zero (the reference's unseeded random_disturbance, navier.rs:304,
 * is left to the caller: set temp/ux/uy with rbc_set_field + rbc_forward). */
int rbc_create_confined(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic,
                        rbc_navier** out);
/* Navier2D::new_periodic(nx, ny, ra, pr, dt, aspect)  navier.rs:384-467 (nx even). */
int rbc_create_periodic(int nx, int ny, double ra, double pr, double dt, double aspect, rbc_navier** out);
/* drop(Navier2D) */
int rbc_destroy(rbc_navier* h);

/* Shape of field `which` in `space`; *is_complex = 1 for complex spectral arrays. */
int rbc_field_shape(rbc_navier* h, int which, int space, int* n0, int* n1, int* is_complex);
/* Read / write Field2::v or Field2::vhat (host buffers, reference layout). */
int rbc_get_field(rbc_navier* h, int which, int space, double* host_out);
int rbc_set_field(rbc_navier* h, int which, int space, const double* host_in);
/* Field2::forward (v -> vhat) / Field2::backward (vhat -> v)  src/field.rs:103-110 */
int rbc_forward(rbc_navier* h, int which);
int rbc_backward(rbc_navier* h, int which);

/* Navier2D::set_temp_bc (navier.rs:517-519): replaces fieldbc.vhat (shape of RBC_FIELDBC,
 * RBC_SPEC) and every quantity derived from it. */
int rbc_set_temp_bc(rbc_navier* h, const double* host_vhat);

/* Navier2D::set_velocity(amp, m, n) / set_temperature(amp, m, n)  navier.rs:927-937 */
int rbc_set_velocity(rbc_navier* h, double amp, double m, double n);
int rbc_set_temperature(rbc_navier* h, double amp, double m, double n);

/* Integrate::update  navier.rs:737-765, `nsteps` times (queued on the handle's stream). */
int rbc_update(rbc_navier* h, int nsteps);
/* Wait until queued work of the handle has finished. */
int rbc_synchronize(rbc_navier* h);
/* Integrate::get_time / get_dt  navier.rs:767-773; time is settable (Navier2D::time, pub). */
double rbc_get_time(rbc_navier* h);
double rbc_get_dt(rbc_navier* h);
int rbc_set_time(rbc_navier* h, double time);

/* Diagnostics, FP64 on the device:
 *   RBC_NU     eval_nu     navier.rs:890-893 -> functions.rs:12-36
 *   RBC_NUVOL  eval_nuvol  navier.rs:899-909 -> functions.rs:42-75
 *   RBC_RE     eval_re     navier.rs:912-921 -> functions.rs:82-101
 *   RBC_DIV    |divergence|_2 as printed by callback / tested by exit (navier.rs:816-861) */
#define RBC_NU 0
#define RBC_NUVOL 1
#define RBC_RE 2
#define RBC_DIV 3
int rbc_eval(rbc_navier* h, int what, double* out);
/* navier.solid = Some([mask, value]) (navier.rs:189-190; the builders of
 * src/navier/solid_masks.rs:34-162 are host code, rustpde_amd/solid_masks.py): mask and
 * value are physical [nx, ny] arrays; mask = NULL removes the obstacle (None).  The
 * penalisation (eta = 1e-2) enters the convective terms of ux, uy and temp
 * (navier.rs:551-560, 579-584, 603-608).  A periodic handle with a mask runs the batched FP64
 * step with the terms, a confined one the field-by-field FP64 step (the confined field triple
 * does not take them); the FP32 convective pipeline has no penalised form: setting a mask
 * while RBC_OPT_MIXED is on, or RBC_OPT_MIXED while a mask is set, returns RBC_ERR_ARG.
 * Not available for ensemble instances. */
int rbc_set_solid(rbc_navier* h, const double* mask, const double* value);

/* Statistics (src/navier/statistics.rs:10-153; Navier2D::statistics, navier.rs:193-194),
 * FP64 on the device.  Four arrays of the `field` space (confined: [nx, ny] real; periodic:
 * [nx/2 + 1, ny] complex; physical [nx, ny]):
 *   RBC_STAT_TEMP     t_avg    running average of temp.to_ortho() + fieldbc.to_ortho()
 *   RBC_STAT_UX       ux_avg   ux.to_ortho() of the latest update (assigned, not averaged,
 *   RBC_STAT_UY       uy_avg   uy.to_ortho()  as statistics.rs:145-146 do)
 *   RBC_STAT_NUSSELT  nusselt  the Nusselt field (dT/dy + uy T / ka) 2 of the latest update
 *                              (statistics.rs:220-247)
 * rbc_stats_enable is Statistics::new(navier, save_stat, write_stat) (statistics.rs:54-88:
 * zero arrays, avg_time 0, tot_time = the handle's time, num_save 0; calling it again
 * restarts the statistics).  rbc_stats_update is the statistics.update call of callback
 * (navier.rs:796-807, statistics.rs:130-153): it changes nothing and sets *updated = 0 when the
 * handle's time is below tot_time (statistics.rs:134-140).  The save/write cadence
 * (navier.rs:798-799, 809-810) is the caller's, as in Integrate::callback.
 * rbc_stats_get reads .vhat (RBC_SPEC) or .v (RBC_PHYS, the field space's backward
 * transform that Statistics::write applies, statistics.rs:176-179); rbc_stats_set and
 * rbc_stats_set_info restore them (Statistics::read, statistics.rs:197-212).  Not available
 * for ensemble instances. */
#define RBC_STAT_TEMP 0
#define RBC_STAT_UX 1
#define RBC_STAT_UY 2
#define RBC_STAT_NUSSELT 3
int rbc_stats_enable(rbc_navier* h, double save_stat, double write_stat);
int rbc_stats_update(rbc_navier* h, int* updated);
int rbc_stats_get(rbc_navier* h, int which, int space, double* host_out);
int rbc_stats_set(rbc_navier* h, int which, const double* host_vhat);
int rbc_stats_info(rbc_navier* h, double* save_stat, double* write_stat, double* avg_time, double* tot_time,
                   long long* num_save);
int rbc_stats_set_info(rbc_navier* h, double avg_time, double tot_time, long long num_save);

/* Integrate::exit  navier.rs:855-862: *out = 1 if |div|_2 is NaN. */
int rbc_exit(rbc_navier* h, int* out);

/* Tensor solver set-up (FdmaTensor pub fields lam/fwd/bwd, src/solver/fdma_tensor.rs:73-91)
 * of solver s (0 ux, 1 uy, 2 temp, 3 pres; navier.rs:165).  *has_fb = 0 when axis 0 is
 * diagonal (Fourier); n is the number of eigenvalues, fwd/bwd are n x n. */
int rbc_tensor_info(rbc_navier* h, int s, int* n, int* has_fb);
int rbc_get_tensor(rbc_navier* h, int s, double* lam, double* fwd, double* bwd);
/* Overwrite lam (and fwd/bwd if has_fb) of solver s, as assigning the pub fields. */
int rbc_set_tensor(rbc_navier* h, int s, const double* lam, const double* fwd, const double* bwd);

/* Execution options of a handle (no counterpart in the reference; they select between
 * equivalent implementations of the same step, for A/B checks):
 *   RBC_OPT_GRAPH  replay update() as a captured HIP graph (default 1)
 *   RBC_OPT_MFMA   dense transforms on the FP64 matrix cores (default 1; 0 = FMA GEMM)
 *   RBC_OPT_FFT    Fourier x transforms by FFT kernels (default 1; 0 = dense DFT)
 *   RBC_OPT_BANDS  stencil / preconditioner as banded kernels (default 1; 0 = dense)
 *   RBC_OPT_DCT    Chebyshev y transforms by FFT-based DCT-I kernels when ny - 1 is a
 *                  power of two (default 1; 0 = dense transform GEMMs)
 *   RBC_OPT_BATCH  periodic step: the transforms of ux, uy and temp and the ux/uy solves
 *                  run as batched launches (default 1; needs the FFT, DCT and band paths);
 *                  confined step (and ensemble groups): ux, uy and temp's transforms,
 *                  products and ADI solves as launches of three times the batch (with
 *                  RBC_OPT_BANDS off: ux and uy as launches of twice the batch)
 *   RBC_OPT_FUSE   periodic batched step: the convective terms in two launches (values and
 *                  x-derivatives of ux, uy, temp from one x-transform launch, then one
 *                  y-transform launch forming d/dy q uy + dq/dx ux) and divergence, Poisson
 *                  rows, projection and pressure update as one kernel per x-wavenumber
 *                  (default 1)
 *   RBC_OPT_MIXED  mixed precision: the convective terms (their x/y transforms and
 *                  products) in FP32, everything else in FP64 (default 0).  Periodic
 *                  batched step: FFT/DCT kernels in FP32; confined step with RBC_OPT_BATCH
 *                  on (and ensemble groups, rbc_ensemble_set_option): the dense transforms
 *                  as FP32 matrix-core GEMMs (with RBC_OPT_BATCH off the confined step
 *                  stays FP64).  Fields agree with the FP64 step to ~1e-9 after 10 steps.
 *   RBC_OPT_CONV3  periodic batched mixed step with RBC_OPT_FUSE: the y-stage of the FP32
 *                  convective terms (physical ux, uy, the three terms and their dealiased
 *                  forward y-transforms) as one launch per x-row: 1 (default) when nx is at
 *                  least four times the GPU's CU count, 2 always, 0 never (three launches:
 *                  y-backward of ux, uy, the convective-term launch, the forward launch).
 *                  The FP64 step always runs the three launches. */
#define RBC_OPT_GRAPH 0
#define RBC_OPT_MFMA 1
#define RBC_OPT_FFT 2
#define RBC_OPT_BANDS 3
#define RBC_OPT_DCT 4
#define RBC_OPT_BATCH 5
#define RBC_OPT_FUSE 6
#define RBC_OPT_MIXED 7
#define RBC_OPT_CONV3 8
int rbc_set_option(rbc_navier* h, int option, int value);

/* Kernel timing of a handle (no counterpart in the reference, whose timing protocol is
 * the external criterion bench benches/benchmark_navier.rs:23-40).  When enabled, HIP
 * events are recorded on the handle's stream around each launch of a kernel class;
 * rbc_timing_read returns the launches, their summed duration and summed algorithmic
 * work (flops for GEMMs, bytes for the memory-bound classes) since enabling.  Each class
 * but RBC_K_ELEMENTWISE is one kernel:
 *   RBC_K_GEMM gemm_mfma_kernel (dense transforms), RBC_K_ROWSOLVE fdma_rows_lds_kernel
 *   (Hholtz/Poisson rows), RBC_K_FFT irfft_cols_kernel, RBC_K_RFFT rfft_cols_kernel,
 *   RBC_K_DCT cheb_y_bwd_kernel, RBC_K_DCT_FWD cheb_y_fwd_kernel, RBC_K_DIFF_Y
 *   cheb_y_diff_kernel, RBC_K_TDMA tdma_rows_lds_kernel, RBC_K_BAND band_y_kernel;
 *   RBC_K_PRESSURE pressure_rows_kernel (the fused pressure stage); RBC_K_CONV
 *   cheb_y_conv_kernel (the convective term's y transforms and products, one launch);
 *   RBC_K_GEMM32 gemm_mfma_kernel in FP32 (the confined mixed step's convective terms);
 *   RBC_K_CONV3 cheb_y_conv3_kernel (the whole y-stage of the convective terms, one launch);
 *   RBC_K_ELEMENTWISE the remaining timed kernels (Fourier derivative, fallbacks). */
#define RBC_K_GEMM 0
#define RBC_K_ROWSOLVE 1
#define RBC_K_FFT 2
#define RBC_K_ELEMENTWISE 3
#define RBC_K_DCT 4
#define RBC_K_RFFT 5
#define RBC_K_DCT_FWD 6
#define RBC_K_DIFF_Y 7
#define RBC_K_TDMA 8
#define RBC_K_BAND 9
#define RBC_K_PRESSURE 10
#define RBC_K_CONV 11
#define RBC_K_GEMM32 12
#define RBC_K_CONV3 13
int rbc_timing_enable(rbc_navier* h, int on);
int rbc_timing_read(rbc_navier* h, int kernel_class, long long* launches, double* total_ms, double* work);

/* --------------------------------------------------------------- Navier2DAdjoint
 * The adjoint descent for steady states (src/navier/navier_adjoint.rs:128-1068; Farazmand,
 * J. Fluid Mech. 795, 2016), on the device.  Navier2DAdjoint::new / new_periodic
 * (navier_adjoint.rs:187-289, 305-447): the adjoint step of size dt descends the residual
 * (NS(q) - q) / dt_navier of a Navier2D step with dt_navier = 1e-2 (:231-232), smoothed by
 * the Hholtz operator (1 - lap / scale^2) and negated (:727-750); velocity and temperature
 * are updated explicitly (:581-648) and the velocity projected (:651-724).
 * rbc_adjoint_part gives views owned by the adjoint: part 0 the adjoint fields (temp, ux, uy,
 * pres[0], pres[1], fieldbc; .v / .vhat, forward / backward, set_velocity / set_temperature,
 * rbc_eval diagnostics eval_nu / eval_nuvol / eval_re (:936-972), the tensors of its solvers:
 * s = 0, 1, 2 the ux, uy, temp smoothers, 3 the pressure Poisson solver), part 1 the
 * Navier2D step of the residual (its fields and pressure).  rbc_update and rbc_destroy refuse
 * the views. */
typedef struct rbc_adjoint rbc_adjoint;
int rbc_adjoint_create_confined(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic,
                                rbc_adjoint** out);
int rbc_adjoint_create_periodic(int nx, int ny, double ra, double pr, double dt, double aspect, rbc_adjoint** out);
int rbc_adjoint_destroy(rbc_adjoint* h);
int rbc_adjoint_part(rbc_adjoint* h, int part, rbc_navier** out);
/* Integrate::update of the adjoint (navier_adjoint.rs:760-790), nsteps times. */
int rbc_adjoint_update(rbc_adjoint* h, int nsteps);
/* Integrate::get_time / get_dt; time is settable (reset_time, :987-991). */
double rbc_adjoint_get_time(rbc_adjoint* h);
double rbc_adjoint_get_dt(rbc_adjoint* h);
int rbc_adjoint_set_time(rbc_adjoint* h, double time);
/* Residual of field which (RBC_TEMP, RBC_UX, RBC_UY): kind 0 the smoothed residual temp[1] /
 * ux[1] / uy[1] (spectral, the field's shape), kind 1 fields_unsmoothed (ortho space of the
 * `field` space: [nx, ny] real or [nx/2 + 1, ny] complex). */
int rbc_adjoint_residual(rbc_adjoint* h, int which, int kind, double* host_out);
/* out[6] = |ux1|, |uy1|, |temp1|, |unsmoothed ux|, |uy|, |temp| (L2 of the coefficients, as
 * callback prints them, :846-880). */
int rbc_adjoint_norms(rbc_adjoint* h, double* out);
/* Integrate::exit (:882-897): *out = 1 if |div| is NaN or |ux1| + |uy1| + |temp1| < 1e-8. */
int rbc_adjoint_exit(rbc_adjoint* h, int* out);

/* ------------------------------------------------------------- Space2 (funspace) */
typedef struct fs_space2 fs_space2;
/* Base kinds: funspace constructors chebyshev, cheb_dirichlet, cheb_neumann,
 * cheb_dirichlet_bc, fourier_r2c (funspace/src/lib.rs:230-345). */
#define FS_CHEBYSHEV 0
#define FS_CHEB_DIRICHLET 1
#define FS_CHEB_NEUMANN 2
#define FS_CHEB_DIRICHLET_BC 3
#define FS_FOURIER_R2C 4

/* Space2::new(base0(n0), base1(n1))  funspace/src/space2.rs:35-83 */
int fs_space2_create(int kind0, int n0, int kind1, int n1, fs_space2** out);
int fs_space2_destroy(fs_space2* sp);
/* shapes: physical [p0,p1], spectral [s0,s1], orthogonal [o0,o1]; complex if kind0 is Fourier */
int fs_space2_shape(fs_space2* sp, int* p0, int* p1, int* s0, int* s1, int* o0, int* o1, int* is_complex);
/* BaseSpace ops (funspace/src/space_traits.rs:71-311) on DEVICE buffers, batched over
 * `batch` contiguous fields (reference layouts, complex interleaved).  Synchronous. */
int fs_forward(fs_space2* sp, const double* d_phys, double* d_spec, int batch);     /* space2.rs:323-333 */
int fs_backward(fs_space2* sp, const double* d_spec, double* d_phys, int batch);    /* space2.rs:346-356 */
int fs_to_ortho(fs_space2* sp, const double* d_spec, double* d_ortho, int batch);   /* space2.rs:181-201 */
int fs_from_ortho(fs_space2* sp, const double* d_ortho, double* d_spec, int batch); /* space2.rs:203-223 */
int fs_gradient(fs_space2* sp, const double* d_spec, double* d_ortho, int d0, int d1, double scale0,
                double scale1, int batch); /* space2.rs:247-264 */

/* ------------------------------------------------------- SolverField (batched solves) */
typedef struct rbc_solver rbc_solver;
/* SolverField variants (src/solver.rs:88-95) */
#define RBC_SOLVER_HHOLTZ_ADI 0 /* HholtzAdi::new  hholtz_adi.rs:44-63 */
#define RBC_SOLVER_HHOLTZ 1     /* Hholtz::new     hholtz.rs:42-76     */
#define RBC_SOLVER_POISSON 2    /* Poisson::new    poisson.rs:50-90    */
/* Solver of `solver` kind built from the field Space2(base0(n0), base1(n1)) with
 * constants c = [c0, c1] (axis 1 cheb_dirichlet or cheb_neumann). */
int rbc_solver_create(int solver, int kind0, int n0, int kind1, int n1, double c0, double c1, rbc_solver** out);
int rbc_solver_destroy(rbc_solver* h);
/* input shape [o0, o1] (orthogonal space), output [s0, s1]; complex if base0 is Fourier */
int rbc_solver_shape(rbc_solver* h, int* o0, int* o1, int* s0, int* s1, int* is_complex);
/* Solve::solve (hholtz.rs:159-197, hholtz_adi.rs:106-131, poisson.rs:122-150) of `batch`
 * contiguous right-hand sides on DEVICE buffers (reference layouts).  Synchronous. */
int rbc_solver_solve(rbc_solver* h, const double* d_rhs, double* d_out, int batch);
/* Mixed precision for solvers whose axis 0 is Chebyshev (Hholtz, Poisson; no counterpart in
 * the reference, which solves in FP64): iters > 0 runs the eigen-transforms fwd = Q^-1 C_x^-1
 * and bwd = Q (fdma_tensor.rs:195-234) as FP32 matrix-core GEMMs with FP64 row solves, and
 * chains iters such solves as iterative refinement against the FP64 residual of the tensor
 * equation (A_x (x) C_y + C_x (x) (A_y + alpha C_y)) u = P_x f P_y^T (fdma_tensor.rs:16-70,
 * all matrices banded); iters = 0 restores the FP64 solve.  rbc_solver_residual gives the
 * max-norm of that residual for solutions d_sol of right-hand sides d_rhs, relative to the
 * max-norm of P_x f P_y^T (synchronous). */
int rbc_solver_set_mixed(rbc_solver* h, int iters);
int rbc_solver_residual(rbc_solver* h, const double* d_rhs, const double* d_sol, int batch, double* out);
/* rbc_solver_solve without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code that must not call hipMalloc, such as
 * tight serving loops or harnesses that forbid it).  The work arrays live in a fixed
 * per-device scratch of RBC_NOALLOC_SCRATCH_BYTES that the library holds until the process
 * ends: 2 W + (Fourier axis 0: I + O) doubles, W = batch * max(o0', s0') * o1 (o0', s0' the
 * row counts 2 o0, 2 s0 of complex arrays), I = batch * o0' * o1, O = batch * s0' * s1, each
 * rounded up to a multiple of 32; a call that needs more returns RBC_ERR_ARG.  Results are
 * those of rbc_solver_solve, bit for bit.  Calls on one GPU share its scratch: calls that
 * overlap on one GPU (e.g. from two threads) race.  Synchronous.  Refused (RBC_ERR_ARG) while
 * the mixed form (rbc_solver_set_mixed) is on. */
#define RBC_NOALLOC_SCRATCH_BYTES (32ull << 20)
int rbc_solver_solve_noalloc(rbc_solver* h, const double* d_rhs, double* d_out, int batch);
/* the pub FdmaTensor fields lam/fwd/bwd of a Hholtz/Poisson solver (fdma_tensor.rs:73-91) */
int rbc_solver_tensor_info(rbc_solver* h, int* n, int* has_fb);
int rbc_solver_get_tensor(rbc_solver* h, double* lam, double* fwd, double* bwd);
int rbc_solver_set_tensor(rbc_solver* h, const double* lam, const double* fwd, const double* bwd);

/* ----------------------------------------------------------------- ensembles
 * Many independent confined instances advanced together (no counterpart in the reference,
 * which runs one Navier2D per process; each instance is exactly Navier2D::new with its own
 * parameters).  Instances with equal (nx, ny, pr, dt, aspect, adiabatic) are advanced by
 * one set of batched launches; Ra (and the seeded fields) may differ per instance.
 * rbc_ensemble_instance returns a handle owned by the ensemble: every per-handle entry
 * point (fields, forward/backward, set_velocity, diagnostics, tensors) applies to that
 * instance; rbc_update and rbc_destroy refuse it (use rbc_ensemble_update / _destroy). */
typedef struct rbc_desc {
  int nx, ny;
  double ra, pr, dt, aspect;
  int adiabatic;
} rbc_desc;
typedef struct rbc_ensemble rbc_ensemble;
int rbc_create_ensemble(int n, const rbc_desc* desc, rbc_ensemble** out);
int rbc_ensemble_destroy(rbc_ensemble* h);
int rbc_ensemble_size(rbc_ensemble* h, int* n, int* groups);
int rbc_ensemble_instance(rbc_ensemble* h, int inst, rbc_navier** out);
/* Integrate::update of every instance, nsteps times (blocks until done). */
int rbc_ensemble_update(rbc_ensemble* h, int nsteps);
/* rbc_set_option for every instance of the ensemble (rbc_set_option refuses ensemble
 * instances); RBC_OPT_MIXED selects the mixed-precision confined step for all groups. */
int rbc_ensemble_set_option(rbc_ensemble* h, int option, int value);
int rbc_ensemble_synchronize(rbc_ensemble* h);
/* Kernel timing of all groups (as rbc_timing_enable / rbc_timing_read of a handle, summed
 * over the groups; no counterpart in the reference). */
int rbc_ensemble_timing_enable(rbc_ensemble* h, int on);
int rbc_ensemble_timing_read(rbc_ensemble* h, int kernel_class, long long* launches, double* total_ms, double* work);

/* ------------------------------------------------------------ device memory helpers
 * For callers without their own HIP runtime binding (e.g. ctypes). */
int rbc_dev_malloc(void** out, unsigned long long bytes);
int rbc_dev_free(void* p);
int rbc_memcpy_h2d(void* dst, const void* src, unsigned long long bytes);
int rbc_memcpy_d2h(void* dst, const void* src, unsigned long long bytes);
int rbc_memcpy_d2d(void* dst, const void* src, unsigned long long bytes);

#ifdef __cplusplus
}
#endif
#endif /* RUSTPDE_AMD_H */
