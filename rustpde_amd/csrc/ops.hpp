// Launch wrappers of the HIP kernels of the RBC step (ops.hip).  All device arrays are
// FP64 and row-major.  Complex spectral fields are stored as split planes: a complex
// [m, c] array is the real array [2m, c] whose first m rows hold the real parts.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

namespace rbc {

// Live kernel timing: when armed (per host thread), the next RBC_LAUNCH records the two
// events at the start and the end of that kernel's dispatch (hipExtLaunchKernelGGL), so
// the measured duration is the kernel's own, as a profiler's kernel trace reports it.
struct LaunchTiming {
  hipEvent_t start = nullptr, stop = nullptr;
};
LaunchTiming& launch_timing();

#define RBC_LAUNCH(kernel, grid, block, shm, stream, ...)                                       \
  do {                                                                                         \
    ::rbc::LaunchTiming& lt_ = ::rbc::launch_timing();                                         \
    if (lt_.start) {                                                                           \
      hipEvent_t a_ = lt_.start, b_ = lt_.stop;                                                \
      lt_.start = lt_.stop = nullptr;                                                          \
      hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, a_, b_, 0, __VA_ARGS__);         \
    } else {                                                                                   \
      hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);                       \
    }                                                                                          \
  } while (0)

// C[M,N] = alpha * (A[M,K] B[K,N]) (.* U if U) + beta * C
void gemm(hipStream_t s, int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C,
          int ldc, double alpha, double beta, const double* U = nullptr, int ldu = 0);

// Same product on the FP64 matrix cores (v_mfma_f64_16x16x4f64), tile shape per launch shape
// picked by an autotuner.
void gemm_mfma(hipStream_t s, int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C,
               int ldc, double alpha, double beta, const double* U = nullptr, int ldu = 0);

// While a GemmNoTune object lives on the calling thread, the GEMMs take the tile shape
// already measured for their launch shape, or the rule-of-thumb shape, and never time
// variants (the timing allocates a scratch output and events).  Every tile shape gives the
// same result bits.
struct GemmNoTune {
  GemmNoTune();
  ~GemmNoTune();
  GemmNoTune(const GemmNoTune&) = delete;
  GemmNoTune& operator=(const GemmNoTune&) = delete;
};

// Batch-index moduli of A, B, U (0: none): entry z reads A + sA (z % a) etc.
// Entries z >= split read A from a2 and B from b2 instead (when given; same types, strides
// and moduli): two operators in one launch.
// zero00: every entry's C(0, 0) is set to 0 after the product (the pseudo-pressure pin).
struct BatchMod {
  int a = 0, b = 0, u = 0;
  int split = 0;
  const void* a2 = nullptr;
  const void* b2 = nullptr;
  int zero00 = 0;
};
// Batched: operand b of batch index z starts at A + z sA, B + z sB, C + z sC, U + z sU
// (with the moduli and the operator switch of bm applied to A, B, U).
void gemm_mfma_batched(hipStream_t s, int M, int N, int K, const double* A, int lda, long sA, const double* B, int ldb,
                       long sB, double* C, int ldc, long sC, int batch, double alpha, double beta,
                       const double* U = nullptr, int ldu = 0, long sU = 0, BatchMod bm = BatchMod{});
// The same product in FP32 on the matrix cores (v_mfma_f32_16x16x4f32): A and U FP32, B FP32
// or FP64 (b_f64, converted as it is staged), C FP32 or FP64 (c_f64).
void gemm_mfma_f32_batched(hipStream_t s, int M, int N, int K, const float* A, int lda, long sA, const void* B,
                           bool b_f64, int ldb, long sB, void* C, bool c_f64, int ldc, long sC, int batch, double alpha,
                           double beta, const float* U, int ldu, long sU, BatchMod bm = BatchMod{});

// Banded 4-diagonal solves (offsets -2, 0, 2, 4; pre-swept as fdma.rs:73-82) of every
// row r of X[nrows, m] with diagonal set (r % nsets); diag arrays are [nsets, m].
void fdma_rows(hipStream_t s, double* X, int ld, int nrows, int m, int nsets, const double* low,
               const double* dia, const double* up1, const double* up2);

// Same solves, one wavefront per row with a scan over slice maps (rows up to 2048).
bool fdma_rows_wave_supported(int m);
void fdma_rows_wave(hipStream_t s, double* X, int ld, int nrows, int m, int nsets, const double* low,
                    const double* dia, const double* up1, const double* up2, int batch = 1, long sx = 0);

// Banded operator with taps at input offsets {0, -2, +2, +4} (summed in that order),
// coefficients coef[k * mo + i] and presence bits mask[i] (bit k: tap k); mo outputs from mi
// inputs, taps the union of the presence bits.  out = op(in) [+ beta out].
struct BandRef {
  const double* coef = nullptr;
  const unsigned char* mask = nullptr;
  int mo = 0, mi = 0, taps = 0;
};
// band_y acts along axis 1 of in[rows, *] (leading dims ldi/ldo), band_x along axis 0
// of in[*, cols].
void band_y(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, const BandRef& b, double beta,
            int batch = 1, long si = 0, long so = 0);
void band_x(hipStream_t s, const double* in, double* out, int cols, const BandRef& b, double beta, int batch = 1,
            long si = 0, long so = 0);
// band_x (x stencil bx) then band_y (y stencil by, input row length by.mi at stride ldi) in
// one launch with the two passes' arithmetic; out [bx.mo, by.mo].  Epilogue (same layout as
// out): + add when given, or with p0 the pressure update p0 = (p0 - div nu) + result inv_dt
// (nu = nuarr[batch entry] when given; out unused).
// Batch entries >= split (when split > 0) use the second x stencil x2.
struct BandXYEpi {
  const double* add = nullptr;
  double* p0 = nullptr;
  const double* div = nullptr;
  const double* nuarr = nullptr;
  double nu = 0.0, inv_dt = 0.0;
  BandRef x2{};
  int split = 0;
};
void band_xy(hipStream_t s, const double* in, int ldi, double* out, const BandRef& bx, const BandRef& by,
             const BandXYEpi& ep, int batch, long si, long so);

// Chebyshev d/dy of every row of in[rows, *] (ortho length n), with the composite
// stencil of band coef/mask fused on load (coef == nullptr: input already ortho);
// out = d / sc [+ beta out].
bool cheb_diff_y_supported(int n);
void cheb_diff_y(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                 const double* coef, const unsigned char* mask, double inv_sc, double beta, int batch = 1, long si = 0,
                 long so = 0);
// from_ortho of every row (composite length m): d = S^T p then the tridiagonal solve with
// the precomputed w and den of linalg.rs:14-57.
bool tdma_y_supported(int m);
void tdma_y(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, int m, const double* diag,
            const double* low2, const double* off, const double* w, const double* den, int batch = 1, long si = 0,
            long so = 0);

// Fourier derivative on split planes: X[2m, cols] -> Y, (i k)^times / sc^times
// (r2c.rs:88-99; num-complex division by (sc, 0) as space2.rs:256-262).
void fourier_diff(hipStream_t s, const double* X, double* Y, int m, int cols, int ld, int times, double sc);

// Element-wise combinations of the step (n = number of doubles, all contiguous)
void ew_add(hipStream_t s, double* out, const double* a, const double* b, long n);          // a + b
void ew_rhs_vel(hipStream_t s, double* out, const double* a, const double* g, const double* buoy,
                const double* c, double dt, long n);
// the same for a field pair: entry 0 (no buoyancy) at offset 0, entry 1 (+ buoy dt) at po
void ew_rhs_vel_pair(hipStream_t s, double* out, const double* a, const double* g, const double* buoy, const double* c,
                     double dt, long n, long po);  // ((a - g dt) [+ buoy dt]) - c dt
void ew_rhs_temp(hipStream_t s, double* out, const double* a, const double* bx, const double* by,
                 const double* c, double dt, long n);  // ((a + bx) + by) - c dt
void ew_acc_prod(hipStream_t s, double* out, const double* a, const double* b, long n);     // out + a b
void ew_acc_prod2(hipStream_t s, double* out, const double* a, const double* b, const double* c, const double* d,
                  long n);  // (out + a b) + c d
void ew_acc_prod2_f32(hipStream_t s, float* out, const float* a, const double* b, const float* c, const double* d,
                      long n);  // the same in FP32 (b, d rounded to FP32)
// The right-hand sides of ux, uy, temp in slots 0, 1, 2 (slot stride F) in one launch:
// slot 0 ((a - g dt)) - c dt, slot 1 ((a - g dt) + buoy dt) - c dt, slot 2 ((a + bx) + by) - c dt;
// buoy_bc: buoy holds to_ortho(fieldbc) and the buoyancy is a's slot 2 + buoy
void ew_rhs3(hipStream_t s, double* out, const double* a, const double* g, const double* buoy, const double* c,
             const double* bx, const double* by, double dt, long n, long F, bool buoy_bc = false);
// The same with the to_ortho stencil of ux, uy, temp (U: slots of [2 mx, my], taps
// stencil[0, ny) and [ny, 2 ny)), the buoyancy temp + fieldbc (thbc) and d/dx pres (from P0,
// Fourier x) formed on the fly; g1 = d/dy pres; all other arrays [2 mx, ny], slot stride F.
// thbc, gxx, gyy are read in rows < bcr only: every later row must hold +0.0; cv is read for
// wavenumbers < cutx and y-modes < cuty only (zero beyond: the dealiasing cuts).
void rhs3_fused(hipStream_t s, const double* U, int my, const double* stencil, const double* P0, const double* g1,
                const double* thbc, const double* cv, const double* gxx, const double* gyy, double* out, int mx, int ny,
                double kscale, double dt, long F, int bcr, int cutx, int cuty);
// the same with FP32 convective terms (the mixed-precision step)
void rhs3_fused_cv32(hipStream_t s, const double* U, int my, const double* stencil, const double* P0, const double* g1,
                     const double* thbc, const float* cv, const double* gxx, const double* gyy, double* out, int mx,
                     int ny, double kscale, double dt, long F, int bcr, int cutx, int cuty);
void ew_project(hipStream_t s, double* u, const double* f, double c, long n);              // f (-c) + u
// ((((a - g dt) + c dt) + b dt) + (lx dt) k) + (ly dt) k; g, b may be null
void ew_adj_rhs(hipStream_t s, double* out, const double* a, const double* g, const double* c, const double* b,
                const double* lx, const double* ly, double dt, double k, long n);
void ew_residual(hipStream_t s, double* out, const double* a, const double* b, double dt, long n);  // (a - b) / dt
void ew_axpy(hipStream_t s, double* out, const double* a, const double* b, double f, long n);       // a + b f
void ew_update_pres(hipStream_t s, double* p, const double* div, const double* phi, double nu, double inv_dt,
                    long n);  // (p - div nu) + phi inv_dt
void ew_scale(hipStream_t s, double* out, const double* a, double f, long n);              // a f
// x[i0] = x[i1] = 0 (n_idx = 1: only x[i0]), in each of `batch` slots `stride` apart
void set_zero_entries(hipStream_t s, double* x, int n_idx, long i0, long i1, int batch = 1, long stride = 0);
// pressure update of an ensemble group: nu of entry i is nu[i / stride]
void ew_update_pres_group(hipStream_t s, double* p, const double* div, const double* phi, const double* nu,
                          double inv_dt, long n, long stride);

// Weighted column sums: out[j] = sum_i v[i, j] * w[i]  (i sequential, as average.rs:25-34)
void col_weighted_sum(hipStream_t s, const double* v, int n0, int n1, const double* w, double* out);
// out[0] = sum_j a[j] * w[j] if a, out[1] = sum of squares of b[0..nb) if b; out must
// hold >= 258 doubles (out[2..258) receives per-block partial sums)
void dot_and_norm(hipStream_t s, const double* a, const double* w, int n, const double* b, long nb, double* out);
// v = sqrt(a^2 + b^2) * f
void ew_solid_damp(hipStream_t s, double* conv, const double* mask, const double* q, const double* bc,
                   const double* val, double c, long n);
void ew_running_avg(hipStream_t s, double* avg, const double* x, double w, long n);
void ew_speed(hipStream_t s, double* out, const double* a, const double* b, double f, long n);
// v = ((d + uyt / ka) * 2) * sy   with uyt = t * uy
void ew_nuvol(hipStream_t s, double* out, const double* d, const double* t, const double* uy, double ka, double sy,
              long n);

// Batched real FFTs of length n (power of two) along axis 0 of a [n, cols] array:
// rfft_cols: real -> split planes [2(n/2+1), cols], rows k >= cut set to zero;
// irfft_cols: split planes -> real, with X_k first multiplied by (i k)^times / sc.
bool fft_supported(int n);
// FP32 variants for the mixed-precision convective terms: irfft from FP64 spectra to FP32
// values, rfft from FP32 values to FP64 spectra (arithmetic in FP32).
void rfft_cols_f32(hipStream_t s, const float* X, double* Y, int n, int cols, int cut, int batch = 1, long sx = 0,
                   long sy = 0);
// nzc >= 0: the columns j >= nzc are zero on input (a y-dealiasing cut); their blocks only
// write the zero spectra
void rfft_cols_f32f(hipStream_t s, const float* X, float* Y, int n, int cols, int cut, int batch, long sx, long sy,
                    int nzc = -1);
void irfft_cols_f32(hipStream_t s, const double* X, float* Y, int n, int cols, int times, double sc, int batch = 1,
                    long sx = 0, long sy = 0);
void rfft_cols(hipStream_t s, const double* X, double* Y, int n, int cols, int cut, int batch = 1, long sx = 0,
               long sy = 0, int nzc = -1);
void irfft_cols(hipStream_t s, const double* X, double* Y, int n, int cols, int times, double sc, int batch = 1,
                long sx = 0, long sy = 0);
// values (entries [0, nin): times 0, sc 1 -> Y0) and (i k)^times1 / sc1 (-> Y1) of the same
// nin inputs in one launch of 2 nin batch entries
void irfft_cols_pair(hipStream_t s, const double* X, double* Y0, double* Y1, int n, int cols, int times1, double sc1,
                     int nin, long sx, long sy);
void irfft_cols_pair_f32(hipStream_t s, const double* X, float* Y0, float* Y1, int n, int cols, int times1, double sc1,
                         int nin, long sx, long sy);

// Banded matvec (taps at offsets 0, -2, +2, +4) then the row solves of fdma_rows, one
// wavefront per row, R [nrows, nr] -> X [nrows, m]; `bandp` and `diagp` hold the taps and
// the swept diagonal sets (low, 1 / dia, up1, up2) in slice order (fdma_rows_lds_per(m)
// values per lane and chain).  Without a band R must hold m values per row (in place allowed).
int fdma_rows_lds_per(int m);
bool fdma_rows_lds_supported(int m, int nr);
void fdma_rows_lds(hipStream_t s, const double* R, int ldr, int nr, double* X, int ld, int nrows, int m, int nsets,
                   const double* bandp, const double* diagp, int batch = 1, long sr = 0, long sx = 0);
// The same for up to 4 batch entries with tables of their own (entry k: bandps[k], diagps[k]).
void fdma_rows_lds_multi(hipStream_t s, const double* R, int ldr, int nr, double* X, int ld, int nrows, int m,
                         int nsets, const double* const* bandps, const double* const* diagps, int batch, long sr,
                         long sx);

// from_ortho of every row as tdma_y, one wavefront per row with LDS staging and the
// tables (diag, low2, off_{i-2}, 1 / den, w) in slice order: R [rows, m + 2] -> X [rows, m].
bool tdma_rows_lds_supported(int m);
// With project the result x is subtracted instead: X = x (-1) + X.
void tdma_rows_lds(hipStream_t s, const double* R, int ldr, double* X, int ld, int rows, int m, const double* tabp,
                   int batch = 1, long sr = 0, long sx = 0, bool project = false);

// The pressure stage of the periodic step fused per x-wavenumber (one workgroup per k):
// divergence, Poisson rows (pseu(0,0) = 0), projection of ux and uy, pressure update.
// Fields are split-plane spectral arrays: ux, uy [2 mx, mu], pres [2 mx, ny], pseu
// [2 mx, mu]; su / sp are the stencil taps [2][ny] of the velocity / pseudo-pressure y
// bases; pybp / pdiag the Poisson solver's P_y taps and diagonal sets and ttab the tdma
// tables of the velocity y basis, all in slice order.  per and bl are set by the call.
struct PressureArgs {
  double *UX, *UY, *P0, *P1;
  const double *su, *sp, *pybp, *pdiag, *ttab;
  int mx, ny, mu, per, bl;
  double kscale, inv_scy, nu, inv_dt;
  double isc = 0.0;  // 1 / kscale when kscale is a power of two (set by pressure_rows)
  // stencils of the velocity and pressure bases: 0 from su / sp, 1 Dirichlet, 2 Neumann
  // (formed in the kernel; the tables' values are the same)
  int su_kind = 0, sp_kind = 0;
  // when set: d/dy of the updated pres (Chebyshev derivative / scale_y, rows as P0) for the
  // next step's right-hand sides
  double* dP = nullptr;
  // when rows_nsets > 0: rows_nsets further blocks of the launch solve the banded rows of
  // another field, as fdma_rows_lds with one block per set (2 rows_nsets rows of rows_nr
  // right-hand side values at stride rows_ldr, rows_m results at stride rows_ld); their
  // arrays must not overlap the stage's
  const double* rows_R = nullptr;
  double* rows_X = nullptr;
  const double *rows_band = nullptr, *rows_diag = nullptr;
  int rows_ldr = 0, rows_nr = 0, rows_ld = 0, rows_m = 0, rows_nsets = 0;
  int rows_per = 0, rows_bl = 0;  // set by pressure_rows
};
bool pressure_rows_supported(int ny, int mu);
void pressure_rows(hipStream_t s, const PressureArgs& a);

// Chebyshev transforms along axis 1 (rows of length n, n - 1 a power of two in [16, 1024])
// by FFT-based DCT-I.  cheb_y_backward: in [rows, mi] coefficients (composite basis when
// coef / mask give its stencil taps, else Chebyshev, mi <= n), `times` (0 or 1)
// derivatives scaled by inv_sc, out [rows, n] = x .* U + beta out (U may be null).
// cheb_y_forward: in [rows, n] values -> out [rows, n] Chebyshev coefficients, k >= cut zero.
bool cheb_y_supported(int n);
// Optional extra products on batch entry xslot: out += e1 f1, then out += e2 f2 ([rows, ldu],
// shared by the batch).
void cheb_y_backward(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                     const double* coef, const unsigned char* mask, int times, double inv_sc, const double* U, int ldu,
                     double beta, int batch = 1, long si = 0, long so = 0, long su = 0, const double* e1 = nullptr,
                     const double* f1 = nullptr, const double* e2 = nullptr, const double* f2 = nullptr,
                     int xslot = -1);
void cheb_y_forward(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, int n, int cut,
                    int batch = 1, long si = 0, long so = 0);
// FP32 variants (FP32 in, arithmetic and out; tables in FP64, rounded on load)
void cheb_y_backward_f32(hipStream_t s, const float* in, int ldi, int mi, float* out, int ldo, int rows, int n,
                         const double* coef, int times, double inv_sc, const float* U, int ldu, double beta,
                         int batch = 1, long si = 0, long so = 0, long su = 0, const float* e1 = nullptr,
                         const double* f1 = nullptr, const float* e2 = nullptr, const double* f2 = nullptr,
                         int xslot = -1);
void cheb_y_forward_f32(hipStream_t s, const float* in, int ldi, float* out, int ldo, int rows, int n, int cut,
                        int batch = 1, long si = 0, long so = 0);
// cheb_y_diff: d/dy in Chebyshev space (ortho.rs:107-125) with the composite stencil
// fused on load: out [rows, n] = d / sc + beta out, by the same LDS suffix scans.
// convective term of `batch` fields from their x-derivatives and values (x physical, y composite
// spectral, stencil coef): out = d/dy(val) .* uy + dx .* ux, one complex transform per row pair
// (f1, f2 with row stride ldf: -1 = ldu, 0 = one row for all; f1 null: no x term)
void cheb_y_conv(hipStream_t s, const double* dx, const double* val, int ldi, int mi, double* out, int ldo, int rows,
                 int n, const double* coef, double inv_sc, const double* ux, const double* uy, int ldu, int batch,
                 long si, long so, const double* e1 = nullptr, const double* f1 = nullptr, const double* e2 = nullptr,
                 const double* f2 = nullptr, int xslot = -1, long ldf = -1);
void cheb_y_conv_f32(hipStream_t s, const float* dx, const float* val, int ldi, int mi, float* out, int ldo, int rows,
                     int n, const double* coef, double inv_sc, const float* ux, const float* uy, int ldu, int batch,
                     long si, long so, const float* e1 = nullptr, const double* f1 = nullptr,
                     const float* e2 = nullptr, const double* f2 = nullptr, int xslot = -1, long ldf = -1);
// the y-stage of the three FP32 convective terms (ux, uy, temp in slots 0, 1, 2 of dx / val,
// stride si; Dirichlet y basis, mi = n - 2 coefficients) in one launch: the physical
// velocity from val's slots 0, 1, then out_f = d/dy(val_f) .* uy + dx_f .* ux (+ ux f1 +
// uy f2 for temp: the fieldbc gradient, row stride ldf, 0 when constant in x; f1 null when
// zero) and their dealiased forward transforms (rows k >= cut zero) into out [3 slots of
// stride so, rows, n]
void cheb_y_conv3_f32(hipStream_t s, const float* dx, const float* val, int ldi, int mi, float* out, int ldo, int rows,
                      int n, double inv_sc, const double* f1, const double* f2, long ldf, int cut, long si, long so);
void cheb_y_diff(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                 const double* coef, double inv_sc, double beta, int batch = 1, long si = 0, long so = 0);

// complex layout conversion: interleaved (re, im) [m, c] <-> split planes [2m, c]
void interleaved_to_split(hipStream_t s, const double* in, double* out, int m, int c);
void split_to_interleaved(hipStream_t s, const double* in, double* out, int m, int c);

}  // namespace rbc
