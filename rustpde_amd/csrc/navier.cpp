// Navier2D on one MI355X: device state, operators and the time step, plus the C ABI.
//
// The step follows Integrate::update (src/navier/navier.rs:737-765) operation by
// operation.  Each 1-D operator along x (axis 0, strided) is applied as a left GEMM
// X[out, in] * A[in, cols]; each operator along y (axis 1, contiguous) as a right GEMM
// A[rows, in] * Y^T[in, out].  Operators are pre-composed on the host (setup.cpp), so a
// 2-D transform is two GEMMs and the convective product u * d(q) is the epilogue of the
// second one.  Complex spectral arrays are split planes [2m, cols].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rustpde_amd.h"
#include "noalloc.hpp"
#include "ops.hpp"
#include "setup.hpp"

namespace rbc {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NumericError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                       \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw rbc::DeviceError(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DBuf {
  double* p = nullptr;
  size_t n = 0;
  bool owned = true;  // false: a view into another buffer
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    n = 0;
    owned = true;
  }
  void view(double* at, size_t count) {
    release();
    p = at;
    n = count;
    owned = false;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    HIPCHECK(hipMalloc(&p, count * sizeof(double)));
    n = count;
    HIPCHECK(hipMemset(p, 0, count * sizeof(double)));
  }
  void upload(const std::vector<double>& v) {
    if (v.size() > n) alloc(v.size());
    HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  std::vector<double> download(size_t count) const {
    std::vector<double> v(count);
    HIPCHECK(hipMemcpy(v.data(), p, count * sizeof(double), hipMemcpyDeviceToHost));
    return v;
  }
};

// Device matrix: x-operators are stored as [out, in], y-operators transposed [in, out].
struct DOp {
  DBuf b;
  DBuf f32;  // FP32 copy for the mixed-precision products (filled on first use)
  int rows = 0, cols = 0;
};

struct DBytes {
  unsigned char* p = nullptr;
  DBytes() = default;
  DBytes(const DBytes&) = delete;
  DBytes& operator=(const DBytes&) = delete;
  ~DBytes() { if (p) (void)hipFree(p); }
  void upload(const std::vector<unsigned char>& v) {
    if (p) (void)hipFree(p);
    HIPCHECK(hipMalloc(&p, std::max<size_t>(v.size(), 1)));
    HIPCHECK(hipMemcpy(p, v.data(), v.size(), hipMemcpyHostToDevice));
  }
};

// Banded operator with taps at offsets {0, -2, +2, +4} (see band_y / band_x)
struct DBand {
  DBuf coef;
  DBytes mask;
  int mo = 0, mi = 0;
  int taps = 0;  // union of the presence bits
  BandRef ref() const { return BandRef{coef.p, mask.p, mo, mi, taps}; }
};
static int tap_union(const std::vector<unsigned char>& mk) {
  int t = 0;
  for (unsigned char b : mk) t |= b;
  return t;
}

// to_ortho stencil of a composite basis (composite_stencil.rs:207-229): n outputs from m inputs
static void make_stencil_band(const Basis& b, DBand& band) {
  const int n = b.n, m = b.m;
  std::vector<double> c(4 * (size_t)n, 0.0);
  std::vector<unsigned char> mk(n, 0);
  for (int i = 0; i < n; ++i) {
    if (i < m) { c[i] = b.diag[i]; mk[i] |= 1; }
    if (i >= 2 && i - 2 < m) { c[(size_t)n + i] = b.low2[i - 2]; mk[i] |= 2; }
  }
  band.mo = n;
  band.mi = m;
  band.taps = tap_union(mk);
  band.coef.upload(c);
  band.mask.upload(mk);
}

// MatVecFdma of an (m x n) matrix (matvec.rs:142-193): taps as the reference's conditions
static void make_matvec_band(const LMat& a, DBand& band, std::vector<double>* taps = nullptr) {
  const int m = a.r;
  std::vector<double> c(4 * (size_t)m, 0.0);
  std::vector<unsigned char> mk(m, 0);
  for (int i = 0; i < m; ++i) {
    c[i] = (double)a(i, i);
    mk[i] |= 1;
    if (i > 1) { c[(size_t)m + i] = (double)a(i, i - 2); mk[i] |= 2; }
    if (i < m - 2) { c[2 * (size_t)m + i] = (double)a(i, i + 2); mk[i] |= 4; }
    if (i < m - 4) { c[3 * (size_t)m + i] = (double)a(i, i + 4); mk[i] |= 8; }
  }
  band.mo = m;
  band.mi = a.c;
  band.taps = tap_union(mk);
  band.coef.upload(c);
  band.mask.upload(mk);
  if (taps) *taps = c;  // absent taps are zero
}

// Per-set arrays in the slice order of fdma_rows_lds: element i = p + 2 t of chain p,
// t = lane * per + q, stored at [array][p][q][lane]; `get(a, i, t, p)` gives the value of
// array a for element i (zero for absent entries).
template <class F>
static std::vector<double> slice_order(int nsets, int m, int narrays, F get) {
  const int per = fdma_rows_lds_per(m), len[2] = {(m + 1) / 2, m / 2};
  // tail padding: the kernel loads q < 16 for every array of every set
  std::vector<double> out((size_t)nsets * narrays * 2 * per * 64 + 16 * 64, 0.0);
  for (int s = 0; s < nsets; ++s)
    for (int a = 0; a < narrays; ++a)
      for (int p = 0; p < 2; ++p)
        for (int q = 0; q < per; ++q)
          for (int l = 0; l < 64; ++l) {
            const int t = l * per + q;
            if (t >= len[p]) continue;
            out[(size_t)s * narrays * 2 * per * 64 + ((size_t)(a * 2 + p) * per + q) * 64 + l] = get(s, a, p + 2 * t, t, p);
          }
  return out;
}

// from_ortho tables of a composite basis: stencil and the tdma factors of linalg.rs:14-57
struct DTdma {
  DBuf diag, low2, off, w, den;
  DBuf tabp;  // diag, low2, off_{i-2}, 1 / den, w in the slice order of tdma_rows_lds
  int m = 0;
};
static void make_tdma(const Basis& b, DTdma& t) {
  const int m = b.m;
  std::vector<double> mainv(m), off(std::max(m - 2, 1)), w(m, 0.0), den(m, 0.0);
  for (int i = 0; i < m; ++i) mainv[i] = b.diag[i] * b.diag[i] + b.low2[i] * b.low2[i];
  for (int i = 0; i < m - 2; ++i) off[i] = b.diag[i + 2] * b.low2[i];
  w[0] = off[0] / mainv[0];
  den[0] = mainv[0];
  if (m - 2 > 1) w[1] = off[1] / mainv[1];
  den[1] = mainv[1];
  for (int i = 2; i < m; ++i) {
    den[i] = mainv[i] - off[i - 2] * w[i - 2];
    if (i < m - 2) w[i] = off[i] / den[i];
  }
  t.m = m;
  t.diag.upload(b.diag);
  t.low2.upload(b.low2);
  t.off.upload(off);
  t.w.upload(w);
  t.den.upload(den);
  t.tabp.upload(slice_order(1, m, 5, [&](int, int a, int i, int tt, int) -> double {
    switch (a) {
      case 0: return b.diag[i];
      case 1: return b.low2[i];
      case 2: return tt >= 1 ? off[i - 2] : 0.0;
      case 3: return 1.0 / den[i];
      default: return i < m - 2 ? w[i] : 0.0;
    }
  }));
}

static ld ipow(ld x, int e) {
  ld r = 1.0L;
  for (int i = 0; i < e; ++i) r *= x;
  return r;
}

static LMat build_axis_op(const Basis& b, const std::string& name, ld sc) { return axis_operator(b, name, sc); }

// Host application of axis operators in long double (set-up of constant fields).
static LMat host_fourier_diff(const LMat& in, int m, int times, ld sc) {
  LMat out = in;
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < in.c; ++j) {
      ld a = in(k, j), b = in(m + k, j);
      for (int t = 0; t < times; ++t) {
        ld na = -b * (ld)k, nb = a * (ld)k;
        a = na;
        b = nb;
      }
      out(k, j) = a / ipow(sc, times);
      out(m + k, j) = b / ipow(sc, times);
    }
  return out;
}
static LMat host_x(const Basis& b, const std::string& name, const LMat& in, ld sc) {
  if (b.k == FO) {
    if (name == "S" || name == "FO") return in;
    if (name == "DS") return host_fourier_diff(in, b.m, 1, sc);
    if (name == "DDS") return host_fourier_diff(in, b.m, 2, sc);
    if (name == "BDS") return dft_c2r(b.n) * host_fourier_diff(in, b.m, 1, sc);
  }
  return build_axis_op(b, name, sc) * in;
}
static LMat host_y(const Basis& b, const std::string& name, const LMat& in) {
  return in * build_axis_op(b, name, 1.0L).t();
}

class OpCache {
 public:
  OpCache(bool transposed, double scale) : transposed_(transposed), scale_(scale) {}
  const DOp& get(const Basis& b, const std::string& name) {
    std::string key = std::to_string((int)b.k) + ":" + std::to_string(b.n) + ":" + name;
    auto it = ops_.find(key);
    if (it != ops_.end()) return *it->second;
    LMat m = build_axis_op(b, name, (ld)scale_);
    if (transposed_) m = m.t();
    auto op = std::make_unique<DOp>();
    op->rows = m.r;
    op->cols = m.c;
    op->b.upload(m.to_double());
    const DOp& ref = *op;
    ops_[key] = std::move(op);
    return ref;
  }

 private:
  bool transposed_;
  double scale_;
  std::map<std::string, std::unique_ptr<DOp>> ops_;
};

static void upload_op(DOp& op, const LMat& m) {
  op.rows = m.r;
  op.cols = m.c;
  op.b.upload(m.to_double());
}

// ------------------------------------------------------------------------- solvers
enum SKind { ADI = 0, HHOLTZ = 1, POISSON = 2 };

struct Solver {
  SKind kind;
  Basis bx, by;
  int nx_in = 0;     // ortho rows of the rhs (confined) or 2m (periodic)
  // ADI: u = (My^-1 Py) o (Mx^-1 Px) f  (hholtz_adi.rs:106-131)
  DOp hx, hy;
  // tensor (hholtz.rs:159-197, fdma_tensor.rs:195-234)
  bool has_fb = false;
  std::vector<double> lam, fwd, bwd;  // fwd/bwd m x m (row-major) when has_fb
  std::vector<double> px;             // P_x (m x n) in double, for composing fwd * P_x
  std::vector<double> ax, cx;         // A_x, C_x (m x m) when axis 0 is not diagonal
  DOp pf, pb, py;
  DBand pyb;  // P_y as a banded matvec
  std::vector<double> pyb_taps;  // its taps on the host, [4][m]
  DBuf low, dia, up1, up2;
  DBuf diagp, pybp;  // the same in slice order for fdma_rows_lds (low, 1 / dia, up1, up2)
  int nsets = 0, m = 0;
  Band4 fa, fc;
  double alpha = 0.0;

  Solver(SKind k, const Basis& b0, const Basis& b1) : kind(k), bx(b0), by(b1) {}

  void build_rows() {  // per-row systems (A_y + (lam_i + alpha) C_y), swept once
    nsets = (int)lam.size();
    m = fa.n;
    std::vector<double> L((size_t)nsets * m, 0.0), D((size_t)nsets * m, 0.0), U1((size_t)nsets * m, 0.0),
        U2((size_t)nsets * m, 0.0);
    for (int i = 0; i < nsets; ++i) {
      const double l = lam[i] + alpha;
      Band4 f = fa.add_scaled(fc, l);
      f.sweep();
      for (int j = 0; j < m - 2; ++j) L[(size_t)i * m + j] = f.low[j];
      for (int j = 0; j < m; ++j) D[(size_t)i * m + j] = f.dia[j];
      for (int j = 0; j < m - 2; ++j) U1[(size_t)i * m + j] = f.up1[j];
      for (int j = 0; j < m - 4; ++j) U2[(size_t)i * m + j] = f.up2[j];
    }
    low.upload(L);
    dia.upload(D);
    up1.upload(U1);
    up2.upload(U2);
    const int len[2] = {(m + 1) / 2, m / 2};
    diagp.upload(slice_order(nsets, m, 4, [&](int st, int a, int i, int t, int p) -> double {
      const size_t o = (size_t)st * m;
      switch (a) {
        case 0: return t >= 1 ? L[o + i - 2] : 0.0;
        case 1: return 1.0 / D[o + i];
        case 2: return t + 1 < len[p] ? U1[o + i] : 0.0;
        default: return t + 2 < len[p] ? U2[o + i] : 0.0;
      }
    }));
    if (!pyb_taps.empty()) {
      const int mo = pyb.mo;
      pybp.upload(slice_order(1, m, 4, [&](int, int a, int i, int, int) { return pyb_taps[(size_t)a * mo + i]; }));
    }
    if (has_fb) {
      const int mm = (int)lam.size();
      LMat F(mm, mm), P((int)px.size() / bx.n, bx.n), Q(mm, mm);
      for (size_t i = 0; i < F.a.size(); ++i) F.a[i] = fwd[i];
      for (size_t i = 0; i < P.a.size(); ++i) P.a[i] = px[i];
      for (size_t i = 0; i < Q.a.size(); ++i) Q.a[i] = bwd[i];
      upload_op(pf, F * P);
      upload_op(pb, Q);
    }
  }
};

static std::vector<double> to_vec(const LMat& m) { return m.to_double(); }

// Build the solver of the field with bases (b0, b1) and constants c (navier.rs:250-265, 410-425)
static std::unique_ptr<Solver> make_solver(SKind kind, const Basis& b0, const Basis& b1, const double c[2]) {
  auto s = std::make_unique<Solver>(kind, b0, b1);
  Ingredients g0 = ingredients(b0), g1 = ingredients(b1);
  if (kind == ADI) {  // hholtz_adi.rs:44-63 : Fdma(mat_a - mat_b c) and MatVecFdma(precond)
    for (int ax = 0; ax < 2; ++ax) {
      const Ingredients& g = ax == 0 ? g0 : g1;
      const int m = g.a.r;
      LMat M(m, m);
      for (size_t i = 0; i < M.a.size(); ++i) M.a[i] = (ld)((double)g.a.a[i] - (double)g.b.a[i] * c[ax]);
      LMat H = inverse(M) * g.p;  // (m x n)
      if (ax == 0) upload_op(s->hx, H);
      else upload_op(s->hy, H.t());
    }
    return s;
  }
  // tensor solvers: A = laplacian (scaled mat_b), C = mass (mat_a)
  auto lap = [&](const Ingredients& g, double ci) {
    std::vector<double> a(g.b.a.size());
    for (size_t i = 0; i < a.size(); ++i)
      a[i] = kind == HHOLTZ ? (-1.0 * (double)g.b.a[i]) * ci : (double)g.b.a[i] * ci;
    return a;
  };
  std::vector<double> A0 = lap(g0, c[0]), A1 = lap(g1, c[1]);
  const int m0 = g0.a.r, m1 = g1.a.r;
  s->alpha = kind == HHOLTZ ? 1.0 : 0.0;
  if (g0.is_diag) {  // fdma_tensor.rs:117-120
    s->has_fb = false;
    s->lam.resize(m0);
    for (int i = 0; i < m0; ++i) s->lam[i] = A0[(size_t)i * m0 + i];
  } else {  // fdma_tensor.rs:121-126 : xmat = inv(C) A ; eig ; fwd = Q^-1 inv(C) ; bwd = Q
    LMat C(m0, m0), Am(m0, m0);
    for (size_t i = 0; i < C.a.size(); ++i) C.a[i] = g0.a.a[i];
    for (size_t i = 0; i < Am.a.size(); ++i) Am.a[i] = A0[i];
    s->ax = A0;
    s->cx.resize(C.a.size());
    for (size_t i = 0; i < C.a.size(); ++i) s->cx[i] = (double)g0.a.a[i];
    LMat Ci = inverse(C);
    std::vector<double> xmat = (Ci * Am).to_double();
    std::vector<double> q, qi;
    std::string msg;
    if (!eig_real(xmat, m0, s->lam, q, qi, msg)) throw NumericError(msg);
    LMat Qi(m0, m0);
    for (size_t i = 0; i < Qi.a.size(); ++i) Qi.a[i] = qi[i];
    s->fwd = (Qi * Ci).to_double();
    s->bwd = q;
    s->has_fb = true;
  }
  if (kind == POISSON && std::fabs(s->lam[0]) < 1e-10)  // poisson.rs:80-83
    for (auto& v : s->lam) v -= 1e-10;
  s->fa = Band4::from_matrix(A1, m1);
  std::vector<double> C1(g1.a.a.size());
  for (size_t i = 0; i < C1.size(); ++i) C1[i] = (double)g1.a.a[i];
  s->fc = Band4::from_matrix(C1, m1);
  if (g0.has_p) s->px = to_vec(g0.p);
  upload_op(s->py, g1.p.t());  // MatVecFdma(precond) along y as a right GEMM
  make_matvec_band(g1.p, s->pyb, &s->pyb_taps);
  s->build_rows();
  return s;
}

// ------------------------------------------------------------------ kernel timing
// HIP events recorded on the handle's stream around each launch of a kernel class, with
// the launch's algorithmic work (flops for GEMMs, bytes for memory-bound kernels).
class KernelTimer {
 public:
  static constexpr int NCAT = 14;  // the RBC_K_* classes of the C ABI
  bool on = false;
  struct Rec { int cat; double work; hipEvent_t a, b; };
  ~KernelTimer() { clear(); for (auto e : pool_) (void)hipEventDestroy(e); }
  // arms the launch hook: the next launch records the events in its own dispatch
  void begin(hipStream_t st, int cat, double work) {
    (void)st;
    if (!on) return;
    Rec r{cat, work, get(), get()};
    launch_timing() = LaunchTiming{r.a, r.b};
    recs_.push_back(r);
  }
  void end(hipStream_t st) {
    if (!on) return;
    LaunchTiming& t = launch_timing();
    if (t.start) {  // nothing was launched: an empty interval
      HIPCHECK(hipEventRecord(t.start, st));
      HIPCHECK(hipEventRecord(t.stop, st));
      t = LaunchTiming{};
    }
  }
  // accumulate finished records; returns per category (launches, ms, work)
  void collect() {
    for (auto& r : recs_) {
      float ms = 0.f;
      HIPCHECK(hipEventSynchronize(r.b));
      HIPCHECK(hipEventElapsedTime(&ms, r.a, r.b));
      n_[r.cat] += 1;
      ms_[r.cat] += ms;
      work_[r.cat] += r.work;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    recs_.clear();
  }
  void read(int cat, long long& n, double& ms, double& work) {
    collect();
    n = n_[cat];
    ms = ms_[cat];
    work = work_[cat];
  }
  void reset() { collect(); for (int i = 0; i < NCAT; ++i) { n_[i] = 0; ms_[i] = 0; work_[i] = 0; } }
  void clear() { recs_.clear(); }

 private:
  hipEvent_t get() {
    if (!pool_.empty()) { hipEvent_t e = pool_.back(); pool_.pop_back(); return e; }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> pool_;
  long long n_[NCAT] = {};
  double ms_[NCAT] = {}, work_[NCAT] = {};
};

// ----------------------------------------------------------------------- Navier2D
class Navier {
 public:
  bool periodic;
  int nx, ny;
  double ra, pr, dt, aspect, nu, ka, time = 0.0;
  double scale[2];
  hipStream_t st = nullptr;
  Basis bx_u, bx_t, bx_p, bx_phi, bx_f, by_u, by_p, by_phi, by_bc;
  OpCache X, Y;
  std::shared_ptr<Solver> solver[4];
  // state
  DBuf T, UX, UY, P0, P1, BC;        // spectral (split planes when periodic)
  DBuf Tv, UXv, UYv, P0v, P1v, BCv;  // physical
  // constants derived from fieldbc
  DBuf THAT_BC, DTBCX, DTBCY, GXX, GYY;
  // the physical fieldbc gradient: d/dx identically zero, d/dy the same in every x row
  bool bc_dx_zero = false, bc_dy_xconst = false;
  // rows of THAT_BC, GXX and GYY (spectral, [rows, ny]) before which every value is +0.0
  int bc_rows = 0;
  // workspaces
  DBuf THAT, UXP, UYP, A, G, CONV, CV, RHS, DIV, W1, W2, W3, W4, W9, RED;
  std::vector<double> wx_dx, wy_dx;  // cell widths of `field` (unscaled, field.rs:135-163)
  double len0 = 0, len1 = 0;
  DBuf DX0, DX1;
  KernelTimer timer;
  bool use_fft = true;
  bool use_mfma = true;   // FP64 GEMMs on the matrix cores
  bool use_graph = true;  // replay update() as a captured HIP graph
  bool use_dct = true;    // Chebyshev y transforms by FFT-based DCT-I kernels
  bool use_batch = true;  // periodic step: the three fields' transforms and solves batched
  bool use_fuse = true;   // periodic step: the pressure stage as one fused kernel
  // periodic fused mixed step: the y-stage of the convective terms in one launch (0 off, 1 when
  // its blocks fill every CU four times, 2 always)
  int use_conv3 = 1;
  // ensemble group: this handle (the group's leader) advances nb instances of the same size
  // whose buffers are slots nb * bs apart; per-instance nu and ADI operators in slabs
  int nb = 1;
  long bs = 0;
  DBuf NUARR;
  DBuf ADIH[3][2];
  long adi_s[3][2] = {};
  // Field batching (confined step): inside a FieldPair scope fb = 2 (or 3) and every launch
  // covers ux and uy (and temp) of every instance, batch entry z = f nb + b (field f,
  // instance b).  Each work buffer of the velocity part then holds partner slots pair_off()
  // apart: the slots of a standalone handle (FS apart), or the parts of the group's slabs.
  int fb = 1;
  int nbt() const { return nb * fb; }                            // batch entries of a launch
  long dstr() const { return nb > 1 ? bs : (long)FS; }           // data stride between entries
  long pair_off() const { return nb > 1 ? (long)nb * bs : (long)FS; }
  long nbn(long n) const { return (long)(nbt() - 1) * dstr() + n; }  // element-wise extent
  // fields = 3: ux, uy and temp (entries 2 nb .. 3 nb - 1 take temp's x operators)
  struct FieldPair {
    Navier& n;
    int prev;
    explicit FieldPair(Navier& nv, int fields = 2) : n(nv), prev(nv.fb) { n.fb = fields; }
    ~FieldPair() { n.fb = prev; }
  };
  bool mixed = false;     // periodic batched step: convective terms in FP32, the rest FP64
  DBuf MIX;               // FP32 slots of the mixed convective pipeline (11 slots of FS floats)
  // confined mixed step: FP32 work buffers of the convective terms, kind k (0: x-transform
  // output, 1: physical ux and uy, 2: convective products, 3: dealiased y-transform) at
  // k * mixc_kind(), each with the partner slots of a field triple
  DBuf MIXC;
  long mixc_kind() const { return 3 * pair_off(); }
  float* mixc(int k) { return reinterpret_cast<float*>(MIXC.p) + (size_t)k * mixc_kind(); }
  void alloc_mixc() {
    const size_t floats = 4 * (size_t)mixc_kind();
    if (MIXC.n < (floats + 1) / 2) MIXC.alloc((floats + 1) / 2);
  }
  bool vel_same = true;   // the ux and uy solvers hold identical tables
  // G's second slot holds d/dy of the current pres: formed by the fused pressure stage of the
  // previous batched step (or zero with pres at the start); a host write of a field or a step
  // on another path clears it, and the next batched step forms it first
  bool gpres_valid = true;
  size_t FS = 0;          // slot stride of the batched buffers
  DBuf STATE, PHYS2;      // [ux, uy, temp] spectral and [ux, uy] physical, one slot each
  hipGraphExec_t gexec = nullptr;    // one captured step
  hipGraphExec_t gexec_k = nullptr;  // GRAPH_STEPS captured steps in one graph
  // Consecutive replays of a graph are separated by its launch (about 8 us between the
  // last kernel of one replay and the first of the next at 1024x1025, against none between
  // the kernels inside one): runs of steps replay GRAPH_STEPS steps per launch.
  static constexpr int GRAPH_STEPS = 8;
  int eager_steps = 0;     // eager steps since the captured step was last dropped
  bool graph_gpres = false;  // the captured step reads d/dy pres left by the step before it
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (gexec_k) (void)hipGraphExecDestroy(gexec_k);
    gexec = nullptr;
    gexec_k = nullptr;
    eager_steps = 0;
  }
  // capture k steps into one graph (the captured launches do not run)
  hipGraphExec_t capture_steps(int k) {
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const double t_saved = time;
    try {
      for (int i = 0; i < k; ++i) update();
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      time = t_saved;
      throw;
    }
    time = t_saved;
    HIPCHECK(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw DeviceError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ex;
  }
  // the one-step graph and the GRAPH_STEPS-step graph; every captured step after the first
  // reads the d/dy pres its predecessor formed, as the one-step graph's replays do
  void capture_step() {
    const bool gp_saved = gpres_valid;
    gexec = capture_steps(1);
    graph_gpres = gpres_valid;  // set by the captured step when its fused pressure stage forms d/dy pres
    gpres_valid = gp_saved;
    try {
      gexec_k = capture_steps(GRAPH_STEPS);
    } catch (...) {
      drop_graph();
      throw;
    }
    gpres_valid = graph_gpres;
  }
  // Integrate::update nsteps times.  After a set-up change the first step runs eagerly (it
  // tunes the GEMM tiles and, after a host write, forms d/dy pres on its own); every later
  // step -- in this call or in a later one, so that one step per call as integrate() makes
  // them reaches it from the second call on -- replays a captured step, GRAPH_STEPS of them
  // per launch while that many remain.
  void update_n(int nsteps) {
    int done = 0;
    if (gexec && graph_gpres && !gpres_valid) drop_graph();  // d/dy pres was not left by a step
    const bool graph = use_graph && !timer.on;
    if (graph && !gexec && eager_steps == 0 && done < nsteps) {
      update();
      ++eager_steps;
      ++done;
    }
    if (graph && !gexec && done < nsteps) capture_step();
    while (done < nsteps) {
      if (graph && gexec) {
        const int k = (gexec_k && nsteps - done >= GRAPH_STEPS) ? GRAPH_STEPS : 1;
        HIPCHECK(hipGraphLaunch(k > 1 ? gexec_k : gexec, st));
        for (int i = 0; i < k; ++i) time += dt;  // as k single steps add it
        done += k;
      } else {
        update();
        ++eager_steps;
        ++done;
      }
    }
  }  // Fourier x transforms by the FFT kernels (dense DFT GEMMs otherwise)

  // instrumented launches
  // sA / sB: DATA for an operand that is a field (one per batch entry), 0 for an operator
  // shared by the batch, or the stride of the group's per-instance operators (indexed by
  // the instance, so ux and uy share them).  U is per instance.
  static constexpr long DATA = -1;
  // A2 / B2: temp's operators for the entries of the third field (fb = 3)
  void GEMM(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc, double beta = 0.0,
            const double* U = nullptr, int ldu = 0, long sA = 0, long sB = 0, const double* A2 = nullptr,
            const double* B2 = nullptr) {
    const int nz = nbt();
    if ((A2 || B2) && fb != 3) throw std::logic_error("GEMM: second operators need fb = 3");
    timer.begin(st, 0, 2.0 * M * N * K * nz);
    if (gemm_zero00 && !A2 && !B2 && (use_mfma || nz > 1)) {  // the product and the pin of C(0, 0)
      const long d = dstr();
      BatchMod bm;
      bm.a = sA > 0 ? nb : 0;
      bm.b = sB > 0 ? nb : 0;
      bm.u = nb;
      bm.zero00 = 1;
      gemm_mfma_batched(st, M, N, K, A, lda, sA == DATA ? d : sA, B, ldb, sB == DATA ? d : sB, C, ldc, d, nz, 1.0,
                        beta, U, ldu, U ? bs : 0, bm);
      timer.end(st);
      return;
    }
    if (gemm_zero00) throw std::logic_error("GEMM: pinned output needs the matrix-core path");
    if (nz > 1) {
      const long d = dstr();
      BatchMod bm;
      bm.a = sA > 0 ? nb : 0;
      bm.b = sB > 0 ? nb : 0;
      bm.u = nb;
      bm.split = (A2 || B2) ? 2 * nb : 0;
      bm.a2 = A2;
      bm.b2 = B2;
      if (use_mfma || nb > 1) {
        gemm_mfma_batched(st, M, N, K, A, lda, sA == DATA ? d : sA, B, ldb, sB == DATA ? d : sB, C, ldc, d, nz, 1.0,
                          beta, U, ldu, U ? bs : 0, bm);
      } else {  // one instance, a field pair or triple, FMA GEMMs
        for (int z = 0; z < nz; ++z) {
          const bool second = bm.split > 0 && z >= bm.split;
          const double* Az = second && A2 ? A2 : A;
          const double* Bz = second && B2 ? B2 : B;
          gemm(st, M, N, K, Az + (sA == DATA ? z * d : 0), lda, Bz + (sB == DATA ? z * d : 0), ldb, C + z * d, ldc,
               1.0, beta, U, ldu);
        }
      }
    } else if (use_mfma) {
      gemm_mfma(st, M, N, K, A, lda, B, ldb, C, ldc, 1.0, beta, U, ldu);
    } else {
      gemm(st, M, N, K, A, lda, B, ldb, C, ldc, 1.0, beta, U, ldu);
    }
    timer.end(st);
  }
  // the same in FP32 (A and U FP32; B and C FP32 or FP64)
  void GEMM32(int M, int N, int K, const float* A, int lda, const void* B, bool b64, int ldb, void* C, bool c64,
              int ldc, double beta, const float* U, int ldu, long sA, long sB, const float* A2 = nullptr) {
    const int nz = nbt();
    const long d = dstr();
    if (A2 && fb != 3) throw std::logic_error("GEMM32: second operators need fb = 3");
    BatchMod bm;
    bm.a = sA > 0 ? nb : 0;
    bm.b = sB > 0 ? nb : 0;
    bm.u = nb;
    bm.split = A2 ? 2 * nb : 0;
    bm.a2 = A2;
    timer.begin(st, 12, 2.0 * M * N * K * nz);
    gemm_mfma_f32_batched(st, M, N, K, A, lda, sA == DATA ? d : sA, B, b64, ldb, sB == DATA ? d : sB, C, c64, ldc, d,
                          nz, 1.0, beta, U, ldu, U ? bs : 0, bm);
    timer.end(st);
  }
  // FP32 copy of an operator (made on first use, outside graph capture: the first step of
  // update_n runs eagerly)
  const float* op32(const DOp& op) {
    DOp& o = const_cast<DOp&>(op);
    if (!o.f32.p) {
      const size_t n = (size_t)o.rows * o.cols;
      std::vector<double> h = o.b.download(n);
      std::vector<float> f(h.begin(), h.end());
      o.f32.alloc((n + 1) / 2);
      HIPCHECK(hipMemcpy(o.f32.p, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    return reinterpret_cast<const float*>(o.f32.p);
  }
  void ROWS(double* X, int ld, int nrows, int m, int nsets, const Solver& s) {
    // bytes: the rows read and written once, the 4 diagonals of every set read once
    timer.begin(st, 3, 8.0 * nbt() * (2.0 * nrows * m + 4.0 * (double)nsets * m));
    if (fdma_rows_wave_supported(m)) {
      fdma_rows_wave(st, X, ld, nrows, m, nsets, s.low.p, s.dia.p, s.up1.p, s.up2.p, nbt(), dstr());
    } else {
      for (int b = 0; b < nbt(); ++b)
        fdma_rows(st, X + b * dstr(), ld, nrows, m, nsets, s.low.p, s.dia.p, s.up1.p, s.up2.p);
    }
    timer.end(st);
  }
  // P_y matvec then the per-row solves: R [nrows, ny] -> X [nrows, my]
  void ROWSF(const double* R, int ldr, double* X, int ld, int nrows, const Solver& s) {
    const int m = s.m, nr = s.pyb.mi;
    if (use_bands && fdma_rows_lds_supported(m, nr)) {
      timer.begin(st, 1, 8.0 * nbt() * ((double)nrows * (nr + m) + 4.0 * (double)s.nsets * m));
      fdma_rows_lds(st, R, ldr, nr, X, ld, nrows, m, s.nsets, s.pybp.p, s.diagp.p, nbt(), dstr(), dstr());
      timer.end(st);
      return;
    }
    BANDY(s.pyb, R, ldr, X, ld, nrows, 0.0);
    ROWS(X, ld, nrows, m, s.nsets, s);
  }
  void BANDY(const DBand& b, const double* in, int ldi, double* out, int ldo, int rows, double beta) {
    timer.begin(st, 9, 8.0 * nbt() * (double)rows * ((double)b.mi + b.mo * (beta != 0.0 ? 2.0 : 1.0)));
    band_y(st, in, ldi, out, ldo, rows, b.ref(), beta, nbt(), dstr(), dstr());
    timer.end(st);
  }
  void BANDX(const DBand& b, const double* in, double* out, int cols, double beta) {
    timer.begin(st, 3, 8.0 * nbt() * (double)cols * ((double)b.mi + b.mo * (beta != 0.0 ? 2.0 : 1.0)));
    band_x(st, in, out, cols, b.ref(), beta, nbt(), dstr(), dstr());
    timer.end(st);
  }
  // cached stencil bands per basis
  std::map<std::string, std::unique_ptr<DBand>> bands;
  const DBand& stencil_band(const Basis& b) {
    std::string key = std::to_string((int)b.k) + ":" + std::to_string(b.n);
    auto it = bands.find(key);
    if (it != bands.end()) return *it->second;
    auto band = std::make_unique<DBand>();
    make_stencil_band(b, *band);
    const DBand& ref = *band;
    bands[key] = std::move(band);
    return ref;
  }
  bool use_bands = true;
  std::map<std::string, std::unique_ptr<DTdma>> tdmas;
  const DTdma& tdma_tables(const Basis& b) {
    std::string key = std::to_string((int)b.k) + ":" + std::to_string(b.n);
    auto it = tdmas.find(key);
    if (it != tdmas.end()) return *it->second;
    auto t = std::make_unique<DTdma>();
    make_tdma(b, *t);
    const DTdma& ref = *t;
    tdmas[key] = std::move(t);
    return ref;
  }
  void FDIFF(const double* X, double* Y, int m, int cols, int ld, int times, double sc) {
    timer.begin(st, 3, 8.0 * 4.0 * m * cols * nbt());
    for (int b = 0; b < nbt(); ++b) fourier_diff(st, X + b * dstr(), Y + b * dstr(), m, cols, ld, times, sc);
    timer.end(st);
  }

  int mx() const { return nx / 2 + 1; }
  int rows_spec(const Basis& b) const { return periodic ? 2 * b.m : b.m; }
  int rows_ortho(const Basis& b) const { return periodic ? 2 * b.m : b.n; }

  // Ensemble groups: `slabs` places every field and work buffer of this instance in slot `slot`
  // of group-owned slabs (slot stride = the buffer size), `ext` is the group's stream and
  // `share` an instance of the same shape whose Poisson solver (Ra-independent) is reused.
  // The field batches' buffers have partner slots B apart: the ux slab holds ux, uy and temp
  // (3 B slots), the physical values ux and uy (2 B), A, CONV, CV, RHS, W1 and W2 three
  // fields' partners (3 B) and G the two pressure gradients (2 B).
  struct Slabs {
    static constexpr int NK = 29;
    size_t F = 0;
    int B = 0;
    DBuf slab[NK];
    static int parts(int k) { return k == 0 || (k >= 5 && k <= 10) ? 3 : (k == 3 || k == 11) ? 2 : 1; }
    static int part_of(int k, int& base) {  // uy, temp: parts 1, 2 of the ux slab; uy values: part 1
      base = k;
      if (k == 1 || k == 2) { base = 0; return k; }
      if (k == 4) { base = 3; return 1; }
      return 0;
    }
    void alloc(size_t f, int b) {
      F = f;
      B = b;
      for (int k = 0; k < NK; ++k) {
        int base = k;
        if (part_of(k, base) == 0) slab[k].alloc(F * (size_t)b * parts(k));
      }
    }
    double* at(int k, int slot) {
      int base = k;
      const int part = part_of(k, base);
      return slab[base].p + ((size_t)part * B + slot) * F;
    }
  };
  bool owns_stream = true;

  Navier(bool per, int nx_, int ny_, double ra_, double pr_, double dt_, double aspect_, bool adiabatic,
         Slabs* slabs = nullptr, int slot = 0, hipStream_t ext = nullptr, const Navier* share = nullptr)
      : periodic(per), nx(nx_), ny(ny_), ra(ra_), pr(pr_), dt(dt_), aspect(aspect_),
        bx_u(per ? FO : CD, nx_), bx_t(per ? FO : (adiabatic ? CN : CD), nx_), bx_p(per ? FO : CHEB, nx_),
        bx_phi(per ? FO : CN, nx_), bx_f(per ? FO : CHEB, nx_), by_u(CD, ny_), by_p(CHEB, ny_), by_phi(CN, ny_),
        by_bc(CDBC, ny_), X(false, aspect_), Y(true, 1.0) {
    if (nx < 8 || ny < 8) throw std::invalid_argument("nx and ny must be >= 8");
    if (per && (nx % 2) != 0) throw std::invalid_argument("periodic nx must be even");
    scale[0] = aspect;
    scale[1] = 1.0;
    // get_nu / get_ka with height 2 (navier.rs:46-55, 231-232)
    nu = std::sqrt(pr / (ra / std::pow(scale[1] * 2.0, 3.0)));
    ka = std::sqrt(1.0 / ((ra / std::pow(scale[1] * 2.0, 3.0)) * pr));
    if (ext) {
      st = ext;
      owns_stream = false;
    } else {
      HIPCHECK(hipStreamCreate(&st));
    }
    const int R = std::max(nx, periodic ? 2 * mx() : nx);
    const size_t full = (size_t)R * ny;
    FS = full;
    if (slabs) {  // ensemble member: one slot of every group slab
      if (per) throw std::invalid_argument("ensemble instances are confined");
      if (slabs->F != full) throw std::invalid_argument("ensemble slab stride mismatch");
      DBuf* all[Slabs::NK] = {&UX, &UY, &T, &UXP, &UYP, &A, &CONV, &CV, &RHS, &W1, &W2, &G, &P0, &P1, &Tv,
                              &UXv, &UYv, &P0v, &P1v, &THAT_BC, &DTBCX, &DTBCY, &GXX, &GYY, &THAT, &DIV, &W3, &W4, &W9};
      for (int k = 0; k < Slabs::NK; ++k) all[k]->view(slabs->at(k, slot), full);
    } else {
      // slot-strided buffers of the batched step (slot k at p + k FS)
      STATE.alloc(3 * full);
      UX.view(STATE.p, full);
      UY.view(STATE.p + full, full);
      T.view(STATE.p + 2 * full, full);
      PHYS2.alloc(2 * full);
      UXP.view(PHYS2.p, full);
      UYP.view(PHYS2.p + full, full);
      for (DBuf* b : {&A, &CONV, &CV, &RHS, &W1, &W2}) b->alloc(3 * full);
      G.alloc(2 * full);
      for (DBuf* b : {&P0, &P1, &Tv, &UXv, &UYv, &P0v, &P1v, &THAT_BC, &DTBCX, &DTBCY, &GXX, &GYY, &THAT, &DIV, &W3,
                      &W4, &W9})
        b->alloc(full);
    }
    BC.alloc((size_t)R * 2);
    BCv.alloc((size_t)nx * ny);
    RED.alloc(std::max(ny, 8) + 300);
    // solvers
    const double cv[2] = {dt * nu / std::pow(scale[0], 2.0), dt * nu / std::pow(scale[1], 2.0)};
    const double ck[2] = {dt * ka / std::pow(scale[0], 2.0), dt * ka / std::pow(scale[1], 2.0)};
    const double cp[2] = {1.0 / std::pow(scale[0], 2.0), 1.0 / std::pow(scale[1], 2.0)};
    const SKind hk = per ? HHOLTZ : ADI;
    solver[0] = make_solver(hk, bx_u, by_u, cv);
    solver[1] = make_solver(hk, bx_u, by_u, cv);
    solver[2] = make_solver(hk, bx_t, by_u, ck);
    solver[3] = share ? share->solver[3] : std::shared_ptr<Solver>(make_solver(POISSON, bx_phi, by_phi, cp));
    // averaging weights of `field` (its coordinates are not rescaled, navier.rs:500-515)
    wx_dx = cell_widths(bx_f.coords(), bx_f.fourier());
    wy_dx = cell_widths(by_p.coords(), false);
    std::vector<double> x0 = bx_f.coords(), x1 = by_p.coords();
    len0 = std::fabs(x0.back() - x0[0]);
    len1 = std::fabs(x1.back() - x1[0]);
    DX0.upload(wx_dx);
    DX1.upload(wy_dx);
    bc_rbc();
  }
  ~Navier() {
    drop_graph();
    if (st && owns_stream) (void)hipStreamDestroy(st);
  }

  static std::vector<double> cell_widths(const std::vector<double>& x, bool periodic) {  // field.rs:135-163
    const int n = (int)x.size();
    std::vector<double> d(n);
    if (periodic) {
      for (int i = 0; i < n; ++i) d[i] = x[2] - x[1];
      return d;
    }
    for (int i = 0; i < n; ++i) {
      double l = i == 0 ? x[0] : (x[i] + x[i - 1]) / 2.0;
      double r = i == n - 1 ? x[n - 1] : (x[i + 1] + x[i]) / 2.0;
      d[i] = r - l;
    }
    return d;
  }

  // -------------------------------------------------------------- axis application
  // x (axis 0): in [rows_in, cols] -> out [rows_out, cols]; identity returns `in`.
  const double* xop(const Basis& b, const std::string& name, const double* in, int cols, double* out, double* tmp) {
    if (b.k == FO && use_fft && fft_supported(b.n)) {  // batched real FFT kernels
      if (name == "S" || name == "FO") return in;
      if (name == "DS") { FDIFF(in, out, b.m, cols, cols, 1, scale[0]); return out; }
      if (name == "DDS") { FDIFF(in, out, b.m, cols, cols, 2, scale[0] * scale[0]); return out; }
      const double work = 8.0 * cols * ((double)b.n + 2.0 * b.m);
      timer.begin(st, (name == "BS" || name == "BDS") ? 2 : 5, work);
      if (name == "BS") irfft_cols(st, in, out, b.n, cols, 0, 1.0);
      else if (name == "BDS") irfft_cols(st, in, out, b.n, cols, 1, scale[0]);
      else if (name == "FS") rfft_cols(st, in, out, b.n, cols, b.m);
      else if (name == "Fdl") rfft_cols(st, in, out, b.n, cols, b.m * 2 / 3);
      else throw std::runtime_error("no Fourier operator " + name);
      timer.end(st);
      return out;
    }
    if (b.k == FO) {
      const int m = b.m;
      if (name == "S" || name == "FO") return in;
      if (name == "DS") { FDIFF(in, out, m, cols, cols, 1, scale[0]); return out; }
      if (name == "DDS") { FDIFF(in, out, m, cols, cols, 2, scale[0] * scale[0]); return out; }
      if (name == "BDS") {
        FDIFF(in, tmp, m, cols, cols, 1, scale[0]);
        const DOp& op = X.get(b, "BS");
        GEMM(op.rows, cols, op.cols, op.b.p, op.cols, tmp, cols, out, cols, 0.0, nullptr, 0, 0, DATA);
        return out;
      }
    } else if (b.k == CHEB && (name == "S" || name == "FO")) {
      return in;
    }
    if (use_bands && name == "S" && (b.k == CD || b.k == CN)) {
      BANDX(stencil_band(b), in, out, cols, 0.0);
      return out;
    }
    const DOp& op = X.get(b, name);
    GEMM(op.rows, cols, op.cols, op.b.p, op.cols, in, cols, out, cols, 0.0, nullptr, 0, 0, DATA);
    return out;
  }
  void xmat(const DOp& op, const double* in, int cols, double* out, const double* per_inst = nullptr,
            long sop = 0) {  // per_inst: the group's operators, sop apart
    GEMM(op.rows, cols, op.cols, per_inst ? per_inst : op.b.p, op.cols, in, cols, out, cols, 0.0, nullptr, 0,
         per_inst ? sop : 0, DATA);
  }
  // y (axis 1): in [rows, cin] -> out [rows, cout]; out = (in Y^T) .* U + beta out
  const double* yop(const Basis& b, const std::string& name, const double* in, int rows, double* out,
                    const double* U = nullptr, double beta = 0.0) {
    if (b.k == CHEB && (name == "S" || name == "FO") && !U && beta == 0.0) return in;
    if (use_bands && name == "S" && !U && (b.k == CD || b.k == CN)) {
      BANDY(stencil_band(b), in, b.m, out, b.n, rows, beta);
      return out;
    }
    if (use_dct && use_bands && name == "DS" && !U && (b.k == CD || b.k == CN || b.k == CHEB) &&
        cheb_y_supported(b.n)) {
      const DBand* band = b.k != CHEB ? &stencil_band(b) : nullptr;
      timer.begin(st, 7, 8.0 * nbt() * (double)rows * ((double)b.m + b.n * (beta != 0.0 ? 2.0 : 1.0)));
      cheb_y_diff(st, in, b.m, b.m, out, b.n, rows, b.n, band ? band->coef.p : nullptr, 1.0 / scale[1], beta, nbt(),
                  dstr(), dstr());
      timer.end(st);
      return out;
    }
    if (use_bands && name == "DS" && !U && (b.k == CD || b.k == CN || b.k == CHEB) && cheb_diff_y_supported(b.n)) {
      const bool comp = b.k != CHEB;
      const DBand* band = comp ? &stencil_band(b) : nullptr;
      timer.begin(st, 3, 8.0 * nbt() * (double)rows * ((double)b.m + b.n * (beta != 0.0 ? 2.0 : 1.0)));
      cheb_diff_y(st, in, b.m, b.m, out, b.n, rows, b.n, comp ? band->coef.p : nullptr, comp ? band->mask.p : nullptr,
                  1.0, beta, nbt(), dstr(), dstr());
      timer.end(st);
      return out;
    }
    if (use_dct && (b.k == CHEB || b.k == CD || b.k == CN) && cheb_y_supported(b.n)) {
      if (name == "BS" || name == "BDS") {
        const DBand* band = b.k == CHEB ? nullptr : &stencil_band(b);
        const int times = name == "BDS" ? 1 : 0;
        // the multiplier U is per instance: one launch per field of a field pair
        const int nf = U ? fb : 1, nz = nbt() / nf;
        const long d = dstr(), po = pair_off();
        for (int f = 0; f < nf; ++f) {
          timer.begin(st, 4,
                      8.0 * nz * (double)rows * ((double)b.m + b.n * (1.0 + (U ? 1.0 : 0.0) + (beta != 0.0 ? 1.0 : 0.0))));
          cheb_y_backward(st, in + f * po, b.m, b.m, out + f * po, b.n, rows, b.n, band ? band->coef.p : nullptr,
                          band ? band->mask.p : nullptr, times, 1.0 / scale[1], U, b.n, beta, nz, d, d, U ? bs : 0);
          timer.end(st);
        }
        return out;
      }
      if ((name == "Fdl" || name == "FS") && b.k == CHEB && !U && beta == 0.0) {
        timer.begin(st, 6, 8.0 * nbt() * (double)rows * 2.0 * b.n);
        cheb_y_forward(st, in, b.n, out, b.n, rows, b.n, name == "Fdl" ? b.n * 2 / 3 : b.n, nbt(), dstr(), dstr());
        timer.end(st);
        return out;
      }
    }
    if (use_bands && name == "FO" && !U && beta == 0.0 && (b.k == CD || b.k == CN) && tdma_rows_lds_supported(b.m)) {
      const DTdma& t = tdma_tables(b);
      timer.begin(st, 8, 8.0 * nbt() * (double)rows * ((double)b.n + b.m));
      tdma_rows_lds(st, in, b.n, out, b.m, rows, b.m, t.tabp.p, nbt(), dstr(), dstr());
      timer.end(st);
      return out;
    }
    if (use_bands && name == "FO" && !U && beta == 0.0 && (b.k == CD || b.k == CN) && tdma_y_supported(b.m)) {
      const DTdma& t = tdma_tables(b);
      timer.begin(st, 3, 8.0 * nbt() * (double)rows * ((double)b.n + b.m));
      tdma_y(st, in, b.n, out, b.m, rows, b.m, t.diag.p, t.low2.p, t.off.p, t.w.p, t.den.p, nbt(), dstr(), dstr());
      timer.end(st);
      return out;
    }
    const DOp& op = Y.get(b, name);
    GEMM(rows, op.cols, op.rows, in, op.rows, op.b.p, op.cols, out, op.cols, beta, U, op.cols, DATA, 0);
    return out;
  }
  void ymat(const DOp& op, const double* in, int rows, double* out, const double* per_inst = nullptr, long sop = 0) {
    GEMM(rows, op.cols, op.rows, in, op.rows, per_inst ? per_inst : op.b.p, op.cols, out, op.cols, 0.0, nullptr, 0, DATA,
         per_inst ? sop : 0);
  }

  // FP32 dense transforms of the confined mixed step: x: out [rows_out, cols] = op in; y:
  // out [rows, cout] = (in Y^T) .* U + beta out
  void xop32(const Basis& b, const std::string& name, const void* in, bool in64, int cols, void* out, bool out64) {
    const DOp& op = X.get(b, name);
    GEMM32(op.rows, cols, op.cols, op32(op), op.cols, in, in64, cols, out, out64, cols, 0.0, nullptr, 0, 0, DATA);
  }
  void yop32(const Basis& b, const std::string& name, const float* in, int rows, void* out, bool out64,
             const float* U = nullptr, double beta = 0.0) {
    const DOp& op = Y.get(b, name);
    GEMM32(rows, op.cols, op.rows, in, op.rows, op32(op), false, op.cols, out, out64, op.cols, beta, U, op.cols, DATA,
           0);
  }
  // ux, uy (a field pair) in physical space, FP32
  void backward_vel32() {
    xop32(bx_u, "BS", UX.p, true, by_u.m, mixc(0), false);
    yop32(by_u, "BS", mixc(0), nx, mixc(1), false);
  }
  // convective term of q (FP64 spectral) into the FP32 products: dq/dx ux + dq/dy uy
  // (conv_term.rs:35-41) with the FP32 physical velocity
  void conv32(const Basis& bxq, const double* q) {
    const int my = by_u.m;
    float* ux = mixc(1);
    float* uy = ux + pair_off();
    xop32(bxq, "BDS", q, true, my, mixc(0), false);
    yop32(by_u, "BS", mixc(0), nx, mixc(2), false, ux, 0.0);
    xop32(bxq, "BS", q, true, my, mixc(0), false);
    yop32(by_u, "BDS", mixc(0), nx, mixc(2), false, uy, 1.0);
  }
  // dealiased forward transforms of the FP32 products into CV (FP64)
  void forward_dealias32() {
    yop32(by_p, "Fdl", mixc(2), nx, mixc(3), false);
    xop32(bx_f, "Fdl", mixc(3), false, ny, CV.p, true);
  }

  // ------------------------------------------- the field triple (ux, uy, temp; confined)
  // The three fields share the y operators and the dealiasing; temp's x operators (bx_t)
  // and ADI operators enter as the second operators of the same launches.
  bool triple_ok() {
    if (!pair_ok()) return false;
    const Solver &s0 = *solver[0], &s2 = *solver[2];
    if (s2.kind != ADI || bx_t.m != bx_u.m || bx_t.n != bx_u.n) return false;
    if (s0.hx.rows != s2.hx.rows || s0.hx.cols != s2.hx.cols || s0.hy.rows != s2.hy.rows || s0.hy.cols != s2.hy.cols)
      return false;
    return nb == 1 || (adi_s[0][0] == adi_s[2][0] && adi_s[0][1] == adi_s[2][1]);
  }
  // x-transform `name` of the triple [m_x, cols] -> [rows, cols] (FP64 GEMM, fb = 3)
  void xop3(const std::string& name, const double* in, int cols, double* out) {
    const DOp &ou = X.get(bx_u, name), &ot = X.get(bx_t, name);
    if (ou.rows != ot.rows || ou.cols != ot.cols) throw std::logic_error("xop3: operator shapes differ");
    GEMM(ou.rows, cols, ou.cols, ou.b.p, ou.cols, in, cols, out, cols, 0.0, nullptr, 0, 0, DATA, ot.b.p);
  }
  void xop3_32(const std::string& name, const double* in, int cols, float* out) {
    const DOp &ou = X.get(bx_u, name), &ot = X.get(bx_t, name);
    if (ou.rows != ot.rows || ou.cols != ot.cols) throw std::logic_error("xop3: operator shapes differ");
    GEMM32(ou.rows, cols, ou.cols, op32(ou), ou.cols, in, true, cols, out, false, cols, 0.0, nullptr, 0, 0, DATA,
           op32(ot));
  }
  // convective terms of ux, uy, temp into CV: dq/dx ux + dq/dy uy (conv_term.rs:35-41), temp
  // + u . grad(fieldbc) (navier.rs:547-550), then the dealiased forward transforms
  void conv_triple() {
    const int my = by_u.m;
    const long np = (long)nx * ny, po = pair_off();
    if (mixed) {
      float* ux = mixc(1);
      float* uy = ux + po;
      {
        FieldPair ft(*this, 3);
        xop3_32("BDS", UX.p, my, mixc(0));
        yop32(by_u, "BS", mixc(0), nx, mixc(2), false, ux, 0.0);
        xop3_32("BS", UX.p, my, mixc(0));
        yop32(by_u, "BDS", mixc(0), nx, mixc(2), false, uy, 1.0);
      }
      ew_acc_prod2_f32(st, mixc(2) + 2 * po, ux, DTBCX.p, uy, DTBCY.p, nbn(np));
      FieldPair ft(*this, 3);
      forward_dealias32();
      return;
    }
    {
      FieldPair ft(*this, 3);
      xop3("BDS", UX.p, my, W1.p);
      yop(by_u, "BS", W1.p, nx, CONV.p, UXP.p, 0.0);
      xop3("BS", UX.p, my, W1.p);
      yop(by_u, "BDS", W1.p, nx, CONV.p, UYP.p, 1.0);
    }
    ew_acc_prod2(st, CONV.p + 2 * po, UXP.p, DTBCX.p, UYP.p, DTBCY.p, nbn(np));
    FieldPair ft(*this, 3);
    forward_dealias_conv();
  }
  // the implicit ADI solves of ux, uy (solver 0) and temp (solver 2) in one pair of launches
  void solve_triple() {
    FieldPair ft(*this, 3);
    const Solver &s0 = *solver[0], &s2 = *solver[2];
    const bool grp = nb > 1;
    const double* ax = grp ? ADIH[0][0].p : s0.hx.b.p;
    const double* ax2 = grp ? ADIH[2][0].p : s2.hx.b.p;
    const double* ay = grp ? ADIH[0][1].p : s0.hy.b.p;
    const double* ay2 = grp ? ADIH[2][1].p : s2.hy.b.p;
    const DOp &hx = s0.hx, &hy = s0.hy;
    GEMM(hx.rows, ny, hx.cols, ax, hx.cols, RHS.p, ny, W1.p, ny, 0.0, nullptr, 0, grp ? adi_s[0][0] : 0, DATA, ax2);
    GEMM(hx.rows, hy.cols, hy.rows, W1.p, hy.rows, ay, hy.cols, UX.p, hy.cols, 0.0, nullptr, 0, DATA,
         grp ? adi_s[0][1] : 0, nullptr, ay2);
  }
  // one confined step with the field triple (navier.rs:737-765; temp's solve moves ahead of
  // the projection, which does not read temp)
  void update_triple() {
    {
      FieldPair fp(*this);
      if (mixed) backward_vel32();
      else backward_vel(UX.p, UXP.p);
    }
    const long no = (long)rows_ortho(bx_f) * ny, po = pair_off();
    {
      FieldPair ft(*this, 3);
      const DBand &su = stencil_band(bx_u), &stt = stencil_band(bx_t), &sy = stencil_band(by_u);
      BandXYEpi ep;
      ep.x2 = stt.ref();
      ep.split = 2 * nb;
      timer.begin(st, 3, 8.0 * nbt() * ((double)bx_u.m * by_u.m + (double)bx_u.n * by_u.n));
      band_xy(st, UX.p, by_u.m, A.p, su.ref(), sy.ref(), ep, nbt(), dstr(), dstr());
      timer.end(st);
    }
    xop(bx_p, "DS", P0.p, ny, G.p, W9.p);                                  // d/dx pres
    const double* gy = yop(by_p, "DS", P0.p, rows_ortho(bx_p), G.p + po);  // d/dy pres
    if (gy != G.p + po) throw std::logic_error("triple: d/dy pres expected in its slot");
    conv_triple();
    // buoyancy to_ortho(temp) + fieldbc (navier.rs:740-742) from temp's slot of A
    ew_rhs3(st, RHS.p, A.p, G.p, THAT_BC.p, CV.p, GXX.p, GYY.p, dt, nbn(no), po, true);
    solve_triple();
    divergence(DIV.p);
    solve_pres();
    project_velocity_pair();
    update_pres();
    time += dt;
  }

  // ------------------------------------------------------------------ set-up pieces
  Basis xb_field() const { return bx_f; }

  // bc_rbc (navier.rs:314-332, 474-492), computed in long double on the host
  void bc_rbc() {
    Basis xb = bx_f;
    LMat bc(nx, 2);
    for (int i = 0; i < nx; ++i) { bc(i, 0) = 0.5L; bc(i, 1) = -0.5L; }
    LMat vh = periodic ? dft_r2c(nx) * bc : cheb_forward(nx) * bc;
    // backward then forward of the Space2(x, cheb_dirichlet_bc(ny))
    LMat v = host_y(by_bc, "BS", host_x(xb, "BS", vh, (ld)scale[0]));
    LMat vh2 = host_x(xb, "FS", host_y(by_bc, "FS", v), (ld)scale[0]);
    std::vector<double> out = vh2.to_double();
    if (periodic) {
      // The boundary values are the same in every x row, and a real FFT of a constant row has
      // exact zeros at k != 0 (its butterflies subtract equal values), as in the reference's
      // and the oracle's fieldbc; the dense long-double DFT above leaves ~1e-19 there instead
      for (size_t i = 2; i < out.size(); ++i) out[i] = 0.0;
    }
    set_bc_spec(out);
  }

  // fieldbc.vhat (device layout [rows_ortho(x), 2]) and all derived constants
  void set_bc_spec(const std::vector<double>& vh) {
    drop_graph();
    const int R = rows_ortho(bx_f);
    if ((int)vh.size() != R * 2) throw std::invalid_argument("fieldbc size mismatch");
    BC.upload(vh);
    LMat V(R, 2);
    for (size_t i = 0; i < V.a.size(); ++i) V.a[i] = vh[i];
    const ld sx = (ld)scale[0];
    const Basis& xb = bx_f;
    // to_ortho (that contribution, navier.rs:740-742)
    LMat that = host_y(by_bc, "S", V);
    // gradient [1,0] / [0,1] then field.backward (conv_term.rs:35-41)
    LMat gx = host_y(by_bc, "S", host_x(xb, "DS", V, sx));
    LMat gy = host_y(by_bc, "DS", V);
    Basis by_cheb(CHEB, ny);
    LMat dxp = host_y(by_cheb, "BS", host_x(xb, "BS", gx, sx));
    LMat dyp = host_y(by_cheb, "BS", host_x(xb, "BS", gy, sx));
    // gradient [2,0], [0,2] * dt * ka (navier.rs:665-668)
    LMat gxx = host_y(by_bc, "S", host_x(xb, "DDS", V, sx));
    LMat gyy = host_y(by_bc, "DDS", V);
    std::vector<double> a = gxx.to_double(), b = gyy.to_double();
    for (auto& v : a) v = (v * dt) * ka;
    for (auto& v : b) v = (v * dt) * ka;
    std::vector<double> t = that.to_double(), hdx = dxp.to_double(), hdy = dyp.to_double();
    // a Fourier x basis and a fieldbc with only the k = 0 row (constant in x): every derived
    // array is zero beyond that row and the x-gradient is zero; they are stored as +0.0, the
    // zeros the reference's transforms form for them (up to the sign of a zero)
    bool xconst = periodic;
    for (size_t i = 2; i < vh.size() && xconst; ++i) xconst = vh[i] == 0.0;
    if (xconst) {
      for (size_t i = (size_t)ny; i < t.size(); ++i) t[i] = a[i] = b[i] = 0.0;
      std::fill(hdx.begin(), hdx.end(), 0.0);
    }
    THAT_BC.upload(t);
    DTBCX.upload(hdx);
    DTBCY.upload(hdy);
    bc_dx_zero = std::all_of(hdx.begin(), hdx.end(), [](double v) { return v == 0.0; });
    bc_dy_xconst = true;
    for (size_t i = (size_t)ny; i < hdy.size() && bc_dy_xconst; ++i) bc_dy_xconst = hdy[i] == hdy[i % (size_t)ny];
    GXX.upload(a);
    GYY.upload(b);
    {
      const size_t rows = t.size() / (size_t)ny;
      bc_rows = 0;
      for (size_t r = 0; r < rows; ++r)
        for (size_t i = 0; i < (size_t)ny; ++i) {
          const size_t j = r * ny + i;
          for (double v : {t[j], a[j], b[j]})
            if (v != 0.0 || std::signbit(v)) bc_rows = (int)r + 1;
        }
    }
    // physical fieldbc (Field2::v after bc_rbc's backward)
    LMat v = host_y(by_bc, "BS", host_x(xb, "BS", V, sx));
    BCv.upload(v.to_double());
  }

  // ------------------------------------------------------------------------ step
  bool gemm_zero00 = false;  // the next GEMM pins C(0, 0) of every entry to 0
  // zero00: out(0, 0) = 0 of every entry after a confined Poisson solve (formed in its last GEMM)
  void solve(int which, double* rhs, double* out, bool zero00 = false) {
    Solver& s = *solver[which];
    const int my = s.by.m;
    if (s.kind == ADI) {
      const bool grp = nb > 1 && which < 3;  // the group's per-instance operators
      xmat(s.hx, rhs, ny, W1.p, grp ? ADIH[which][0].p : nullptr, adi_s[which][0]);         // [mx, ny]
      ymat(s.hy, W1.p, s.hx.rows, out, grp ? ADIH[which][1].p : nullptr, adi_s[which][1]);  // [mx, my]
      return;
    }
    if (!s.has_fb) {  // Fourier x: P_y then the per-wavenumber row systems
      ROWSF(rhs, ny, out, my, 2 * s.bx.m, s);
      return;
    }
    xmat(s.pf, rhs, ny, W1.p);              // fwd * P_x : [m0, ny]
    ROWSF(W1.p, ny, W2.p, my, s.pf.rows, s);  // P_y and the row systems : [m0, my]
    gemm_zero00 = zero00;
    try {
      xmat(s.pb, W2.p, my, out);  // bwd : [m0, my]
    } catch (...) {
      gemm_zero00 = false;
      throw;
    }
    gemm_zero00 = false;
  }

  // convective term of q with x-basis bxq (y basis cheb_dirichlet), into CONV (physical)
  void conv_pair(const Basis& bxq, const double* q) {
    const int my = by_u.m;
    const double* a = xop(bxq, "BDS", q, my, W1.p, W9.p);  // d/dx, x-backward  [nx, my]
    yop(by_u, "BS", a, nx, CONV.p, UXP.p, 0.0);             // y-backward .* ux
    const double* b = xop(bxq, "BS", q, my, W1.p, W9.p);
    yop(by_u, "BDS", b, nx, CONV.p, UYP.p, 1.0);            // d/dy .* uy, accumulate
  }
  void forward_dealias_conv() {  // CONV [nx, ny] -> CV [rows_ortho, ny]
    yop(by_p, "Fdl", CONV.p, nx, W2.p);
    xop(bx_f, "Fdl", W2.p, ny, CV.p, W9.p);
  }
  bool stencil2d_ok(const Basis& bx, const Basis& by) const {
    return use_bands && (bx.k == CD || bx.k == CN) && (by.k == CD || by.k == CN);
  }
  // both stencils in one pass with the epilogue ep
  void STENCIL2D(const Basis& bx, const Basis& by, const double* q, double* out, const BandXYEpi& ep) {
    const DBand &sx = stencil_band(bx), &sy = stencil_band(by);
    const double n = (double)bx.n * by.n, extra = ep.p0 ? 3.0 : ep.add ? 2.0 : 1.0;
    timer.begin(st, 3, 8.0 * nbt() * ((double)bx.m * by.m + n * extra));
    band_xy(st, q, by.m, out, sx.ref(), sy.ref(), ep, nbt(), dstr(), dstr());
    timer.end(st);
  }
  // out = to_ortho(q) (+ add, same layout, when given)
  void to_ortho(const Basis& bx, const Basis& by, const double* q, double* out, const double* add = nullptr) {
    const long n = (long)rows_ortho(bx) * by.n;
    if (stencil2d_ok(bx, by)) {
      BandXYEpi ep;
      ep.add = add;
      STENCIL2D(bx, by, q, out, ep);
      return;
    }
    const double* a = xop(bx, "S", q, by.m, W1.p, W9.p);
    const double* r = yop(by, "S", a, rows_ortho(bx), out);
    if (r != out)
      HIPCHECK(hipMemcpyAsync(out, r, sizeof(double) * nbn(n), hipMemcpyDeviceToDevice, st));
    if (add) ew_add(st, out, out, add, nbn(n));
  }

  void solve_vel(int which, double* U, const double* buoy) {  // navier.rs:622-654
    const long no = (long)rows_ortho(bx_f) * ny;
    to_ortho(bx_u, by_u, U, A.p);
    if (which == 0) xop(bx_p, "DS", P0.p, ny, G.p, W9.p);   // d/dx pres
    else yop(by_p, "DS", P0.p, rows_ortho(bx_p), G.p);      // d/dy pres
    conv_pair(bx_u, U);
    if (solid_on) ew_solid_damp(st, CONV.p, SOLID_M.p, which == 0 ? UXP.p : UYP.p, nullptr, nullptr, -1.0 / 1e-2,
                                (long)nx * ny);  // navier.rs:579-584, :603-608 (eta = 1e-2)
    forward_dealias_conv();
    ew_rhs_vel(st, RHS.p, A.p, G.p, buoy, CV.p, dt, nbn(no));
    solve(which, RHS.p, U);
  }
  // solve_vel of ux (no buoyancy) and uy (buoyancy THAT) with their transforms, products
  // and implicit solves batched (the two use the same operators; each reads only the
  // physical ux, uy and pres of the step start, so the result equals the two calls in turn)
  bool pair_ok() const {
    return !solid_on && use_batch && !periodic && vel_same && solver[0]->kind == ADI && solver[1]->kind == ADI;
  }
  // confined mixed step: the convective terms (their dense transforms and products) in FP32,
  // everything else FP64; it runs with the field pairs
  bool mixed_confined() const { return mixed && !periodic && pair_ok(); }
  void solve_vel_pair() {
    const long no = (long)rows_ortho(bx_f) * ny, po = pair_off();
    {
      FieldPair fp(*this);
      to_ortho(bx_u, by_u, UX.p, A.p);
    }
    xop(bx_p, "DS", P0.p, ny, G.p, W9.p);                    // d/dx pres
    const double* gy = yop(by_p, "DS", P0.p, rows_ortho(bx_p), G.p + po);  // d/dy pres
    if (gy != G.p + po) throw std::logic_error("pair: d/dy pres expected in its slot");
    {
      FieldPair fp(*this);
      if (mixed) {
        conv32(bx_u, UX.p);
        forward_dealias32();
      } else {
        conv_pair(bx_u, UX.p);
        forward_dealias_conv();
      }
    }
    ew_rhs_vel_pair(st, RHS.p, A.p, G.p, THAT.p, CV.p, dt, nbn(no), po);
    FieldPair fp(*this);
    solve(0, RHS.p, UX.p);
  }
  // project_velocity with the from_ortho transforms and the subtraction of both fields batched
  void project_velocity_pair() {
    const int myp = by_phi.m, ro = rows_ortho(bx_phi);
    const long ns = (long)rows_spec(bx_u) * by_u.m, po = pair_off();
    const double* a = xop(bx_phi, "DS", P1.p, myp, W1.p, W9.p);
    const double* g = yop(by_phi, "S", a, ro, W2.p);
    a = xop(bx_phi, "S", P1.p, myp, W1.p, W9.p);
    const double* g2 = yop(by_phi, "DS", a, ro, W2.p + po);
    if (g != W2.p || g2 != W2.p + po) throw std::logic_error("pair: projection gradients expected in their slots");
    FieldPair fp(*this);
    const double* f = xop(bx_u, "FO", W2.p, ny, W1.p, W9.p);
    if (use_bands && (by_u.k == CD || by_u.k == CN) && tdma_rows_lds_supported(by_u.m)) {
      // from_ortho along y with the subtraction u = x (-1) + u of ew_project in its store
      const DTdma& t = tdma_tables(by_u);
      const int rows = rows_spec(bx_u);
      timer.begin(st, 8, 8.0 * nbt() * (double)rows * ((double)by_u.n + 2.0 * by_u.m));
      tdma_rows_lds(st, f, by_u.n, UX.p, by_u.m, rows, by_u.m, t.tabp.p, nbt(), dstr(), dstr(), true);
      timer.end(st);
      return;
    }
    yop(by_u, "FO", f, rows_spec(bx_u), RHS.p);
    ew_project(st, UX.p, RHS.p, 1.0, nbn(ns));
  }
  void divergence(double* out) {  // navier.rs:698-703
    const int my = by_u.m, ro = rows_ortho(bx_u);
    const double* a = xop(bx_u, "DS", UX.p, my, W1.p, W9.p);
    yop(by_u, "S", a, ro, out);
    const double* b = xop(bx_u, "S", UY.p, my, W1.p, W9.p);
    yop(by_u, "DS", b, ro, out, nullptr, 1.0);
  }
  void project_velocity() {  // navier.rs:683-695 with c = 1
    const int myp = by_phi.m, ro = rows_ortho(bx_phi);
    const long ns = (long)rows_spec(bx_u) * by_u.m;
    // ux
    const double* a = xop(bx_phi, "DS", P1.p, myp, W1.p, W9.p);
    const double* g = yop(by_phi, "S", a, ro, W2.p);
    const double* f = xop(bx_u, "FO", g, ny, W3.p, W9.p);
    yop(by_u, "FO", f, rows_spec(bx_u), W4.p);
    ew_project(st, UX.p, W4.p, 1.0, nbn(ns));
    // uy
    a = xop(bx_phi, "S", P1.p, myp, W1.p, W9.p);
    g = yop(by_phi, "DS", a, ro, W2.p);
    f = xop(bx_u, "FO", g, ny, W3.p, W9.p);
    yop(by_u, "FO", f, rows_spec(bx_u), W4.p);
    ew_project(st, UY.p, W4.p, 1.0, nbn(ns));
  }
  void solve_pres() {  // navier.rs:710-715
    if (!periodic && solver[3]->has_fb && (use_mfma || nb > 1)) {  // pseu(0, 0) = 0 in the solve's last GEMM
      solve(3, DIV.p, P1.p, true);
      return;
    }
    solve(3, DIV.p, P1.p);
    if (periodic) set_zero_entries(st, P1.p, 2, 0, (long)bx_phi.m * by_phi.m, nb, bs);
    else set_zero_entries(st, P1.p, 1, 0, 0, nb, bs);
  }
  void update_pres() {  // navier.rs:717-721
    const long no = (long)rows_ortho(bx_p) * ny;
    if (!periodic && stencil2d_ok(bx_phi, by_phi) && bx_phi.n == rows_ortho(bx_p) && by_phi.n == ny) {
      BandXYEpi ep;  // to_ortho(pseu) and the update in one pass
      ep.p0 = P0.p;
      ep.div = DIV.p;
      ep.nuarr = nb > 1 ? NUARR.p : nullptr;
      ep.nu = nu;
      ep.inv_dt = 1.0 / dt;
      STENCIL2D(bx_phi, by_phi, P1.p, W3.p, ep);
      return;
    }
    to_ortho(bx_phi, by_phi, P1.p, W3.p);
    if (nb > 1) ew_update_pres_group(st, P0.p, DIV.p, W3.p, NUARR.p, 1.0 / dt, nbn(no), bs);
    else ew_update_pres(st, P0.p, DIV.p, W3.p, nu, 1.0 / dt, no);
  }
  void solve_temp() {  // navier.rs:660-674
    const long no = (long)rows_ortho(bx_f) * ny;
    const long np = (long)nx * ny;
    to_ortho(bx_t, by_u, T.p, A.p);
    if (mixed_confined()) {
      conv32(bx_t, T.p);
      float* ux = mixc(1);
      ew_acc_prod2_f32(st, mixc(2), ux, DTBCX.p, ux + pair_off(), DTBCY.p, nbn(np));
      forward_dealias32();
    } else {
      conv_pair(bx_t, T.p);
      ew_acc_prod2(st, CONV.p, UXP.p, DTBCX.p, UYP.p, DTBCY.p, nbn(np));
      if (solid_on) {  // navier.rs:551-560: temp.backward(), then the damp of T + T_bc - value
        backward(RBC_TEMP);
        ew_solid_damp(st, CONV.p, SOLID_M.p, Tv.p, BCv.p, SOLID_V.p, -1.0 / 1e-2, np);
      }
      forward_dealias_conv();
    }
    ew_rhs_temp(st, RHS.p, A.p, GXX.p, GYY.p, CV.p, dt, nbn(no));
    solve(2, RHS.p, T.p);
  }
  void backward_vel(const double* U, double* out) {
    const double* a = xop(bx_u, "BS", U, by_u.m, W1.p, W9.p);
    yop(by_u, "BS", a, nx, out);
  }
  // ----------------------------------------------------------- batched periodic step
  bool batch_ok() const {
    return nb == 1 && use_batch && periodic && vel_same && use_fft && fft_supported(nx) && use_dct && cheb_y_supported(ny) &&
           use_bands && fdma_rows_lds_supported(by_u.m, by_u.n) && tdma_rows_lds_supported(by_u.m) &&
           by_phi.n == by_u.n;
  }
  // irfft of `batch` slots [2 mx, cols] -> [nx, cols], (i k)^times / sc first
  void IRFFT_B(const double* in, double* out, int cols, int times, double sc, int batch) {
    timer.begin(st, 2, 8.0 * batch * cols * ((double)nx + 2.0 * mx()));
    irfft_cols(st, in, out, nx, cols, times, sc, batch, (long)FS, (long)FS);
    timer.end(st);
  }
  // nzc >= 0: the input columns j >= nzc are zero (the y-dealiasing cut), only written as zeros
  void RFFT_B(const double* in, double* out, int cols, int cut, int batch, int nzc = -1) {
    const double tc = nzc >= 0 ? (double)nzc : (double)cols;  // transformed columns
    timer.begin(st, 5, 8.0 * batch * (tc * nx + cols * 2.0 * mx()));
    rfft_cols(st, in, out, nx, cols, cut, batch, (long)FS, (long)FS, nzc);
    timer.end(st);
  }
  // Chebyshev y backward of composite by (times derivatives): out = x .* U + beta out
  // extra = true: temp's slot (2) also gets + ux dTbc/dx + uy dTbc/dy (navier.rs:547-550)
  void CHEBY_BWD_B(const Basis& b, const double* in, double* out, int rows, int times, const double* U, double beta,
                   int batch, bool extra = false) {
    const DBand& band = stencil_band(b);
    const double xb = extra ? 8.0 * 4.0 * rows * b.n : 0.0;
    timer.begin(st, 4,
                8.0 * batch * rows * ((double)b.m + b.n * (1.0 + (U ? 1.0 : 0.0) + (beta != 0.0 ? 1.0 : 0.0))) + xb);
    cheb_y_backward(st, in, b.m, b.m, out, b.n, rows, b.n, band.coef.p, band.mask.p, times, 1.0 / scale[1], U, b.n,
                    beta, batch, (long)FS, (long)FS, 0, extra ? UXP.p : nullptr, DTBCX.p, extra ? UYP.p : nullptr,
                    DTBCY.p, extra ? 2 : -1);
    timer.end(st);
  }
  void CHEBY_FWD_B(const double* in, double* out, int rows, int cut, int batch) {
    timer.begin(st, 6, 8.0 * batch * rows * 2.0 * ny);
    cheb_y_forward(st, in, ny, out, ny, rows, ny, cut, batch, (long)FS, (long)FS);
    timer.end(st);
  }
  void update_batched() {  // navier.rs:737-765 with the three fields' work batched
    const int my = by_u.m;
    const size_t F = FS;
    if (mixed) {
      conv_mixed();
    } else {
      conv_fp64();
    }
    // d/dy pres, then the right-hand sides (navier.rs:622-674) with to_ortho of ux, uy, temp,
    // the buoyancy temp + fieldbc (navier.rs:740-742) and d/dx pres formed in the same kernel
    if (!gpres_valid || !pressure_fused_ok()) yop(by_p, "DS", P0.p, rows_ortho(bx_p), G.p + F);
    // bytes: ux, uy, temp coefficients, pres, d/dy pres read, three right-hand sides
    // written; fieldbc, gxx, gyy in their bc_rows nonzero rows; the three convective terms
    // inside the dealiasing cuts (bx_f.m 2/3 wavenumbers, ny 2/3 y-modes: zero beyond)
    const double bcb = 8.0 * 3.0 * bc_rows * ny;
    const int cutx = bx_f.m * 2 / 3, cuty = ny * 2 / 3;
    const double cvn = 3.0 * 2.0 * cutx * cuty;
    if (mixed) {  // FP32 convective terms (conv_mixed)
      timer.begin(st, 3, 8.0 * 2.0 * mx() * (3.0 * my + 5.0 * ny) + 4.0 * cvn + bcb);
      rhs3_fused_cv32(st, UX.p, my, stencil_band(by_u).coef.p, P0.p, G.p + F, THAT_BC.p,
                      reinterpret_cast<const float*>(MIX.p), GXX.p, GYY.p, RHS.p, mx(), ny, scale[0], dt, (long)F,
                      bc_rows, cutx, cuty);
    } else {
      timer.begin(st, 3, 8.0 * 2.0 * mx() * (3.0 * my + 5.0 * ny) + 8.0 * cvn + bcb);
      rhs3_fused(st, UX.p, my, stencil_band(by_u).coef.p, P0.p, G.p + F, THAT_BC.p, CV.p, GXX.p, GYY.p, RHS.p, mx(), ny,
                 scale[0], dt, (long)F, bc_rows, cutx, cuty);
    }
    timer.end(st);
    solves_and_pressure();
  }
  // convective terms of ux, uy, temp into CV (FP64): dq/dx ux + dq/dy uy (conv_term.rs:35-41),
  // dealiased forward transforms
  void conv_fp64() {
    const int my = by_u.m;
    if (use_fuse) {
      // values of ux, uy, temp (W2) and their x-derivatives (W1) in one launch
      // bytes: the spectra of ux, uy, temp read once, six physical arrays written
      timer.begin(st, 2, 8.0 * my * (3.0 * 2.0 * mx() + 6.0 * nx));
      irfft_cols_pair(st, UX.p, W2.p, W1.p, nx, my, 1, scale[0], 3, (long)FS, (long)FS);
      timer.end(st);
      CHEBY_BWD_B(by_u, W2.p, UXP.p, nx, 0, nullptr, 0.0, 2);  // ux, uy in physical space
      // conv = d/dy q uy + dq/dx ux (conv_term.rs:35-41), + u . grad(fieldbc) for temp
      const DBand& band = stencil_band(by_u);
      // bytes, distinct arrays once: x-derivatives and values of the three fields, ux, uy
      // read, three convective terms written, the fieldbc gradient (one row when constant
      // in x, the x part left out when zero)
      const double fbr = conv3_ldf() ? (double)nx : 1.0;
      timer.begin(st, 11, 8.0 * nx * (6.0 * my + 5.0 * ny) + 8.0 * fbr * ny * (bc_dx_zero ? 1.0 : 2.0));
      cheb_y_conv(st, W1.p, W2.p, my, my, CONV.p, ny, nx, ny, band.coef.p, 1.0 / scale[1], UXP.p, UYP.p, ny, 3,
                  (long)FS, (long)FS, UXP.p, bc_dx_zero ? nullptr : DTBCX.p, UYP.p, DTBCY.p, 2, conv3_ldf());
      timer.end(st);
      if (solid_on) {
        CHEBY_BWD_B(by_u, W2.p + 2 * FS, Tv.p, nx, 0, nullptr, 0.0, 1);  // temp in physical space
        solid_damp_batched();
      }
      CHEBY_FWD_B(CONV.p, W2.p, nx, ny * 2 / 3, 3);  // dealiased forward transforms
      RFFT_B(W2.p, CV.p, ny, bx_f.m * 2 / 3, 3, ny * 2 / 3);  // y-modes >= the cut are zero
      return;
    }
    // ux, uy in physical space
    IRFFT_B(UX.p, W1.p, my, 0, 1.0, 2);
    CHEBY_BWD_B(by_u, W1.p, UXP.p, nx, 0, nullptr, 0.0, 2);
    // convection of ux, uy, temp: dq/dx ux + dq/dy uy (conv_term.rs:35-41)
    IRFFT_B(UX.p, W1.p, my, 1, scale[0], 3);
    CHEBY_BWD_B(by_u, W1.p, CONV.p, nx, 0, UXP.p, 0.0, 3);
    IRFFT_B(UX.p, W1.p, my, 0, 1.0, 3);
    CHEBY_BWD_B(by_u, W1.p, CONV.p, nx, 1, UYP.p, 1.0, 3, true);  // + u . grad(fieldbc) for temp
    if (solid_on) {
      backward(RBC_TEMP);  // temp in physical space
      solid_damp_batched();
    }
    CHEBY_FWD_B(CONV.p, W2.p, nx, ny * 2 / 3, 3);  // dealiased forward transforms
    RFFT_B(W2.p, CV.p, ny, bx_f.m * 2 / 3, 3);
  }
  // the solid obstacle's penalisation of the three physical convective terms (CONV slots
  // 0, 1, 2), as the field-by-field step applies it: -mask ux / eta, -mask uy / eta
  // (navier.rs:579-584, 603-608) and -mask (temp + fieldbc - value) / eta with temp in
  // physical space (Tv, formed by the caller; navier.rs:551-560), eta = 1e-2
  void solid_damp_batched() {
    const long np = (long)nx * ny;
    ew_solid_damp(st, CONV.p, SOLID_M.p, UXP.p, nullptr, nullptr, -1.0 / 1e-2, np);
    ew_solid_damp(st, CONV.p + FS, SOLID_M.p, UYP.p, nullptr, nullptr, -1.0 / 1e-2, np);
    ew_solid_damp(st, CONV.p + 2 * FS, SOLID_M.p, Tv.p, BCv.p, SOLID_V.p, -1.0 / 1e-2, np);
  }
  // the same pipeline with FP32 values and arithmetic (FP64 spectra in, FP32 convective
  // terms out in the first three FP32 slots of MIX)
  void conv_mixed() {
    const int my = by_u.m, mx2 = 2 * mx();
    const long F = (long)FS;
    if (MIX.n < (11 * FS + 1) / 2) throw std::logic_error("mixed buffers not allocated");
    float* W1f = reinterpret_cast<float*>(MIX.p);
    float* UXPf = W1f + 3 * F;
    float* UYPf = UXPf + F;
    float* CONVf = UXPf + 2 * F;
    float* W2f = CONVf + 3 * F;
    const DBand& band = stencil_band(by_u);
    const double fx = 4.0 * nx + 8.0 * mx2;  // bytes per column: FP64 spectrum in, FP32 values out
    auto irfft = [&](int times, double sc, int batch) {
      timer.begin(st, 2, (double)batch * my * fx);
      irfft_cols_f32(st, UX.p, W1f, nx, my, times, sc, batch, F, F);
      timer.end(st);
    };
    auto bwd = [&](float* out, int times, const float* U, double beta, int batch, bool extra) {
      const double xb = extra ? 24.0 * nx * ny : 0.0;
      timer.begin(st, 4,
                  4.0 * batch * nx * ((double)my + ny * (1.0 + (U ? 1.0 : 0.0) + (beta != 0.0 ? 1.0 : 0.0))) + xb);
      cheb_y_backward_f32(st, W1f, my, my, out, ny, nx, ny, band.coef.p, times, 1.0 / scale[1], U, ny, beta, batch, F,
                          F, 0, extra ? UXPf : nullptr, DTBCX.p, extra ? UYPf : nullptr, DTBCY.p, extra ? 2 : -1);
      timer.end(st);
    };
    if (use_fuse) {
      // values (W2f) and x-derivatives (W1f) of ux, uy, temp in one launch, then ux, uy
      // in physical space and the convective terms in one launch each
      timer.begin(st, 2, my * (8.0 * 3.0 * mx2 + 4.0 * 6.0 * nx));  // spectra once, six FP32 outputs
      irfft_cols_pair_f32(st, UX.p, W2f, W1f, nx, my, 1, scale[0], 3, F, F);
      timer.end(st);
      if (conv3_ok()) {
        // the y-stage in one launch (CONVf: y-spectral FP32 convective terms), then the
        // x-transforms into W1f's slots
        timer.begin(st, 13, conv3_bytes(4.0, 4.0));
        cheb_y_conv3_f32(st, W1f, W2f, my, my, CONVf, ny, nx, ny, 1.0 / scale[1],
                         bc_dx_zero ? nullptr : DTBCX.p, DTBCY.p, conv3_ldf(), ny * 2 / 3, F, F);
        timer.end(st);
        const int nzc = ny * 2 / 3;  // y-modes >= the cut are zero: written, not transformed
        timer.begin(st, 5, 3.0 * (4.0 * nx * nzc + 4.0 * mx2 * ny));
        rfft_cols_f32f(st, CONVf, W1f, nx, ny, bx_f.m * 2 / 3, 3, F, F, nzc);
        timer.end(st);
        return;
      }
      timer.begin(st, 4, 4.0 * 2.0 * nx * ((double)my + ny));
      cheb_y_backward_f32(st, W2f, my, my, UXPf, ny, nx, ny, band.coef.p, 0, 1.0 / scale[1], nullptr, ny, 0.0, 2, F, F,
                          0, nullptr, nullptr, nullptr, nullptr, -1);
      timer.end(st);
      // FP32 fields, the FP64 fieldbc gradient (one row when constant in x, no x part when zero)
      const double fbr = conv3_ldf() ? (double)nx : 1.0;
      timer.begin(st, 11, 4.0 * nx * (6.0 * my + 5.0 * ny) + 8.0 * fbr * ny * (bc_dx_zero ? 1.0 : 2.0));
      cheb_y_conv_f32(st, W1f, W2f, my, my, CONVf, ny, nx, ny, band.coef.p, 1.0 / scale[1], UXPf, UYPf, ny, 3, F, F,
                      UXPf, bc_dx_zero ? nullptr : DTBCX.p, UYPf, DTBCY.p, 2, conv3_ldf());
      timer.end(st);
    } else {
      irfft(0, 1.0, 2);
      bwd(UXPf, 0, nullptr, 0.0, 2, false);
      irfft(1, scale[0], 3);
      bwd(CONVf, 0, UXPf, 0.0, 3, false);
      irfft(0, 1.0, 3);
      bwd(CONVf, 1, UYPf, 1.0, 3, true);  // + u . grad(fieldbc) for temp
    }
    timer.begin(st, 6, 4.0 * 3.0 * nx * 2.0 * ny);
    cheb_y_forward_f32(st, CONVf, ny, W2f, ny, nx, ny, ny * 2 / 3, 3, F, F);
    timer.end(st);
    // FP32 convective terms into W1f's slots (free after the y transforms)
    const int nzc = ny * 2 / 3;  // y-modes >= the cut are zero: written, not transformed
    timer.begin(st, 5, 3.0 * (4.0 * nx * nzc + 4.0 * mx2 * ny));
    rfft_cols_f32f(st, W2f, W1f, nx, ny, bx_f.m * 2 / 3, 3, F, F, nzc);
    timer.end(st);
  }
  // the one-launch y-stage of the FP32 convective terms: DCT rows and the Dirichlet stencil
  // for the three fields (the periodic Rayleigh-Benard bases).  The FP64 step keeps its
  // three launches: in FP64 the one-launch kernel holds 168 VGPRs and 69 KB of LDS, two
  // blocks per CU, and at 1024 x 1025 takes 69.6 us against 68.2 for the three launches.
  bool conv3_ok() const {
    if (!use_conv3 || !cheb_y_supported(ny) || by_u.k != CD || by_u.n != ny || by_u.m != ny - 2 ||
        !dirichlet_taps(by_u))
      return false;
    // one block per x-row runs six transforms in turn, four blocks per CU: below four blocks
    // per CU the three separate launches are faster (measured at 256 x 257: 14,530 against
    // 13,880 steps/s; at 1024 x 1025 the one launch wins, 5,558 against 5,465)
    return use_conv3 == 2 || nx >= 4 * device_cus();
  }
  static int device_cus() {
    static int cus = 0;
    if (!cus) {
      int dev = 0, v = 0;
      if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
        cus = v;
      else
        cus = 256;
    }
    return cus;
  }
  // the stencil of b is the Dirichlet one the conv3 kernel forms in place (d = 1, l = -1)
  static bool dirichlet_taps(const Basis& b) {
    for (int i = 0; i < b.m; ++i)
      if (b.diag[i] != 1.0 || b.low2[i] != -1.0) return false;
    return true;
  }
  long conv3_ldf() const { return bc_dy_xconst && bc_dx_zero ? 0 : (long)ny; }
  // bytes, distinct arrays once: x-derivatives and values of the three fields read, the
  // three y-spectral convective terms written (element sizes bi in, bo out), temp's
  // fieldbc gradient (FP64; one row when constant in x, x part skipped when zero)
  double conv3_bytes(double bi, double bo) const {
    const double frows = conv3_ldf() ? (double)nx : 1.0;
    return bi * nx * 6.0 * by_u.m + bo * nx * 3.0 * ny + 8.0 * frows * ny * (bc_dx_zero ? 1.0 : 2.0);
  }
  void solves_and_pressure() {
    const int my = by_u.m, R2 = 2 * mx();
    const size_t F = FS;
    // implicit solves of ux, uy (one solver for both) and temp in one launch; with the fused
    // pressure stage temp's solve runs in further blocks of the stage's launch instead (the
    // stage does not touch temp), where they fill the workgroup slots its mx blocks leave free
    const Solver &sv = *solver[0], &stt = *solver[2];
    const bool fold_temp = pressure_fused_ok() && stt.nsets == mx() && !stt.has_fb;
    const int nsolve = fold_temp ? 2 : 3;
    const double* bands[3] = {sv.pybp.p, sv.pybp.p, stt.pybp.p};
    const double* diags[3] = {sv.diagp.p, sv.diagp.p, stt.diagp.p};
    // bytes: right-hand sides read and results written per row, the diagonal sets once per solver
    const double temp_bytes = 8.0 * ((double)R2 * (ny + my) + 4.0 * (double)stt.nsets * my);
    timer.begin(st, 1, 8.0 * (2.0 * (double)R2 * (ny + my) + 4.0 * (double)sv.nsets * my) +
                           (fold_temp ? 0.0 : temp_bytes));
    fdma_rows_lds_multi(st, RHS.p, ny, ny, UX.p, my, R2, my, sv.nsets, bands, diags, nsolve, (long)F, (long)F);
    timer.end(st);
    // projection (navier.rs:683-721)
    if (pressure_fused_ok()) {
      const Solver& sp = *solver[3];
      PressureArgs pa{UX.p, UY.p, P0.p, P1.p, stencil_band(by_u).coef.p, stencil_band(by_phi).coef.p, sp.pybp.p,
                      sp.diagp.p, tdma_tables(by_u).tabp.p, mx(), ny, by_u.m, 0, 0, scale[0], 1.0 / scale[1], nu,
                      1.0 / dt};
      pa.su_kind = by_u.k == CD ? 1 : by_u.k == CN ? 2 : 0;
      pa.sp_kind = by_phi.k == CD ? 1 : by_phi.k == CN ? 2 : 0;
      pa.dP = G.p + FS;  // d/dy of the new pres for the next step
      if (fold_temp) {
        pa.rows_R = RHS.p + 2 * F;
        pa.rows_X = UX.p + 2 * F;
        pa.rows_band = stt.pybp.p;
        pa.rows_diag = stt.diagp.p;
        pa.rows_ldr = ny;
        pa.rows_nr = ny;
        pa.rows_ld = my;
        pa.rows_m = my;
        pa.rows_nsets = stt.nsets;
      }
      const double rows = 2.0 * mx();
      // bytes, distinct arrays once per row: ux, uy, pres read; ux, uy, pseu, pres and d/dy
      // pres written; the four diagonal arrays of the row's Poisson set (and temp's solve)
      timer.begin(st, 10, 8.0 * rows * (2.0 * by_u.m + ny) + 8.0 * rows * (3.0 * by_u.m + 2.0 * ny) +
                              8.0 * 4.0 * mx() * (double)by_phi.m + (fold_temp ? temp_bytes : 0.0));
      pressure_rows(st, pa);
      timer.end(st);
      gpres_valid = true;
    } else {
      gpres_valid = false;
      divergence(DIV.p);
      solve_pres();
      project_velocity_batched();
      update_pres();
    }
    time += dt;
  }
  bool pressure_fused_ok() {
    return use_fuse && pressure_rows_supported(ny, by_u.m) && by_phi.m == by_u.m && !solver[3]->has_fb &&
           solver[3]->nsets == mx();
  }
  // ux -= from_ortho(d/dx S pseu), uy -= from_ortho(d/dy S pseu) (navier.rs:683-695, c = 1)
  void project_velocity_batched() {
    const int myp = by_phi.m, ro = rows_ortho(bx_phi);
    const size_t F = FS;
    const double* a = xop(bx_phi, "DS", P1.p, myp, W1.p, W9.p);
    const double* g = yop(by_phi, "S", a, ro, W2.p);
    if (g != W2.p) throw std::logic_error("project: stencil output expected in place");
    const double* g2 = yop(by_phi, "DS", P1.p, ro, W2.p + F);
    if (g2 != W2.p + F) throw std::logic_error("project: derivative output expected in place");
    const DTdma& t = tdma_tables(by_u);
    timer.begin(st, 8, 8.0 * 2.0 * (double)ro * ((double)by_u.n + 2.0 * by_u.m));
    tdma_rows_lds(st, W2.p, by_u.n, UX.p, by_u.m, ro, by_u.m, t.tabp.p, 2, (long)F, (long)F, true);
    timer.end(st);
  }

  void update() {  // navier.rs:737-765
    if (batch_ok()) {
      update_batched();
      return;
    }
    gpres_valid = false;
    if (!periodic && triple_ok() && stencil2d_ok(bx_t, by_u) && stencil2d_ok(bx_u, by_u)) {
      update_triple();
      return;
    }
    to_ortho(bx_t, by_u, T.p, THAT.p, THAT_BC.p);  // buoyancy temp + fieldbc (navier.rs:740-742)
    if (pair_ok()) {  // ux and uy as one batch of twice the size
      {
        FieldPair fp(*this);
        if (mixed) backward_vel32();
        else backward_vel(UX.p, UXP.p);
      }
      solve_vel_pair();
      divergence(DIV.p);
      solve_pres();
      project_velocity_pair();
    } else {
      backward_vel(UX.p, UXP.p);
      backward_vel(UY.p, UYP.p);
      solve_vel(0, UX.p, nullptr);
      solve_vel(1, UY.p, THAT.p);
      divergence(DIV.p);
      solve_pres();
      project_velocity();
    }
    update_pres();
    solve_temp();
    time += dt;
  }

  // ----------------------------------------------------------------- field access
  struct FieldRef {
    DBuf* spec;
    DBuf* phys;
    Basis bx, by;
  };
  FieldRef field(int which) {
    switch (which) {
      case RBC_TEMP: return {&T, &Tv, bx_t, by_u};
      case RBC_UX: return {&UX, &UXv, bx_u, by_u};
      case RBC_UY: return {&UY, &UYv, bx_u, by_u};
      case RBC_PRES: return {&P0, &P0v, bx_p, by_p};
      case RBC_PSEU: return {&P1, &P1v, bx_phi, by_phi};
      case RBC_FIELDBC: return {&BC, &BCv, bx_f, by_bc};
    }
    throw std::invalid_argument("unknown field index");
  }
  void shape(int which, int space, int& n0, int& n1, int& cplx) {
    FieldRef f = field(which);
    if (space == RBC_PHYS) { n0 = nx; n1 = ny; cplx = 0; return; }
    n0 = f.bx.m;
    n1 = f.by.m;
    cplx = periodic ? 1 : 0;
  }
  void forward(int which) {  // Space2::forward: axis 1 then axis 0
    gpres_valid = false;
    FieldRef f = field(which);
    const double* a = yop(f.by, "FS", f.phys->p, nx, W1.p);
    const double* r = xop(f.bx, "FS", a, f.by.m, W2.p, W9.p);
    HIPCHECK(hipMemcpyAsync(f.spec->p, r, sizeof(double) * rows_spec(f.bx) * f.by.m, hipMemcpyDeviceToDevice, st));
  }
  void backward(int which) {  // Space2::backward: axis 0 then axis 1
    FieldRef f = field(which);
    const double* a = xop(f.bx, "BS", f.spec->p, f.by.m, W1.p, W9.p);
    const double* r = yop(f.by, "BS", a, nx, W2.p);
    HIPCHECK(hipMemcpyAsync(f.phys->p, r, sizeof(double) * nx * ny, hipMemcpyDeviceToDevice, st));
  }
  void get(int which, int space, double* host) {
    int n0, n1, c;
    shape(which, space, n0, n1, c);
    FieldRef f = field(which);
    HIPCHECK(hipStreamSynchronize(st));
    if (space == RBC_PHYS || !c) {
      std::vector<double> v = (space == RBC_PHYS ? f.phys : f.spec)->download((size_t)n0 * n1);
      std::memcpy(host, v.data(), v.size() * sizeof(double));
      return;
    }
    std::vector<double> v = f.spec->download((size_t)2 * n0 * n1);
    const size_t pl = (size_t)n0 * n1;
    for (size_t i = 0; i < pl; ++i) { host[2 * i] = v[i]; host[2 * i + 1] = v[pl + i]; }
  }
  void set(int which, int space, const double* host) {
    gpres_valid = false;
    int n0, n1, c;
    shape(which, space, n0, n1, c);
    FieldRef f = field(which);
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<double> v((size_t)(c ? 2 : 1) * n0 * n1);
    if (space == RBC_PHYS || !c) {
      std::memcpy(v.data(), host, v.size() * sizeof(double));
    } else {
      const size_t pl = (size_t)n0 * n1;
      for (size_t i = 0; i < pl; ++i) { v[i] = host[2 * i]; v[pl + i] = host[2 * i + 1]; }
    }
    if (which == RBC_FIELDBC && space == RBC_SPEC) { set_bc_spec(v); return; }
    HIPCHECK(hipMemcpy((space == RBC_PHYS ? f.phys : f.spec)->p, v.data(), v.size() * sizeof(double),
                       hipMemcpyHostToDevice));
  }

  // set_velocity / set_temperature (navier.rs:927-937, 1035-1076): host trig, device forward
  void apply_trig(int which, double amp, double m, double n, bool sin_cos) {
    std::vector<double> x = field(which).bx.coords(), y = field(which).by.coords();
    for (auto& v : x) v *= scale[0];
    for (auto& v : y) v *= scale[1];
    std::vector<double> xs(x.size()), ys(y.size());
    for (size_t i = 0; i < x.size(); ++i) xs[i] = (x[i] - x[0]) / (x.back() - x[0]);
    for (size_t j = 0; j < y.size(); ++j) ys[j] = (y[j] - y[0]) / (y.back() - y[0]);
    const double ax = M_PI * m, ay = M_PI * n;
    std::vector<double> v((size_t)nx * ny);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        v[(size_t)i * ny + j] = sin_cos ? amp * std::sin(ax * xs[i]) * std::cos(ay * ys[j])
                                        : amp * std::cos(ax * xs[i]) * std::sin(ay * ys[j]);
    set(which, RBC_PHYS, v.data());
    forward(which);
  }

  // -------------------------------------------------------------- diagnostics
  // field.backward of an ortho-space array (Space2(x field basis, chebyshev)) into W2
  const double* field_backward(const double* ortho) {
    const double* a = xop(bx_f, "BS", ortho, ny, W1.p, W9.p);
    return yop(by_p, "BS", a, nx, W2.p);
  }
  void temp_total_ortho(double* out) {
    const long no = (long)rows_ortho(bx_f) * ny;
    to_ortho(bx_t, by_u, T.p, W3.p);
    ew_add(st, out, W3.p, THAT_BC.p, no);
  }
  double eval_nu() {  // functions.rs:12-36
    const long no = (long)rows_ortho(bx_f) * ny;
    temp_total_ortho(A.p);
    yop(by_p, "DS", A.p, rows_ortho(bx_f), G.p);
    ew_scale(st, G.p, G.p, -1.0, no);
    ew_scale(st, G.p, G.p, 1.0 / (scale[1] / 2.0), no);
    const double* v = field_backward(G.p);
    col_weighted_sum(st, v, nx, ny, DX0.p, RED.p);
    std::vector<double> avg = RED.download(ny);
    const double a0 = avg[0] / len0, a1 = avg[ny - 1] / len0;
    return (a1 + a0) / 2.0;
  }
  double eval_nuvol() {  // functions.rs:42-75
    const long no = (long)rows_ortho(bx_f) * ny, np = (long)nx * ny;
    temp_total_ortho(A.p);
    const double* tv = field_backward(A.p);
    HIPCHECK(hipMemcpyAsync(CONV.p, tv, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    backward_vel(UY.p, UYv.p);
    yop(by_p, "DS", A.p, rows_ortho(bx_f), G.p);
    ew_scale(st, G.p, G.p, 1.0 / (scale[1] * -1.0), no);
    const double* dv = field_backward(G.p);
    ew_nuvol(st, CV.p, dv, CONV.p, UYv.p, ka, scale[1], np);
    return weighted_average(CV.p);
  }
  double eval_re() {  // functions.rs:82-101
    const long np = (long)nx * ny;
    backward_vel(UX.p, UXv.p);
    backward_vel(UY.p, UYv.p);
    ew_speed(st, CV.p, UXv.p, UYv.p, 2.0 * scale[1] / nu, np);
    return weighted_average(CV.p);
  }
  double weighted_average(const double* v) {
    col_weighted_sum(st, v, nx, ny, DX0.p, RED.p);
    std::vector<double> s = RED.download(ny);
    double tot = 0.0;
    for (int j = 0; j < ny; ++j) tot += (s[j] / len0) * wy_dx[j] / len1;
    return tot;
  }
  // ------------------------------------------------------------ solid obstacle
  // navier.solid = Some([mask, value]) (navier.rs:189-190): the penalisation terms enter the
  // physical convective terms of the FP64 step (the batched periodic step and the
  // field-by-field step; the confined field triple and the FP32 convective pipeline do not
  // take them, so a confined handle with a mask runs field by field and a mixed one refuses it)
  bool solid_on = false;
  DBuf SOLID_M, SOLID_V;
  void set_solid(const double* mask, const double* value) {
    HIPCHECK(hipStreamSynchronize(st));
    drop_graph();
    gpres_valid = false;
    if (!mask) {
      solid_on = false;
      return;
    }
    if (!value) throw std::invalid_argument("solid: value array required with the mask");
    if (mixed)
      throw std::invalid_argument("solid: the penalised step is an FP64 step; switch RBC_OPT_MIXED off first");
    const size_t np = (size_t)nx * ny;
    SOLID_M.upload(std::vector<double>(mask, mask + np));
    SOLID_V.upload(std::vector<double>(value, value + np));
    solid_on = true;
  }
  // -------------------------------------------------------------- statistics
  // Statistics (src/navier/statistics.rs:10-153): the running average of the total
  // temperature, the latest ortho velocities and the latest Nusselt field, each an array of
  // the `field` space (ortho x basis of the field, Chebyshev y), FP64 on the device
  bool st_on = false;
  double st_save = 0.0, st_write = 0.0, st_avg_time = 0.0, st_tot_time = 0.0;
  long long st_num = 0;
  DBuf STA[4];  // t_avg, ux_avg, uy_avg, nusselt (.vhat)
  long stats_len() const { return (long)rows_ortho(bx_f) * ny; }
  void stats_enable(double save_stat, double write_stat) {  // Statistics::new, statistics.rs:54-88
    const long no = stats_len();
    for (DBuf& b : STA) {
      if (b.n < (size_t)no) b.alloc(no);
      HIPCHECK(hipMemsetAsync(b.p, 0, sizeof(double) * no, st));
    }
    st_save = save_stat;
    st_write = write_stat;
    st_avg_time = 0.0;
    st_tot_time = time;
    st_num = 0;
    st_on = true;
  }
  void stats_need() const {
    if (!st_on) throw std::invalid_argument("statistics are not enabled (rbc_stats_enable)");
  }
  // Statistics::update(that, ux.to_ortho(), uy.to_ortho(), time) as callback calls it
  // (navier.rs:800-806, statistics.rs:130-153); false when time < tot_time (no update)
  bool stats_update() {
    stats_need();
    if (time < st_tot_time) return false;
    const long no = stats_len(), np = (long)nx * ny;
    temp_total_ortho(A.p);  // that = temp.to_ortho() + fieldbc.to_ortho()
    ew_running_avg(st, STA[0].p, A.p, (double)st_num, no);
    to_ortho(bx_u, by_u, UX.p, STA[1].p);
    to_ortho(bx_u, by_u, UY.p, STA[2].p);
    // nusselt (statistics.rs:220-247): uy and that to physical space through the field
    // space, dT/dy / (-scale_y) (scale_y = 1: the product with its reciprocal is exact),
    // (dT/dy + uy T / ka) 2 scale_y, then the field space's forward transform
    const double* u = field_backward(STA[2].p);
    HIPCHECK(hipMemcpyAsync(CV.p, u, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    const double* t = field_backward(A.p);
    HIPCHECK(hipMemcpyAsync(CONV.p, t, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    yop(by_p, "DS", A.p, rows_ortho(bx_f), G.p);
    ew_scale(st, G.p, G.p, 1.0 / (scale[1] * -1.0), no);
    const double* dv = field_backward(G.p);
    ew_nuvol(st, RHS.p, dv, CONV.p, CV.p, ka, scale[1], np);
    const double* a = yop(by_p, "FS", RHS.p, nx, W1.p);
    const double* r = xop(bx_f, "FS", a, ny, W2.p, W9.p);
    HIPCHECK(hipMemcpyAsync(STA[3].p, r, sizeof(double) * no, hipMemcpyDeviceToDevice, st));
    ++st_num;
    st_avg_time += time - st_tot_time;
    st_tot_time = time;
    return true;
  }
  // .vhat (RBC_SPEC) or, after the field space's backward transform, .v (RBC_PHYS) of
  // statistics array `which` (statistics.rs:176-179)
  void stats_get(int which, int space, double* host) {
    stats_need();
    if (which < 0 || which > 3) throw std::invalid_argument("unknown statistics field");
    const long no = stats_len(), np = (long)nx * ny;
    if (space == RBC_PHYS) {
      const double* v = field_backward(STA[which].p);
      HIPCHECK(hipStreamSynchronize(st));
      HIPCHECK(hipMemcpy(host, v, sizeof(double) * np, hipMemcpyDeviceToHost));
      return;
    }
    if (space != RBC_SPEC) throw std::invalid_argument("unknown space");
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<double> v = STA[which].download(no);
    if (!periodic) {
      std::memcpy(host, v.data(), v.size() * sizeof(double));
      return;
    }
    const size_t pl = (size_t)no / 2;
    for (size_t i = 0; i < pl; ++i) { host[2 * i] = v[i]; host[2 * i + 1] = v[pl + i]; }
  }
  void stats_set(int which, const double* host) {  // Statistics::read, statistics.rs:197-212
    stats_need();
    if (which < 0 || which > 3) throw std::invalid_argument("unknown statistics field");
    const long no = stats_len();
    std::vector<double> v(no);
    if (!periodic) {
      std::memcpy(v.data(), host, v.size() * sizeof(double));
    } else {
      const size_t pl = (size_t)no / 2;
      for (size_t i = 0; i < pl; ++i) { v[i] = host[2 * i]; v[pl + i] = host[2 * i + 1]; }
    }
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipMemcpy(STA[which].p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  double div_norm() {
    divergence(G.p);
    const long no = (long)rows_ortho(bx_u) * ny;
    dot_and_norm(st, nullptr, nullptr, 0, G.p, no, RED.p);
    std::vector<double> r = RED.download(2);
    return std::sqrt(r[1]);
  }
};

// ----------------------------------------------------------------- funspace Space2
class Space {
 public:
  Basis b0, b1;
  bool cplx;
  hipStream_t st = nullptr;
  OpCache X, Y;
  DBuf W1, W2, W3, W9;
  Space(int k0, int n0, int k1, int n1) : b0((Kind)k0, n0), b1((Kind)k1, n1), cplx(k0 == FO), X(false, 1.0), Y(true, 1.0) {
    if (k1 == FO) throw std::invalid_argument("axis 1 must be a Chebyshev basis");
    if (n0 < 6 || n1 < 6) throw std::invalid_argument("sizes must be >= 6");
    HIPCHECK(hipStreamCreate(&st));
  }
  ~Space() {
    if (st) (void)hipStreamDestroy(st);
  }
  int r_phys() const { return b0.n; }
  int r_spec() const { return cplx ? 2 * b0.m : b0.m; }
  int r_ortho() const { return cplx ? 2 * b0.m : b0.n; }
  void ensure(size_t n) {
    if (W1.n < n) { W1.alloc(n); W2.alloc(n); W3.alloc(n); W9.alloc(n); }
  }
  // x-op with the scale used for gradients
  const double* xop(const std::string& name, const double* in, int cols, double* out, double sc) {
    if (b0.k == FO) {
      if (name == "S" || name == "FO") return in;
      if (fft_supported(b0.n) && (name == "BS" || name == "FS")) {  // real FFT kernels
        if (name == "BS") irfft_cols(st, in, out, b0.n, cols, 0, 1.0);
        else rfft_cols(st, in, out, b0.n, cols, b0.m);
        return out;
      }
      if (name == "DS" || name == "DDS") {
        fourier_diff(st, in, out, b0.m, cols, cols, name == "DS" ? 1 : 2, sc);
        return out;
      }
    } else if (b0.k == CHEB && (name == "S" || name == "FO")) {
      return in;
    }
    const DOp& op = X.get(b0, name);
    gemm(st, op.rows, cols, op.cols, op.b.p, op.cols, in, cols, out, cols, 1.0, 0.0);
    return out;
  }
  std::unique_ptr<DBand> band;  // stencil taps of a composite axis-1 basis
  const double* yop(const std::string& name, const double* in, int rows, double* out) {
    if (b1.k == CHEB && (name == "S" || name == "FO")) return in;
    if ((b1.k == CHEB || b1.k == CD || b1.k == CN) && cheb_y_supported(b1.n)) {  // FFT-based DCT-I
      if (name == "BS") {
        if (b1.k != CHEB && !band) {
          band = std::make_unique<DBand>();
          make_stencil_band(b1, *band);
        }
        cheb_y_backward(st, in, b1.m, b1.m, out, b1.n, rows, b1.n, band ? band->coef.p : nullptr,
                        band ? band->mask.p : nullptr, 0, 1.0, nullptr, 0, 0.0);
        return out;
      }
      if (name == "FS" && b1.k == CHEB) {
        cheb_y_forward(st, in, b1.n, out, b1.n, rows, b1.n, b1.n);
        return out;
      }
    }
    const DOp& op = Y.get(b1, name);
    gemm(st, rows, op.cols, op.rows, in, op.rows, op.b.p, op.cols, out, op.cols, 1.0, 0.0);
    return out;
  }
  // generic 2-D application on device interleaved/real buffers
  void run(int op, const double* in, double* out, int batch, int d0, int d1, double s0, double s1) {
    const size_t big = (size_t)std::max(r_ortho(), b0.n) * b1.n + 16;
    ensure(big);
    const bool in_c = cplx && op != 0, out_c = cplx && op != 1;
    int in_r, in_c1, out_r, out_c1;
    switch (op) {
      case 0: in_r = b0.n; in_c1 = b1.n; out_r = r_spec(); out_c1 = b1.m; break;
      case 1: in_r = r_spec(); in_c1 = b1.m; out_r = b0.n; out_c1 = b1.n; break;
      case 2: in_r = r_spec(); in_c1 = b1.m; out_r = r_ortho(); out_c1 = b1.n; break;
      case 3: in_r = r_ortho(); in_c1 = b1.n; out_r = r_spec(); out_c1 = b1.m; break;
      default: in_r = r_spec(); in_c1 = b1.m; out_r = r_ortho(); out_c1 = b1.n; break;
    }
    const size_t in_sz = (size_t)in_r * in_c1, out_sz = (size_t)out_r * out_c1;
    for (int b = 0; b < batch; ++b) {
      const double* src = in + b * in_sz;
      double* dst = out + b * out_sz;
      const double* x = src;
      if (in_c) { interleaved_to_split(st, src, W3.p, in_r / 2, in_c1); x = W3.p; }
      const double* r = nullptr;
      // gradient scale divides by s0^d0 s1^d1 (space2.rs:256-262); folded into the x part
      const double sc = std::pow(s0, d0) * std::pow(s1, d1);
      if (op == 0) {  // forward: axis 1 then axis 0
        const double* a = yop("FS", x, b0.n, W1.p);
        r = xop("FS", a, b1.m, W2.p, 1.0);
      } else if (op == 1) {  // backward: axis 0 then axis 1
        const double* a = xop("BS", x, b1.m, W1.p, 1.0);
        r = yop("BS", a, b0.n, W2.p);
      } else if (op == 2) {
        const double* a = xop("S", x, b1.m, W1.p, 1.0);
        r = yop("S", a, r_ortho(), W2.p);
      } else if (op == 3) {
        const double* a = xop("FO", x, b1.n, W1.p, 1.0);
        r = yop("FO", a, r_spec(), W2.p);
      } else {
        if (d0 > 2 || d1 > 2) throw std::invalid_argument("derivative order > 2");
        const char* nx_ = d0 == 0 ? "S" : d0 == 1 ? "DS" : "DDS";
        const char* ny_ = d1 == 0 ? "S" : d1 == 1 ? "DS" : "DDS";
        const double* a = xop(nx_, x, b1.m, W1.p, 1.0);
        r = yop(ny_, a, r_ortho(), W2.p);
        if (sc != 1.0) {
          if (r != W2.p) { HIPCHECK(hipMemcpyAsync(W2.p, r, sizeof(double) * r_ortho() * b1.n, hipMemcpyDeviceToDevice, st)); r = W2.p; }
          if (b0.k == FO) {
            // complex division by (sc, 0): (a sc) / sc^2 per component
            fourier_diff(st, W2.p, W2.p, b0.m, b1.n, b1.n, 0, sc);
          } else {
            ew_scale(st, W2.p, W2.p, 1.0 / sc, (long)r_ortho() * b1.n);
          }
        }
      }
      if (out_c) split_to_interleaved(st, r, dst, out_r / 2, out_c1);
      else HIPCHECK(hipMemcpyAsync(dst, r, sizeof(double) * out_sz, hipMemcpyDeviceToDevice, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
  }
};

// ------------------------------------------------------------ SolverField (batched)
// Hholtz / HholtzAdi / Poisson built from the bases of a field (src/solver.rs:88-95,
// hholtz.rs:42-76, hholtz_adi.rs:44-63, poisson.rs:50-90), solving `batch` right-hand
// sides per call.  Input: ortho-space arrays [batch, o0, o1]; output [batch, m0, m1].
class SolverDev {
 public:
  Basis b0, b1;
  bool cplx;
  std::unique_ptr<Solver> s;
  hipStream_t st = nullptr;
  DBuf W1, W2, IN, OUT;
  SolverDev(int skind, int k0, int n0, int k1, int n1, double c0, double c1) : b0((Kind)k0, n0), b1((Kind)k1, n1), cplx(k0 == FO) {
    if (k1 == FO || k1 == CHEB || k1 == CDBC) throw std::invalid_argument("axis 1 must be cheb_dirichlet or cheb_neumann");
    if (k0 == CDBC) throw std::invalid_argument("axis 0 cannot be cheb_dirichlet_bc");
    if (skind == ADI && k0 == FO) throw std::invalid_argument("HholtzAdi with a Fourier axis is not supported");
    if (n0 < 8 || n1 < 8) throw std::invalid_argument("sizes must be >= 8");
    const double c[2] = {c0, c1};
    s = make_solver((SKind)skind, b0, b1, c);
    HIPCHECK(hipStreamCreate(&st));
  }
  ~SolverDev() {
    if (st) (void)hipStreamDestroy(st);
  }
  int rows_in() const { return cplx ? 2 * b0.m : b0.ortho(); }
  int rows_out() const { return cplx ? 2 * b0.m : b0.m; }
  void ensure(DBuf& b, size_t n) {
    if (b.n < n) b.alloc(n);
  }
  // work arrays of a solve of `batch` right-hand sides: W1, W2 (big * batch each) and, with a
  // Fourier axis 0, the split-plane input and output (in_sz * batch, out_sz * batch)
  struct Work { double *w1, *w2, *in, *out; };
  size_t big() const { return (size_t)std::max(rows_in(), rows_out()) * b1.ortho(); }
  size_t in_size() const { return (size_t)rows_in() * b1.ortho(); }
  size_t out_size() const { return (size_t)rows_out() * b1.m; }
  void solve(const double* rhs, double* out, int batch) {
    if (mixed_iters > 0) {
      solve_mixed(rhs, out, batch);
      return;
    }
    ensure(W1, big() * batch);
    ensure(W2, big() * batch);
    if (cplx) {
      ensure(IN, in_size() * batch);
      ensure(OUT, out_size() * batch);
    }
    solve_ws(rhs, out, batch, Work{W1.p, W2.p, IN.p, OUT.p});
  }
  // ---- mixed precision (a Chebyshev axis 0: Hholtz, Poisson).  The eigen-transforms
  // fwd = Q^-1 C_x^-1 and bwd = Q run as FP32 matrix-core GEMMs (FP64 in, FP64 out, FP32
  // products), the per-row solves stay FP64, and `mixed_iters` such solves are chained as
  // iterative refinement against the FP64 residual of the tensor equation
  //     (A_x (x) C_y + C_x (x) (A_y + alpha C_y)) u = P_x f P_y^T    (fdma_tensor.rs:16-70),
  // all of whose matrices are banded (so the residual is O(n^2) band passes).
  int mixed_iters = 0;
  bool mixed_ready = false;
  DBuf F32, B32;                   // fwd, bwd as FP32 (m0 x m0)
  DBand pxb, axb, cxb, cyb, aycb;  // P_x; A_x, C_x; C_y and A_y + alpha C_y
  DBuf MR, MX, MV, MT, ML;         // P f P^T, transform buffer, solution, band temporaries
  size_t mixed_batch = 0;
  static void band_from(const std::vector<double>& a, int r, int c, DBand& b) {
    LMat m(r, c);
    for (size_t i = 0; i < m.a.size(); ++i) m.a[i] = a[i];
    make_matvec_band(m, b);
  }
  static std::vector<double> dense_of(const Band4& f) {
    const int n = f.n;
    std::vector<double> a((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) a[(size_t)i * n + i] = f.dia[i];
    for (int i = 0; i + 2 < n; ++i) a[(size_t)(i + 2) * n + i] = f.low[i];
    for (int i = 0; i + 2 < n; ++i) a[(size_t)i * n + i + 2] = f.up1[i];
    for (int i = 0; i + 4 < n; ++i) a[(size_t)i * n + i + 4] = f.up2[i];
    return a;
  }
  void set_mixed(int iters) {
    if (iters < 0 || iters > 8) throw std::invalid_argument("mixed iterations must be in [0, 8]");
    if (iters > 0 && (s->kind == ADI || !s->has_fb))
      throw std::invalid_argument("the mixed form needs eigen-transforms (Hholtz / Poisson with a Chebyshev axis 0)");
    mixed_iters = iters;
    if (!iters || mixed_ready) return;
    Solver& sv = *s;
    const int m0 = b0.m, m1 = b1.m;
    if (!fdma_rows_wave_supported(m1)) throw std::invalid_argument("mixed form: row length not supported");
    std::vector<float> f(sv.fwd.begin(), sv.fwd.end()), b(sv.bwd.begin(), sv.bwd.end());
    F32.alloc((f.size() + 1) / 2);
    B32.alloc((b.size() + 1) / 2);
    HIPCHECK(hipMemcpy(F32.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(B32.p, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
    band_from(sv.px, m0, b0.ortho(), pxb);
    band_from(sv.ax, m0, m0, axb);
    band_from(sv.cx, m0, m0, cxb);
    band_from(dense_of(sv.fc), m1, m1, cyb);
    band_from(dense_of(sv.fa.add_scaled(sv.fc, sv.alpha)), m1, m1, aycb);
    mixed_ready = true;
  }
  // out = R - L(V) for `batch` entries [m0, m1] (L the tensor operator above)
  void mixed_residual(const double* R, const double* V, double* out, int batch) {
    const int m0 = b0.m, m1 = b1.m;
    const long e = (long)m0 * m1;
    band_x(st, V, MT.p, m1, axb.ref(), 0.0, batch, e, e);
    band_y(st, MT.p, m1, ML.p, m1, m0, cyb.ref(), 0.0, batch, e, e);
    band_x(st, V, MT.p, m1, cxb.ref(), 0.0, batch, e, e);
    band_y(st, MT.p, m1, ML.p, m1, m0, aycb.ref(), 1.0, batch, e, e);
    ew_axpy(st, out, R, ML.p, -1.0, e * batch);
  }
  void solve_mixed(const double* f, double* u, int batch) {
    Solver& sv = *s;
    const int m0 = b0.m, m1 = b1.m, o0 = b0.ortho(), o1 = b1.ortho();
    const long e = (long)m0 * m1;
    if (mixed_batch < (size_t)batch) {
      for (DBuf* b : {&MR, &MX, &MV, &MT, &ML}) b->alloc((size_t)std::max(e, (long)m0 * o1) * batch);
      mixed_batch = batch;
    }
    // R = P_x f P_y^T
    band_x(st, f, MT.p, o1, pxb.ref(), 0.0, batch, (long)o0 * o1, (long)m0 * o1);
    band_y(st, MT.p, o1, MR.p, m1, m0, sv.pyb.ref(), 0.0, batch, (long)m0 * o1, e);
    for (int it = 0; it < mixed_iters; ++it) {
      const double* src = it == 0 ? MR.p : ML.p;  // the residual of the previous iterate
      gemm_mfma_f32_batched(st, m0, m1, m0, reinterpret_cast<const float*>(F32.p), m0, 0, src, true, m1, e, MX.p,
                            true, m1, e, batch, 1.0, 0.0, nullptr, 0, 0);
      fdma_rows_wave(st, MX.p, m1, m0, m1, m0, sv.low.p, sv.dia.p, sv.up1.p, sv.up2.p, batch, e);
      double* dst = it == 0 ? u : MT.p;
      gemm_mfma_f32_batched(st, m0, m1, m0, reinterpret_cast<const float*>(B32.p), m0, 0, MX.p, true, m1, e, dst,
                            true, m1, e, batch, 1.0, 0.0, nullptr, 0, 0);
      if (it > 0) ew_add(st, u, u, MT.p, e * batch);
      if (it + 1 < mixed_iters) mixed_residual(MR.p, u, ML.p, batch);
    }
    HIPCHECK(hipStreamSynchronize(st));
  }
  // max-norm of P_x f P_y^T - L(u) relative to that of P_x f P_y^T, per batch of solutions u
  double residual(const double* f, const double* u, int batch) {
    if (!s->has_fb || s->kind == ADI) throw std::invalid_argument("residual: tensor solvers with a Chebyshev axis 0");
    const int saved = mixed_iters;
    set_mixed(1);  // tables
    mixed_iters = saved;
    Solver& sv = *s;
    const int m0 = b0.m, m1 = b1.m, o0 = b0.ortho(), o1 = b1.ortho();
    const long e = (long)m0 * m1;
    if (mixed_batch < (size_t)batch) {
      for (DBuf* b : {&MR, &MX, &MV, &MT, &ML}) b->alloc((size_t)std::max(e, (long)m0 * o1) * batch);
      mixed_batch = batch;
    }
    band_x(st, f, MT.p, o1, pxb.ref(), 0.0, batch, (long)o0 * o1, (long)m0 * o1);
    band_y(st, MT.p, o1, MR.p, m1, m0, sv.pyb.ref(), 0.0, batch, (long)m0 * o1, e);
    mixed_residual(MR.p, u, MV.p, batch);
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<double> r = MR.download((size_t)e * batch), d = MV.download((size_t)e * batch);
    double rm = 0.0, dm = 0.0;
    for (size_t i = 0; i < r.size(); ++i) {
      rm = std::max(rm, std::fabs(r[i]));
      dm = std::max(dm, std::fabs(d[i]));
    }
    return dm / std::max(rm, 1e-300);
  }
  // the same solve in caller-provided work arrays (no allocation): the doubles needed, each
  // array starting on a 256-byte boundary
  static size_t round32(size_t n) { return (n + 31) / 32 * 32; }
  size_t work_doubles(int batch) const {
    size_t n = 2 * round32(big() * batch);
    if (cplx) n += round32(in_size() * batch) + round32(out_size() * batch);
    return n;
  }
  void solve_in(const double* rhs, double* out, int batch, double* scratch) {
    Work w{};
    w.w1 = scratch;
    w.w2 = w.w1 + round32(big() * batch);
    w.in = w.w2 + round32(big() * batch);
    w.out = cplx ? w.in + round32(in_size() * batch) : nullptr;
    if (!cplx) w.in = nullptr;
    solve_ws(rhs, out, batch, w);
  }
  void solve_ws(const double* rhs, double* out, int batch, const Work& w) {
    const int o1 = b1.ortho(), m1 = b1.m, m0 = b0.m, ro = rows_out();
    const size_t in_sz = in_size(), out_sz = out_size();
    const double* x = rhs;
    if (cplx) {  // interleaved -> split planes, per right-hand side
      for (int b = 0; b < batch; ++b) interleaved_to_split(st, rhs + 2 * (size_t)m0 * o1 * b, w.in + in_sz * b, m0, o1);
      x = w.in;
    }
    double* y = out;
    if (cplx) y = w.out;
    Solver& sv = *s;
    if (sv.kind == ADI) {
      gemm_mfma_batched(st, sv.hx.rows, o1, sv.hx.cols, sv.hx.b.p, sv.hx.cols, 0, x, o1, (long)in_sz, w.w1, o1,
                        (long)sv.hx.rows * o1, batch, 1.0, 0.0);
      gemm_mfma(st, batch * sv.hx.rows, m1, sv.hy.rows, w.w1, sv.hy.rows, sv.hy.b.p, sv.hy.cols, y, m1, 1.0, 0.0);
    } else if (!sv.has_fb) {
      rows_launch(x, o1, y, m1, batch * ro, sv);
    } else {
      gemm_mfma_batched(st, sv.pf.rows, o1, sv.pf.cols, sv.pf.b.p, sv.pf.cols, 0, x, o1, (long)in_sz, w.w1, o1,
                        (long)sv.pf.rows * o1, batch, 1.0, 0.0);
      rows_launch(w.w1, o1, w.w2, m1, batch * sv.pf.rows, sv);
      gemm_mfma_batched(st, sv.pb.rows, m1, sv.pb.cols, sv.pb.b.p, sv.pb.cols, 0, w.w2, m1, (long)sv.pf.rows * m1, y,
                        m1, (long)out_sz, batch, 1.0, 0.0);
    }
    if (cplx)
      for (int b = 0; b < batch; ++b) split_to_interleaved(st, y + out_sz * b, out + 2 * (size_t)m0 * m1 * b, m0, m1);
    HIPCHECK(hipStreamSynchronize(st));
  }
  // P_y matvec and the row solves: R [nrows, ldr] -> X [nrows, ld]
  void rows_launch(const double* R, int ldr, double* X, int ld, int nrows, const Solver& sv) {
    const int m = sv.m, nr = sv.pyb.mi;
    if (fdma_rows_lds_supported(m, nr)) {
      fdma_rows_lds(st, R, ldr, nr, X, ld, nrows, m, sv.nsets, sv.pybp.p, sv.diagp.p);
      return;
    }
    band_y(st, R, ldr, X, ld, nrows, sv.pyb.ref(), 0.0);
    if (fdma_rows_wave_supported(m)) fdma_rows_wave(st, X, ld, nrows, m, sv.nsets, sv.low.p, sv.dia.p, sv.up1.p, sv.up2.p);
    else fdma_rows(st, X, ld, nrows, m, sv.nsets, sv.low.p, sv.dia.p, sv.up1.p, sv.up2.p);
  }
};

// ----------------------------------------------------------------------- ensembles
// Independent confined instances (C4-style ensembles).  Instances with the same shape, Pr,
// dt, aspect and sidewall condition form a group: their fields and work arrays are slots of
// group slabs and the group's first instance (the leader) advances all of them with one
// set of launches (batched GEMMs over shared operators, per-instance ADI operators and nu;
// batched lane kernels; element-wise kernels over the whole slab), captured once as a HIP
// graph.  Groups run on their own streams.
class Ensemble {
 public:
  struct Group {
    std::unique_ptr<Navier::Slabs> slabs;
    std::vector<std::unique_ptr<Navier>> navs;  // navs[0] is the leader
    hipStream_t st = nullptr;                   // one of the ensemble's streams
    ~Group() {
      if (st) (void)hipStreamSynchronize(st);
      navs.clear();
      slabs.reset();
    }
  };
  // Streams the groups are spread over.  The groups are latency-bound chains of small
  // launches, so they overlap well on separate hardware queues; with more groups than
  // queues (HIP's default is 4 per process, one of them taken by the null stream) the
  // groups are packed onto NSTREAMS streams by longest-processing-time-first.
  static constexpr int NSTREAMS = 3;
  std::vector<hipStream_t> streams;
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<Navier*> inst;
  std::vector<int> group_of;

  // estimated time of one step of a group of B instances of nx x ny: a launch-latency floor
  // plus the dense transforms' work (measured on MI355X: ~450 us + 3.6e-6 us per B N^3)
  static double group_cost(int B, int nx, int ny) {
    return 450.0 + 3.6e-6 * (double)B * (double)nx * (double)ny * 0.5 * ((double)nx + ny);
  }
  ~Ensemble() {
    for (hipStream_t st : streams) (void)hipStreamSynchronize(st);
    groups.clear();
    for (hipStream_t st : streams) (void)hipStreamDestroy(st);
  }

  Ensemble(int n, const rbc_desc* d) {
    if (n < 1 || !d) throw std::invalid_argument("ensemble: no instances");
    std::map<std::tuple<int, int, double, double, double, int>, std::vector<int>> keys;
    for (int i = 0; i < n; ++i) {
      if (d[i].nx < 8 || d[i].ny < 8) throw std::invalid_argument("ensemble: nx and ny must be >= 8");
      keys[std::make_tuple(d[i].nx, d[i].ny, d[i].pr, d[i].dt, d[i].aspect, d[i].adiabatic != 0)].push_back(i);
    }
    inst.assign(n, nullptr);
    group_of.assign(n, -1);
    // groups in decreasing cost, each onto the stream with the least work so far
    std::vector<std::pair<double, const std::vector<int>*>> order;
    for (auto& kv : keys) {
      const rbc_desc& d0 = d[kv.second[0]];
      order.emplace_back(group_cost((int)kv.second.size(), d0.nx, d0.ny), &kv.second);
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const auto& x, const auto& y) { return x.first > y.first; });
    const int ns = std::min<int>(NSTREAMS, (int)order.size());
    streams.assign(ns, nullptr);
    for (int k = 0; k < ns; ++k) HIPCHECK(hipStreamCreate(&streams[k]));
    try {
      build_groups(d, order);
    } catch (...) {
      for (hipStream_t st : streams) (void)hipStreamSynchronize(st);
      groups.clear();
      for (hipStream_t st : streams) (void)hipStreamDestroy(st);
      streams.clear();
      throw;
    }
  }
  void build_groups(const rbc_desc* d, const std::vector<std::pair<double, const std::vector<int>*>>& order) {
    std::vector<double> load(streams.size(), 0.0);
    for (auto& oc : order) {
      const std::vector<int>& mem = *oc.second;
      const int B = (int)mem.size();
      const rbc_desc& d0 = d[mem[0]];
      auto g = std::make_unique<Group>();
      const int si = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[si] += oc.first;
      g->st = streams[si];
      g->slabs = std::make_unique<Navier::Slabs>();
      g->slabs->alloc((size_t)d0.nx * d0.ny, B);
      for (int b = 0; b < B; ++b) {
        const rbc_desc& di = d[mem[b]];
        g->navs.push_back(std::make_unique<Navier>(false, di.nx, di.ny, di.ra, di.pr, di.dt, di.aspect,
                                                   di.adiabatic != 0, g->slabs.get(), b, g->st,
                                                   b > 0 ? g->navs[0].get() : nullptr));
        inst[mem[b]] = g->navs.back().get();
        group_of[mem[b]] = (int)groups.size();
      }
      Navier& L = *g->navs[0];
      if (B > 1) {
        L.nb = B;
        L.bs = (long)g->slabs->F;
        std::vector<double> nus(B);
        for (int b = 0; b < B; ++b) nus[b] = g->navs[b]->nu;
        L.NUARR.upload(nus);
        for (int w = 0; w < 3; ++w) {  // ADI operators of ux, uy, temp per instance
          for (int ax = 0; ax < 2; ++ax) {
            const DOp& op0 = ax == 0 ? L.solver[w]->hx : L.solver[w]->hy;
            const size_t sz = (size_t)op0.rows * op0.cols;
            L.ADIH[w][ax].alloc(sz * B);
            L.adi_s[w][ax] = (long)sz;
            for (int b = 0; b < B; ++b) {
              const DOp& op = ax == 0 ? g->navs[b]->solver[w]->hx : g->navs[b]->solver[w]->hy;
              HIPCHECK(hipMemcpy(L.ADIH[w][ax].p + b * sz, op.b.p, sz * sizeof(double), hipMemcpyDeviceToDevice));
            }
          }
        }
      }
      groups.push_back(std::move(g));
    }
  }
  // every instance advances nsteps steps; the groups run concurrently
  void update(int nsteps) {
    for (auto& g : groups) g->navs[0]->update_n(nsteps);
    synchronize();
    for (auto& g : groups) {
      const double t = g->navs[0]->time;
      for (auto& nv : g->navs) nv->time = t;
    }
  }
  void synchronize() {
    for (hipStream_t st : streams) HIPCHECK(hipStreamSynchronize(st));
  }
};

// ------------------------------------------------------------------ adjoint descent
// Navier2DAdjoint (src/navier/navier_adjoint.rs:128-1068): Farazmand's adjoint descent for
// steady states.  `ns` is the Navier-Stokes step whose residual (NS(q) - q) / dt_navier is
// descended (dt_navier = 1e-2, navier_adjoint.rs:231-232); `ad` is a Navier of the adjoint's
// dt whose fields T, UX, UY, P0, P1 are the adjoint fields [0] and whose transforms, work
// arrays, fieldbc, pressure solver (Poisson, c = 1 / scale^2, the adjoint's) and diagnostics
// serve the adjoint; its three other solver slots hold the Hholtz smoothers (1 - c lap),
// c = 1 / scale^2 (:247-268).  Both share one stream, and every array stays on the device.
class Adjoint {
 public:
  std::unique_ptr<Navier> ad, ns;
  DBuf F1[3];        // smoothed residuals [1] of temp, ux, uy (spectral, the fields' layout)
  DBuf UNS[3];       // unsmoothed residuals (fields_unsmoothed: ux, uy, temp; ortho space)
  DBuf U, V, TP;     // physical ux, uy, temp of the step start
  DBuf B, LXX, LYY;  // buoyancy and the residuals' second derivatives (ortho space)
  double time = 0.0, dt = 0.0, res_tol = 1e-8;  // RES_TOL (:114)

  Adjoint(bool per, int nx, int ny, double ra, double pr, double dt_, double aspect, bool adiabatic) : dt(dt_) {
    ad = std::make_unique<Navier>(per, nx, ny, ra, pr, dt_, aspect, adiabatic);
    ns = std::make_unique<Navier>(per, nx, ny, ra, pr, 1e-2, aspect, adiabatic, nullptr, 0, ad->st);
    Navier& n = *ad;
    const double w = 1e0;  // weight_laplacian
    const double cs[2] = {w / std::pow(n.scale[0], 2.0), w / std::pow(n.scale[1], 2.0)};
    n.solver[0] = std::shared_ptr<Solver>(make_solver(HHOLTZ, n.bx_u, n.by_u, cs));
    n.solver[1] = std::shared_ptr<Solver>(make_solver(HHOLTZ, n.bx_u, n.by_u, cs));
    n.solver[2] = std::shared_ptr<Solver>(make_solver(HHOLTZ, n.bx_t, n.by_u, cs));
    for (DBuf* b : {&F1[0], &F1[1], &F1[2], &UNS[0], &UNS[1], &UNS[2], &U, &V, &TP, &B, &LXX, &LYY}) b->alloc(n.FS);
  }
  ~Adjoint() {
    if (ad) (void)hipStreamSynchronize(ad->st);
    ns.reset();  // shares ad's stream
  }
  long spec_len(const Basis& bx) const { return (long)ad->rows_spec(bx) * ad->by_u.m; }
  long ortho_len() const { return (long)ad->rows_ortho(ad->bx_f) * ad->ny; }
  // CONV = M .* field.backward(gradient(F, [d0, d1])) + beta CONV  (conv_term.rs:22-42)
  void conv_acc(const Basis& bx, const double* F, int d0, int d1, const double* M, double beta) {
    Navier& n = *ad;
    const double* a = n.xop(bx, d0 ? "DS" : "S", F, n.by_u.m, n.W1.p, n.W9.p);
    const double* g = n.yop(n.by_u, d1 ? "DS" : "S", a, n.rows_ortho(bx), n.W2.p);
    const double* b = n.xop(n.bx_f, "BS", g, n.ny, n.W3.p, n.W9.p);
    n.yop(n.by_p, "BS", b, n.nx, n.CONV.p, M, beta);
  }
  // lap parts of a residual with x basis bx: LXX = d2/dx2, LYY = d2/dy2 (ortho space)
  void laplace_parts(const Basis& bx, const double* F) {
    Navier& n = *ad;
    const int ro = n.rows_ortho(bx);
    const double* a = n.xop(bx, "DDS", F, n.by_u.m, n.W1.p, n.W9.p);
    const double* r = n.yop(n.by_u, "S", a, ro, LXX.p);
    if (r != LXX.p) HIPCHECK(hipMemcpyAsync(LXX.p, r, sizeof(double) * ortho_len(), hipMemcpyDeviceToDevice, n.st));
    a = n.xop(bx, "S", F, n.by_u.m, n.W1.p, n.W9.p);
    r = n.yop(n.by_u, "DDS", a, ro, LYY.p);
    if (r != LYY.p) HIPCHECK(hipMemcpyAsync(LYY.p, r, sizeof(double) * ortho_len(), hipMemcpyDeviceToDevice, n.st));
  }
  // q = from_ortho(RHS) in the basis (bx, cheb_dirichlet)
  void from_ortho_into(const Basis& bx, double* q) {
    Navier& n = *ad;
    const double* a = n.xop(bx, "FO", n.RHS.p, n.ny, n.W1.p, n.W9.p);
    const double* r = n.yop(n.by_u, "FO", a, n.rows_spec(bx), q);
    if (r != q) HIPCHECK(hipMemcpyAsync(q, r, sizeof(double) * spec_len(bx), hipMemcpyDeviceToDevice, n.st));
  }
  void update_residual() {  // navier_adjoint.rs:727-750
    Navier &n = *ad, &s = *ns;
    const long lu = spec_len(n.bx_u), lt = spec_len(n.bx_t);
    HIPCHECK(hipMemcpyAsync(s.UX.p, n.UX.p, sizeof(double) * lu, hipMemcpyDeviceToDevice, n.st));
    HIPCHECK(hipMemcpyAsync(s.UY.p, n.UY.p, sizeof(double) * lu, hipMemcpyDeviceToDevice, n.st));
    HIPCHECK(hipMemcpyAsync(s.T.p, n.T.p, sizeof(double) * lt, hipMemcpyDeviceToDevice, n.st));
    s.update_n(1);
    ew_residual(n.st, s.UX.p, s.UX.p, n.UX.p, s.dt, lu);
    ew_residual(n.st, s.UY.p, s.UY.p, n.UY.p, s.dt, lu);
    ew_residual(n.st, s.T.p, s.T.p, n.T.p, s.dt, lt);
    s.to_ortho(s.bx_u, s.by_u, s.UX.p, UNS[0].p);
    s.to_ortho(s.bx_u, s.by_u, s.UY.p, UNS[1].p);
    s.to_ortho(s.bx_t, s.by_u, s.T.p, UNS[2].p);
    // smoothed and negated: ux1, uy1, temp1 (F1[1], F1[2], F1[0])
    n.solve(0, UNS[0].p, F1[1].p);
    n.solve(1, UNS[1].p, F1[2].p);
    n.solve(2, UNS[2].p, F1[0].p);
    ew_scale(n.st, F1[1].p, F1[1].p, -1.0, lu);
    ew_scale(n.st, F1[2].p, F1[2].p, -1.0, lu);
    ew_scale(n.st, F1[0].p, F1[0].p, -1.0, lt);
  }
  // solve_ux / solve_uy (navier_adjoint.rs:581-626)
  void solve_vel(int which) {
    Navier& n = *ad;
    double* q = which ? n.UY.p : n.UX.p;
    const double* f = which ? F1[2].p : F1[1].p;
    n.to_ortho(n.bx_u, n.by_u, q, n.A.p);
    const double* gp = which ? n.yop(n.by_p, "DS", n.P0.p, n.rows_ortho(n.bx_p), n.G.p)
                             : n.xop(n.bx_p, "DS", n.P0.p, n.ny, n.G.p, n.W9.p);
    if (!which) {  // + ux dux1/dx + uy dux1/dy + ux dux1/dx + uy duy1/dx + T dT1/dx + T_bc dT1/dx
      conv_acc(n.bx_u, F1[1].p, 1, 0, U.p, 0.0);
      conv_acc(n.bx_u, F1[1].p, 0, 1, V.p, 1.0);
      conv_acc(n.bx_u, F1[1].p, 1, 0, U.p, 1.0);
      conv_acc(n.bx_u, F1[2].p, 1, 0, V.p, 1.0);
      conv_acc(n.bx_t, F1[0].p, 1, 0, TP.p, 1.0);
      conv_acc(n.bx_t, F1[0].p, 1, 0, n.BCv.p, 1.0);
    } else {  // + ux duy1/dx + uy duy1/dy + ux dux1/dy + uy duy1/dy + T dT1/dy + T_bc dT1/dy
      conv_acc(n.bx_u, F1[2].p, 1, 0, U.p, 0.0);
      conv_acc(n.bx_u, F1[2].p, 0, 1, V.p, 1.0);
      conv_acc(n.bx_u, F1[1].p, 0, 1, U.p, 1.0);
      conv_acc(n.bx_u, F1[2].p, 0, 1, V.p, 1.0);
      conv_acc(n.bx_t, F1[0].p, 0, 1, TP.p, 1.0);
      conv_acc(n.bx_t, F1[0].p, 0, 1, n.BCv.p, 1.0);
    }
    n.forward_dealias_conv();
    laplace_parts(n.bx_u, f);
    ew_adj_rhs(n.st, n.RHS.p, n.A.p, gp, n.CV.p, nullptr, LXX.p, LYY.p, dt, n.nu, ortho_len());
    from_ortho_into(n.bx_u, q);
  }
  void solve_temp() {  // navier_adjoint.rs:629-648
    Navier& n = *ad;
    n.to_ortho(n.bx_t, n.by_u, n.T.p, n.A.p);
    conv_acc(n.bx_t, F1[0].p, 1, 0, U.p, 0.0);
    conv_acc(n.bx_t, F1[0].p, 0, 1, V.p, 1.0);
    n.forward_dealias_conv();
    n.to_ortho(n.bx_u, n.by_u, F1[2].p, B.p);  // adjoint buoyancy: uy1
    laplace_parts(n.bx_t, F1[0].p);
    ew_adj_rhs(n.st, n.RHS.p, n.A.p, nullptr, n.CV.p, B.p, LXX.p, LYY.p, dt, n.ka, ortho_len());
    from_ortho_into(n.bx_t, n.T.p);
  }
  void update() {  // Integrate::update (navier_adjoint.rs:760-790)
    Navier& n = *ad;
    n.backward_vel(n.UX.p, U.p);
    n.backward_vel(n.UY.p, V.p);
    {
      const double* a = n.xop(n.bx_t, "BS", n.T.p, n.by_u.m, n.W1.p, n.W9.p);
      const double* r = n.yop(n.by_u, "BS", a, n.nx, TP.p);
      if (r != TP.p) HIPCHECK(hipMemcpyAsync(TP.p, r, sizeof(double) * n.nx * n.ny, hipMemcpyDeviceToDevice, n.st));
    }
    update_residual();
    solve_vel(0);
    solve_vel(1);
    n.divergence(n.DIV.p);
    n.solve_pres();
    n.project_velocity();
    // update_pres (:720-724): pres += to_ortho(pseu) / dt, no viscous term
    n.to_ortho(n.bx_phi, n.by_phi, n.P1.p, n.W3.p);
    ew_axpy(n.st, n.P0.p, n.P0.p, n.W3.p, 1.0 / dt, (long)n.rows_ortho(n.bx_p) * n.ny);
    solve_temp();
    n.gpres_valid = false;
    time += dt;
  }
  static double norm(const std::vector<double>& v) {  // norm_l2_f64 / norm_l2_c64 (:902-913)
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
  }
  // [|ux1|, |uy1|, |temp1|, |unsmoothed ux|, |uy|, |temp|] (callback / exit, :846-897)
  void norms(double* out) {
    Navier& n = *ad;
    HIPCHECK(hipStreamSynchronize(n.st));
    out[0] = norm(F1[1].download(spec_len(n.bx_u)));
    out[1] = norm(F1[2].download(spec_len(n.bx_u)));
    out[2] = norm(F1[0].download(spec_len(n.bx_t)));
    for (int j = 0; j < 3; ++j) out[3 + j] = norm(UNS[j].download(ortho_len()));
  }
  bool exit() {  // :882-897
    Navier& n = *ad;
    n.divergence(n.DIV.p);
    HIPCHECK(hipStreamSynchronize(n.st));
    if (std::isnan(norm(n.DIV.download(ortho_len())))) return true;
    double r[6];
    norms(r);
    return r[0] + r[1] + r[2] < res_tol;
  }
};

}  // namespace rbc

// =================================================================== C ABI
using rbc::Navier;
using rbc::Space;

struct rbc_navier {
  std::unique_ptr<Navier> own;
  Navier* n = nullptr;
  bool member = false;  // an ensemble instance: owned and advanced by its ensemble
};
struct rbc_ensemble {
  std::unique_ptr<rbc::Ensemble> e;
  std::vector<std::unique_ptr<rbc_navier>> inst;
};
struct fs_space2 {
  std::unique_ptr<Space> s;
};
struct rbc_solver {
  std::unique_ptr<rbc::SolverDev> s;
};
struct rbc_adjoint {
  std::unique_ptr<rbc::Adjoint> a;
  rbc_navier parts[2];  // views of the adjoint fields and of the Navier-Stokes step (owned here)
};

static thread_local std::string g_err;

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return RBC_OK;
  } catch (const rbc::DeviceError& e) {
    g_err = e.what();
    return RBC_ERR_DEVICE;
  } catch (const rbc::NumericError& e) {
    g_err = e.what();
    return RBC_ERR_NUMERIC;
  } catch (const std::exception& e) {
    g_err = e.what();
    return RBC_ERR_ARG;
  } catch (...) {
    g_err = "unknown error";
    return RBC_ERR_ARG;
  }
}

#define NEED(h)                                        \
  do {                                                 \
    if (!(h) || !(h)->n) throw std::invalid_argument("null handle"); \
  } while (0)

extern "C" {

const char* rbc_last_error(void) { return g_err.c_str(); }
const char* rbc_version(void) { return "rustpde_amd 0.2 (gfx950; FP64 and mixed FP32/FP64 steps)"; }

int rbc_create_confined(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic,
                        rbc_navier** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null out");
    auto* h = new rbc_navier();
    try {
      h->own = std::make_unique<Navier>(false, nx, ny, ra, pr, dt, aspect, adiabatic != 0);
      h->n = h->own.get();
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int rbc_create_periodic(int nx, int ny, double ra, double pr, double dt, double aspect, rbc_navier** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null out");
    auto* h = new rbc_navier();
    try {
      h->own = std::make_unique<Navier>(true, nx, ny, ra, pr, dt, aspect, true);
      h->n = h->own.get();
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int rbc_destroy(rbc_navier* h) {
  return guarded([&] {
    if (h && h->member) throw std::invalid_argument("an ensemble instance or adjoint part is destroyed with its owner");
    delete h;
  });
}
int rbc_field_shape(rbc_navier* h, int which, int space, int* n0, int* n1, int* is_complex) {
  return guarded([&] {
    NEED(h);
    int a, b, c;
    h->n->shape(which, space, a, b, c);
    if (n0) *n0 = a;
    if (n1) *n1 = b;
    if (is_complex) *is_complex = c;
  });
}
int rbc_get_field(rbc_navier* h, int which, int space, double* out) {
  return guarded([&] {
    NEED(h);
    if (!out) throw std::invalid_argument("null buffer");
    h->n->get(which, space, out);
  });
}
int rbc_set_field(rbc_navier* h, int which, int space, const double* in) {
  return guarded([&] {
    NEED(h);
    if (!in) throw std::invalid_argument("null buffer");
    h->n->set(which, space, in);
  });
}
int rbc_forward(rbc_navier* h, int which) {
  return guarded([&] {
    NEED(h);
    h->n->forward(which);
    HIPCHECK(hipStreamSynchronize(h->n->st));
  });
}
int rbc_backward(rbc_navier* h, int which) {
  return guarded([&] {
    NEED(h);
    h->n->backward(which);
    HIPCHECK(hipStreamSynchronize(h->n->st));
  });
}
int rbc_set_temp_bc(rbc_navier* h, const double* vhat) {
  return guarded([&] {
    NEED(h);
    if (!vhat) throw std::invalid_argument("null buffer");
    h->n->set(RBC_FIELDBC, RBC_SPEC, vhat);
  });
}
int rbc_set_velocity(rbc_navier* h, double amp, double m, double n) {
  return guarded([&] {
    NEED(h);
    h->n->apply_trig(RBC_UX, amp, m, n, true);
    h->n->apply_trig(RBC_UY, -amp, m, n, false);
    HIPCHECK(hipStreamSynchronize(h->n->st));
  });
}
int rbc_set_temperature(rbc_navier* h, double amp, double m, double n) {
  return guarded([&] {
    NEED(h);
    h->n->apply_trig(RBC_TEMP, -amp, m, n, false);
    HIPCHECK(hipStreamSynchronize(h->n->st));
  });
}
int rbc_update(rbc_navier* h, int nsteps) {
  return guarded([&] {
    NEED(h);
    if (h->member) throw std::invalid_argument("ensemble instances and adjoint parts advance with their owner");
    if (nsteps < 0) throw std::invalid_argument("nsteps < 0");
    h->n->update_n(nsteps);
  });
}
int rbc_synchronize(rbc_navier* h) {
  return guarded([&] {
    NEED(h);
    HIPCHECK(hipStreamSynchronize(h->n->st));
  });
}
double rbc_get_time(rbc_navier* h) { return (h && h->n) ? h->n->time : NAN; }
double rbc_get_dt(rbc_navier* h) { return (h && h->n) ? h->n->dt : NAN; }
int rbc_set_time(rbc_navier* h, double t) {
  return guarded([&] {
    NEED(h);
    h->n->time = t;
  });
}
int rbc_eval(rbc_navier* h, int what, double* out) {
  return guarded([&] {
    NEED(h);
    if (!out) throw std::invalid_argument("null out");
    Navier& n = *h->n;
    switch (what) {
      case RBC_NU: *out = n.eval_nu(); break;
      case RBC_NUVOL: *out = n.eval_nuvol(); break;
      case RBC_RE: *out = n.eval_re(); break;
      case RBC_DIV: *out = n.div_norm(); break;
      default: throw std::invalid_argument("unknown diagnostic");
    }
  });
}
int rbc_set_solid(rbc_navier* h, const double* mask, const double* value) {
  return guarded([&] {
    NEED(h);
    if (h->member) throw std::invalid_argument("solid masks of ensemble instances are not supported");
    h->n->set_solid(mask, value);
  });
}
int rbc_stats_enable(rbc_navier* h, double save_stat, double write_stat) {
  return guarded([&] {
    NEED(h);
    if (h->member) throw std::invalid_argument("statistics of ensemble instances are not supported");
    h->n->stats_enable(save_stat, write_stat);
  });
}
int rbc_stats_update(rbc_navier* h, int* updated) {
  return guarded([&] {
    NEED(h);
    const bool u = h->n->stats_update();
    if (updated) *updated = u ? 1 : 0;
  });
}
int rbc_stats_get(rbc_navier* h, int which, int space, double* out) {
  return guarded([&] {
    NEED(h);
    if (!out) throw std::invalid_argument("null out");
    h->n->stats_get(which, space, out);
  });
}
int rbc_stats_set(rbc_navier* h, int which, const double* vhat) {
  return guarded([&] {
    NEED(h);
    if (!vhat) throw std::invalid_argument("null vhat");
    h->n->stats_set(which, vhat);
  });
}
int rbc_stats_info(rbc_navier* h, double* save_stat, double* write_stat, double* avg_time, double* tot_time,
                   long long* num_save) {
  return guarded([&] {
    NEED(h);
    const Navier& n = *h->n;
    n.stats_need();
    if (save_stat) *save_stat = n.st_save;
    if (write_stat) *write_stat = n.st_write;
    if (avg_time) *avg_time = n.st_avg_time;
    if (tot_time) *tot_time = n.st_tot_time;
    if (num_save) *num_save = n.st_num;
  });
}
int rbc_stats_set_info(rbc_navier* h, double avg_time, double tot_time, long long num_save) {
  return guarded([&] {
    NEED(h);
    Navier& n = *h->n;
    n.stats_need();
    if (num_save < 0) throw std::invalid_argument("num_save < 0");
    n.st_avg_time = avg_time;
    n.st_tot_time = tot_time;
    n.st_num = num_save;
  });
}
int rbc_exit(rbc_navier* h, int* out) {
  return guarded([&] {
    NEED(h);
    if (!out) throw std::invalid_argument("null out");
    *out = std::isnan(h->n->div_norm()) ? 1 : 0;
  });
}
int rbc_tensor_info(rbc_navier* h, int s, int* n, int* has_fb) {
  return guarded([&] {
    NEED(h);
    if (s < 0 || s > 3) throw std::invalid_argument("solver index");
    const rbc::Solver& sv = *h->n->solver[s];
    if (sv.kind == rbc::ADI) { *n = 0; *has_fb = 0; return; }
    *n = (int)sv.lam.size();
    *has_fb = sv.has_fb ? 1 : 0;
  });
}
int rbc_get_tensor(rbc_navier* h, int s, double* lam, double* fwd, double* bwd) {
  return guarded([&] {
    NEED(h);
    if (s < 0 || s > 3) throw std::invalid_argument("solver index");
    const rbc::Solver& sv = *h->n->solver[s];
    if (sv.kind == rbc::ADI) throw std::invalid_argument("solver is HholtzAdi (no tensor)");
    std::memcpy(lam, sv.lam.data(), sv.lam.size() * sizeof(double));
    if (sv.has_fb) {
      std::memcpy(fwd, sv.fwd.data(), sv.fwd.size() * sizeof(double));
      std::memcpy(bwd, sv.bwd.data(), sv.bwd.size() * sizeof(double));
    }
  });
}
int rbc_set_tensor(rbc_navier* h, int s, const double* lam, const double* fwd, const double* bwd) {
  return guarded([&] {
    NEED(h);
    if (s < 0 || s > 3) throw std::invalid_argument("solver index");
    rbc::Solver& sv = *h->n->solver[s];
    if (sv.kind == rbc::ADI) throw std::invalid_argument("solver is HholtzAdi (no tensor)");
    HIPCHECK(hipStreamSynchronize(h->n->st));
    h->n->drop_graph();
    std::memcpy(sv.lam.data(), lam, sv.lam.size() * sizeof(double));
    if (sv.has_fb) {
      if (!fwd || !bwd) throw std::invalid_argument("fwd/bwd required");
      std::memcpy(sv.fwd.data(), fwd, sv.fwd.size() * sizeof(double));
      std::memcpy(sv.bwd.data(), bwd, sv.bwd.size() * sizeof(double));
    }
    sv.build_rows();
    const rbc::Solver &s0 = *h->n->solver[0], &s1 = *h->n->solver[1];
    h->n->vel_same = s0.lam == s1.lam && s0.fwd == s1.fwd && s0.bwd == s1.bwd;
  });
}

namespace {
// alloc = false: an ensemble member (its group's leader holds the work buffers)
void apply_option(Navier& n, int option, int value, bool alloc = true) {
  HIPCHECK(hipStreamSynchronize(n.st));
  n.drop_graph();
  n.gpres_valid = false;
  switch (option) {
    case RBC_OPT_GRAPH: n.use_graph = value != 0; break;
    case RBC_OPT_MFMA: n.use_mfma = value != 0; break;
    case RBC_OPT_FFT: n.use_fft = value != 0; break;
    case RBC_OPT_BANDS: n.use_bands = value != 0; break;
    case RBC_OPT_DCT: n.use_dct = value != 0; break;
    case RBC_OPT_BATCH: n.use_batch = value != 0; break;
    case RBC_OPT_FUSE: n.use_fuse = value != 0; break;
    case RBC_OPT_CONV3:
      if (value < 0 || value > 2) throw std::invalid_argument("RBC_OPT_CONV3: 0, 1 or 2");
      n.use_conv3 = value;
      break;
    case RBC_OPT_MIXED:
      if (value && n.solid_on)
        throw std::invalid_argument("RBC_OPT_MIXED: a handle with a solid obstacle runs an FP64 step; "
                                    "remove the obstacle first");
      n.mixed = value != 0;
      if (alloc && n.mixed && n.periodic && n.MIX.n < (11 * n.FS + 1) / 2) n.MIX.alloc((11 * n.FS + 1) / 2);
      if (alloc && n.mixed && !n.periodic) n.alloc_mixc();
      break;
    default: throw std::invalid_argument("unknown option");
  }
}
}  // namespace

int rbc_set_option(rbc_navier* h, int option, int value) {
  return guarded([&] {
    NEED(h);
    if (h->member) throw std::invalid_argument("options of ensemble instances are set with rbc_ensemble_set_option");
    apply_option(*h->n, option, value);
  });
}

int rbc_timing_enable(rbc_navier* h, int on) {
  return guarded([&] {
    NEED(h);
    HIPCHECK(hipStreamSynchronize(h->n->st));
    h->n->drop_graph();
    h->n->timer.reset();
    h->n->timer.on = on != 0;
  });
}
int rbc_timing_read(rbc_navier* h, int kernel_class, long long* launches, double* total_ms, double* work) {
  return guarded([&] {
    NEED(h);
    if (kernel_class < 0 || kernel_class >= rbc::KernelTimer::NCAT) throw std::invalid_argument("kernel class");
    HIPCHECK(hipStreamSynchronize(h->n->st));
    long long n;
    double ms, w;
    h->n->timer.read(kernel_class, n, ms, w);
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (work) *work = w;
  });
}

int fs_space2_create(int kind0, int n0, int kind1, int n1, fs_space2** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null out");
    if (kind0 < 0 || kind0 > 4 || kind1 < 0 || kind1 > 4) throw std::invalid_argument("unknown base kind");
    auto* s = new fs_space2();
    try {
      s->s = std::make_unique<Space>(kind0, n0, kind1, n1);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}
int fs_space2_destroy(fs_space2* sp) {
  return guarded([&] { delete sp; });
}
int fs_space2_shape(fs_space2* sp, int* p0, int* p1, int* s0, int* s1, int* o0, int* o1, int* is_complex) {
  return guarded([&] {
    if (!sp || !sp->s) throw std::invalid_argument("null space");
    Space& s = *sp->s;
    *p0 = s.b0.n; *p1 = s.b1.n;
    *s0 = s.b0.m; *s1 = s.b1.m;
    *o0 = s.b0.ortho(); *o1 = s.b1.ortho();
    *is_complex = s.cplx ? 1 : 0;
  });
}
#define FS_OP(name, code)                                                                    \
  int name(fs_space2* sp, const double* in, double* out, int batch) {                        \
    return guarded([&] {                                                                     \
      if (!sp || !sp->s || !in || !out || batch < 1) throw std::invalid_argument("bad argument"); \
      sp->s->run(code, in, out, batch, 0, 0, 1.0, 1.0);                                      \
    });                                                                                      \
  }
FS_OP(fs_forward, 0)
FS_OP(fs_backward, 1)
FS_OP(fs_to_ortho, 2)
FS_OP(fs_from_ortho, 3)
int fs_gradient(fs_space2* sp, const double* in, double* out, int d0, int d1, double scale0, double scale1, int batch) {
  return guarded([&] {
    if (!sp || !sp->s || !in || !out || batch < 1 || d0 < 0 || d1 < 0) throw std::invalid_argument("bad argument");
    sp->s->run(4, in, out, batch, d0, d1, scale0, scale1);
  });
}


int rbc_solver_create(int solver, int kind0, int n0, int kind1, int n1, double c0, double c1, rbc_solver** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null out");
    if (solver < 0 || solver > 2) throw std::invalid_argument("unknown solver kind");
    if (kind0 < 0 || kind0 > 4 || kind1 < 0 || kind1 > 4) throw std::invalid_argument("unknown base kind");
    auto* h = new rbc_solver();
    try {
      h->s = std::make_unique<rbc::SolverDev>(solver, kind0, n0, kind1, n1, c0, c1);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int rbc_solver_destroy(rbc_solver* h) {
  return guarded([&] { delete h; });
}
int rbc_solver_shape(rbc_solver* h, int* o0, int* o1, int* s0, int* s1, int* is_complex) {
  return guarded([&] {
    if (!h || !h->s) throw std::invalid_argument("null handle");
    rbc::SolverDev& s = *h->s;
    *o0 = s.b0.ortho(); *o1 = s.b1.ortho(); *s0 = s.b0.m; *s1 = s.b1.m; *is_complex = s.cplx ? 1 : 0;
  });
}
int rbc_solver_solve(rbc_solver* h, const double* d_rhs, double* d_out, int batch) {
  return guarded([&] {
    if (!h || !h->s || !d_rhs || !d_out || batch < 1) throw std::invalid_argument("bad argument");
    h->s->solve(d_rhs, d_out, batch);
  });
}
static int adjoint_create(bool per, int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic,
                          rbc_adjoint** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null out");
    auto h = std::make_unique<rbc_adjoint>();
    h->a = std::make_unique<rbc::Adjoint>(per, nx, ny, ra, pr, dt, aspect, adiabatic != 0);
    h->parts[0].n = h->a->ad.get();
    h->parts[1].n = h->a->ns.get();
    h->parts[0].member = h->parts[1].member = true;
    *out = h.release();
  });
}
int rbc_adjoint_create_confined(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic,
                                rbc_adjoint** out) {
  return adjoint_create(false, nx, ny, ra, pr, dt, aspect, adiabatic, out);
}
int rbc_adjoint_create_periodic(int nx, int ny, double ra, double pr, double dt, double aspect, rbc_adjoint** out) {
  return adjoint_create(true, nx, ny, ra, pr, dt, aspect, 1, out);
}
int rbc_adjoint_destroy(rbc_adjoint* h) {
  return guarded([&] { delete h; });
}
int rbc_adjoint_part(rbc_adjoint* h, int part, rbc_navier** out) {
  return guarded([&] {
    if (!h || !h->a || !out || part < 0 || part > 1) throw std::invalid_argument("bad argument");
    *out = &h->parts[part];
  });
}
int rbc_adjoint_update(rbc_adjoint* h, int nsteps) {
  return guarded([&] {
    if (!h || !h->a || nsteps < 0) throw std::invalid_argument("bad argument");
    for (int k = 0; k < nsteps; ++k) h->a->update();
  });
}
double rbc_adjoint_get_time(rbc_adjoint* h) { return h && h->a ? h->a->time : 0.0; }
double rbc_adjoint_get_dt(rbc_adjoint* h) { return h && h->a ? h->a->dt : 0.0; }
int rbc_adjoint_set_time(rbc_adjoint* h, double time) {
  return guarded([&] {
    if (!h || !h->a) throw std::invalid_argument("null handle");
    h->a->time = time;
  });
}
int rbc_adjoint_residual(rbc_adjoint* h, int which, int kind, double* host_out) {
  return guarded([&] {
    if (!h || !h->a || !host_out || which < 0 || which > 2 || kind < 0 || kind > 1)
      throw std::invalid_argument("bad argument");
    rbc::Adjoint& A = *h->a;
    Navier& n = *A.ad;
    HIPCHECK(hipStreamSynchronize(n.st));
    // kind 0: the smoothed residual [1] in the field's spectral layout; kind 1: the
    // unsmoothed one in the ortho space of the `field` space (fields_unsmoothed: ux, uy, temp)
    const rbc::Basis& bx = kind == 0 ? (which == RBC_TEMP ? n.bx_t : n.bx_u) : n.bx_f;
    const int r = kind == 0 ? n.rows_spec(bx) : n.rows_ortho(bx), c = kind == 0 ? n.by_u.m : n.ny;
    const int j = which == RBC_TEMP ? 2 : which - 1;
    std::vector<double> v = (kind == 0 ? A.F1[which] : A.UNS[j]).download((size_t)r * c);
    if (!n.periodic) {
      std::memcpy(host_out, v.data(), v.size() * sizeof(double));
      return;
    }
    const size_t pl = v.size() / 2;  // split planes -> interleaved complex
    for (size_t i = 0; i < pl; ++i) { host_out[2 * i] = v[i]; host_out[2 * i + 1] = v[pl + i]; }
  });
}
int rbc_adjoint_norms(rbc_adjoint* h, double* out) {
  return guarded([&] {
    if (!h || !h->a || !out) throw std::invalid_argument("bad argument");
    h->a->norms(out);
  });
}
int rbc_adjoint_exit(rbc_adjoint* h, int* out) {
  return guarded([&] {
    if (!h || !h->a || !out) throw std::invalid_argument("bad argument");
    *out = h->a->exit() ? 1 : 0;
  });
}
int rbc_solver_solve_noalloc(rbc_solver* h, const double* d_rhs, double* d_out, int batch) {
  return guarded([&] {
    if (!h || !h->s || !d_rhs || !d_out || batch < 1) throw std::invalid_argument("bad argument");
    if (h->s->mixed_iters > 0)
      throw std::invalid_argument("rbc_solver_solve_noalloc: not available with the mixed form (rbc_solver_set_mixed)");
    const size_t need = h->s->work_doubles(batch);
    if (need > rbc::noalloc::kScratchDoubles)
      throw std::invalid_argument("rbc_solver_solve_noalloc: the work arrays (" + std::to_string(need) +
                                  " doubles) exceed the fixed scratch of " +
                                  std::to_string(rbc::noalloc::kScratchDoubles) + " doubles");
    double* scratch = nullptr;
    try {
      scratch = rbc::noalloc::scratch_ptr();
    } catch (const std::runtime_error& e) {
      throw rbc::DeviceError(e.what());
    }
    rbc::GemmNoTune no_tune;
    h->s->solve_in(d_rhs, d_out, batch, scratch);
  });
}
int rbc_solver_set_mixed(rbc_solver* h, int iters) {
  return guarded([&] {
    if (!h || !h->s) throw std::invalid_argument("null handle");
    h->s->set_mixed(iters);
  });
}
int rbc_solver_residual(rbc_solver* h, const double* d_rhs, const double* d_sol, int batch, double* out) {
  return guarded([&] {
    if (!h || !h->s || !d_rhs || !d_sol || !out || batch < 1) throw std::invalid_argument("bad argument");
    *out = h->s->residual(d_rhs, d_sol, batch);
  });
}
int rbc_solver_tensor_info(rbc_solver* h, int* n, int* has_fb) {
  return guarded([&] {
    if (!h || !h->s) throw std::invalid_argument("null handle");
    const rbc::Solver& sv = *h->s->s;
    if (sv.kind == rbc::ADI) { *n = 0; *has_fb = 0; return; }
    *n = (int)sv.lam.size();
    *has_fb = sv.has_fb ? 1 : 0;
  });
}
int rbc_solver_get_tensor(rbc_solver* h, double* lam, double* fwd, double* bwd) {
  return guarded([&] {
    if (!h || !h->s) throw std::invalid_argument("null handle");
    const rbc::Solver& sv = *h->s->s;
    if (sv.kind == rbc::ADI) throw std::invalid_argument("solver is HholtzAdi (no tensor)");
    std::memcpy(lam, sv.lam.data(), sv.lam.size() * sizeof(double));
    if (sv.has_fb) {
      std::memcpy(fwd, sv.fwd.data(), sv.fwd.size() * sizeof(double));
      std::memcpy(bwd, sv.bwd.data(), sv.bwd.size() * sizeof(double));
    }
  });
}
int rbc_solver_set_tensor(rbc_solver* h, const double* lam, const double* fwd, const double* bwd) {
  return guarded([&] {
    if (!h || !h->s) throw std::invalid_argument("null handle");
    rbc::Solver& sv = *h->s->s;
    if (sv.kind == rbc::ADI) throw std::invalid_argument("solver is HholtzAdi (no tensor)");
    std::memcpy(sv.lam.data(), lam, sv.lam.size() * sizeof(double));
    if (sv.has_fb) {
      if (!fwd || !bwd) throw std::invalid_argument("fwd/bwd required");
      std::memcpy(sv.fwd.data(), fwd, sv.fwd.size() * sizeof(double));
      std::memcpy(sv.bwd.data(), bwd, sv.bwd.size() * sizeof(double));
    }
    sv.build_rows();
  });
}

int rbc_create_ensemble(int n, const rbc_desc* desc, rbc_ensemble** out) {
  return guarded([&] {
    if (!out || !desc) throw std::invalid_argument("null argument");
    auto h = std::make_unique<rbc_ensemble>();
    h->e = std::make_unique<rbc::Ensemble>(n, desc);
    for (int i = 0; i < n; ++i) {
      auto nv = std::make_unique<rbc_navier>();
      nv->n = h->e->inst[i];
      nv->member = true;
      h->inst.push_back(std::move(nv));
    }
    *out = h.release();
  });
}
int rbc_ensemble_destroy(rbc_ensemble* h) {
  return guarded([&] { delete h; });
}
int rbc_ensemble_size(rbc_ensemble* h, int* n, int* groups) {
  return guarded([&] {
    if (!h || !h->e) throw std::invalid_argument("null handle");
    if (n) *n = (int)h->inst.size();
    if (groups) *groups = (int)h->e->groups.size();
  });
}
int rbc_ensemble_instance(rbc_ensemble* h, int inst, rbc_navier** out) {
  return guarded([&] {
    if (!h || !h->e || !out) throw std::invalid_argument("null argument");
    if (inst < 0 || inst >= (int)h->inst.size()) throw std::invalid_argument("instance index");
    *out = h->inst[inst].get();
  });
}
int rbc_ensemble_update(rbc_ensemble* h, int nsteps) {
  return guarded([&] {
    if (!h || !h->e) throw std::invalid_argument("null handle");
    if (nsteps < 0) throw std::invalid_argument("nsteps < 0");
    h->e->update(nsteps);
  });
}
int rbc_ensemble_synchronize(rbc_ensemble* h) {
  return guarded([&] {
    if (!h || !h->e) throw std::invalid_argument("null handle");
    h->e->synchronize();
  });
}
int rbc_ensemble_set_option(rbc_ensemble* h, int option, int value) {
  return guarded([&] {
    if (!h || !h->e) throw std::invalid_argument("null handle");
    h->e->synchronize();
    for (auto& g : h->e->groups)  // the leader launches the group's kernels
      for (size_t k = 0; k < g->navs.size(); ++k) apply_option(*g->navs[k], option, value, k == 0);
  });
}
int rbc_ensemble_timing_enable(rbc_ensemble* h, int on) {
  return guarded([&] {
    if (!h || !h->e) throw std::invalid_argument("null handle");
    h->e->synchronize();
    for (auto& g : h->e->groups) {
      Navier& L = *g->navs[0];  // the leader launches the group's kernels
      L.drop_graph();
      L.timer.reset();
      L.timer.on = on != 0;
    }
  });
}
int rbc_ensemble_timing_read(rbc_ensemble* h, int kernel_class, long long* launches, double* total_ms, double* work) {
  return guarded([&] {
    if (!h || !h->e) throw std::invalid_argument("null handle");
    if (kernel_class < 0 || kernel_class >= rbc::KernelTimer::NCAT) throw std::invalid_argument("kernel class");
    h->e->synchronize();
    long long nt = 0;
    double mt = 0.0, wt = 0.0;
    for (auto& g : h->e->groups) {
      long long n;
      double ms, w;
      g->navs[0]->timer.read(kernel_class, n, ms, w);
      nt += n;
      mt += ms;
      wt += w;
    }
    if (launches) *launches = nt;
    if (total_ms) *total_ms = mt;
    if (work) *work = wt;
  });
}

int rbc_dev_malloc(void** out, unsigned long long bytes) {
  return guarded([&] { HIPCHECK(hipMalloc(out, bytes)); });
}
int rbc_dev_free(void* p) {
  return guarded([&] { HIPCHECK(hipFree(p)); });
}
int rbc_memcpy_h2d(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] { HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int rbc_memcpy_d2h(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] { HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int rbc_memcpy_d2d(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] { HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}

}  // extern "C"
