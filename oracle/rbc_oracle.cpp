// CPU FP64 restatement of rustpde's 2-D Rayleigh-Benard time step.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle: the correctness checker for the HIP
// path and the CPU baseline timed beside it.  Nothing in the product library links,
// loads or calls this file; only tests/, the smoke check and bench.py's cpu_baseline
// leg do, through oracle/oracle.py.
//
// Every routine cites the reference file:line it restates (paths relative to the
// reference repository root).  Arrays are row-major [n0, n1] (ndarray C order);
// complex values are interleaved (re, im) like num_complex::Complex<f64>.
//
// Arithmetic is plain IEEE double without FMA contraction (-ffp-contract=off), the
// order of operations inside each lane follows the reference.  The DCT-I and real
// FFT are this file's own (radix-2 FFT when the length is a power of two, a direct
// sum with exact argument reduction otherwise), so they agree with rustdct/rustfft
// to rounding, not bitwise.
//
// Parity status (each function is pinned by a test in tests/test_oracle_*.py):
//  - transforms, stencils, derivatives, Fourier: reference doc/unit known answers
//  - Fdma, MatVecFdma, HholtzAdi, Hholtz, Poisson: pypde golden vectors, closed
//    forms, residual tests, and the numpy/LAPACK restatement
//  - eig: reconstruction residual and LAPACK (numpy.linalg.eig) eigenvalues
//  - Navier2D::update: numpy restatement fixtures (tests/golden) + invariants
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
using vecd = std::vector<double>;

// ------------------------------------------------------------------ helpers
template <class T> struct Arr {
  int n0 = 0, n1 = 0;
  std::vector<T> d;
  Arr() = default;
  Arr(int a, int b) : n0(a), n1(b), d((size_t)a * b, T(0)) {}
  T& operator()(int i, int j) { return d[(size_t)i * n1 + j]; }
  const T& operator()(int i, int j) const { return d[(size_t)i * n1 + j]; }
};

inline double re_of(double x) { return x; }
inline double re_of(const cd& x) { return x.real(); }

// real scalar * generic (componentwise for complex, as num_complex Mul<T>)
inline double mulr(double a, double s) { return a * s; }
inline cd mulr(const cd& a, double s) { return cd(a.real() * s, a.imag() * s); }
inline double divr(double a, double s) { return a / s; }
inline cd divr(const cd& a, double s) { return cd(a.real() / s, a.imag() / s); }

// ------------------------------------------------------------------ FFT / DCT
bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// In-place iterative radix-2 complex FFT, sign = -1 forward, +1 backward (unnormalised).
void fft_pow2(std::vector<cd>& a, int sign) {
  const int n = (int)a.size();
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len / 2;
    for (int k = 0; k < half; ++k) {
      // twiddle computed directly (no recurrence) to keep errors ~1 ulp
      double ang = sign * 2.0 * M_PI * (double)k / (double)len;
      cd w(std::cos(ang), std::sin(ang));
      for (int s = 0; s < n; s += len) {
        cd u = a[s + k], v = a[s + k + half] * w;
        a[s + k] = u + v;
        a[s + k + half] = u - v;
      }
    }
  }
}

// cos(pi * r / q) and sin(pi * r / q) for r in [0, 2q), tabulated once per q (per thread):
// the direct sums below take the same values as calling std::cos / std::sin each time
struct TrigTable {
  std::vector<double> c, s;
};
const TrigTable& trig_table(long long q) {
  thread_local std::vector<std::pair<long long, std::unique_ptr<TrigTable>>> cache;
  for (auto& e : cache)
    if (e.first == q) return *e.second;
  auto t = std::make_unique<TrigTable>();
  t->c.resize(2 * q);
  t->s.resize(2 * q);
  for (long long r = 0; r < 2 * q; ++r) {
    t->c[r] = std::cos(M_PI * (double)r / (double)q);
    t->s[r] = std::sin(M_PI * (double)r / (double)q);
  }
  cache.emplace_back(q, std::move(t));
  return *cache.back().second;
}
// cos(pi * p / q) with exact integer reduction of p modulo 2q
inline double cospi_ratio(long long p, long long q) {
  long long r = p % (2 * q);
  if (r < 0) r += 2 * q;
  return trig_table(q).c[r];
}
inline double sinpi_ratio(long long p, long long q) {
  long long r = p % (2 * q);
  if (r < 0) r += 2 * q;
  return trig_table(q).s[r];
}

// DCT-I as used by ndrustfft::nddct1 (rustdct 0.6):
//   y_k = x_0 + (-1)^k x_{n-1} + 2 * sum_{j=1}^{n-2} x_j cos(pi j k / (n-1))
void dct1(const double* x, double* y, int n) {
  const int N = 2 * (n - 1);
  if (n >= 3 && is_pow2(N)) {
    std::vector<cd> a(N);
    for (int j = 0; j < n; ++j) a[j] = cd(x[j], 0.0);
    for (int j = n; j < N; ++j) a[j] = cd(x[N - j], 0.0);
    fft_pow2(a, -1);
    for (int k = 0; k < n; ++k) y[k] = a[k].real();
    return;
  }
  const TrigTable& tb = trig_table(n - 1);
  const long long q2 = 2LL * (n - 1);
  for (int k = 0; k < n; ++k) {
    double s = 0.0;
    long long r = 0;  // r = jk mod 2(n-1): cos(pi jk / (n-1)) = tb.c[r]
    for (int j = 1; j < n - 1; ++j) {
      r += k;
      if (r >= q2) r -= q2;
      s += x[j] * tb.c[r];
    }
    y[k] = x[0] + ((k & 1) ? -x[n - 1] : x[n - 1]) + 2.0 * s;
  }
}

// r2c: X_k = sum_j v_j e^{-2 pi i jk/n}, k = 0..n/2 (realfft, unnormalised)
void rfft(const double* v, cd* X, int n) {
  const int m = n / 2 + 1;
  if (is_pow2(n)) {
    std::vector<cd> a(n);
    for (int j = 0; j < n; ++j) a[j] = cd(v[j], 0.0);
    fft_pow2(a, -1);
    for (int k = 0; k < m; ++k) X[k] = a[k];
    return;
  }
  const TrigTable& tb = trig_table(n);
  for (int k = 0; k < m; ++k) {
    double sr = 0.0, si = 0.0;
    const long long step = (2LL * k) % (2LL * n);
    long long r = 0;  // r = 2jk mod 2n: cos / sin(2 pi jk / n)
    for (int j = 0; j < n; ++j) {
      sr += v[j] * tb.c[r];
      si -= v[j] * tb.s[r];
      r += step;
      if (r >= 2LL * n) r -= 2LL * n;
    }
    X[k] = cd(sr, si);
  }
}

// c2r normalised by 1/n (ndrustfft ndifft_r2c); imaginary parts of the k = 0 and
// (even n) k = n/2 entries are ignored, the half-spectrum convention.
void irfft(const cd* X, double* v, int n) {
  const int m = n / 2 + 1;
  std::vector<cd> a(n);
  for (int k = 0; k < m; ++k) a[k] = X[k];
  a[0] = cd(X[0].real(), 0.0);
  if (n % 2 == 0) a[n / 2] = cd(X[n / 2].real(), 0.0);
  for (int k = m; k < n; ++k) a[k] = std::conj(a[n - k]);
  if (is_pow2(n)) {
    fft_pow2(a, +1);
    for (int j = 0; j < n; ++j) v[j] = a[j].real() * (1.0 / n);
    return;
  }
  const TrigTable& tb = trig_table(n);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    const long long step = (2LL * j) % (2LL * n);
    long long r = 0;  // r = 2jk mod 2n
    for (int k = 0; k < n; ++k) {
      s += a[k].real() * tb.c[r] - a[k].imag() * tb.s[r];
      r += step;
      if (r >= 2LL * n) r -= 2LL * n;
    }
    v[j] = s * (1.0 / n);
  }
}

// ------------------------------------------------------------------ 1-D bases
enum Kind { CHEB = 0, CD = 1, CN = 2, CDBC = 3, FO = 4 };

struct Base {
  Kind k;
  int n, m;           // physical length, spectral length
  vecd diag, low2;    // composite stencil  composite_stencil.rs:117-162
  vecd main, off;     // tdma normal equations  composite_stencil.rs:165-176
  double t0[2], t1[2];// boundary stencil  composite_stencil.rs:287-296
  vecd corr_fwd, corr_bwd;  // ortho.rs:51-66

  Base(Kind kind, int n_) : k(kind), n(n_) {
    switch (k) {
      case CHEB: m = n; break;
      case CD:
      case CN: m = n - 2; break;
      case CDBC: m = 2; break;
      case FO: m = n / 2 + 1; break;
    }
    if (k == CD || k == CN) {
      diag.assign(m, 1.0);
      low2.assign(m, 0.0);
      for (int i = 0; i < m; ++i) {
        if (k == CD) low2[i] = -1.0;
        else low2[i] = -1.0 * (double)((long long)i * i) / (double)((long long)(i + 2) * (i + 2));
      }
      main.assign(m, 0.0);
      off.assign(m - 2, 0.0);
      for (int i = 0; i < m; ++i) main[i] = diag[i] * diag[i] + low2[i] * low2[i];
      for (int i = 0; i < m - 2; ++i) off[i] = diag[i + 2] * low2[i];
    }
    if (k == CDBC) { t0[0] = 0.5; t0[1] = 0.5; t1[0] = -0.5; t1[1] = 0.5; }
    if (k != FO) {
      corr_fwd.resize(n);
      corr_bwd.resize(n);
      for (int i = 0; i < n; ++i) {
        double c = std::pow(-1.0, (double)i);
        corr_fwd[i] = c * (1.0 / (double)(n - 1));
        corr_bwd[i] = c / 2.0;
      }
    }
  }
  bool fourier() const { return k == FO; }
  int ortho_len() const { return k == FO ? m : n; }

  // nodes: ortho.rs:71-80 ; c2c.rs:63-66
  vecd coords() const {
    vecd x(n);
    if (k == FO) {
      double step = 2.0 * M_PI / (double)n;
      for (int i = 0; i < n; ++i) x[i] = 0.0 + step * (double)i;
    } else {
      double mm = (double)(n - 1);
      for (int i = 0; i < n; ++i) x[i] = -std::sin(M_PI * (mm - 2.0 * (double)i) / (2.0 * mm));
    }
    return x;
  }

  // Dense matrices for solver ingredients (field.rs:194-252)
  Arr<double> mass() const {
    if (k == CHEB) { Arr<double> a(n, n); for (int i = 0; i < n; ++i) a(i, i) = 1.0; return a; }
    if (k == FO) { Arr<double> a(m, m); for (int i = 0; i < m; ++i) a(i, i) = 1.0; return a; }
    Arr<double> s(n, m);
    if (k == CDBC) { s(0, 0) = t0[0]; s(0, 1) = t0[1]; s(1, 0) = t1[0]; s(1, 1) = t1[1]; return s; }
    for (int i = 0; i < m; ++i) { s(i, i) = diag[i]; s(i + 2, i) = low2[i]; }
    return s;
  }
  Arr<double> laplace() const {  // dmsuite.rs:24-49 ; r2c.rs:372-378
    if (k == FO) { Arr<double> a(m, m); for (int i = 0; i < m; ++i) a(i, i) = -(double)i * (double)i; return a; }
    Arr<double> d(n, n);
    for (int p = 0; p < n; ++p)
      for (int q = p + 2; q < n; ++q)
        if ((p + q) % 2 == 0) d(p, q) = (double)((long long)q * ((long long)q * q - (long long)p * p));
    for (int q = 0; q < n; ++q) d(0, q) *= 0.5;
    return d;
  }
  Arr<double> laplace_inv() const {  // ortho.rs:147-174 ; r2c.rs:381-387
    if (k == FO) { Arr<double> a = laplace(); for (int i = 1; i < m; ++i) a(i, i) = 1.0 / a(i, i); return a; }
    Arr<double> p(n, n);
    p(2, 0) = 0.25;
    for (int i = 3; i < n; ++i) p(i, i - 2) = 1.0 / (double)(4LL * i * (i - 1));
    for (int i = 2; i < n - 2; ++i) p(i, i) = -1.0 / (double)(2LL * ((long long)i * i - 1));
    for (int i = 2; i < n - 4; ++i) p(i, i + 2) = 1.0 / (double)(4LL * i * (i + 1));
    return p;
  }
  Arr<double> laplace_inv_eye() const {  // ortho.rs:177-180 ; r2c.rs:391-394
    int sh = (k == FO) ? 1 : 2, len = (k == FO) ? m : n;
    Arr<double> e(len - sh, len);
    for (int i = 0; i < len - sh; ++i) e(i, i + sh) = 1.0;
    return e;
  }
};

// ---- lane operations (each acts on one contiguous lane) ------------------------
// Chebyshev forward ortho.rs:235-258
void cheb_forward_lane(const Base& b, const double* v, double* c) {
  dct1(v, c, b.n);
  for (int i = 0; i < b.n; ++i) c[i] *= b.corr_fwd[i];
  c[0] *= 0.5;
  c[b.n - 1] *= 0.5;
}
// Chebyshev backward ortho.rs:288-312
void cheb_backward_lane(const Base& b, const double* c, double* v) {
  std::vector<double> buf(c, c + b.n);
  for (int i = 0; i < b.n; ++i) buf[i] *= b.corr_bwd[i];
  buf[0] *= 2.0;
  buf[b.n - 1] *= 2.0;
  dct1(buf.data(), v, b.n);
}
// to_ortho  composite_stencil.rs:207-229, :355-360
template <class T> void to_ortho_lane(const Base& b, const T* c, T* p) {
  const int n = b.n;
  if (b.k == CHEB || b.k == FO) { for (int i = 0; i < b.m; ++i) p[i] = c[i]; return; }
  for (int i = 0; i < n; ++i) p[i] = T(0);
  if (b.k == CDBC) {
    p[0] = mulr(c[0], b.t0[0]) + mulr(c[1], b.t0[1]);
    p[1] = mulr(c[0], b.t1[0]) + mulr(c[1], b.t1[1]);
    return;
  }
  p[0] = mulr(c[0], b.diag[0]);
  p[1] = mulr(c[1], b.diag[1]);
  for (int i = 2; i < n - 2; ++i) p[i] = mulr(c[i], b.diag[i]) + mulr(c[i - 2], b.low2[i - 2]);
  p[n - 2] = mulr(c[n - 4], b.low2[n - 4]);
  p[n - 1] = mulr(c[n - 3], b.low2[n - 3]);
}
// tdma  funspace/src/chebyshev/linalg.rs:14-57
template <class T> void tdma_lane(const vecd& a, const vecd& bb, const vecd& c, T* d, int n) {
  std::vector<T> x(n), g(n);
  vecd w(n - 2);
  w[0] = c[0] / bb[0];
  g[0] = divr(d[0], bb[0]);
  if (c.size() > 1) w[1] = c[1] / bb[1];
  g[1] = divr(d[1], bb[1]);
  for (int i = 2; i < n - 2; ++i) w[i] = c[i] / (bb[i] - a[i - 2] * w[i - 2]);
  for (int i = 2; i < n; ++i) g[i] = divr(d[i] - mulr(g[i - 2], a[i - 2]), bb[i] - a[i - 2] * w[i - 2]);
  x[n - 1] = g[n - 1];
  x[n - 2] = g[n - 2];
  for (int i = n - 2; i >= 1; --i) x[i - 1] = g[i - 1] - mulr(x[i + 1], w[i - 1]);
  for (int i = 0; i < n; ++i) d[i] = x[i];
}
// from_ortho  composite_stencil.rs:250-276, :366-388
template <class T> void from_ortho_lane(const Base& b, const T* p, T* c) {
  if (b.k == CHEB || b.k == FO) { for (int i = 0; i < b.m; ++i) c[i] = p[i]; return; }
  if (b.k == CDBC) {
    T c0 = mulr(p[0], b.t0[0]) + mulr(p[1], b.t1[0]);
    T c1 = mulr(p[0], b.t0[1]) + mulr(p[1], b.t1[1]);
    double a = b.t0[0] * b.t0[0] + b.t1[0] * b.t1[0];
    double bq = b.t0[0] * b.t0[1] + b.t1[0] * b.t1[1];
    double cq = b.t0[1] * b.t0[0] + b.t1[1] * b.t1[0];
    double dq = b.t0[1] * b.t0[1] + b.t1[1] * b.t1[1];
    double det = 1.0 / (a * dq - bq * cq);
    c[0] = mulr(mulr(c0, dq) - mulr(c1, bq), det);
    c[1] = mulr(mulr(c1, a) - mulr(c0, cq), det);
    return;
  }
  for (int i = 0; i < b.m; ++i) c[i] = mulr(p[i], b.diag[i]) + mulr(p[i + 2], b.low2[i]);
  tdma_lane(b.off, b.main, b.off, c, b.m);
}
// Chebyshev differentiation recurrence  ortho.rs:107-125
template <class T> void cheb_diff_lane(T* d, int n, int times) {
  for (int t = 0; t < times; ++t) {
    d[0] = d[1];
    for (int i = 1; i < n - 1; ++i) d[i] = mulr(d[i + 1], 2.0 * (double)(i + 1));
    d[n - 1] = T(0);
    for (int i = n - 3; i >= 1; --i) d[i] = d[i] + d[i + 2];
    d[0] = d[0] + divr(d[2], 2.0);
  }
}
// Fourier differentiation  r2c.rs:88-99 : multiply by (i k), n times
void fo_diff_lane(cd* d, int m, int times) {
  for (int t = 0; t < times; ++t)
    for (int k = 0; k < m; ++k) {
      double a = d[k].real(), bq = d[k].imag(), kk = (double)k;
      d[k] = cd(a * 0.0 - bq * kk, a * kk + bq * 0.0);
    }
}

// apply f(lane_in, lane_out) along `axis` of a 2-D array (ndarray lanes)
template <class Tin, class Tout, class F>
Arr<Tout> along(const Arr<Tin>& a, int axis, int nout, F f) {
  Arr<Tout> out = axis == 0 ? Arr<Tout>(nout, a.n1) : Arr<Tout>(a.n0, nout);
  const int nl = axis == 0 ? a.n1 : a.n0;
  const int nin = axis == 0 ? a.n0 : a.n1;
#pragma omp parallel
  {
    std::vector<Tin> li(nin);
    std::vector<Tout> lo(nout);
#pragma omp for schedule(static)
    for (int l = 0; l < nl; ++l) {
      if (axis == 0) for (int i = 0; i < nin; ++i) li[i] = a(i, l);
      else for (int i = 0; i < nin; ++i) li[i] = a(l, i);
      f(li.data(), lo.data());
      if (axis == 0) for (int i = 0; i < nout; ++i) out(i, l) = lo[i];
      else for (int i = 0; i < nout; ++i) out(l, i) = lo[i];
    }
  }
  return out;
}

// ------------------------------------------------------------------ Space2
// funspace/src/space2.rs:136-356.  T = spectral scalar (double or cd).
template <class T> struct Space2 {
  Base b0, b1;
  Space2(const Base& a, const Base& b) : b0(a), b1(b) {}
  int nphys0() const { return b0.n; }
  int nphys1() const { return b1.n; }
  int nspec0() const { return b0.m; }
  int nspec1() const { return b1.m; }
};

// axis-0 helpers that depend on whether axis 0 is Fourier
Arr<double> fwd_axis0(const Base& b, const Arr<double>& in, double*) {
  return along<double, double>(in, 0, b.m, [&](const double* x, double* y) {
    std::vector<double> p(b.n);
    cheb_forward_lane(b, x, p.data());
    from_ortho_lane<double>(b, p.data(), y);
  });
}
Arr<cd> fwd_axis0(const Base& b, const Arr<double>& in, cd*) {
  return along<double, cd>(in, 0, b.m, [&](const double* x, cd* y) { rfft(x, y, b.n); });
}
Arr<double> bwd_axis0(const Base& b, const Arr<double>& in) {
  return along<double, double>(in, 0, b.n, [&](const double* x, double* y) {
    std::vector<double> p(b.n);
    to_ortho_lane<double>(b, x, p.data());
    cheb_backward_lane(b, p.data(), y);
  });
}
Arr<double> bwd_axis0(const Base& b, const Arr<cd>& in) {
  return along<cd, double>(in, 0, b.n, [&](const cd* x, double* y) { irfft(x, y, b.n); });
}

template <class T> Arr<T> space_forward(const Space2<T>& s, const Arr<double>& v) {
  const Base& b1 = s.b1;
  Arr<double> buf = along<double, double>(v, 1, b1.m, [&](const double* x, double* y) {
    std::vector<double> p(b1.n);
    cheb_forward_lane(b1, x, p.data());
    from_ortho_lane<double>(b1, p.data(), y);
  });
  return fwd_axis0(s.b0, buf, (T*)nullptr);
}
template <class T> Arr<double> space_backward(const Space2<T>& s, const Arr<T>& vh) {
  const Base& b1 = s.b1;
  Arr<double> buf = bwd_axis0(s.b0, vh);
  return along<double, double>(buf, 1, b1.n, [&](const double* x, double* y) {
    std::vector<double> p(b1.n);
    to_ortho_lane<double>(b1, x, p.data());
    cheb_backward_lane(b1, p.data(), y);
  });
}
template <class T> Arr<T> space_to_ortho(const Space2<T>& s, const Arr<T>& vh) {
  Arr<T> buf = along<T, T>(vh, 0, s.b0.ortho_len(), [&](const T* x, T* y) { to_ortho_lane<T>(s.b0, x, y); });
  return along<T, T>(buf, 1, s.b1.ortho_len(), [&](const T* x, T* y) { to_ortho_lane<T>(s.b1, x, y); });
}
template <class T> Arr<T> space_from_ortho(const Space2<T>& s, const Arr<T>& p) {
  Arr<T> buf = along<T, T>(p, 0, s.b0.m, [&](const T* x, T* y) { from_ortho_lane<T>(s.b0, x, y); });
  return along<T, T>(buf, 1, s.b1.m, [&](const T* x, T* y) { from_ortho_lane<T>(s.b1, x, y); });
}
template <class T> void diff_lane(const Base& b, const T* x, T* y, int times);
template <> void diff_lane<double>(const Base& b, const double* x, double* y, int times) {
  to_ortho_lane<double>(b, x, y);
  cheb_diff_lane<double>(y, b.n, times);
}
template <> void diff_lane<cd>(const Base& b, const cd* x, cd* y, int times) {
  if (b.k == FO) {
    for (int i = 0; i < b.m; ++i) y[i] = x[i];
    fo_diff_lane(y, b.m, times);
    return;
  }
  to_ortho_lane<cd>(b, x, y);
  cheb_diff_lane<cd>(y, b.n, times);
}
inline void scale_div(double& v, double sc) { v = v / sc; }
inline void scale_div(cd& v, double sc) {  // num-complex 0.4 Complex / Complex(sc, 0)
  double a = v.real(), b = v.imag();
  double ns = sc * sc + 0.0 * 0.0;
  v = cd((a * sc + b * 0.0) / ns, (b * sc - a * 0.0) / ns);
}
// gradient  space2.rs:247-264
template <class T> Arr<T> space_gradient(const Space2<T>& s, const Arr<T>& vh, int d0, int d1, const double* scale) {
  Arr<T> buf = along<T, T>(vh, 0, s.b0.ortho_len(), [&](const T* x, T* y) { diff_lane<T>(s.b0, x, y, d0); });
  Arr<T> out = along<T, T>(buf, 1, s.b1.ortho_len(), [&](const T* x, T* y) { diff_lane<T>(s.b1, x, y, d1); });
  if (scale) {
    double sc = std::pow(scale[0], d0) * std::pow(scale[1], d1);
    for (auto& v : out.d) scale_div(v, sc);
  }
  return out;
}

// ------------------------------------------------------------------ dense linalg
Arr<double> matmul(const Arr<double>& a, const Arr<double>& b) {
  if (a.n1 != b.n0) throw std::runtime_error("matmul shape");
  Arr<double> c(a.n0, b.n1);
  for (int i = 0; i < a.n0; ++i)
    for (int k = 0; k < a.n1; ++k) {
      double aik = a(i, k);
      if (aik == 0.0) continue;
      for (int j = 0; j < b.n1; ++j) c(i, j) += aik * b(k, j);
    }
  return c;
}
template <class T> Arr<T> matmul_left(const Arr<double>& a, const Arr<T>& b) {  // a (real) * b
  Arr<T> c(a.n0, b.n1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < a.n0; ++i)
    for (int k = 0; k < a.n1; ++k) {
      double aik = a(i, k);
      for (int j = 0; j < b.n1; ++j) c(i, j) += mulr(b(k, j), aik);
    }
  return c;
}
// inverse via Gauss-Jordan with partial pivoting
Arr<double> inv(const Arr<double>& a0) {
  const int n = a0.n0;
  Arr<double> a = a0, r(n, n);
  for (int i = 0; i < n; ++i) r(i, i) = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i) if (std::fabs(a(i, c)) > std::fabs(a(p, c))) p = i;
    if (a(p, c) == 0.0) throw std::runtime_error("singular matrix in inv");
    if (p != c) for (int j = 0; j < n; ++j) { std::swap(a(p, j), a(c, j)); std::swap(r(p, j), r(c, j)); }
    double piv = a(c, c);
    for (int j = 0; j < n; ++j) { a(c, j) /= piv; r(c, j) /= piv; }
    for (int i = 0; i < n; ++i) {
      if (i == c) continue;
      double f = a(i, c);
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { a(i, j) -= f * a(c, j); r(i, j) -= f * r(c, j); }
    }
  }
  return r;
}

// Eigenvalues of a real upper-Hessenberg block by the implicit double-shift (Francis)
// QR iteration (Golub & Van Loan, Alg. 7.5.1-7.5.2), eigenvalues only.
void hessenberg(Arr<double>& h, Arr<double>* q = nullptr) {  // Householder reduction in place, H = Q^T A Q
  const int n = h.n0;
  if (q) { *q = Arr<double>(n, n); for (int i = 0; i < n; ++i) (*q)(i, i) = 1.0; }
  for (int k = 0; k < n - 2; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += h(i, k) * h(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (h(k + 1, k) > 0) alpha = -alpha;
    vecd v(n, 0.0);
    v[k + 1] = h(k + 1, k) - alpha;
    for (int i = k + 2; i < n; ++i) v[i] = h(i, k);
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    for (int j = 0; j < n; ++j) {  // H = (I - 2vv'/v'v) H
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * h(i, j);
      s = 2.0 * s / vn;
      for (int i = k + 1; i < n; ++i) h(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {  // H = H (I - 2vv'/v'v)
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += h(i, j) * v[j];
      s = 2.0 * s / vn;
      for (int j = k + 1; j < n; ++j) h(i, j) -= s * v[j];
    }
    if (q)
      for (int i = 0; i < n; ++i) {  // Q = Q (I - 2vv'/v'v)
        double s = 0.0;
        for (int j = k + 1; j < n; ++j) s += (*q)(i, j) * v[j];
        s = 2.0 * s / vn;
        for (int j = k + 1; j < n; ++j) (*q)(i, j) -= s * v[j];
      }
    for (int i = k + 2; i < n; ++i) h(i, k) = 0.0;
  }
}

// Eigenvector of an upper-Hessenberg H for eigenvalue lam by inverse iteration; the LU
// with partial pivoting of (H - lam I) only ever swaps adjacent rows, O(n^2).
vecd hessenberg_inverse_iteration(const Arr<double>& h, double lam, double anorm) {
  const int n = h.n0;
  Arr<double> u = h;
  for (int i = 0; i < n; ++i) u(i, i) -= lam;
  std::vector<int> swp(n, 0);
  vecd mult(n, 0.0);
  const double tiny = 2.220446049250313e-16 * std::max(anorm, (double)1e-300);
  for (int k = 0; k < n; ++k) {
    if (k + 1 < n && std::fabs(u(k + 1, k)) > std::fabs(u(k, k))) {
      for (int j = k; j < n; ++j) std::swap(u(k, j), u(k + 1, j));
      swp[k] = 1;
    }
    if (std::fabs(u(k, k)) < tiny) u(k, k) = (u(k, k) >= 0 ? tiny : -tiny);
    if (k + 1 < n) {
      const double f = u(k + 1, k) / u(k, k);
      mult[k] = f;
      u(k + 1, k) = 0.0;
      for (int j = k + 1; j < n; ++j) u(k + 1, j) -= f * u(k, j);
    }
  }
  vecd x(n);
  for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.01 * std::sin(1.0 + i);
  for (int it = 0; it < 3; ++it) {
    for (int k = 0; k + 1 < n; ++k) {  // apply L^-1 P
      if (swp[k]) std::swap(x[k], x[k + 1]);
      x[k + 1] -= mult[k] * x[k];
    }
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[i] -= u(i, j) * x[j];
      x[i] /= u(i, i);
    }
    double nr = 0.0;
    for (double v : x) nr += v * v;
    nr = std::sqrt(nr);
    for (double& v : x) v /= nr;
  }
  return x;
}

void apply_reflector3(Arr<double>& h, int k, int nr, const double* v, int lo, int hi, int rlo, int rhi) {
  // left: rows k..k+nr-1, cols lo..hi ; right: rows rlo..rhi, cols k..k+nr-1
  double vv = 0.0;
  for (int i = 0; i < nr; ++i) vv += v[i] * v[i];
  if (vv == 0.0) return;
  for (int j = lo; j <= hi; ++j) {
    double s = 0.0;
    for (int i = 0; i < nr; ++i) s += v[i] * h(k + i, j);
    s = 2.0 * s / vv;
    for (int i = 0; i < nr; ++i) h(k + i, j) -= s * v[i];
  }
  for (int i = rlo; i <= rhi; ++i) {
    double s = 0.0;
    for (int j = 0; j < nr; ++j) s += h(i, k + j) * v[j];
    s = 2.0 * s / vv;
    for (int j = 0; j < nr; ++j) h(i, k + j) -= s * v[j];
  }
}

void house(const double* x, int nr, double* v) {
  double a = 0.0;
  for (int i = 0; i < nr; ++i) a += x[i] * x[i];
  a = std::sqrt(a);
  for (int i = 0; i < nr; ++i) v[i] = x[i];
  if (a == 0.0) return;
  v[0] += (x[0] >= 0 ? a : -a);
}

vecd hqr_eigenvalues(Arr<double> h, bool& complex_found) {
  const int n = h.n0;
  vecd wr(n, 0.0);
  complex_found = false;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iter = 0;
  while (hi >= 0) {
    int l = hi;
    for (; l > 0; --l) {
      double s = std::fabs(h(l - 1, l - 1)) + std::fabs(h(l, l));
      if (std::fabs(h(l, l - 1)) <= eps * s) { h(l, l - 1) = 0.0; break; }
    }
    if (l == hi) { wr[hi] = h(hi, hi); --hi; iter = 0; continue; }
    if (l == hi - 1) {
      double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
      double p = 0.5 * (a - d), disc = p * p + b * c;
      if (disc < 0) { complex_found = true; wr[hi - 1] = wr[hi] = 0.5 * (a + d); }
      else {
        double z = std::sqrt(disc);
        z = p >= 0 ? p + z : p - z;
        wr[hi - 1] = d + z;
        wr[hi] = (z != 0.0) ? d - b * c / z : d + z;
      }
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++iter > 200) throw std::runtime_error("hqr: no convergence");
    double s, t;
    if (iter % 11 == 10) {  // exceptional shift
      double w = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
      s = 1.5 * w;
      t = w * w;
    } else {
      s = h(hi - 1, hi - 1) + h(hi, hi);
      t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
    }
    double x[3];
    x[0] = h(l, l) * h(l, l) + h(l, l + 1) * h(l + 1, l) - s * h(l, l) + t;
    x[1] = h(l + 1, l) * (h(l, l) + h(l + 1, l + 1) - s);
    x[2] = h(l + 1, l) * h(l + 2, l + 1);
    for (int k = l; k <= hi - 2; ++k) {
      double v[3];
      house(x, 3, v);
      int q = std::max(l, k - 1);
      int r = std::min(k + 3, hi);
      apply_reflector3(h, k, 3, v, q, hi, l, r);
      x[0] = h(k + 1, k);
      x[1] = h(k + 2, k);
      if (k < hi - 2) x[2] = h(k + 3, k);
    }
    double v[3];
    house(x, 2, v);
    apply_reflector3(h, hi - 1, 2, v, hi - 2, hi, l, hi);
  }
  return wr;
}

// Eigenvector of A for eigenvalue lam by inverse iteration (LU with partial pivoting).
vecd inverse_iteration(const Arr<double>& a, double lam, double anorm) {
  const int n = a.n0;
  Arr<double> lu = a;
  double shift = lam;
  for (int i = 0; i < n; ++i) lu(i, i) -= shift;
  std::vector<int> piv(n);
  const double tiny = 2.220446049250313e-16 * std::max(anorm, (double)1e-300);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i) if (std::fabs(lu(i, c)) > std::fabs(lu(p, c))) p = i;
    piv[c] = p;
    if (p != c) for (int j = 0; j < n; ++j) std::swap(lu(p, j), lu(c, j));
    if (std::fabs(lu(c, c)) < tiny) lu(c, c) = (lu(c, c) >= 0 ? tiny : -tiny);
    for (int i = c + 1; i < n; ++i) {
      double f = lu(i, c) / lu(c, c);
      lu(i, c) = f;
      for (int j = c + 1; j < n; ++j) lu(i, j) -= f * lu(c, j);
    }
  }
  vecd x(n);
  for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.01 * std::sin(1.0 + i);
  for (int it = 0; it < 3; ++it) {
    for (int c = 0; c < n; ++c) std::swap(x[c], x[piv[c]]);
    for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) x[i] -= lu(i, j) * x[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[i] -= lu(i, j) * x[j];
      x[i] /= lu(i, i);
    }
    double nr = 0.0;
    for (double v : x) nr += v * v;
    nr = std::sqrt(nr);
    for (double& v : x) v /= nr;
  }
  return x;
}

// eig as in src/solver/utils.rs:66-98: real parts, sorted descending (reverse of a
// stable ascending argsort), Q^-1 = inv(Q).  Columns whose off-diagonal entries are
// exactly zero give exact unit eigenvectors e_j, as LAPACK's balancing isolates them.
void eig_sorted(const Arr<double>& a, vecd& lam, Arr<double>& q, Arr<double>& qi) {
  const int n = a.n0;
  double anorm = 0.0;
  for (double v : a.d) anorm = std::max(anorm, std::fabs(v));
  Arr<double> h = a, hq;
  hessenberg(h, &hq);
  bool cplx = false;
  vecd w = hqr_eigenvalues(h, cplx);
  if (cplx) throw std::runtime_error("eig: complex eigenvalues");
  std::vector<int> iso(n, 0);
  for (int j = 0; j < n; ++j) {
    bool z = true;
    for (int i = 0; i < n && z; ++i) if (i != j && a(i, j) != 0.0) z = false;
    iso[j] = z;
  }
  // match isolated exact eigenvalues to computed ones
  std::vector<int> used(n, 0);
  Arr<double> qq(n, n);
  vecd lam0(n);
  int col = 0;
  for (int j = 0; j < n; ++j) {
    if (!iso[j]) continue;
    int best = -1;
    for (int k = 0; k < n; ++k)
      if (!used[k] && (best < 0 || std::fabs(w[k] - a(j, j)) < std::fabs(w[best] - a(j, j)))) best = k;
    used[best] = 1;
    lam0[col] = a(j, j);
    qq(j, col) = 1.0;
    ++col;
  }
  for (int k = 0; k < n; ++k) {
    if (used[k]) continue;
    vecd y = hessenberg_inverse_iteration(h, w[k], anorm);
    vecd x(n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) x[i] += hq(i, j) * y[j];
    lam0[col] = w[k];
    for (int i = 0; i < n; ++i) qq(i, col) = x[i];
    ++col;
  }
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int i, int j) { return lam0[i] < lam0[j]; });
  std::reverse(perm.begin(), perm.end());
  lam.resize(n);
  q = Arr<double>(n, n);
  for (int c = 0; c < n; ++c) {
    lam[c] = lam0[perm[c]];
    for (int i = 0; i < n; ++i) q(i, c) = qq(i, perm[c]);
  }
  qi = inv(q);
}

// ------------------------------------------------------------------ solvers
vecd diag_of(const Arr<double>& a, int off) {  // utils.rs:17-43
  int n = a.n0, m = std::abs(off);
  vecd d(n - m);
  for (int i = 0; i < n - m; ++i) d[i] = off >= 0 ? a(i, i + m) : a(i + m, i);
  return d;
}

struct Fdma {  // fdma.rs:11-118
  int n = 0;
  vecd low, dia, up1, up2;
  Fdma() = default;
  explicit Fdma(const Arr<double>& a) : n(a.n0), low(diag_of(a, -2)), dia(diag_of(a, 0)), up1(diag_of(a, 2)), up2(diag_of(a, 4)) {}
  Fdma add_scaled(const Fdma& o, double l) const {  // fdma.rs:187-227
    Fdma r;
    r.n = n;
    r.low.resize(low.size()); r.dia.resize(n); r.up1.resize(up1.size()); r.up2.resize(up2.size());
    for (size_t i = 0; i < low.size(); ++i) r.low[i] = low[i] + o.low[i] * l;
    for (int i = 0; i < n; ++i) r.dia[i] = dia[i] + o.dia[i] * l;
    for (size_t i = 0; i < up1.size(); ++i) r.up1[i] = up1[i] + o.up1[i] * l;
    for (size_t i = 0; i < up2.size(); ++i) r.up2[i] = up2[i] + o.up2[i] * l;
    return r;
  }
  void sweep() {  // fdma.rs:73-82
    for (int i = 2; i < n; ++i) {
      low[i - 2] /= dia[i - 2];
      dia[i] -= low[i - 2] * up1[i - 2];
      if (i < n - 2) up1[i] -= low[i - 2] * up2[i - 2];
    }
  }
  template <class T> void solve(T* x) const {  // fdma.rs:101-118
    for (int i = 2; i < n; ++i) x[i] = x[i] - mulr(x[i - 2], low[i - 2]);
    x[n - 1] = divr(x[n - 1], dia[n - 1]);
    x[n - 2] = divr(x[n - 2], dia[n - 2]);
    x[n - 3] = divr(x[n - 3] - mulr(x[n - 1], up1[n - 3]), dia[n - 3]);
    x[n - 4] = divr(x[n - 4] - mulr(x[n - 2], up1[n - 4]), dia[n - 4]);
    for (int i = n - 5; i >= 0; --i)
      x[i] = divr(x[i] - mulr(x[i + 2], up1[i]) - mulr(x[i + 4], up2[i]), dia[i]);
  }
};

struct MatVecFdma {  // matvec.rs:142-193
  int m = 0;
  vecd low, dia, up1, up2;
  MatVecFdma() = default;
  explicit MatVecFdma(const Arr<double>& a) : m(a.n0), low(m, 0.0), dia(m, 0.0), up1(m, 0.0), up2(m, 0.0) {
    for (int i = 0; i < m; ++i) {
      dia[i] = a(i, i);
      if (i > 1) low[i] = a(i, i - 2);
      if (i < m - 2) up1[i] = a(i, i + 2);
      if (i < m - 4) up2[i] = a(i, i + 4);
    }
  }
  template <class T> void apply(const T* in, T* out) const {
    for (int i = 0; i < m; ++i) {
      T o = mulr(in[i], dia[i]);
      if (i > 1) o = o + mulr(in[i - 2], low[i]);
      if (i < m - 2) o = o + mulr(in[i + 2], up1[i]);
      if (i < m - 4) o = o + mulr(in[i + 4], up2[i]);
      out[i] = o;
    }
  }
};

template <class T> Arr<T> matvec_axis(const MatVecFdma& mv, const Arr<T>& a, int axis) {
  return along<T, T>(a, axis, mv.m, [&](const T* x, T* y) { mv.apply<T>(x, y); });
}
template <class T> Arr<T> fdma_axis(const Fdma& f, const Arr<T>& a, int axis) {
  return along<T, T>(a, axis, f.n, [&](const T* x, T* y) {
    for (int i = 0; i < f.n; ++i) y[i] = x[i];
    f.solve<T>(y);
  });
}

struct Ingredients { Arr<double> a, b, p; bool has_p, is_diag; };
Ingredients ingredients(const Base& b) {  // field.rs:194-252
  Ingredients r;
  Arr<double> mass = b.mass();
  Arr<double> peye = b.laplace_inv_eye();
  Arr<double> pinv = matmul(peye, b.laplace_inv());
  if (b.k == CHEB) {
    Arr<double> ms(b.n, b.n - 2);
    for (int i = 0; i < b.n; ++i) for (int j = 0; j < b.n - 2; ++j) ms(i, j) = mass(i, j + 2);
    r.a = matmul(pinv, ms);
    r.b = matmul(peye, ms);
  } else if (b.k == FO) {
    r.a = mass;
    r.b = b.laplace();
  } else {
    r.a = matmul(pinv, mass);
    r.b = matmul(peye, mass);
  }
  r.has_p = b.k != FO;
  if (r.has_p) r.p = pinv;
  r.is_diag = b.k == FO;
  return r;
}
Arr<double> scaled(const Arr<double>& a, double s) { Arr<double> r = a; for (auto& v : r.d) v = v * s; return r; }

// FdmaTensor for N = 2  fdma_tensor.rs:105-153, 185-234
struct FdmaTensor2 {
  vecd lam;
  Arr<double> fwd, bwd;
  bool has_fb = false;
  Fdma fa, fc;
  double alpha = 0.0;
  std::vector<Fdma> rows;  // row systems; the reference sweeps them per solve, the values are identical
  void init(const Arr<double>& a0, const Arr<double>& a1, const Arr<double>& c0, const Arr<double>& c1, bool diag0, double alph) {
    alpha = alph;
    if (diag0) {
      lam = diag_of(a0, 0);
      has_fb = false;
    } else {
      Arr<double> cinv = inv(c0);
      Arr<double> x = matmul(cinv, a0);
      Arr<double> q, p;
      eig_sorted(x, lam, q, p);
      fwd = matmul(p, cinv);
      bwd = q;
      has_fb = true;
    }
    fa = Fdma(a1);
    fc = Fdma(c1);
  }
  void build_rows() {
    rows.resize(lam.size());
    for (size_t i = 0; i < lam.size(); ++i) {
      double l = lam[i] + alpha;
      rows[i] = fa.add_scaled(fc, l);
      rows[i].sweep();
    }
  }
  template <class T> Arr<T> solve(const Arr<T>& rhs) const {
    Arr<T> out = has_fb ? matmul_left<T>(fwd, rhs) : rhs;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < out.n0; ++i) rows[i].solve<T>(&out(i, 0));
    if (has_fb) out = matmul_left<T>(bwd, out);
    return out;
  }
};

enum SolverKind { S_HHOLTZ_ADI = 0, S_HHOLTZ = 1, S_POISSON = 2 };

struct Solver2 {  // hholtz_adi.rs, hholtz.rs, poisson.rs
  SolverKind kind;
  MatVecFdma mv[2];
  bool has_mv[2] = {false, false};
  Fdma adi[2];
  FdmaTensor2 t;
  Solver2(SolverKind k, const Base& b0, const Base& b1, const double* c) : kind(k) {
    Ingredients ing[2] = {ingredients(b0), ingredients(b1)};
    for (int ax = 0; ax < 2; ++ax) {
      has_mv[ax] = ing[ax].has_p;
      if (has_mv[ax]) mv[ax] = MatVecFdma(ing[ax].p);
    }
    if (k == S_HHOLTZ_ADI) {  // hholtz_adi.rs:44-63
      for (int ax = 0; ax < 2; ++ax) {
        Arr<double> mat = ing[ax].a;
        for (size_t i = 0; i < mat.d.size(); ++i) mat.d[i] = ing[ax].a.d[i] - ing[ax].b.d[i] * c[ax];
        adi[ax] = Fdma(mat);
        adi[ax].sweep();
      }
    } else if (k == S_HHOLTZ) {  // hholtz.rs:42-76
      Arr<double> l0 = scaled(scaled(ing[0].b, -1.0), c[0]);
      Arr<double> l1 = scaled(scaled(ing[1].b, -1.0), c[1]);
      t.init(l0, l1, ing[0].a, ing[1].a, ing[0].is_diag, 1.0);
      t.build_rows();
    } else {  // poisson.rs:50-90
      Arr<double> l0 = scaled(ing[0].b, c[0]);
      Arr<double> l1 = scaled(ing[1].b, c[1]);
      t.init(l0, l1, ing[0].a, ing[1].a, ing[0].is_diag, 0.0);
      if (std::fabs(t.lam[0]) < 1e-10) for (auto& v : t.lam) v -= 1e-10;
      t.build_rows();
    }
  }
  template <class T> Arr<T> solve(const Arr<T>& in) const {
    Arr<T> r = has_mv[0] ? matvec_axis<T>(mv[0], in, 0) : in;
    if (has_mv[1]) r = matvec_axis<T>(mv[1], r, 1);
    if (kind == S_HHOLTZ_ADI) {
      Arr<T> o = fdma_axis<T>(adi[0], r, 0);
      return fdma_axis<T>(adi[1], o, 1);
    }
    return t.solve<T>(r);
  }
};

// ------------------------------------------------------------------ Field / Navier2D
template <class T> struct Field2 {  // src/field.rs:67-163
  Space2<T> space;
  Arr<double> v;
  Arr<T> vhat;
  vecd x[2], dx[2];
  explicit Field2(const Space2<T>& s) : space(s), v(s.nphys0(), s.nphys1()), vhat(s.nspec0(), s.nspec1()) {
    x[0] = s.b0.coords();
    x[1] = s.b1.coords();
    dx[0] = get_dx(x[0], s.b0.fourier());
    dx[1] = get_dx(x[1], s.b1.fourier());
  }
  static vecd get_dx(const vecd& x, bool periodic) {
    const int n = (int)x.size();
    vecd d(n);
    if (periodic) { for (int i = 0; i < n; ++i) d[i] = x[2] - x[1]; return d; }
    for (int i = 0; i < n; ++i) {
      double l = i == 0 ? x[0] : (x[i] + x[i - 1]) / 2.0;
      double r = i == n - 1 ? x[n - 1] : (x[i + 1] + x[i]) / 2.0;
      d[i] = r - l;
    }
    return d;
  }
  void forward() { vhat = space_forward<T>(space, v); }
  void backward() { v = space_backward<T>(space, vhat); }
  Arr<T> to_ortho() const { return space_to_ortho<T>(space, vhat); }
  void from_ortho(const Arr<T>& p) { vhat = space_from_ortho<T>(space, p); }
  Arr<T> gradient(int d0, int d1, const double* scale) const { return space_gradient<T>(space, vhat, d0, d1, scale); }
  vecd average_axis(int axis) const {  // average.rs:25-34
    double length = std::fabs(x[axis].back() - x[axis][0]);
    const int nl = axis == 0 ? v.n1 : v.n0, nn = axis == 0 ? v.n0 : v.n1;
    vecd out(nl, 0.0);
    for (int l = 0; l < nl; ++l) {
      double s = 0.0;
      for (int i = 0; i < nn; ++i) {
        double vv = axis == 0 ? v(i, l) : v(l, i);
        s += vv * dx[axis][i] / length;
      }
      out[l] = s;
    }
    return out;
  }
  double average() const {  // average.rs:51-57
    double length = std::fabs(x[1].back() - x[1][0]);
    vecd ax = average_axis(0);
    double s = 0.0;
    for (size_t j = 0; j < ax.size(); ++j) s += ax[j] * dx[1][j] / length;
    return s;
  }
};

template <class T> Arr<T> operator_add(const Arr<T>& a, const Arr<T>& b) { Arr<T> r = a; for (size_t i = 0; i < r.d.size(); ++i) r.d[i] = a.d[i] + b.d[i]; return r; }

template <class T> struct Navier2D {  // src/navier/navier.rs:153-195
  double nu, ka, ra, pr, dt, time = 0.0, scale[2];
  bool periodic, dealias = true;
  std::unique_ptr<Field2<T>> temp, ux, uy, pres0, pres1, field, fieldbc;
  std::unique_ptr<Solver2> solver[4];

  static Base xb(bool periodic, Kind k, int nx) { return periodic ? Base(FO, nx) : Base(k, nx); }

  Navier2D(int nx, int ny, double ra_, double pr_, double dt_, double aspect, bool adiabatic, bool per)
      : ra(ra_), pr(pr_), dt(dt_), periodic(per) {
    scale[0] = aspect;
    scale[1] = 1.0;
    auto get_nu = [](double r, double p, double h) { return std::sqrt(p / (r / std::pow(h, 3.0))); };
    auto get_ka = [](double r, double p, double h) { return std::sqrt(1.0 / ((r / std::pow(h, 3.0)) * p)); };
    nu = get_nu(ra, pr, scale[1] * 2.0);
    ka = get_ka(ra, pr, scale[1] * 2.0);
    Base cdy(CD, ny), cny(CN, ny), chy(CHEB, ny);
    ux.reset(new Field2<T>(Space2<T>(xb(per, CD, nx), cdy)));
    uy.reset(new Field2<T>(Space2<T>(xb(per, CD, nx), cdy)));
    temp.reset(new Field2<T>(Space2<T>(xb(per, (adiabatic && !per) ? CN : CD, nx), cdy)));
    pres0.reset(new Field2<T>(Space2<T>(xb(per, CHEB, nx), chy)));
    pres1.reset(new Field2<T>(Space2<T>(xb(per, CN, nx), cny)));
    field.reset(new Field2<T>(Space2<T>(xb(per, CHEB, nx), chy)));
    double cv[2] = {dt * nu / std::pow(scale[0], 2.0), dt * nu / std::pow(scale[1], 2.0)};
    double ck[2] = {dt * ka / std::pow(scale[0], 2.0), dt * ka / std::pow(scale[1], 2.0)};
    double cp[2] = {1.0 / std::pow(scale[0], 2.0), 1.0 / std::pow(scale[1], 2.0)};
    SolverKind hk = per ? S_HHOLTZ : S_HHOLTZ_ADI;  // navier.rs:250-261 / :410-421
    solver[0].reset(new Solver2(hk, ux->space.b0, ux->space.b1, cv));
    solver[1].reset(new Solver2(hk, uy->space.b0, uy->space.b1, cv));
    solver[2].reset(new Solver2(hk, temp->space.b0, temp->space.b1, ck));
    solver[3].reset(new Solver2(S_POISSON, pres1->space.b0, pres1->space.b1, cp));
    for (Field2<T>* f : {temp.get(), ux.get(), uy.get(), pres0.get()}) {  // _scale navier.rs:500-515
      for (int ax = 0; ax < 2; ++ax) {
        for (auto& v : f->x[ax]) v *= scale[ax];
        for (auto& v : f->dx[ax]) v *= scale[ax];
      }
    }
    bc_rbc(nx, ny);
  }

  void bc_rbc(int nx, int ny);  // navier.rs:314-332 / :474-492

  Arr<double> conv_term(const Field2<T>& f, const Arr<double>& u, int d0, int d1) {  // conv_term.rs:22-42
    field->vhat = f.gradient(d0, d1, scale);
    field->backward();
    Arr<double> r(u.n0, u.n1);
    for (size_t i = 0; i < r.d.size(); ++i) r.d[i] = u.d[i] * field->v.d[i];
    return r;
  }
  static void addto(Arr<double>& a, const Arr<double>& b) { for (size_t i = 0; i < a.d.size(); ++i) a.d[i] = a.d[i] + b.d[i]; }
  Arr<T> finish_conv(const Arr<double>& conv) {  // navier.rs:562-568, dealias :1022-1032
    field->v = conv;
    field->forward();
    if (dealias) {
      int nx2 = field->vhat.n0 * 2 / 3, ny2 = field->vhat.n1 * 2 / 3;
      for (int i = 0; i < field->vhat.n0; ++i)
        for (int j = 0; j < field->vhat.n1; ++j)
          if (i >= nx2 || j >= ny2) field->vhat(i, j) = T(0);
    }
    return field->vhat;
  }
  // solid obstacle: navier.solid = Some([mask, value]) (navier.rs:189-190), physical [nx, ny]
  bool solid_on = false;
  Arr<double> solid_m, solid_v;
  Arr<T> conv_temp(const Arr<double>& u, const Arr<double>& v) {
    Arr<double> c = conv_term(*temp, u, 1, 0);
    addto(c, conv_term(*temp, v, 0, 1));
    addto(c, conv_term(*fieldbc, u, 1, 0));
    addto(c, conv_term(*fieldbc, v, 0, 1));
    if (solid_on) {  // navier.rs:551-560 (eta = 1e-2)
      const double eta = 1e-2;
      temp->backward();
      for (size_t i = 0; i < c.d.size(); ++i) {
        const double damp = (-1.0 / eta * solid_m.d[i]) * ((temp->v.d[i] + fieldbc->v.d[i]) - solid_v.d[i]);
        c.d[i] = c.d[i] - damp;
      }
    }
    return finish_conv(c);
  }
  // conv_ux / conv_uy (navier.rs:564-616); w is the velocity component of the field f
  Arr<T> conv_vel(const Field2<T>& f, const Arr<double>& u, const Arr<double>& v, const Arr<double>& w) {
    Arr<double> c = conv_term(f, u, 1, 0);
    addto(c, conv_term(f, v, 0, 1));
    if (solid_on) {  // navier.rs:579-584, :603-608
      const double eta = 1e-2;
      for (size_t i = 0; i < c.d.size(); ++i) c.d[i] = c.d[i] - (-1.0 / eta * solid_m.d[i]) * w.d[i];
    }
    return finish_conv(c);
  }
  void solve_ux(const Arr<double>& u, const Arr<double>& v) {  // navier.rs:622-633
    Arr<T> rhs = ux->to_ortho();
    Arr<T> gp = pres0->gradient(1, 0, scale);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] - mulr(gp.d[i], dt);
    Arr<T> cv = conv_vel(*ux, u, v, u);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] - mulr(cv.d[i], dt);
    ux->vhat = solver[0]->solve<T>(rhs);
  }
  void solve_uy(const Arr<double>& u, const Arr<double>& v, const Arr<T>& buoy) {  // navier.rs:636-654
    Arr<T> rhs = uy->to_ortho();
    Arr<T> gp = pres0->gradient(0, 1, scale);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] - mulr(gp.d[i], dt);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(buoy.d[i], dt);
    Arr<T> cv = conv_vel(*uy, u, v, v);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] - mulr(cv.d[i], dt);
    uy->vhat = solver[1]->solve<T>(rhs);
  }
  void solve_temp(const Arr<double>& u, const Arr<double>& v) {  // navier.rs:660-674
    Arr<T> rhs = temp->to_ortho();
    Arr<T> gx = fieldbc->gradient(2, 0, scale), gy = fieldbc->gradient(0, 2, scale);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(mulr(gx.d[i], dt), ka);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(mulr(gy.d[i], dt), ka);
    Arr<T> cv = conv_temp(u, v);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] - mulr(cv.d[i], dt);
    temp->vhat = solver[2]->solve<T>(rhs);
  }
  Arr<T> divergence() {  // navier.rs:698-703
    Arr<T> a = ux->gradient(1, 0, scale), b = uy->gradient(0, 1, scale);
    return operator_add(a, b);
  }
  void solve_pres(const Arr<T>& div) {  // navier.rs:710-715
    pres1->vhat = solver[3]->solve<T>(div);
    pres1->vhat(0, 0) = T(0);
  }
  void project_velocity(double c) {  // navier.rs:683-695
    Arr<T> dpdx = pres1->gradient(1, 0, scale), dpdy = pres1->gradient(0, 1, scale);
    Arr<T> uxo = ux->vhat, uyo = uy->vhat;
    ux->from_ortho(dpdx);
    uy->from_ortho(dpdy);
    for (size_t i = 0; i < uxo.d.size(); ++i) ux->vhat.d[i] = mulr(ux->vhat.d[i], -c) + uxo.d[i];
    for (size_t i = 0; i < uyo.d.size(); ++i) uy->vhat.d[i] = mulr(uy->vhat.d[i], -c) + uyo.d[i];
  }
  void update_pres(const Arr<T>& div) {  // navier.rs:717-721
    for (size_t i = 0; i < div.d.size(); ++i) pres0->vhat.d[i] = pres0->vhat.d[i] - mulr(div.d[i], nu);
    Arr<T> po = pres1->to_ortho();
    double inv_dt = 1.0 / dt;
    for (size_t i = 0; i < po.d.size(); ++i) pres0->vhat.d[i] = pres0->vhat.d[i] + mulr(po.d[i], inv_dt);
  }
  void update() {  // navier.rs:737-765
    Arr<T> that = operator_add(temp->to_ortho(), fieldbc->to_ortho());
    ux->backward();
    uy->backward();
    Arr<double> u = ux->v, v = uy->v;
    solve_ux(u, v);
    solve_uy(u, v, that);
    Arr<T> div = divergence();
    solve_pres(div);
    project_velocity(1.0);
    update_pres(div);
    solve_temp(u, v);
    time += dt;
  }
  // diagnostics  src/navier/functions.rs:12-101 ; navier.rs:869-879
  double eval_nu() {
    field->vhat = operator_add(temp->to_ortho(), fieldbc->to_ortho());
    Arr<T> g = field->gradient(0, 1, nullptr);
    double f = 1.0 / (scale[1] / 2.0);
    for (auto& v : g.d) v = mulr(mulr(v, -1.0), f);
    field->vhat = g;
    field->backward();
    vecd xa = field->average_axis(0);
    return (xa.back() + xa[0]) / 2.0;
  }
  double eval_nuvol() {
    field->vhat = operator_add(temp->to_ortho(), fieldbc->to_ortho());
    field->backward();
    uy->backward();
    Arr<double> uyt(uy->v.n0, uy->v.n1);
    for (size_t i = 0; i < uyt.d.size(); ++i) uyt.d[i] = field->v.d[i] * uy->v.d[i];
    Arr<T> g = field->gradient(0, 1, nullptr);
    for (auto& v : g.d) v = divr(v, scale[1] * -1.0);
    field->vhat = g;
    field->backward();
    for (size_t i = 0; i < uyt.d.size(); ++i) field->v.d[i] = (field->v.d[i] + uyt.d[i] / ka) * 2.0 * scale[1];
    return field->average();
  }
  double eval_re() {
    ux->backward();
    uy->backward();
    double f = 2.0 * scale[1] / nu;
    for (size_t i = 0; i < field->v.d.size(); ++i) {
      double a = ux->v.d[i], b = uy->v.d[i];
      field->v.d[i] = std::sqrt(a * a + b * b) * f;
    }
    return field->average();
  }
  // Statistics  src/navier/statistics.rs:10-153, :220-247 (arrays of the field space)
  bool st_on = false;
  double st_save = 0.0, st_write = 0.0, st_avg_time = 0.0, st_tot_time = 0.0;
  long long st_num = 0;
  Arr<T> st_arr[4];  // t_avg, ux_avg, uy_avg, nusselt (.vhat)
  void stats_new(double save_stat, double write_stat) {  // statistics.rs:54-88
    for (auto& a : st_arr) a = Arr<T>(field->vhat.n0, field->vhat.n1);
    st_save = save_stat;
    st_write = write_stat;
    st_avg_time = 0.0;
    st_tot_time = time;
    st_num = 0;
    st_on = true;
  }
  // callback's update (navier.rs:800-806) -> Statistics::update (statistics.rs:130-153)
  bool stats_update() {
    if (time < st_tot_time) return false;
    Arr<T> that = operator_add(temp->to_ortho(), fieldbc->to_ortho());
    Arr<T> uxhat = ux->to_ortho(), uyhat = uy->to_ortho();
    const double w = (double)st_num;
    for (size_t i = 0; i < that.d.size(); ++i) st_arr[0].d[i] = divr(mulr(st_arr[0].d[i], w) + that.d[i], w + 1.0);
    st_arr[1] = uxhat;
    st_arr[2] = uyhat;
    // nusselt (statistics.rs:220-247)
    field->vhat = uyhat;
    field->backward();
    Arr<double> uy_v = field->v;
    field->vhat = that;
    field->backward();
    Arr<double> uy_temp(uy_v.n0, uy_v.n1);
    for (size_t i = 0; i < uy_temp.d.size(); ++i) uy_temp.d[i] = field->v.d[i] * uy_v.d[i];
    Arr<T> g = field->gradient(0, 1, nullptr);
    for (auto& v : g.d) v = divr(v, scale[1] * -1.0);
    field->vhat = g;
    field->backward();
    for (size_t i = 0; i < uy_temp.d.size(); ++i)
      field->v.d[i] = (field->v.d[i] + uy_temp.d[i] / ka) * 2.0 * scale[1];
    field->forward();
    st_arr[3] = field->vhat;
    st_num += 1;
    st_avg_time += time - st_tot_time;
    st_tot_time = time;
    return true;
  }
  // statistics array `which` in the field space (spec) or after its backward (phys)
  Arr<double> stats_phys(int which) {
    field->vhat = st_arr[which];
    field->backward();
    return field->v;
  }
  double div_norm() {
    Arr<T> d = divergence();
    double s = 0.0;
    for (auto& v : d.d) { double r = re_of(v); s += r * r; if constexpr (std::is_same<T, cd>::value) s += v.imag() * v.imag(); }
    return std::sqrt(s);
  }
  // set_velocity / set_temperature  navier.rs:927-937, apply_* :1035-1076
  void apply_trig(Field2<T>& f, double amp, double m, double n, bool sin_cos) {
    const vecd& x0 = f.x[0];
    const vecd& y0 = f.x[1];
    vecd x(x0.size()), y(y0.size());
    for (size_t i = 0; i < x.size(); ++i) x[i] = (x0[i] - x0[0]) / (x0.back() - x0[0]);
    for (size_t j = 0; j < y.size(); ++j) y[j] = (y0[j] - y0[0]) / (y0.back() - y0[0]);
    double ax = M_PI * m, ay = M_PI * n;
    for (int i = 0; i < f.v.n0; ++i)
      for (int j = 0; j < f.v.n1; ++j)
        f.v(i, j) = sin_cos ? amp * std::sin(ax * x[i]) * std::cos(ay * y[j]) : amp * std::cos(ax * x[i]) * std::sin(ay * y[j]);
    f.forward();
  }
};

template <> void Navier2D<double>::bc_rbc(int nx, int ny) {
  Base xbase(CHEB, nx);
  fieldbc.reset(new Field2<double>(Space2<double>(xbase, Base(CDBC, ny))));
  Arr<double> bc(nx, 2);
  for (int i = 0; i < nx; ++i) { bc(i, 0) = 0.5; bc(i, 1) = -0.5; }
  fieldbc->vhat = along<double, double>(bc, 0, nx, [&](const double* a, double* b) { cheb_forward_lane(xbase, a, b); });
  fieldbc->backward();
  fieldbc->forward();
}
template <> void Navier2D<cd>::bc_rbc(int nx, int ny) {
  Base xbase(FO, nx);
  fieldbc.reset(new Field2<cd>(Space2<cd>(xbase, Base(CDBC, ny))));
  Arr<double> bc(nx, 2);
  for (int i = 0; i < nx; ++i) { bc(i, 0) = 0.5; bc(i, 1) = -0.5; }
  fieldbc->vhat = along<double, cd>(bc, 0, xbase.m, [&](const double* a, cd* b) { rfft(a, b, nx); });
  fieldbc->backward();
  fieldbc->forward();
}

// ------------------------------------------------------------------ C interface
thread_local std::string g_err;

// NavierAdjoint (src/navier/navier_adjoint.rs:128-1068): the adjoint descent of Farazmand
// (2016) for steady states.  `navier` is the Navier-Stokes step whose residual is descended
// (dt_navier = 1e-2, navier_adjoint.rs:231-232); `a` holds the adjoint fields [0] (temp, ux,
// uy, pres, pseu), the `field` space, fieldbc and the pressure Poisson solver, which are the
// same objects a Navier2D of the adjoint's dt holds; temp1, ux1, uy1 are the smoothed
// residuals [1] and smoother the Hholtz smoothers (1 - c lap), c = 1 / scale^2 (:247-268).
template <class T> struct NavierAdjoint {
  std::shared_ptr<Navier2D<T>> navier, a;
  std::unique_ptr<Field2<T>> temp1, ux1, uy1;
  std::unique_ptr<Solver2> smoother[3];
  Arr<T> unsm[3];  // fields_unsmoothed (navier_adjoint.rs:151-152)
  double dt, time = 0.0, res_tol = 1e-8;  // RES_TOL (:114)
  NavierAdjoint(int nx, int ny, double ra, double pr, double dt_, double aspect, bool adiabatic, bool per) : dt(dt_) {
    navier.reset(new Navier2D<T>(nx, ny, ra, pr, 1e-2, aspect, adiabatic, per));
    a.reset(new Navier2D<T>(nx, ny, ra, pr, dt_, aspect, adiabatic, per));
    temp1.reset(new Field2<T>(a->temp->space));
    ux1.reset(new Field2<T>(a->ux->space));
    uy1.reset(new Field2<T>(a->uy->space));
    const double w = 1e0;  // weight_laplacian
    const double cs[2] = {w / std::pow(a->scale[0], 2.0), w / std::pow(a->scale[1], 2.0)};
    smoother[0].reset(new Solver2(S_HHOLTZ, ux1->space.b0, ux1->space.b1, cs));
    smoother[1].reset(new Solver2(S_HHOLTZ, uy1->space.b0, uy1->space.b1, cs));
    smoother[2].reset(new Solver2(S_HHOLTZ, temp1->space.b0, temp1->space.b1, cs));
    for (auto& u : unsm) u = Arr<T>(a->field->vhat.n0, a->field->vhat.n1);
  }
  Field2<T>& f1(int which) { return which == 0 ? *temp1 : which == 1 ? *ux1 : *uy1; }
  static void addto(Arr<double>& x, const Arr<double>& y) { for (size_t i = 0; i < x.d.size(); ++i) x.d[i] = x.d[i] + y.d[i]; }
  Arr<T> conv_ux(const Arr<double>& u, const Arr<double>& v, const Arr<double>& t) {  // :488-522
    Arr<double> c = a->conv_term(*ux1, u, 1, 0);
    addto(c, a->conv_term(*ux1, v, 0, 1));
    addto(c, a->conv_term(*ux1, u, 1, 0));
    addto(c, a->conv_term(*uy1, v, 1, 0));
    addto(c, a->conv_term(*temp1, t, 1, 0));
    addto(c, a->conv_term(*temp1, a->fieldbc->v, 1, 0));
    return a->finish_conv(c);
  }
  Arr<T> conv_uy(const Arr<double>& u, const Arr<double>& v, const Arr<double>& t) {  // :525-559
    Arr<double> c = a->conv_term(*uy1, u, 1, 0);
    addto(c, a->conv_term(*uy1, v, 0, 1));
    addto(c, a->conv_term(*ux1, u, 0, 1));
    addto(c, a->conv_term(*uy1, v, 0, 1));
    addto(c, a->conv_term(*temp1, t, 0, 1));
    addto(c, a->conv_term(*temp1, a->fieldbc->v, 0, 1));
    return a->finish_conv(c);
  }
  Arr<T> conv_temp(const Arr<double>& u, const Arr<double>& v) {  // :562-578
    Arr<double> c = a->conv_term(*temp1, u, 1, 0);
    addto(c, a->conv_term(*temp1, v, 0, 1));
    return a->finish_conv(c);
  }
  // solve_ux / solve_uy (:581-626): explicit, rhs = to_ortho(u) - dt dp + dt conv + dt nu lap(u1)
  void solve_vel(int which, const Arr<double>& u, const Arr<double>& v, const Arr<double>& t) {
    Field2<T>& f0 = which == 1 ? *a->ux : *a->uy;
    Field2<T>& g1 = which == 1 ? *ux1 : *uy1;
    Arr<T> rhs = f0.to_ortho();
    Arr<T> gp = a->pres0->gradient(which == 1 ? 1 : 0, which == 1 ? 0 : 1, a->scale);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] - mulr(gp.d[i], dt);
    Arr<T> cv = which == 1 ? conv_ux(u, v, t) : conv_uy(u, v, t);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(cv.d[i], dt);
    Arr<T> lx = g1.gradient(2, 0, a->scale), ly = g1.gradient(0, 2, a->scale);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(mulr(lx.d[i], dt), a->nu);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(mulr(ly.d[i], dt), a->nu);
    f0.from_ortho(rhs);
  }
  void solve_temp(const Arr<double>& u, const Arr<double>& v) {  // :629-648
    Arr<T> rhs = a->temp->to_ortho();
    Arr<T> cv = conv_temp(u, v);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(cv.d[i], dt);
    Arr<T> buoy = uy1->to_ortho();
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(buoy.d[i], dt);
    Arr<T> lx = temp1->gradient(2, 0, a->scale), ly = temp1->gradient(0, 2, a->scale);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(mulr(lx.d[i], dt), a->ka);
    for (size_t i = 0; i < rhs.d.size(); ++i) rhs.d[i] = rhs.d[i] + mulr(mulr(ly.d[i], dt), a->ka);
    a->temp->from_ortho(rhs);
  }
  void update_residual() {  // :727-750
    Field2<T>* nf[3] = {navier->temp.get(), navier->ux.get(), navier->uy.get()};
    Field2<T>* af[3] = {a->temp.get(), a->ux.get(), a->uy.get()};
    for (int k = 0; k < 3; ++k) nf[k]->vhat = af[k]->vhat;
    navier->update();
    for (int k = 0; k < 3; ++k) {
      Arr<T>& x = nf[k]->vhat;
      for (size_t i = 0; i < x.d.size(); ++i) x.d[i] = divr(x.d[i] - af[k]->vhat.d[i], navier->dt);
    }
    // unsmoothed residuals (ux, uy, temp) and the smoothed ones, negated
    const int order[3] = {1, 2, 0};  // fields_unsmoothed[0..3] = ux, uy, temp
    for (int j = 0; j < 3; ++j) {
      const int k = order[j];
      unsm[j] = nf[k]->to_ortho();
      Field2<T>& g = f1(k);
      g.vhat = smoother[j]->solve<T>(unsm[j]);
      for (auto& x : g.vhat.d) x = mulr(x, -1.0);
    }
  }
  void update() {  // Integrate::update (:760-790)
    a->ux->backward();
    a->uy->backward();
    a->temp->backward();
    Arr<double> u = a->ux->v, v = a->uy->v, t = a->temp->v;
    update_residual();
    advance(u, v, t);
  }
  // the update after update_residual: explicit velocity and temperature updates from the
  // physical fields u, v, t of the step start and the residuals [1], with the projection
  void advance(const Arr<double>& u, const Arr<double>& v, const Arr<double>& t) {
    solve_vel(1, u, v, t);
    solve_vel(2, u, v, t);
    Arr<T> div = a->divergence();
    a->solve_pres(div);
    a->project_velocity(1.0);
    Arr<T> po = a->pres1->to_ortho();  // update_pres (:720-724): no viscous term
    const double inv_dt = 1.0 / dt;
    for (size_t i = 0; i < po.d.size(); ++i) a->pres0->vhat.d[i] = a->pres0->vhat.d[i] + mulr(po.d[i], inv_dt);
    solve_temp(u, v);
    time += dt;
  }
  static double norm(const Arr<T>& x) {  // norm_l2_f64 / norm_l2_c64 (:902-913)
    double s = 0.0;
    for (const auto& v : x.d) s += std::norm(std::complex<double>(v));
    return std::sqrt(s);
  }
  // [|ux1|, |uy1|, |temp1|, |unsmoothed ux|, |uy|, |temp|] (callback / exit, :846-880)
  void norms(double* out) {
    out[0] = norm(ux1->vhat);
    out[1] = norm(uy1->vhat);
    out[2] = norm(temp1->vhat);
    for (int j = 0; j < 3; ++j) out[3 + j] = norm(unsm[j]);
  }
  bool exit() {  // :882-897
    Arr<T> div = a->divergence();
    if (std::isnan(norm(div))) return true;
    return norm(ux1->vhat) + norm(uy1->vhat) + norm(temp1->vhat) < res_tol;
  }
};

struct Handle {
  bool periodic;
  std::shared_ptr<Navier2D<double>> r;
  std::shared_ptr<Navier2D<cd>> c;
};
struct AHandle {
  bool periodic;
  std::unique_ptr<NavierAdjoint<double>> r;
  std::unique_ptr<NavierAdjoint<cd>> c;
};

template <class T> Field2<T>* pick(Navier2D<T>& n, int which) {
  switch (which) {
    case 0: return n.temp.get();
    case 1: return n.ux.get();
    case 2: return n.uy.get();
    case 3: return n.pres0.get();
    case 4: return n.pres1.get();
    case 5: return n.field.get();
    case 6: return n.fieldbc.get();
  }
  throw std::runtime_error("bad field index");
}

template <class T> void copy_in(Arr<T>& a, const double* src) { std::memcpy((void*)a.d.data(), src, a.d.size() * sizeof(T)); }
template <class T> void copy_out(const Arr<T>& a, double* dst) { std::memcpy(dst, a.d.data(), a.d.size() * sizeof(T)); }

Base mkbase(int kind, int n) { return Base((Kind)kind, n); }

}  // namespace

#define GUARD(...)                      \
  try {                                 \
    __VA_ARGS__;                        \
    return 0;                           \
  } catch (const std::exception& e) {   \
    g_err = e.what();                   \
    return -1;                          \
  }

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void* orc_navier_create(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic, int periodic) {
  try {
    Handle* h = new Handle();
    h->periodic = periodic != 0;
    if (h->periodic) h->c.reset(new Navier2D<cd>(nx, ny, ra, pr, dt, aspect, adiabatic != 0, true));
    else h->r.reset(new Navier2D<double>(nx, ny, ra, pr, dt, aspect, adiabatic != 0, false));
    return h;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_navier_destroy(void* p) { delete (Handle*)p; }

// shape of a field: space 0 = physical (real), 1 = spectral
int orc_navier_shape(void* p, int which, int space, int* n0, int* n1) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) {
      Field2<cd>* f = pick(*h->c, which);
      *n0 = space ? f->vhat.n0 : f->v.n0;
      *n1 = space ? f->vhat.n1 : f->v.n1;
    } else {
      Field2<double>* f = pick(*h->r, which);
      *n0 = space ? f->vhat.n0 : f->v.n0;
      *n1 = space ? f->vhat.n1 : f->v.n1;
    }
  })
}
int orc_navier_get(void* p, int which, int space, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) {
      Field2<cd>* f = pick(*h->c, which);
      if (space) copy_out(f->vhat, out); else copy_out(f->v, out);
    } else {
      Field2<double>* f = pick(*h->r, which);
      if (space) copy_out(f->vhat, out); else copy_out(f->v, out);
    }
  })
}
int orc_navier_set(void* p, int which, int space, const double* in) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) {
      Field2<cd>* f = pick(*h->c, which);
      if (space) copy_in(f->vhat, in); else copy_in(f->v, in);
    } else {
      Field2<double>* f = pick(*h->r, which);
      if (space) copy_in(f->vhat, in); else copy_in(f->v, in);
    }
  })
}
int orc_navier_forward(void* p, int which) {
  Handle* h = (Handle*)p;
  GUARD({ if (h->periodic) pick(*h->c, which)->forward(); else pick(*h->r, which)->forward(); })
}
int orc_navier_backward(void* p, int which) {
  Handle* h = (Handle*)p;
  GUARD({ if (h->periodic) pick(*h->c, which)->backward(); else pick(*h->r, which)->backward(); })
}
int orc_navier_set_velocity(void* p, double amp, double m, double n) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) { h->c->apply_trig(*h->c->ux, amp, m, n, true); h->c->apply_trig(*h->c->uy, -amp, m, n, false); }
    else { h->r->apply_trig(*h->r->ux, amp, m, n, true); h->r->apply_trig(*h->r->uy, -amp, m, n, false); }
  })
}
int orc_navier_set_temperature(void* p, double amp, double m, double n) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) h->c->apply_trig(*h->c->temp, -amp, m, n, false);
    else h->r->apply_trig(*h->r->temp, -amp, m, n, false);
  })
}
int orc_navier_update(void* p, int nsteps) {
  Handle* h = (Handle*)p;
  GUARD({ for (int s = 0; s < nsteps; ++s) { if (h->periodic) h->c->update(); else h->r->update(); } })
}
// ------------------------------------------------------------ Navier2DAdjoint
void* orc_adjoint_create(int nx, int ny, double ra, double pr, double dt, double aspect, int adiabatic, int periodic) {
  try {
    AHandle* h = new AHandle();
    h->periodic = periodic != 0;
    if (h->periodic) h->c.reset(new NavierAdjoint<cd>(nx, ny, ra, pr, dt, aspect, adiabatic != 0, true));
    else h->r.reset(new NavierAdjoint<double>(nx, ny, ra, pr, dt, aspect, adiabatic != 0, false));
    return h;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_adjoint_destroy(void* p) { delete (AHandle*)p; }
// a Navier2D handle sharing part 0 (the adjoint fields [0], their space, fieldbc and the
// pressure solver) or part 1 (the Navier-Stokes step of the residual) of the adjoint
void* orc_adjoint_part(void* p, int part) {
  AHandle* a = (AHandle*)p;
  Handle* h = new Handle();
  h->periodic = a->periodic;
  if (a->periodic) h->c = part ? a->c->navier : a->c->a;
  else h->r = part ? a->r->navier : a->r->a;
  return h;
}
int orc_adjoint_update(void* p, int nsteps) {
  AHandle* h = (AHandle*)p;
  GUARD({ for (int s = 0; s < nsteps; ++s) { if (h->periodic) h->c->update(); else h->r->update(); } })
}
double orc_adjoint_time(void* p) { AHandle* h = (AHandle*)p; return h->periodic ? h->c->time : h->r->time; }
// the smoothed residual [1] of field which (0 temp, 1 ux, 2 uy), spectral
int orc_adjoint_residual(void* p, int which, double* out) {
  AHandle* h = (AHandle*)p;
  GUARD({ if (h->periodic) copy_out(h->c->f1(which).vhat, out); else copy_out(h->r->f1(which).vhat, out); })
}
// the unsmoothed residual (fields_unsmoothed[j], j = 0 ux, 1 uy, 2 temp), ortho space
int orc_adjoint_unsmoothed(void* p, int j, double* out) {
  AHandle* h = (AHandle*)p;
  GUARD({ if (h->periodic) copy_out(h->c->unsm[j], out); else copy_out(h->r->unsm[j], out); })
}
int orc_adjoint_set_residual(void* p, int which, const double* in) {
  AHandle* h = (AHandle*)p;
  GUARD({ if (h->periodic) copy_in(h->c->f1(which).vhat, in); else copy_in(h->r->f1(which).vhat, in); })
}
// the update without update_residual (the residuals [1] as they are), time += dt
int orc_adjoint_advance(void* p) {
  AHandle* h = (AHandle*)p;
  GUARD({
    auto run = [&](auto& ad) {
      ad.a->ux->backward();
      ad.a->uy->backward();
      ad.a->temp->backward();
      Arr<double> u = ad.a->ux->v, v = ad.a->uy->v, t = ad.a->temp->v;
      ad.advance(u, v, t);
    };
    if (h->periodic) run(*h->c); else run(*h->r);
  })
}
// the adjoint convective term of field which (0 temp, 1 ux, 2 uy) from the current fields
// and residuals, dealiased, in the `field` space
int orc_adjoint_conv(void* p, int which, double* out) {
  AHandle* h = (AHandle*)p;
  GUARD({
    auto run = [&](auto& ad) {
      ad.a->ux->backward();
      ad.a->uy->backward();
      ad.a->temp->backward();
      Arr<double> u = ad.a->ux->v, v = ad.a->uy->v, t = ad.a->temp->v;
      if (which == 0) copy_out(ad.conv_temp(u, v), out);
      else if (which == 1) copy_out(ad.conv_ux(u, v, t), out);
      else copy_out(ad.conv_uy(u, v, t), out);
    };
    if (h->periodic) run(*h->c); else run(*h->r);
  })
}
int orc_adjoint_norms(void* p, double* out) {
  AHandle* h = (AHandle*)p;
  GUARD({ if (h->periodic) h->c->norms(out); else h->r->norms(out); })
}
int orc_adjoint_exit(void* p, int* out) {
  AHandle* h = (AHandle*)p;
  GUARD({ *out = (h->periodic ? h->c->exit() : h->r->exit()) ? 1 : 0; })
}
// smoother s (0 ux, 1 uy, 2 temp): its FdmaTensor lam / fwd / bwd (as orc_solver_tensor)
int orc_adjoint_smoother(void* p, int s, void** out) {
  AHandle* h = (AHandle*)p;
  GUARD({ *out = h->periodic ? (void*)h->c->smoother[s].get() : (void*)h->r->smoother[s].get(); })
}

// The convective term of field `which` (0 temp, 1 ux, 2 uy) from the current state, as the
// step forms it (navier.rs:538-616): ux and uy backward, then conv_temp / conv_ux / conv_uy,
// dealiased, in the `field` space (spectral [nx, ny] real or [nx/2+1, ny] complex).
int orc_navier_conv(void* p, int which, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    auto run = [&](auto& n) {
      n.ux->backward();
      n.uy->backward();
      auto u = n.ux->v;
      auto v = n.uy->v;
      if (which == 0) copy_out(n.conv_temp(u, v), out);
      else if (which == 1) copy_out(n.conv_vel(*n.ux, u, v, u), out);
      else if (which == 2) copy_out(n.conv_vel(*n.uy, u, v, v), out);
      else throw std::invalid_argument("conv: which must be 0, 1 or 2");
    };
    if (h->periodic) run(*h->c); else run(*h->r);
  })
}
// conv_term(field which, u, [d0, d1]) (conv_term.rs:22-42) with the handle's scale: physical
// [nx, ny] in and out.
int orc_navier_conv_term(void* p, int which, int d0, int d1, const double* u, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    auto run = [&](auto& n) {
      auto* f = pick(n, which);
      Arr<double> ua(f->v.n0, f->v.n1);
      std::memcpy(ua.d.data(), u, sizeof(double) * ua.d.size());
      copy_out(n.conv_term(*f, ua, d0, d1), out);
    };
    if (h->periodic) run(*h->c); else run(*h->r);
  })
}
// finish_conv (navier.rs:562-568): forward transform of a physical [nx, ny] array in the
// `field` space followed by dealias (navier.rs:1022-1032) when the handle's dealias is on.
int orc_navier_finish_conv(void* p, const double* phys, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    auto run = [&](auto& n) {
      Arr<double> a(n.field->v.n0, n.field->v.n1);
      std::memcpy(a.d.data(), phys, sizeof(double) * a.d.size());
      copy_out(n.finish_conv(a), out);
    };
    if (h->periodic) run(*h->c); else run(*h->r);
  })
}
double orc_navier_time(void* p) { Handle* h = (Handle*)p; return h->periodic ? h->c->time : h->r->time; }
// 0 Nu, 1 Nuvol, 2 Re, 3 |div|_2
int orc_navier_eval(void* p, int what, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    double v = 0.0;
    if (h->periodic) {
      auto& n = *h->c;
      v = what == 0 ? n.eval_nu() : what == 1 ? n.eval_nuvol() : what == 2 ? n.eval_re() : n.div_norm();
    } else {
      auto& n = *h->r;
      v = what == 0 ? n.eval_nu() : what == 1 ? n.eval_nuvol() : what == 2 ? n.eval_re() : n.div_norm();
    }
    *out = v;
  })
}
// navier.solid = Some([mask, value]) (physical [nx, ny] each); mask == nullptr: None
int orc_navier_set_solid(void* p, const double* mask, const double* value) {
  Handle* h = (Handle*)p;
  GUARD({
    auto apply = [&](auto& n) {
      n.solid_on = mask != nullptr;
      if (!mask) return;
      const int n0 = n.temp->v.n0, n1 = n.temp->v.n1;
      n.solid_m = Arr<double>(n0, n1);
      n.solid_v = Arr<double>(n0, n1);
      std::memcpy(n.solid_m.d.data(), mask, sizeof(double) * n0 * n1);
      std::memcpy(n.solid_v.d.data(), value, sizeof(double) * n0 * n1);
    };
    if (h->periodic) apply(*h->c); else apply(*h->r);
  })
}
// statistics: enable (Statistics::new), update (0/1 in *updated), arrays, scalars
int orc_stats_enable(void* p, double save_stat, double write_stat) {
  Handle* h = (Handle*)p;
  GUARD({ if (h->periodic) h->c->stats_new(save_stat, write_stat); else h->r->stats_new(save_stat, write_stat); })
}
int orc_stats_update(void* p, int* updated) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic ? !h->c->st_on : !h->r->st_on) throw std::invalid_argument("statistics are not enabled");
    *updated = (h->periodic ? h->c->stats_update() : h->r->stats_update()) ? 1 : 0;
  })
}
int orc_stats_get(void* p, int which, int space, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    if (which < 0 || which > 3) throw std::invalid_argument("unknown statistics field");
    if (h->periodic) {
      if (space) copy_out(h->c->st_arr[which], out); else copy_out(h->c->stats_phys(which), out);
    } else {
      if (space) copy_out(h->r->st_arr[which], out); else copy_out(h->r->stats_phys(which), out);
    }
  })
}
// out: avg_time, tot_time, num_save
int orc_stats_info(void* p, double* out) {
  Handle* h = (Handle*)p;
  GUARD({
    if (h->periodic) { out[0] = h->c->st_avg_time; out[1] = h->c->st_tot_time; out[2] = (double)h->c->st_num; }
    else { out[0] = h->r->st_avg_time; out[1] = h->r->st_tot_time; out[2] = (double)h->r->st_num; }
  })
}
int orc_navier_set_time(void* p, double t) {
  Handle* h = (Handle*)p;
  GUARD({ if (h->periodic) h->c->time = t; else h->r->time = t; })
}
// tensor-solver setup of solver s (0 ux, 1 uy, 2 temp, 3 pres): sizes and arrays
int orc_navier_tensor_info(void* p, int s, int* nlam, int* has_fb) {
  Handle* h = (Handle*)p;
  GUARD({
    const Solver2& sv = h->periodic ? *h->c->solver[s] : *h->r->solver[s];
    if (sv.kind == S_HHOLTZ_ADI) { *nlam = 0; *has_fb = 0; return 0; }
    *nlam = (int)sv.t.lam.size();
    *has_fb = sv.t.has_fb ? 1 : 0;
  })
}
int orc_navier_tensor_get(void* p, int s, double* lam, double* fwd, double* bwd) {
  Handle* h = (Handle*)p;
  GUARD({
    const Solver2& sv = h->periodic ? *h->c->solver[s] : *h->r->solver[s];
    if (sv.kind == S_HHOLTZ_ADI) throw std::runtime_error("solver is ADI");
    std::memcpy(lam, sv.t.lam.data(), sv.t.lam.size() * 8);
    if (sv.t.has_fb) {
      std::memcpy(fwd, sv.t.fwd.d.data(), sv.t.fwd.d.size() * 8);
      std::memcpy(bwd, sv.t.bwd.d.data(), sv.t.bwd.d.size() * 8);
    }
  })
}

int orc_navier_tensor_set(void* p, int s, const double* lam, const double* fwd, const double* bwd) {
  Handle* h = (Handle*)p;
  GUARD({
    Solver2& sv = h->periodic ? *h->c->solver[s] : *h->r->solver[s];
    if (sv.kind == S_HHOLTZ_ADI) throw std::runtime_error("solver is ADI");
    std::memcpy(sv.t.lam.data(), lam, sv.t.lam.size() * 8);
    if (sv.t.has_fb) {
      std::memcpy(sv.t.fwd.d.data(), fwd, sv.t.fwd.d.size() * 8);
      std::memcpy(sv.t.bwd.d.data(), bwd, sv.t.bwd.d.size() * 8);
    }
    sv.t.build_rows();
  })
}

// ---- funspace / solver entry points for the unit tests -------------------------
// Space2 of (kind0, n0) x (kind1, n1); complex spectral when kind0 is Fourier.
// op: 0 forward (phys -> spec), 1 backward (spec -> phys), 2 to_ortho, 3 from_ortho,
//     4 gradient (d0, d1, scale).
int orc_space_apply(int k0, int n0, int k1, int n1, int op, int d0, int d1, const double* scale,
                    const double* in, double* out) {
  GUARD({
    Base b0 = mkbase(k0, n0), b1 = mkbase(k1, n1);
    if (b0.fourier()) {
      Space2<cd> s(b0, b1);
      if (op == 0) { Arr<double> v(b0.n, b1.n); copy_in(v, in); copy_out(space_forward<cd>(s, v), out); }
      else if (op == 1) { Arr<cd> c(b0.m, b1.m); copy_in(c, in); copy_out(space_backward<cd>(s, c), out); }
      else if (op == 2) { Arr<cd> c(b0.m, b1.m); copy_in(c, in); copy_out(space_to_ortho<cd>(s, c), out); }
      else if (op == 3) { Arr<cd> c(b0.ortho_len(), b1.ortho_len()); copy_in(c, in); copy_out(space_from_ortho<cd>(s, c), out); }
      else { Arr<cd> c(b0.m, b1.m); copy_in(c, in); copy_out(space_gradient<cd>(s, c, d0, d1, scale), out); }
    } else {
      Space2<double> s(b0, b1);
      if (op == 0) { Arr<double> v(b0.n, b1.n); copy_in(v, in); copy_out(space_forward<double>(s, v), out); }
      else if (op == 1) { Arr<double> c(b0.m, b1.m); copy_in(c, in); copy_out(space_backward<double>(s, c), out); }
      else if (op == 2) { Arr<double> c(b0.m, b1.m); copy_in(c, in); copy_out(space_to_ortho<double>(s, c), out); }
      else if (op == 3) { Arr<double> c(b0.ortho_len(), b1.ortho_len()); copy_in(c, in); copy_out(space_from_ortho<double>(s, c), out); }
      else { Arr<double> c(b0.m, b1.m); copy_in(c, in); copy_out(space_gradient<double>(s, c, d0, d1, scale), out); }
    }
  })
}

// Solve with solver kind (0 HholtzAdi, 1 Hholtz, 2 Poisson) built from the field space
// (kind0, n0) x (kind1, n1) and constants c[2]; input ortho-space array, output spectral.
int orc_solver_apply(int skind, int k0, int n0, int k1, int n1, const double* c, int cplx, const double* in, double* out) {
  GUARD({
    Base b0 = mkbase(k0, n0), b1 = mkbase(k1, n1);
    Solver2 s((SolverKind)skind, b0, b1, c);
    int r0 = b0.ortho_len(), r1 = b1.ortho_len();
    if (cplx) { Arr<cd> a(r0, r1); copy_in(a, in); copy_out(s.solve<cd>(a), out); }
    else { Arr<double> a(r0, r1); copy_in(a, in); copy_out(s.solve<double>(a), out); }
  })
}

// build-once solver handle (same solvers as orc_solver_apply)
void* orc_solver_create(int skind, int k0, int n0, int k1, int n1, const double* c) {
  try {
    return new Solver2((SolverKind)skind, mkbase(k0, n0), mkbase(k1, n1), c);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_solver_destroy(void* h) { delete (Solver2*)h; }
int orc_solver_solve(void* h, int n0in, int n1in, int cplx, const double* in, double* out) {
  GUARD({
    Solver2& s = *(Solver2*)h;
    if (cplx) { Arr<cd> a(n0in, n1in); copy_in(a, in); copy_out(s.solve<cd>(a), out); }
    else { Arr<double> a(n0in, n1in); copy_in(a, in); copy_out(s.solve<double>(a), out); }
  })
}
int orc_solver_tensor(void* h, double* lam, double* fwd, double* bwd, int* n, int* has_fb) {
  GUARD({
    Solver2& s = *(Solver2*)h;
    if (s.kind == S_HHOLTZ_ADI) { *n = 0; *has_fb = 0; return 0; }
    *n = (int)s.t.lam.size();
    *has_fb = s.t.has_fb ? 1 : 0;
    if (lam) std::memcpy(lam, s.t.lam.data(), s.t.lam.size() * 8);
    if (s.t.has_fb && fwd && bwd) {
      std::memcpy(fwd, s.t.fwd.d.data(), s.t.fwd.d.size() * 8);
      std::memcpy(bwd, s.t.bwd.d.data(), s.t.bwd.d.size() * 8);
    }
  })
}

// eig as utils.rs:66-98 ; a is n x n row-major
int orc_eig(int n, const double* a, double* lam, double* q, double* qi) {
  GUARD({
    Arr<double> A(n, n);
    copy_in(A, a);
    vecd l;
    Arr<double> Q, Qi;
    eig_sorted(A, l, Q, Qi);
    std::memcpy(lam, l.data(), n * 8);
    copy_out(Q, q);
    copy_out(Qi, qi);
  })
}

// Fdma from a dense square matrix, sweep, solve one lane (fdma.rs:33-118)
int orc_fdma_solve(int n, const double* a, const double* b, double* x) {
  GUARD({
    Arr<double> A(n, n);
    copy_in(A, a);
    Fdma f(A);
    f.sweep();
    for (int i = 0; i < n; ++i) x[i] = b[i];
    f.solve<double>(x);
  })
}

// 1-D lane primitives for known-answer tests
int orc_dct1(int n, const double* x, double* y) { GUARD(dct1(x, y, n)) }
int orc_cheb_diff(int n, int times, double* d) { GUARD(cheb_diff_lane<double>(d, n, times)) }

}  // extern "C"
