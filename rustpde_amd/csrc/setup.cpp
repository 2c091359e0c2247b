// Host-side set-up of the 1-D operators of the RBC step (extended precision).
// See setup.hpp.  Reference citations are relative to the reference repository root.
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace rbc {

LMat operator*(const LMat& x, const LMat& y) {
  if (x.c != y.r) throw std::runtime_error("LMat product: shape mismatch");
  // nonzero pattern of y's rows, so banded/sparse factors cost O(nnz) not O(n^3)
  std::vector<std::vector<int>> nz(y.r);
  for (int k = 0; k < y.r; ++k)
    for (int j = 0; j < y.c; ++j)
      if (y(k, j) != 0.0L) nz[k].push_back(j);
  LMat z(x.r, y.c);
  for (int i = 0; i < x.r; ++i)
    for (int k = 0; k < x.c; ++k) {
      const ld a = x(i, k);
      if (a == 0.0L) continue;
      for (int j : nz[k]) z(i, j) += a * y(k, j);
    }
  return z;
}

LMat scaled(const LMat& x, ld s) {
  LMat z = x;
  for (auto& v : z.a) v *= s;
  return z;
}

static const ld PI_L = 3.141592653589793238462643383279502884L;

// cos(pi p / q) with exact integer reduction
static ld cospi(long long p, long long q) {
  long long r = p % (2 * q);
  if (r < 0) r += 2 * q;
  return cosl(PI_L * (ld)r / (ld)q);
}
static ld sinpi(long long p, long long q) {
  long long r = p % (2 * q);
  if (r < 0) r += 2 * q;
  return sinl(PI_L * (ld)r / (ld)q);
}

Basis::Basis(Kind kind, int n_) : k(kind), n(n_) {
  switch (k) {
    case CHEB: m = n; break;
    case CD:
    case CN: m = n - 2; break;
    case CDBC: m = 2; break;
    case FO: m = n / 2 + 1; break;
    default: throw std::runtime_error("unknown basis kind");
  }
  if (k == CD || k == CN) {
    diag.assign(m, 1.0);
    low2.assign(m, 0.0);
    for (int i = 0; i < m; ++i)
      low2[i] = (k == CD) ? -1.0 : -1.0 * (double)((long long)i * i) / (double)((long long)(i + 2) * (i + 2));
  }
}

std::vector<double> Basis::coords() const {
  std::vector<double> x(n);
  if (k == FO) {
    const double step = 2.0 * M_PI / (double)n;
    for (int i = 0; i < n; ++i) x[i] = step * (double)i;
  } else {
    const double mm = (double)(n - 1);
    for (int i = 0; i < n; ++i) x[i] = -std::sin(M_PI * (mm - 2.0 * (double)i) / (2.0 * mm));
  }
  return x;
}

// composite_stencil.rs:180-188 (CD/CN), :319-326 (CDBC); identity for CHEB
LMat stencil(const Basis& b) {
  if (b.k == CHEB || b.k == FO) return LMat::eye(b.ortho());
  LMat s(b.n, b.m);
  if (b.k == CDBC) {
    s(0, 0) = 0.5L; s(0, 1) = 0.5L; s(1, 0) = -0.5L; s(1, 1) = 0.5L;
    return s;
  }
  for (int i = 0; i < b.m; ++i) {
    s(i, i) = (ld)b.diag[i];
    s(i + 2, i) = (ld)b.low2[i];
  }
  return s;
}

// ortho.rs:288-312: v_j = sum_k c_k T_k(x_j), x_j = -cos(pi j/(n-1)),
// T_k(x_j) = (-1)^k cos(pi j k/(n-1))
LMat cheb_backward(int n) {
  LMat b(n, n);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) b(j, k) = ((k & 1) ? -1.0L : 1.0L) * cospi((long long)j * k, n - 1);
  return b;
}

// ortho.rs:235-258: c_k = (-1)^k/(n-1) * DCT1(v)_k, c_0 and c_{n-1} halved, with
// DCT1(v)_k = v_0 + (-1)^k v_{n-1} + 2 sum_{j=1}^{n-2} v_j cos(pi j k/(n-1))
LMat cheb_forward(int n) {
  LMat f(n, n);
  for (int k = 0; k < n; ++k) {
    ld h = (k == 0 || k == n - 1) ? 0.5L : 1.0L;
    ld s = ((k & 1) ? -1.0L : 1.0L) / (ld)(n - 1) * h;
    for (int j = 0; j < n; ++j) {
      ld w = (j == 0 || j == n - 1) ? 1.0L : 2.0L;
      f(k, j) = s * w * cospi((long long)j * k, n - 1);
    }
  }
  return f;
}

// ortho.rs:107-125 differentiate_lane (one lane, long double)
static void diff_lane(std::vector<ld>& c) {
  const int n = (int)c.size();
  std::vector<ld> d(n);
  d[0] = c[1];
  for (int i = 1; i < n - 1; ++i) d[i] = 2.0L * (ld)(i + 1) * c[i + 1];
  d[n - 1] = 0.0L;
  for (int i = n - 3; i >= 1; --i) d[i] = d[i] + d[i + 2];
  d[0] = d[0] + d[2] / 2.0L;
  c.swap(d);
}

// D^t applied column by column to `m` (n rows)
static LMat diff_cols(const LMat& m, int times) {
  LMat out = m;
  std::vector<ld> col(m.r);
  for (int j = 0; j < m.c; ++j) {
    for (int i = 0; i < m.r; ++i) col[i] = m(i, j);
    for (int t = 0; t < times; ++t) diff_lane(col);
    for (int i = 0; i < m.r; ++i) out(i, j) = col[i];
  }
  return out;
}

LMat cheb_diff(int n, int times) { return diff_cols(LMat::eye(n), times); }

// B D: (B D)[j, k] = 2k sum_{i < k, i + k odd} B[j, i] / c_i  (c_0 = 2, else 1), O(n^2)
static LMat cheb_backward_diff(int n) {
  LMat b = cheb_backward(n), bd(n, n);
  for (int j = 0; j < n; ++j) {
    ld s[2] = {0.0L, 0.0L};  // running sums over even / odd i
    for (int k = 0; k < n; ++k) {
      bd(j, k) = 2.0L * (ld)k * s[(k + 1) & 1];
      s[k & 1] += b(j, k) / (k == 0 ? 2.0L : 1.0L);
    }
  }
  return bd;
}

// composite_stencil.rs:250-276 + linalg.rs:14-57 on one lane, long double
static void from_ortho_lane(const Basis& b, const ld* p, ld* c) {
  if (b.k == CDBC) {  // composite_stencil.rs:366-388
    const ld t0[2] = {0.5L, 0.5L}, t1[2] = {-0.5L, 0.5L};
    ld c0 = p[0] * t0[0] + p[1] * t1[0], c1 = p[0] * t0[1] + p[1] * t1[1];
    ld a = t0[0] * t0[0] + t1[0] * t1[0], bb = t0[0] * t0[1] + t1[0] * t1[1];
    ld cc = t0[1] * t0[0] + t1[1] * t1[0], d = t0[1] * t0[1] + t1[1] * t1[1];
    ld det = 1.0L / (a * d - bb * cc);
    c[0] = (c0 * d - c1 * bb) * det;
    c[1] = (c1 * a - c0 * cc) * det;
    return;
  }
  const int m = b.m;
  std::vector<ld> mainv(m), off(m - 2), d(m), w(m - 2), g(m);
  for (int i = 0; i < m; ++i) mainv[i] = (ld)b.diag[i] * b.diag[i] + (ld)b.low2[i] * b.low2[i];
  for (int i = 0; i < m - 2; ++i) off[i] = (ld)b.diag[i + 2] * b.low2[i];
  for (int i = 0; i < m; ++i) d[i] = p[i] * (ld)b.diag[i] + p[i + 2] * (ld)b.low2[i];
  w[0] = off[0] / mainv[0];
  g[0] = d[0] / mainv[0];
  if (m - 2 > 1) w[1] = off[1] / mainv[1];
  g[1] = d[1] / mainv[1];
  for (int i = 2; i < m - 2; ++i) w[i] = off[i] / (mainv[i] - off[i - 2] * w[i - 2]);
  for (int i = 2; i < m; ++i) g[i] = (d[i] - g[i - 2] * off[i - 2]) / (mainv[i] - off[i - 2] * w[i - 2]);
  c[m - 1] = g[m - 1];
  c[m - 2] = g[m - 2];
  for (int i = m - 2; i >= 1; --i) c[i - 1] = g[i - 1] - c[i + 1] * w[i - 1];
}

// from_ortho applied to every column of `m` (ortho rows -> composite rows)
static LMat from_ortho_cols(const Basis& b, const LMat& m) {
  LMat out(b.m, m.c);
  std::vector<ld> p(m.r), c(b.m);
  for (int j = 0; j < m.c; ++j) {
    for (int i = 0; i < m.r; ++i) p[i] = m(i, j);
    from_ortho_lane(b, p.data(), c.data());
    for (int i = 0; i < b.m; ++i) out(i, j) = c[i];
  }
  return out;
}

LMat inverse(const LMat& m0) {
  const int n = m0.r;
  if (m0.c != n) throw std::runtime_error("inverse: not square");
  LMat a = m0, r = LMat::eye(n);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i)
      if (fabsl(a(i, c)) > fabsl(a(p, c))) p = i;
    if (a(p, c) == 0.0L) throw std::runtime_error("inverse: singular matrix");
    if (p != c)
      for (int j = 0; j < n; ++j) { std::swap(a(p, j), a(c, j)); std::swap(r(p, j), r(c, j)); }
    ld piv = a(c, c);
    for (int j = 0; j < n; ++j) { a(c, j) /= piv; r(c, j) /= piv; }
    for (int i = 0; i < n; ++i) {
      if (i == c) continue;
      ld f = a(i, c);
      if (f == 0.0L) continue;
      for (int j = 0; j < n; ++j) { a(i, j) -= f * a(c, j); r(i, j) -= f * r(c, j); }
    }
  }
  return r;
}

// composite_stencil.rs:250-276: c = (S^T S)^-1 S^T p (the tridiagonal normal equations)
LMat from_ortho(const Basis& b) {
  if (b.k == CHEB || b.k == FO) return LMat::eye(b.ortho());
  return from_ortho_cols(b, LMat::eye(b.n));
}

// r2c.rs:160-175 (realfft, unnormalised), split planes: rows [0,m) = Re, [m,2m) = Im
LMat dft_r2c(int n) {
  const int m = n / 2 + 1;
  LMat a(2 * m, n);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < n; ++j) {
      a(k, j) = cospi(2LL * j * k, n);
      a(m + k, j) = -sinpi(2LL * j * k, n);
    }
  return a;
}

// r2c.rs:211-226 (c2r, normalised 1/n): v_j = 1/n [Re X_0 + (-1)^j Re X_{n/2}
//   + 2 sum_{0<k<n/2} (Re X_k cos - Im X_k sin)]; Im X_0 and Im X_{n/2} ignored
LMat dft_c2r(int n) {
  const int m = n / 2 + 1;
  LMat a(n, 2 * m);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < m; ++k) {
      bool edge = (k == 0) || (n % 2 == 0 && k == n / 2);
      ld w = edge ? 1.0L : 2.0L;
      a(j, k) = w * cospi(2LL * j * k, n) / (ld)n;
      a(j, m + k) = edge ? 0.0L : -w * sinpi(2LL * j * k, n) / (ld)n;
    }
  return a;
}

// ortho.rs:147-174 (pinv of D2) and r2c.rs:372-394
static LMat laplace_inv(const Basis& b) {
  if (b.k == FO) {
    LMat p(b.m, b.m);
    for (int i = 1; i < b.m; ++i) p(i, i) = 1.0L / (-(ld)i * (ld)i);
    return p;
  }
  const int n = b.n;
  LMat p(n, n);
  p(2, 0) = 0.25L;
  for (int i = 3; i < n; ++i) p(i, i - 2) = (ld)(1.0 / (double)(4LL * i * (i - 1)));
  for (int i = 2; i < n - 2; ++i) p(i, i) = (ld)(-1.0 / (double)(2LL * ((long long)i * i - 1)));
  for (int i = 2; i < n - 4; ++i) p(i, i + 2) = (ld)(1.0 / (double)(4LL * i * (i + 1)));
  return p;
}
static LMat laplace_inv_eye(const Basis& b) {
  const int sh = b.k == FO ? 1 : 2, len = b.ortho();
  LMat e(len - sh, len);
  for (int i = 0; i < len - sh; ++i) e(i, i + sh) = 1.0L;
  return e;
}

// field.rs:194-252 ingredients_for_hholtz / _poisson.  The entries of these matrices are
// the reference's double values (pinv entries are rounded exactly as ortho.rs does).
Ingredients ingredients(const Basis& b) {
  Ingredients g;
  LMat peye = laplace_inv_eye(b);
  LMat pinv = peye * laplace_inv(b);
  if (b.k == CHEB) {
    LMat ms(b.n, b.n - 2);
    for (int i = 0; i < b.n - 2; ++i) ms(i + 2, i) = 1.0L;
    g.a = pinv * ms;
    g.b = peye * ms;
  } else if (b.k == FO) {
    g.a = LMat::eye(b.m);
    g.b = LMat(b.m, b.m);
    for (int i = 0; i < b.m; ++i) g.b(i, i) = -(ld)i * (ld)i;
  } else {
    LMat s = stencil(b);
    g.a = pinv * s;
    g.b = peye * s;
  }
  g.has_p = b.k != FO;
  if (g.has_p) g.p = pinv;
  g.is_diag = b.k == FO;
  // round to the double values the reference holds
  for (LMat* m : {&g.a, &g.b, &g.p})
    for (auto& v : m->a) v = (ld)(double)v;
  return g;
}

Band4 Band4::from_matrix(const std::vector<double>& a, int n) {  // utils.rs:17-43 diag
  Band4 f;
  f.n = n;
  f.low.resize(n - 2);
  f.dia.resize(n);
  f.up1.resize(n - 2);
  f.up2.resize(std::max(n - 4, 0));
  for (int i = 0; i < n - 2; ++i) f.low[i] = a[(size_t)(i + 2) * n + i];
  for (int i = 0; i < n; ++i) f.dia[i] = a[(size_t)i * n + i];
  for (int i = 0; i < n - 2; ++i) f.up1[i] = a[(size_t)i * n + i + 2];
  for (int i = 0; i < n - 4; ++i) f.up2[i] = a[(size_t)i * n + i + 4];
  return f;
}
Band4 Band4::add_scaled(const Band4& o, double l) const {
  Band4 r = *this;
  for (size_t i = 0; i < low.size(); ++i) r.low[i] = low[i] + o.low[i] * l;
  for (size_t i = 0; i < dia.size(); ++i) r.dia[i] = dia[i] + o.dia[i] * l;
  for (size_t i = 0; i < up1.size(); ++i) r.up1[i] = up1[i] + o.up1[i] * l;
  for (size_t i = 0; i < up2.size(); ++i) r.up2[i] = up2[i] + o.up2[i] * l;
  return r;
}
void Band4::sweep() {  // fdma.rs:73-82
  for (int i = 2; i < n; ++i) {
    low[i - 2] /= dia[i - 2];
    dia[i] -= low[i - 2] * up1[i - 2];
    if (i < n - 2) up1[i] -= low[i - 2] * up2[i - 2];
  }
}

// Dense 1-D operator `name` of basis b (axis scale sc), every construction O(n^2):
//   S    to_ortho                      (composite_stencil.rs:207-229)
//   BS   backward                      (composite.rs:430-443 ; r2c.rs:211-226)
//   BDS  d/dx then backward            (conv_term.rs:35-38)
//   DS   d/dx (to ortho)               (space2.rs:247-264, composite.rs:523-536)
//   DDS  d2/dx2 (to ortho)
//   FO   from_ortho                    (composite_stencil.rs:250-276)
//   FS   forward                       (composite.rs:383-396 ; r2c.rs:160-175)
//   Fdl  forward then 2/3 dealias      (navier.rs:562-568, 1022-1032)
LMat axis_operator(const Basis& b, const std::string& name, ld sc) {
  const int n = b.n;
  if (b.k == FO) {
    if (name == "BS") return dft_c2r(n);
    if (name == "FS") return dft_r2c(n);
    if (name == "Fdl") {
      LMat a = dft_r2c(n);
      const int cut = b.m * 2 / 3;
      for (int k = cut; k < b.m; ++k)
        for (int j = 0; j < n; ++j) a(k, j) = a(b.m + k, j) = 0.0L;
      return a;
    }
    throw std::runtime_error("no dense Fourier operator " + name);
  }
  LMat S = stencil(b);
  if (name == "S") return S;
  if (name == "BS") return cheb_backward(n) * S;
  if (name == "BDS") return scaled(cheb_backward_diff(n) * S, 1.0L / sc);
  if (name == "DS") return scaled(diff_cols(S, 1), 1.0L / sc);
  if (name == "DDS") return scaled(diff_cols(S, 2), 1.0L / (sc * sc));
  if (name == "FO") return from_ortho(b);
  if (name == "FS") return b.k == CHEB ? cheb_forward(n) : from_ortho_cols(b, cheb_forward(n));
  if (name == "Fdl") {
    LMat f = cheb_forward(n);
    const int cut = n * 2 / 3;
    for (int k = cut; k < n; ++k)
      for (int j = 0; j < n; ++j) f(k, j) = 0.0L;
    return f;
  }
  throw std::runtime_error("unknown operator " + name);
}

// ---------------------------------------------------------------- eigen-solver
// Real Schur decomposition A = Z T Z^T by Householder-Hessenberg reduction followed by
// the implicit double-shift (Francis) QR iteration with accumulated transformations
// (Golub & Van Loan, Algs. 7.4.2, 7.5.1, 7.5.2), in long double; eigenvectors by
// back-substitution on the triangular T.
namespace {

struct Schur {
  int n;
  std::vector<ld> h, z;
  ld& H(int i, int j) { return h[(size_t)i * n + j]; }
  ld& Z(int i, int j) { return z[(size_t)i * n + j]; }
};

void reflect(Schur& s, int k, int nr, const ld* v, int col_lo, int row_hi) {
  // H <- P H on rows k..k+nr-1 (cols col_lo..n-1); H <- H P on cols k.. (rows 0..row_hi); Z <- Z P
  ld vv = 0.0L;
  for (int i = 0; i < nr; ++i) vv += v[i] * v[i];
  if (vv == 0.0L) return;
  const int n = s.n;
  for (int j = col_lo; j < n; ++j) {
    ld t = 0.0L;
    for (int i = 0; i < nr; ++i) t += v[i] * s.H(k + i, j);
    t = 2.0L * t / vv;
    for (int i = 0; i < nr; ++i) s.H(k + i, j) -= t * v[i];
  }
  for (int i = 0; i <= row_hi; ++i) {
    ld t = 0.0L;
    for (int j = 0; j < nr; ++j) t += s.H(i, k + j) * v[j];
    t = 2.0L * t / vv;
    for (int j = 0; j < nr; ++j) s.H(i, k + j) -= t * v[j];
  }
  for (int i = 0; i < n; ++i) {
    ld t = 0.0L;
    for (int j = 0; j < nr; ++j) t += s.Z(i, k + j) * v[j];
    t = 2.0L * t / vv;
    for (int j = 0; j < nr; ++j) s.Z(i, k + j) -= t * v[j];
  }
}

void house_vec(const ld* x, int nr, ld* v) {
  ld a = 0.0L;
  for (int i = 0; i < nr; ++i) a += x[i] * x[i];
  a = sqrtl(a);
  for (int i = 0; i < nr; ++i) v[i] = x[i];
  if (a != 0.0L) v[0] += (x[0] >= 0 ? a : -a);
}

// returns false if a complex pair is found
bool schur(Schur& s) {
  const int n = s.n;
  // Hessenberg reduction with accumulation
  for (int k = 0; k + 2 < n; ++k) {
    std::vector<ld> x(n - k - 1), v(n - k - 1);
    for (int i = k + 1; i < n; ++i) x[i - k - 1] = s.H(i, k);
    house_vec(x.data(), n - k - 1, v.data());
    reflect(s, k + 1, n - k - 1, v.data(), 0, n - 1);
    for (int i = k + 2; i < n; ++i) s.H(i, k) = 0.0L;
  }
  const ld eps = 1.0842021724855044340e-19L;  // long double epsilon
  int hi = n - 1, iter = 0;
  while (hi >= 1) {
    int l = hi;
    for (; l > 0; --l) {
      ld sc = fabsl(s.H(l - 1, l - 1)) + fabsl(s.H(l, l));
      if (fabsl(s.H(l, l - 1)) <= eps * sc) { s.H(l, l - 1) = 0.0L; break; }
    }
    if (l == hi) { --hi; iter = 0; continue; }
    if (l == hi - 1) {
      // 2x2 block: split by a rotation if real, else complex -> fail
      ld a = s.H(hi - 1, hi - 1), b = s.H(hi - 1, hi), c = s.H(hi, hi - 1), d = s.H(hi, hi);
      ld p = 0.5L * (a - d), disc = p * p + b * c;
      if (disc < 0) return false;
      ld zz = sqrtl(disc);
      zz = p >= 0 ? p + zz : p - zz;
      // mu = d + zz is an eigenvalue of the block; (mu - d, c) = (zz, c) its eigenvector.
      // A reflector whose first column is that eigenvector triangularises the block.
      ld x[2] = {zz, c};
      ld v[2];
      house_vec(x, 2, v);
      // reflect on rows/cols hi-1..hi using full-matrix updates
      reflect(s, hi - 1, 2, v, hi - 1, hi);
      s.H(hi, hi - 1) = 0.0L;
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++iter > 300) throw std::runtime_error("eig: QR iteration did not converge");
    ld sh, tt;
    if (iter % 11 == 10) {
      ld w = fabsl(s.H(hi, hi - 1)) + fabsl(s.H(hi - 1, hi - 2));
      sh = 1.5L * w;
      tt = w * w;
    } else {
      sh = s.H(hi - 1, hi - 1) + s.H(hi, hi);
      tt = s.H(hi - 1, hi - 1) * s.H(hi, hi) - s.H(hi - 1, hi) * s.H(hi, hi - 1);
    }
    ld x[3];
    x[0] = s.H(l, l) * s.H(l, l) + s.H(l, l + 1) * s.H(l + 1, l) - sh * s.H(l, l) + tt;
    x[1] = s.H(l + 1, l) * (s.H(l, l) + s.H(l + 1, l + 1) - sh);
    x[2] = s.H(l + 1, l) * s.H(l + 2, l + 1);
    for (int k = l; k <= hi - 2; ++k) {
      ld v[3];
      house_vec(x, 3, v);
      int q = std::max(l, k - 1);
      int r = std::min(k + 3, hi);
      // left on rows k..k+2 from column q; right on rows 0..r
      reflect(s, k, 3, v, q, r);
      x[0] = s.H(k + 1, k);
      x[1] = s.H(k + 2, k);
      if (k < hi - 2) x[2] = s.H(k + 3, k);
    }
    ld v[2];
    house_vec(x, 2, v);
    reflect(s, hi - 1, 2, v, hi - 2, hi);
  }
  return true;
}

}  // namespace

bool eig_real(const std::vector<double>& a, int n, std::vector<double>& lam, std::vector<double>& q,
              std::vector<double>& qi, std::string& msg) {
  // isolated columns (zero off-diagonal): exact eigenvectors e_j (LAPACK dgebal permutation)
  std::vector<int> iso(n, 0), rest;
  for (int j = 0; j < n; ++j) {
    bool z = true;
    for (int i = 0; i < n && z; ++i)
      if (i != j && a[(size_t)i * n + j] != 0.0) z = false;
    iso[j] = z;
    if (!z) rest.push_back(j);
  }
  const int nr = (int)rest.size();
  std::vector<ld> lam0;
  LMat vecs(n, n);
  int col = 0;
  for (int j = 0; j < n; ++j)
    if (iso[j]) {
      lam0.push_back((ld)a[(size_t)j * n + j]);
      vecs(j, col++) = 1.0L;
    }
  if (nr > 0) {
    Schur s;
    s.n = nr;
    s.h.assign((size_t)nr * nr, 0.0L);
    s.z.assign((size_t)nr * nr, 0.0L);
    for (int i = 0; i < nr; ++i) {
      s.Z(i, i) = 1.0L;
      for (int j = 0; j < nr; ++j) s.H(i, j) = (ld)a[(size_t)rest[i] * n + rest[j]];
    }
    if (!schur(s)) { msg = "eig: complex eigenvalues"; return false; }
    ld tn = 0.0L;
    for (auto v : s.h) tn = std::max(tn, fabsl(v));
    const ld tiny = tn * 1e-30L + 1e-300L;
    for (int k = 0; k < nr; ++k) {
      // eigenvector of upper-triangular T for T(k,k)
      std::vector<ld> x(nr, 0.0L);
      x[k] = 1.0L;
      const ld tk = s.H(k, k);
      for (int i = k - 1; i >= 0; --i) {
        ld sum = 0.0L;
        for (int j = i + 1; j <= k; ++j) sum += s.H(i, j) * x[j];
        ld den = s.H(i, i) - tk;
        if (fabsl(den) < tiny) den = den >= 0 ? tiny : -tiny;
        x[i] = -sum / den;
      }
      std::vector<ld> w(nr, 0.0L);
      for (int i = 0; i < nr; ++i)
        for (int j = 0; j <= k; ++j) w[i] += s.Z(i, j) * x[j];
      // extend to the isolated coordinates: v_I = (mu - d_I)^-1 A_IJ w
      std::vector<ld> v(n, 0.0L);
      for (int i = 0; i < nr; ++i) v[rest[i]] = w[i];
      for (int j = 0; j < n; ++j) {
        if (!iso[j]) continue;
        ld s2 = 0.0L;
        for (int i = 0; i < nr; ++i) s2 += (ld)a[(size_t)j * n + rest[i]] * w[i];
        ld den = tk - (ld)a[(size_t)j * n + j];
        if (fabsl(den) < tiny) den = den >= 0 ? tiny : -tiny;
        v[j] = s2 / den;
      }
      ld nv = 0.0L;
      for (auto t : v) nv += t * t;
      nv = sqrtl(nv);
      for (int i = 0; i < n; ++i) vecs(i, col) = v[i] / nv;
      lam0.push_back(tk);
      ++col;
    }
  }
  // sort descending: reverse of a stable ascending sort (utils.rs:80-96)
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int i, int j) { return (double)lam0[i] < (double)lam0[j]; });
  std::reverse(perm.begin(), perm.end());
  LMat qs(n, n);
  lam.resize(n);
  for (int c = 0; c < n; ++c) {
    lam[c] = (double)lam0[perm[c]];
    for (int i = 0; i < n; ++i) qs(i, c) = vecs(i, perm[c]);
  }
  LMat qinv = inverse(qs);
  q = qs.to_double();
  qi = qinv.to_double();
  return true;
}

}  // namespace rbc
