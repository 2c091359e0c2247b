// Host-side set-up of the 1-D operators of the RBC step (extended precision).
//
// Every 1-D operation the step applies along one axis (composite stencil, Chebyshev
// transform, derivative recurrence, Fourier DFT, normal-equation projection, ADI
// inverse, eigen-transform) is linear; it is built here once as a dense matrix in
// long double and rounded to double, then applied on the GPU by the GEMM kernels.
// Bases mirror funspace (funspace/src/chebyshev/*.rs, funspace/src/fourier/r2c.rs).
#pragma once
#include <string>
#include <vector>

namespace rbc {

using ld = long double;

enum Kind { CHEB = 0, CD = 1, CN = 2, CDBC = 3, FO = 4 };

// Dense row-major matrix in long double.
struct LMat {
  int r = 0, c = 0;
  std::vector<ld> a;
  LMat() = default;
  LMat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0L) {}
  ld& operator()(int i, int j) { return a[(size_t)i * c + j]; }
  ld operator()(int i, int j) const { return a[(size_t)i * c + j]; }
  static LMat eye(int n) { LMat m(n, n); for (int i = 0; i < n; ++i) m(i, i) = 1.0L; return m; }
  LMat t() const { LMat m(c, r); for (int i = 0; i < r; ++i) for (int j = 0; j < c; ++j) m(j, i) = (*this)(i, j); return m; }
  std::vector<double> to_double() const { std::vector<double> d(a.size()); for (size_t i = 0; i < a.size(); ++i) d[i] = (double)a[i]; return d; }
};
LMat operator*(const LMat& x, const LMat& y);
LMat scaled(const LMat& x, ld s);

struct Basis {
  Kind k;
  int n, m;  // physical length, spectral length
  Basis(Kind kind, int n_);
  int ortho() const { return k == FO ? m : n; }
  bool fourier() const { return k == FO; }
  // composite stencil coefficients (composite_stencil.rs:117-162)
  std::vector<double> diag, low2;
  // physical coordinates (ortho.rs:71-80 ; c2c.rs:63-66)
  std::vector<double> coords() const;
};

// ---- Chebyshev operators (size n = physical length) ----
LMat stencil(const Basis& b);         // S : ortho <- composite   (n x m)
LMat cheb_backward(int n);            // B : physical <- ortho    (n x n)
LMat cheb_forward(int n);             // F : ortho <- physical    (n x n)
LMat cheb_diff(int n, int times);     // D^t on ortho coefficients (n x n)
LMat from_ortho(const Basis& b);      // (S^T S)^-1 S^T, composite <- ortho (m x n)
// named dense 1-D operator of basis b along an axis of scale sc (see setup.cpp)
LMat axis_operator(const Basis& b, const std::string& name, ld sc);
// ---- Fourier (r2c) operators on split planes [re; im] ----
LMat dft_r2c(int n);                  // (2m x n)
LMat dft_c2r(int n);                  // (n x 2m), 1/n normalised, Im of k=0, n/2 ignored
// ---- solver ingredients (field.rs:194-252) ----
struct Ingredients { LMat a, b, p; bool has_p, is_diag; };
Ingredients ingredients(const Basis& b);
LMat inverse(const LMat& m);          // Gauss-Jordan with partial pivoting, long double

// Banded 4-diagonal system (offsets -2, 0, 2, 4), swept as fdma.rs:73-82.
struct Band4 {
  std::vector<double> low, dia, up1, up2;  // low/up1 length n-2, up2 length n-4
  int n = 0;
  static Band4 from_matrix(const std::vector<double>& a, int n);  // diagonals of a square matrix
  Band4 add_scaled(const Band4& o, double l) const;                // fdma.rs:187-227 : self + o*l
  void sweep();
};

// Eigen-decomposition of a real matrix with real spectrum, as src/solver/utils.rs:66-98:
// eigenvalues sorted descending, columns of q the eigenvectors (unit 2-norm), qi = q^-1.
// Columns of `a` whose off-diagonal entries are exactly zero yield exact unit vectors.
// Returns false (with msg) if complex eigenvalues appear.
bool eig_real(const std::vector<double>& a, int n, std::vector<double>& lam, std::vector<double>& q,
              std::vector<double>& qi, std::string& msg);

}  // namespace rbc
