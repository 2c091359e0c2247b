// Lane kernels of the RBC step: banded row solves and banded / stencil operators.
//
// fdma_rows_wave: one wavefront per row of X[nrows, m]; the offsets -2, 0, 2, 4 of the
// 4-diagonal systems (fdma.rs:101-118) couple only indices of equal parity, so a row is
// two independent chains.  Every lane owns the same contiguous slice of both chains.  A
// pass is a linear recurrence: each lane composes the affine map of its slice, the maps
// are scanned across the wavefront with shuffles, and each lane then replays its slice
// with the reference's arithmetic from its incoming state.
//
// band_apply: out_i = sum of up to four taps at offsets 0, -2, +2, +4 (in that order of
// summation) along axis 1 (rows) or axis 0 (columns); covers the composite stencil
// to_ortho (composite_stencil.rs:207-229) and the banded matvec MatVecFdma
// (matvec.rs:172-193).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "bufrow.hpp"
#include "ops.hpp"

namespace rbc {

namespace {

constexpr int WAVE = 64;
constexpr int ROWS_PER_BLOCK = 4;
constexpr int SLICE = 16;  // max elements per lane per parity (rows up to 2048)

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__global__ __launch_bounds__(WAVE* ROWS_PER_BLOCK) void fdma_rows_wave_kernel(
    double* __restrict__ X, int ld, int nrows, int m, int nsets, const double* __restrict__ low,
    const double* __restrict__ dia, const double* __restrict__ up1, const double* __restrict__ up2, long sx) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (WAVE - 1);
  const int r = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x / WAVE);
  if (r >= nrows) return;  // uniform per wavefront
  X += sx * blockIdx.y;
  const int set = r % nsets;
  double* x = X + (size_t)r * ld;
  const double* Lw = low + (size_t)set * m;
  const double* D = dia + (size_t)set * m;
  const double* U1 = up1 + (size_t)set * m;
  const double* U2 = up2 + (size_t)set * m;
  const int len0 = (m + 1) / 2;            // chain length of parity 0 (>= parity 1)
  const int per = (len0 + WAVE - 1) / WAVE;  // slice length, <= SLICE
  const int t0 = lane * per;
  double v[2][SLICE];
  const int len[2] = {len0, m / 2};
  int cnt[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    cnt[p] = max(0, min(per, len[p] - t0));
#pragma unroll
    for (int q = 0; q < SLICE; ++q) v[p][q] = (q < cnt[p]) ? x[p + 2 * (t0 + q)] : 0.0;
  }
  // ---- forward elimination: y_t = x_t - y_{t-1} low[i-2]  (i = p + 2t >= 2)
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double A = 1.0, B = 0.0;  // slice map y_out = A y_in + B
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      if (q < cnt[p]) {
        const int t = t0 + q, i = p + 2 * t;
        if (t == 0) {
          A = 0.0;
          B = v[p][q];
        } else {
          const double l = Lw[i - 2];
          A = -A * l;
          B = v[p][q] - B * l;
        }
      }
    }
    double pa = A, pb = B;  // inclusive scan, lower lanes first
    for (int off = 1; off < WAVE; off <<= 1) {
      const double qa = __shfl_up(pa, off), qb = __shfl_up(pb, off);
      if (lane >= off) { pb = pa * qb + pb; pa = pa * qa; }
    }
    double y = __shfl_up(pb, 1);
    if (lane == 0) y = 0.0;
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      if (q < cnt[p]) {
        const int t = t0 + q, i = p + 2 * t;
        y = (t == 0) ? v[p][q] : v[p][q] - y * Lw[i - 2];
        v[p][q] = y;
      }
    }
  }
  // ---- back substitution: x_t = (y_t - x_{t+1} up1[i] - x_{t+2} up2[i]) / dia[i]
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    // slice map on the state s = (x_{t+1}, x_{t+2}):  s' = M s + w
    double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0, w0 = 0.0, w1 = 0.0;
#pragma unroll
    for (int q = SLICE - 1; q >= 0; --q) {
      if (q >= cnt[p]) continue;
      const int t = t0 + q, i = p + 2 * t;
      const double d = D[i];
      const double u1 = (t + 1 < len[p]) ? U1[i] : 0.0;
      const double u2 = (t + 2 < len[p]) ? U2[i] : 0.0;
      const double c0 = v[p][q] / d, c1 = -u1 / d, c2 = -u2 / d;
      // row 0 of the new map: c1 * row0 + c2 * row1 ; row 1: old row 0
      const double n00 = c1 * m00 + c2 * m10, n01 = c1 * m01 + c2 * m11, nw0 = c1 * w0 + c2 * w1 + c0;
      m10 = m00; m11 = m01; w1 = w0;
      m00 = n00; m01 = n01; w0 = nw0;
    }
    // suffix scan: higher lanes are applied first
    double a00 = m00, a01 = m01, a10 = m10, a11 = m11, b0 = w0, b1 = w1;
    for (int off = 1; off < WAVE; off <<= 1) {
      const double q00 = __shfl_down(a00, off), q01 = __shfl_down(a01, off), q10 = __shfl_down(a10, off),
                   q11 = __shfl_down(a11, off), r0 = __shfl_down(b0, off), r1 = __shfl_down(b1, off);
      if (lane + off < WAVE) {  // composite = this o q
        const double n00 = a00 * q00 + a01 * q10, n01 = a00 * q01 + a01 * q11;
        const double n10 = a10 * q00 + a11 * q10, n11 = a10 * q01 + a11 * q11;
        const double nb0 = a00 * r0 + a01 * r1 + b0, nb1 = a10 * r0 + a11 * r1 + b1;
        a00 = n00; a01 = n01; a10 = n10; a11 = n11; b0 = nb0; b1 = nb1;
      }
    }
    double s0 = __shfl_down(b0, 1), s1 = __shfl_down(b1, 1);
    if (lane == WAVE - 1) { s0 = 0.0; s1 = 0.0; }
#pragma unroll
    for (int q = SLICE - 1; q >= 0; --q) {
      if (q >= cnt[p]) continue;
      const int t = t0 + q, i = p + 2 * t;
      double val = v[p][q];
      if (t + 1 < len[p]) val = val - s0 * U1[i];
      if (t + 2 < len[p]) val = val - s1 * U2[i];
      val = val / D[i];
      s1 = s0;
      s0 = val;
      v[p][q] = val;
    }
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < SLICE; ++q)
      if (q < cnt[p]) x[p + 2 * (t0 + q)] = v[p][q];
}

// taps: coef[k * mo + i] for k = 0..3 at input offsets {0, -2, +2, +4}; mask bit k set if
// tap k is present at output index i.  The input is read through a raw buffer over its extent
// (reads outside it return 0): every tap kind in `taps` (the union of the mask bits) is loaded
// for every output, and the output's mask selects among the products, so the loads are issued
// together (a load guarded by the mask would wait for the mask's round trip, then for its own).
__global__ void band_y_kernel(const double* __restrict__ in, int ldi, double* __restrict__ out, int ldo, int rows,
                              int mo, int mi, int taps, const double* __restrict__ coef,
                              const unsigned char* __restrict__ mask, double beta, long si, long so) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)rows * mo) return;
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  const int r = (int)(t / mo), i = (int)(t % mo);
  const BufRow<double> inb(in, (rows - 1) * ldi + mi);
  const unsigned mk = mask[i];
  const int off[4] = {0, -2, 2, 4};
  double x[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = 0.0;
    c[k] = 0.0;
    if (taps & (1 << k)) {
      x[k] = inb[r * ldi + i + off[k]];
      c[k] = coef[k * mo + i];
    }
  }
  double acc = 0.0;
  bool first = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!(mk & (1u << k))) continue;
    const double term = x[k] * c[k];
    acc = first ? term : acc + term;
    first = false;
  }
  double* o = out + (size_t)r * ldo + i;
  *o = (beta != 0.0) ? acc + beta * (*o) : acc;
}

__global__ void band_x_kernel(const double* __restrict__ in, double* __restrict__ out, int cols, int mo, int mi,
                              int taps, const double* __restrict__ coef, const unsigned char* __restrict__ mask,
                              double beta, long si, long so) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)cols * mo) return;
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  const int i = (int)(t / cols), j = (int)(t % cols);
  const BufRow<double> inb(in, mi * cols);
  const unsigned mk = mask[i];
  const int off[4] = {0, -2, 2, 4};
  double x[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = 0.0;
    c[k] = 0.0;
    if (taps & (1 << k)) {
      x[k] = inb[(i + off[k]) * cols + j];
      c[k] = coef[k * mo + i];
    }
  }
  double acc = 0.0;
  bool first = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!(mk & (1u << k))) continue;
    const double term = x[k] * c[k];
    acc = first ? term : acc + term;
    first = false;
  }
  double* o = out + (size_t)i * cols + j;
  *o = (beta != 0.0) ? acc + beta * (*o) : acc;
}

// band_x then band_y in one pass (the 2-D composite stencil to_ortho of a field): each
// output forms the band_x values it needs with band_x's arithmetic and combines them as
// band_y does.  in [mix, ldi] (miy values per row), out [mox, moy].  Epilogue: out = acc,
// out = acc + add (as ew_add), or with p0 the pressure update of ew_update_pres:
// p0 = (p0 - div nu) + acc inv_dt, nu = nuarr[entry] when given.  The inputs of every tap
// pair in tx x ty are loaded as band_y_kernel loads its taps.
__global__ void band_xy_kernel(const double* __restrict__ in, int ldi, double* __restrict__ out, int mox, int moy,
                               int mix, int miy, int tx, int ty, const double* __restrict__ cx,
                               const unsigned char* __restrict__ mkx, const double* __restrict__ cy,
                               const unsigned char* __restrict__ mky, BandXYEpi ep, long si, long so) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)mox * moy) return;
  in += si * blockIdx.y;
  if (ep.split > 0 && (int)blockIdx.y >= ep.split) {  // the second field's x stencil
    cx = ep.x2.coef;
    mkx = ep.x2.mask;
    mix = ep.x2.mi;
    tx = ep.x2.taps;
  }
  const int i = (int)(t / moy), j = (int)(t % moy);
  const BufRow<double> inb(in, (mix - 1) * ldi + miy);
  const unsigned bx = mkx[i], by = mky[j];
  const int off[4] = {0, -2, 2, 4};
  double cxv[4], cyv[4], xin[4][4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    cxv[k] = (tx & (1 << k)) ? cx[k * mox + i] : 0.0;
    cyv[k] = (ty & (1 << k)) ? cy[k * moy + j] : 0.0;
  }
#pragma unroll
  for (int ky = 0; ky < 4; ++ky)
#pragma unroll
    for (int kx = 0; kx < 4; ++kx)
      xin[ky][kx] = ((ty & (1 << ky)) && (tx & (1 << kx))) ? inb[(i + off[kx]) * ldi + j + off[ky]] : 0.0;
  auto xval = [&](int ky) {
    double acc = 0.0;
    bool first = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!(bx & (1u << k))) continue;
      const double term = xin[ky][k] * cxv[k];
      acc = first ? term : acc + term;
      first = false;
    }
    return acc;
  };
  double acc = 0.0;
  bool first = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!(by & (1u << k))) continue;
    const double term = xval(k) * cyv[k];
    acc = first ? term : acc + term;
    first = false;
  }
  const size_t o = so * blockIdx.y + (size_t)i * moy + j;
  if (ep.p0) {
    const double nu = ep.nuarr ? ep.nuarr[blockIdx.y] : ep.nu;
    const double q = ep.p0[o] - ep.div[o] * nu;
    ep.p0[o] = q + acc * ep.inv_dt;
  } else {
    out[o] = ep.add ? acc + ep.add[o] : acc;
  }
}

// ---------------------------------------------------------------------------------
// cheb_diff_y: d/dy of every row (ortho.rs:107-125, once), the composite stencil
// (to_ortho) fused on load.  d_i = w_i + d_{i+2} with w_i = 2 (i+1) p_{i+1} (i >= 1),
// d_{n-1} = 0, d_0 = p_1 + d_2 / 2: a suffix sum along each parity chain.  Output is
// scaled by 1/sc and optionally accumulated: out = d / sc [+ beta out].
__global__ __launch_bounds__(WAVE* ROWS_PER_BLOCK) void cheb_diff_y_kernel(
    const double* __restrict__ in, int ldi, int mi, double* __restrict__ out, int ldo, int rows, int n,
    const double* __restrict__ coef, const unsigned char* __restrict__ mask, double inv_sc, double beta, long si,
    long so) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (WAVE - 1);
  // the row of the wavefront (the same in every lane: scalar, so the row buffers below are)
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x / WAVE));
  if (r >= rows) return;
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  const double* a = in + (size_t)r * ldi;
  double* o = out + (size_t)r * ldo;
  const int per = ((n + 1) / 2 + WAVE - 1) / WAVE;  // slice length per parity
  const int i0 = 2 * lane * per;                      // first index of this lane's slice
  // p_j (ortho coefficients) for j in [i0, i0 + 2 per + 1].  The row (mi values) and the
  // stencil taps are read through raw buffers whose range check returns 0 outside, so every
  // load is issued and the mask only selects among the products (a load guarded by a mask bit
  // would wait for the mask's round trip and then for its own)
  const BufRow<double> ar(a, mi), cr(coef, coef ? 2 * n : 0);
  auto pval = [&](int j) -> double {
    if (!coef) {
      const double x = ar[j];
      return j < n ? x : 0.0;
    }
    const unsigned mk = mask[min(j, n - 1)];
    const double x0 = ar[j], c0 = cr[j], x2 = ar[j - 2], c2 = cr[n + j];
    double acc = 0.0;
    bool first = true;
    if (mk & 1u) { acc = x0 * c0; first = false; }
    if (mk & 2u) { const double t = x2 * c2; acc = first ? t : acc + t; }
    return j < n ? acc : 0.0;
  };
  double w[2][SLICE];
  double s[2] = {0.0, 0.0};  // slice sums, accumulated top-down as the reference does
#pragma unroll
  for (int qq = SLICE - 1; qq >= 0; --qq)
#pragma unroll
    for (int p = 1; p >= 0; --p) {
      const int i = i0 + 2 * qq + p;
      double wi = 0.0;
      if (qq < per && i >= 1 && i <= n - 2) {
        wi = pval(i + 1) * (2.0 * (double)(i + 1));
        s[p] = wi + s[p];
      }
      w[p][qq] = wi;
    }
  // suffix sums of slice sums across lanes (higher lanes first)
  double in0 = s[0], in1 = s[1];
  for (int off = 1; off < WAVE; off <<= 1) {
    const double u0 = __shfl_down(in0, off), u1 = __shfl_down(in1, off);
    if (lane + off < WAVE) { in0 = in0 + u0; in1 = in1 + u1; }
  }
  double d[2] = {__shfl_down(in0, 1), __shfl_down(in1, 1)};
  if (lane == WAVE - 1) { d[0] = 0.0; d[1] = 0.0; }
  double res[2][SLICE];
#pragma unroll
  for (int qq = SLICE - 1; qq >= 0; --qq)
#pragma unroll
    for (int p = 1; p >= 0; --p) {
      const int i = i0 + 2 * qq + p;
      double v = 0.0;
      if (qq < per && i < n) {
        if (i == n - 1) v = 0.0;
        else if (i == n - 2) v = w[p][qq];
        else if (i >= 1) v = w[p][qq] + d[p];
        if (i >= 1) d[p] = v;
      }
      res[p][qq] = v;
    }
  // d_0 = p_1 + d_2 / 2, with d_2 in lane 0 (slices of length >= 2) or lane 1
  const double d2 = (per >= 2) ? res[0][1] : __shfl(res[0][0], 1);
  if (lane == 0) res[0][0] = pval(1) + d2 / 2.0;
#pragma unroll
  for (int qq = 0; qq < SLICE; ++qq)
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int i = i0 + 2 * qq + p;
      if (qq < per && i < n) {
        double v = res[p][qq];
        if (inv_sc != 1.0) v = v * inv_sc;
        o[i] = (beta != 0.0) ? v + beta * o[i] : v;
      }
    }
}

// tdma_y: from_ortho of every row (composite_stencil.rs:250-276 + linalg.rs:14-57):
// d_i = p_i diag_i + p_{i+2} low2_i (i < m), then the tridiagonal (offsets -2, 0, 2)
// normal equations with the host-computed w_i and den_i = main_i - off_{i-2} w_{i-2}:
// g_i = (d_i - g_{i-2} off_{i-2}) / den_i, x_{i} = g_{i} - x_{i+2} w_{i}.
__global__ __launch_bounds__(WAVE* ROWS_PER_BLOCK) void tdma_y_kernel(
    const double* __restrict__ in, int ldi, double* __restrict__ out, int ldo, int rows, int m,
    const double* __restrict__ diag, const double* __restrict__ low2, const double* __restrict__ off,
    const double* __restrict__ wv, const double* __restrict__ den, long si, long so) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (WAVE - 1);
  const int r = blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x / WAVE);
  if (r >= rows) return;
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  const double* a = in + (size_t)r * ldi;
  double* o = out + (size_t)r * ldo;
  const int per = ((m + 1) / 2 + WAVE - 1) / WAVE;
  const int t0 = lane * per;
  const int len[2] = {(m + 1) / 2, m / 2};
  double v[2][SLICE];
  int cnt[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    cnt[p] = max(0, min(per, len[p] - t0));
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      v[p][q] = 0.0;
      if (q < cnt[p]) {
        const int i = p + 2 * (t0 + q);
        v[p][q] = a[i] * diag[i] + a[i + 2] * low2[i];
      }
    }
  }
  // forward: g_i = (d_i - g_{i-2} off_{i-2}) / den_i  (i >= 2), g_i = d_i / den_i (i < 2)
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      if (q < cnt[p]) {
        const int t = t0 + q, i = p + 2 * t;
        if (t == 0) {
          A = 0.0;
          B = v[p][q] / den[i];
        } else {
          A = (-off[i - 2] / den[i]) * A;
          B = (v[p][q] - B * off[i - 2]) / den[i];
        }
      }
    }
    double pa = A, pb = B;
    for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
      const double qa = __shfl_up(pa, o2), qb = __shfl_up(pb, o2);
      if (lane >= o2) { pb = pa * qb + pb; pa = pa * qa; }
    }
    double g = __shfl_up(pb, 1);
    if (lane == 0) g = 0.0;
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      if (q < cnt[p]) {
        const int t = t0 + q, i = p + 2 * t;
        g = (t == 0) ? v[p][q] / den[i] : (v[p][q] - g * off[i - 2]) / den[i];
        v[p][q] = g;
      }
    }
  }
  // back: x_{m-1} = g_{m-1}, x_{m-2} = g_{m-2}, x_i = g_i - x_{i+2} w_i
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double A = 1.0, B = 0.0;  // x_out = A x_in + B over the slice, top-down
#pragma unroll
    for (int q = SLICE - 1; q >= 0; --q) {
      if (q < cnt[p]) {
        const int t = t0 + q, i = p + 2 * t;
        if (i >= m - 2) {
          A = 0.0;
          B = v[p][q];
        } else {
          A = -wv[i] * A;
          B = v[p][q] - B * wv[i];
        }
      }
    }
    double pa = A, pb = B;
    for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
      const double qa = __shfl_down(pa, o2), qb = __shfl_down(pb, o2);
      if (lane + o2 < WAVE) { pb = pa * qb + pb; pa = pa * qa; }
    }
    double x = __shfl_down(pb, 1);
    if (lane == WAVE - 1) x = 0.0;
#pragma unroll
    for (int q = SLICE - 1; q >= 0; --q) {
      if (q < cnt[p]) {
        const int t = t0 + q, i = p + 2 * t;
        x = (i >= m - 2) ? v[p][q] : v[p][q] - x * wv[i];
        v[p][q] = x;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < SLICE; ++q)
      if (q < cnt[p]) o[p + 2 * (t0 + q)] = v[p][q];
}
}  // namespace

bool cheb_diff_y_supported(int n) { return n >= 6 && (n + 1) / 2 <= WAVE * SLICE; }
bool tdma_y_supported(int m) { return m >= 4 && (m + 1) / 2 <= WAVE * SLICE; }

void cheb_diff_y(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                 const double* coef, const unsigned char* mask, double inv_sc, double beta, int batch, long si,
                 long so) {
  if (!cheb_diff_y_supported(n)) throw std::invalid_argument("cheb_diff_y: unsupported length");
  if (mi < 1 || mi > ldi || mi > n || (!coef && mi != n)) throw std::invalid_argument("cheb_diff_y: bad input length");
  const unsigned blocks = (unsigned)((rows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK);
  RBC_LAUNCH(cheb_diff_y_kernel, dim3(blocks, batch), dim3(WAVE * ROWS_PER_BLOCK), 0, s, in, ldi, mi, out, ldo,
                     rows, n, coef, mask, inv_sc, beta, si, so);
  check("cheb_diff_y");
}

void tdma_y(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, int m, const double* diag,
            const double* low2, const double* off, const double* w, const double* den, int batch, long si, long so) {
  if (!tdma_y_supported(m)) throw std::invalid_argument("tdma_y: unsupported length");
  const unsigned blocks = (unsigned)((rows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK);
  RBC_LAUNCH(tdma_y_kernel, dim3(blocks, batch), dim3(WAVE * ROWS_PER_BLOCK), 0, s, in, ldi, out, ldo, rows, m,
                     diag, low2, off, w, den, si, so);
  check("tdma_y");
}

bool fdma_rows_wave_supported(int m) { return m >= 5 && (m + 1) / 2 <= WAVE * SLICE; }

void fdma_rows_wave(hipStream_t s, double* X, int ld, int nrows, int m, int nsets, const double* low, const double* dia,
                    const double* up1, const double* up2, int batch, long sx) {
  if (!fdma_rows_wave_supported(m)) throw std::invalid_argument("fdma_rows_wave: unsupported row length");
  const unsigned blocks = (unsigned)((nrows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK);
  RBC_LAUNCH(fdma_rows_wave_kernel, dim3(blocks, batch), dim3(WAVE * ROWS_PER_BLOCK), 0, s, X, ld, nrows, m,
                     nsets, low, dia, up1, up2, sx);
  check("fdma_rows_wave");
}

// the raw buffers of the band kernels span a batch entry's input: its byte size must fit an int
static void band_extent_check(long elems, const char* what) {
  if (elems < 1 || elems > 2147483647L / 8) throw std::invalid_argument(std::string(what) + ": input extent");
}

void band_y(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, const BandRef& b, double beta,
            int batch, long si, long so) {
  const long n = (long)rows * b.mo;
  if (n <= 0) return;
  if (b.mi < 1 || b.mi > ldi) throw std::invalid_argument("band_y: bad input length");
  band_extent_check((long)(rows - 1) * ldi + b.mi, "band_y");
  RBC_LAUNCH(band_y_kernel, dim3((unsigned)((n + 255) / 256), batch), dim3(256), 0, s, in, ldi, out, ldo, rows,
             b.mo, b.mi, b.taps, b.coef, b.mask, beta, si, so);
  check("band_y");
}

void band_xy(hipStream_t s, const double* in, int ldi, double* out, const BandRef& bx, const BandRef& by,
             const BandXYEpi& ep, int batch, long si, long so) {
  const long n = (long)bx.mo * by.mo;
  if (n <= 0) return;
  if (batch < 1 || batch > 65535) throw std::invalid_argument("band_xy: bad batch");
  if (ep.p0 && !ep.div) throw std::invalid_argument("band_xy: pressure update without div");
  if (ep.split > 0 && (!ep.x2.coef || !ep.x2.mask || ep.x2.mo != bx.mo))
    throw std::invalid_argument("band_xy: second stencil missing");
  if (by.mi < 1 || by.mi > ldi || bx.mi < 1) throw std::invalid_argument("band_xy: bad input lengths");
  band_extent_check((long)(bx.mi - 1) * ldi + by.mi, "band_xy");
  if (ep.split > 0) band_extent_check((long)(ep.x2.mi - 1) * ldi + by.mi, "band_xy");
  RBC_LAUNCH(band_xy_kernel, dim3((unsigned)((n + 255) / 256), batch), dim3(256), 0, s, in, ldi, out, bx.mo, by.mo,
             bx.mi, by.mi, bx.taps, by.taps, bx.coef, bx.mask, by.coef, by.mask, ep, si, so);
  check("band_xy");
}

void band_x(hipStream_t s, const double* in, double* out, int cols, const BandRef& b, double beta, int batch, long si,
            long so) {
  const long n = (long)cols * b.mo;
  if (n <= 0) return;
  if (b.mi < 1) throw std::invalid_argument("band_x: bad input length");
  band_extent_check((long)b.mi * cols, "band_x");
  RBC_LAUNCH(band_x_kernel, dim3((unsigned)((n + 255) / 256), batch), dim3(256), 0, s, in, out, cols, b.mo, b.mi,
             b.taps, b.coef, b.mask, beta, si, so);
  check("band_x");
}

}  // namespace rbc
