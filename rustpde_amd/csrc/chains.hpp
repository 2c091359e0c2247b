// Recurrences along one parity chain of a row, held by one wavefront: lane l owns chain
// positions t in [t0, t0 + cnt), t0 = l * per, in registers v[q] (q = t - t0).  Each pass is
// a linear recurrence: the lane composes the affine map of its slice, the maps are scanned
// across the wavefront with shuffles, and each lane replays its slice with the reference's
// arithmetic from its incoming state.  Used by the row kernels and the fused pressure stage.
#pragma once
#include <hip/hip_runtime.h>

namespace rbc {
namespace chains {

constexpr int WAVE = 64;

// padded LDS index of row element i (stored at pad(i + 2)): one pad double per 16
__host__ __device__ __forceinline__ int pad(int j) { return j + (j >> 4); }

// LDS doubles for a row of nr values with two guard zeros on each side
inline int buf_len(int nr) { return pad(nr + 6) + 1; }

// A row of nr values -> LDS (element i at pad(i + 2), zeros at i in [-2, 0) and
// [nr, nr + 4)) by the nthr threads of a group, thread tid; all loads are issued before the
// first store.  Needs MAXQ * nthr >= nr + 6.
template <int MAXQ>
__device__ __forceinline__ void stage_row(double* buf, const double* __restrict__ row, bool valid, int nr, int tid,
                                          int nthr) {
  double v[MAXQ];
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int i = tid + q * nthr - 2;
    v[q] = (valid && i >= 0 && i < nr) ? row[i] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int j = tid + q * nthr;
    if (j < nr + 6) buf[pad(j)] = v[q];
  }
}

// LDS row -> HBM: row[i] = x_i, or row[i] = x_i (-1) + row[i] with `subtract` (old values
// loaded first).  Needs MAXQ * nthr >= m.
template <int MAXQ>
__device__ __forceinline__ void unstage_row(const double* buf, double* __restrict__ row, bool valid, int m, int tid,
                                            int nthr, bool subtract) {
#pragma clang fp contract(off)
  if (!valid) return;
  double old[MAXQ];
  if (subtract) {
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
      const int i = tid + q * nthr;
      old[q] = i < m ? row[i] : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < MAXQ; ++q) {
    const int i = tid + q * nthr;
    if (i < m) {
      const double x = buf[pad(i + 2)];
      row[i] = subtract ? x * (-1.0) + old[q] : x;
    }
  }
}

// Fdma forward elimination (fdma.rs:101-118): y_t = x_t - y_{t-1} lo_t (lo_t = low[i - 2])
template <int SLICE>
__device__ __forceinline__ void fdma_forward(double (&v)[SLICE], const double (&lo)[SLICE], int lane, int t0,
                                             int cnt) {
#pragma clang fp contract(off)
  double A = 1.0, B = 0.0;  // slice map y_out = A y_in + B
#pragma unroll
  for (int q = 0; q < SLICE; ++q) {
    if (q < cnt) {
      if (t0 + q == 0) {
        A = 0.0;
        B = v[q];
      } else {
        A = -A * lo[q];
        B = v[q] - B * lo[q];
      }
    }
  }
  double pa = A, pb = B;  // inclusive scan, lower lanes first
  for (int off = 1; off < WAVE; off <<= 1) {
    const double qa = __shfl_up(pa, off), qb = __shfl_up(pb, off);
    if (lane >= off) {
      pb = pa * qb + pb;
      pa = pa * qa;
    }
  }
  double y = __shfl_up(pb, 1);
  if (lane == 0) y = 0.0;
#pragma unroll
  for (int q = 0; q < SLICE; ++q) {
    if (q < cnt) {
      y = (t0 + q == 0) ? v[q] : v[q] - y * lo[q];
      v[q] = y;
    }
  }
}

// Fdma back substitution: x_t = (y_t - x_{t+1} up1 - x_{t+2} up2) / dia over a chain of
// length len (g1, g2 zero where t + 1, t + 2 leave the chain)
template <int SLICE>
__device__ __forceinline__ void fdma_backward(double (&v)[SLICE], const double (&dg)[SLICE], const double (&g1)[SLICE],
                                              const double (&g2)[SLICE], int lane, int t0, int cnt, int len) {
#pragma clang fp contract(off)
  // slice map on the state s = (x_{t+1}, x_{t+2}):  s' = M s + w
  double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0, w0 = 0.0, w1 = 0.0;
#pragma unroll
  for (int q = SLICE - 1; q >= 0; --q) {
    if (q >= cnt) continue;
    const double d = dg[q];
    const double c0 = v[q] / d, c1 = -g1[q] / d, c2 = -g2[q] / d;
    const double n00 = c1 * m00 + c2 * m10, n01 = c1 * m01 + c2 * m11, nw0 = c1 * w0 + c2 * w1 + c0;
    m10 = m00;
    m11 = m01;
    w1 = w0;
    m00 = n00;
    m01 = n01;
    w0 = nw0;
  }
  double a00 = m00, a01 = m01, a10 = m10, a11 = m11, b0 = w0, b1 = w1;  // suffix scan
  for (int off = 1; off < WAVE; off <<= 1) {
    const double q00 = __shfl_down(a00, off), q01 = __shfl_down(a01, off), q10 = __shfl_down(a10, off),
                 q11 = __shfl_down(a11, off), r0 = __shfl_down(b0, off), r1 = __shfl_down(b1, off);
    if (lane + off < WAVE) {
      const double n00 = a00 * q00 + a01 * q10, n01 = a00 * q01 + a01 * q11;
      const double n10 = a10 * q00 + a11 * q10, n11 = a10 * q01 + a11 * q11;
      const double nb0 = a00 * r0 + a01 * r1 + b0, nb1 = a10 * r0 + a11 * r1 + b1;
      a00 = n00;
      a01 = n01;
      a10 = n10;
      a11 = n11;
      b0 = nb0;
      b1 = nb1;
    }
  }
  double s0 = __shfl_down(b0, 1), s1 = __shfl_down(b1, 1);
  if (lane == WAVE - 1) {
    s0 = 0.0;
    s1 = 0.0;
  }
#pragma unroll
  for (int q = SLICE - 1; q >= 0; --q) {
    if (q >= cnt) continue;
    const int t = t0 + q;
    double val = v[q];
    if (t + 1 < len) val = val - s0 * g1[q];
    if (t + 2 < len) val = val - s1 * g2[q];
    val = val / dg[q];
    s1 = s0;
    s0 = val;
    v[q] = val;
  }
}

// tdma normal equations (linalg.rs:14-57), forward: g_i = (d_i - g_{i-2} off_{i-2}) / den_i,
// g_i = d_i / den_i at the chain start (of_q = off_{i-2}, dn_q = den_i)
template <int SLICE>
__device__ __forceinline__ void tdma_forward(double (&v)[SLICE], const double (&of)[SLICE], const double (&dn)[SLICE],
                                             int lane, int t0, int cnt) {
#pragma clang fp contract(off)
  double A = 1.0, B = 0.0;
#pragma unroll
  for (int q = 0; q < SLICE; ++q) {
    if (q < cnt) {
      if (t0 + q == 0) {
        A = 0.0;
        B = v[q] / dn[q];
      } else {
        A = (-of[q] / dn[q]) * A;
        B = (v[q] - B * of[q]) / dn[q];
      }
    }
  }
  double pa = A, pb = B;
  for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
    const double qa = __shfl_up(pa, o2), qb = __shfl_up(pb, o2);
    if (lane >= o2) {
      pb = pa * qb + pb;
      pa = pa * qa;
    }
  }
  double g = __shfl_up(pb, 1);
  if (lane == 0) g = 0.0;
#pragma unroll
  for (int q = 0; q < SLICE; ++q) {
    if (q < cnt) {
      g = (t0 + q == 0) ? v[q] / dn[q] : (v[q] - g * of[q]) / dn[q];
      v[q] = g;
    }
  }
}

// tdma back substitution: x_i = g_i for i >= m - 2, else x_i = g_i - x_{i+2} w_i (chain p)
template <int SLICE>
__device__ __forceinline__ void tdma_backward(double (&v)[SLICE], const double (&wv)[SLICE], int lane, int t0, int cnt,
                                              int p, int m) {
#pragma clang fp contract(off)
  double A = 1.0, B = 0.0;
#pragma unroll
  for (int q = SLICE - 1; q >= 0; --q) {
    if (q < cnt) {
      if (p + 2 * (t0 + q) >= m - 2) {
        A = 0.0;
        B = v[q];
      } else {
        A = -wv[q] * A;
        B = v[q] - B * wv[q];
      }
    }
  }
  double pa = A, pb = B;
  for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
    const double qa = __shfl_down(pa, o2), qb = __shfl_down(pb, o2);
    if (lane + o2 < WAVE) {
      pb = pa * qb + pb;
      pa = pa * qa;
    }
  }
  double x = __shfl_down(pb, 1);
  if (lane == WAVE - 1) x = 0.0;
#pragma unroll
  for (int q = SLICE - 1; q >= 0; --q) {
    if (q < cnt) {
      x = (p + 2 * (t0 + q) >= m - 2) ? v[q] : v[q] - x * wv[q];
      v[q] = x;
    }
  }
}

// suffix sums along the chain: v_t <- v_t + v_{t+1} + ... (the derivative recurrence
// d_i = g_i + d_{i+2} of ortho.rs:107-125, summed from the top as the reference does)
template <int SLICE>
__device__ __forceinline__ void suffix_sum(double (&v)[SLICE], int lane, int cnt) {
#pragma clang fp contract(off)
  double tot = 0.0;
#pragma unroll
  for (int q = SLICE - 1; q >= 0; --q)
    if (q < cnt) tot = v[q] + tot;
  double inc = tot;  // inclusive suffix sum over lanes
  for (int off = 1; off < WAVE; off <<= 1) {
    const double o = __shfl_down(inc, off);
    if (lane + off < WAVE) inc = inc + o;
  }
  double acc = __shfl_down(inc, 1);
  if (lane == WAVE - 1) acc = 0.0;
#pragma unroll
  for (int q = SLICE - 1; q >= 0; --q) {
    if (q < cnt) {
      acc = v[q] + acc;
      v[q] = acc;
    }
  }
}

}  // namespace chains
}  // namespace rbc
