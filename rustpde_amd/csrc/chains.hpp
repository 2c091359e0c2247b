// Recurrences along one parity chain of a row, held by one wavefront: lane l owns chain
// positions t in [t0, t0 + cnt), t0 = l * per, in registers v[q] (q = t - t0).  Each pass is
// a linear recurrence: the lane composes the affine map of its slice, the maps are scanned
// across the wavefront with shuffles, and each lane replays its slice with the reference's
// arithmetic from its incoming state (divisions by a diagonal taken as products with its
// tabulated reciprocal).  Used by the row kernels and the fused pressure stage.
#pragma once
#include <hip/hip_runtime.h>

namespace rbc {
namespace chains {

constexpr int WAVE = 64;

// Workgroup barrier that orders LDS accesses only: the fences cover the local address
// space, so the wave's outstanding global loads (prefetched operands) and stores stay in
// flight across it (__syncthreads first waits for every outstanding global access of the
// wave).  For phases that exchange data between the threads of a workgroup through LDS
// only; data passed through global memory needs __syncthreads.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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

// Every pass splits the lane's slice into a lower half L (q < H) and an upper half U
// (q >= H), H = SLICE / 2.  Their maps are composed side by side, the lane map is the
// composition of the two, and after the wavefront scan the entry state of the second half
// in the direction of the recurrence is the first half's map applied to the lane's entry
// state, so both halves are replayed side by side: two independent dependency chains per
// lane instead of one.

// Fdma forward elimination (fdma.rs:101-118): y_t = x_t - y_{t-1} lo_t (lo_t = low[i - 2])
template <int SLICE>
__device__ __forceinline__ void fdma_forward(double (&v)[SLICE], const double (&lo)[SLICE], int lane, int t0,
                                             int cnt) {
#pragma clang fp contract(off)
  static_assert(SLICE >= 4 && SLICE % 2 == 0, "two half slices of at least two entries");
  constexpr int H = SLICE / 2;
  double A1 = 1.0, B1 = 0.0, A2 = 1.0, B2 = 0.0;  // half maps y_out = A y_in + B
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < cnt) {
      if (t0 + j == 0) {
        A1 = 0.0;
        B1 = v[j];
      } else {
        A1 = -A1 * lo[j];
        B1 = v[j] - B1 * lo[j];
      }
    }
    const int q = H + j;
    if (q < cnt) {
      A2 = -A2 * lo[q];
      B2 = v[q] - B2 * lo[q];
    }
  }
  double pa = A2 * A1, pb = A2 * B1 + B2;  // lane map, then inclusive scan, lower lanes first
  for (int off = 1; off < WAVE; off <<= 1) {
    const double qa = __shfl_up(pa, off), qb = __shfl_up(pb, off);
    if (lane >= off) {
      pb = pa * qb + pb;
      pa = pa * qa;
    }
  }
  double y1 = __shfl_up(pb, 1);
  if (lane == 0) y1 = 0.0;
  double y2 = A1 * y1 + B1;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < cnt) {
      y1 = (t0 + j == 0) ? v[j] : v[j] - y1 * lo[j];
      v[j] = y1;
    }
    const int q = H + j;
    if (q < cnt) {
      y2 = v[q] - y2 * lo[q];
      v[q] = y2;
    }
  }
}

// Fdma back substitution: x_t = (y_t - x_{t+1} up1 - x_{t+2} up2) / dia over a chain of
// length len (g1, g2 zero where t + 1, t + 2 leave the chain).  rd holds 1 / dia: the
// division becomes a product with the tabulated reciprocal (one rounding apart).
struct Map2 {  // s' = M s + w on the state s = (x_{t+1}, x_{t+2})
  double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0, w0 = 0.0, w1 = 0.0;
};
__device__ __forceinline__ void map2_step(Map2& a, double v, double r, double g1, double g2) {
#pragma clang fp contract(off)
  const double c0 = v * r, c1 = -(g1 * r), c2 = -(g2 * r);
  const double n00 = c1 * a.m00 + c2 * a.m10, n01 = c1 * a.m01 + c2 * a.m11, nw0 = c1 * a.w0 + c2 * a.w1 + c0;
  a.m10 = a.m00;
  a.m11 = a.m01;
  a.w1 = a.w0;
  a.m00 = n00;
  a.m01 = n01;
  a.w0 = nw0;
}
template <int SLICE>
__device__ __forceinline__ void fdma_backward(double (&v)[SLICE], const double (&rd)[SLICE], const double (&g1)[SLICE],
                                              const double (&g2)[SLICE], int lane, int t0, int cnt, int len) {
#pragma clang fp contract(off)
  constexpr int H = SLICE / 2;
  Map2 L, U;  // top-down: U first, then L
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) map2_step(L, v[j], rd[j], g1[j], g2[j]);
    if (H + j < cnt) map2_step(U, v[H + j], rd[H + j], g1[H + j], g2[H + j]);
  }
  // lane map M = M_L M_U, w = M_L w_U + w_L, then the suffix scan
  double a00 = L.m00 * U.m00 + L.m01 * U.m10, a01 = L.m00 * U.m01 + L.m01 * U.m11;
  double a10 = L.m10 * U.m00 + L.m11 * U.m10, a11 = L.m10 * U.m01 + L.m11 * U.m11;
  double b0 = L.m00 * U.w0 + L.m01 * U.w1 + L.w0, b1 = L.m10 * U.w0 + L.m11 * U.w1 + L.w1;
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
  double u0 = __shfl_down(b0, 1), u1 = __shfl_down(b1, 1);
  if (lane == WAVE - 1) {
    u0 = 0.0;
    u1 = 0.0;
  }
  double l0 = U.m00 * u0 + U.m01 * u1 + U.w0, l1 = U.m10 * u0 + U.m11 * u1 + U.w1;
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) {
      const int t = t0 + j;
      double val = v[j];
      if (t + 1 < len) val = val - l0 * g1[j];
      if (t + 2 < len) val = val - l1 * g2[j];
      val = val * rd[j];
      l1 = l0;
      l0 = val;
      v[j] = val;
    }
    const int q = H + j;
    if (q < cnt) {
      const int t = t0 + q;
      double val = v[q];
      if (t + 1 < len) val = val - u0 * g1[q];
      if (t + 2 < len) val = val - u1 * g2[q];
      val = val * rd[q];
      u1 = u0;
      u0 = val;
      v[q] = val;
    }
  }
}

// tdma normal equations (linalg.rs:14-57), forward: g_i = (d_i - g_{i-2} off_{i-2}) / den_i,
// g_i = d_i / den_i at the chain start (of_q = off_{i-2}, rn_q = 1 / den_i: the divisions
// become products with the tabulated reciprocal)
template <int SLICE>
__device__ __forceinline__ void tdma_forward(double (&v)[SLICE], const double (&of)[SLICE], const double (&rn)[SLICE],
                                             int lane, int t0, int cnt) {
#pragma clang fp contract(off)
  constexpr int H = SLICE / 2;
  double A1 = 1.0, B1 = 0.0, A2 = 1.0, B2 = 0.0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < cnt) {
      if (t0 + j == 0) {
        A1 = 0.0;
        B1 = v[j] * rn[j];
      } else {
        A1 = -(of[j] * rn[j]) * A1;
        B1 = (v[j] - B1 * of[j]) * rn[j];
      }
    }
    const int q = H + j;
    if (q < cnt) {
      A2 = -(of[q] * rn[q]) * A2;
      B2 = (v[q] - B2 * of[q]) * rn[q];
    }
  }
  double pa = A2 * A1, pb = A2 * B1 + B2;
  for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
    const double qa = __shfl_up(pa, o2), qb = __shfl_up(pb, o2);
    if (lane >= o2) {
      pb = pa * qb + pb;
      pa = pa * qa;
    }
  }
  double g1v = __shfl_up(pb, 1);
  if (lane == 0) g1v = 0.0;
  double g2v = A1 * g1v + B1;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < cnt) {
      g1v = (t0 + j == 0) ? v[j] * rn[j] : (v[j] - g1v * of[j]) * rn[j];
      v[j] = g1v;
    }
    const int q = H + j;
    if (q < cnt) {
      g2v = (v[q] - g2v * of[q]) * rn[q];
      v[q] = g2v;
    }
  }
}

// tdma back substitution: x_i = g_i for i >= m - 2, else x_i = g_i - x_{i+2} w_i (chain p)
template <int SLICE>
__device__ __forceinline__ void tdma_backward(double (&v)[SLICE], const double (&wv)[SLICE], int lane, int t0, int cnt,
                                              int p, int m) {
#pragma clang fp contract(off)
  constexpr int H = SLICE / 2;
  double A1 = 1.0, B1 = 0.0, A2 = 1.0, B2 = 0.0;  // L and U, top-down
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) {
      if (p + 2 * (t0 + j) >= m - 2) {
        A1 = 0.0;
        B1 = v[j];
      } else {
        A1 = -wv[j] * A1;
        B1 = v[j] - B1 * wv[j];
      }
    }
    const int q = H + j;
    if (q < cnt) {
      if (p + 2 * (t0 + q) >= m - 2) {
        A2 = 0.0;
        B2 = v[q];
      } else {
        A2 = -wv[q] * A2;
        B2 = v[q] - B2 * wv[q];
      }
    }
  }
  double pa = A1 * A2, pb = A1 * B2 + B1;
  for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
    const double qa = __shfl_down(pa, o2), qb = __shfl_down(pb, o2);
    if (lane + o2 < WAVE) {
      pb = pa * qb + pb;
      pa = pa * qa;
    }
  }
  double x2 = __shfl_down(pb, 1);
  if (lane == WAVE - 1) x2 = 0.0;
  double x1 = A2 * x2 + B2;
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) {
      x1 = (p + 2 * (t0 + j) >= m - 2) ? v[j] : v[j] - x1 * wv[j];
      v[j] = x1;
    }
    const int q = H + j;
    if (q < cnt) {
      x2 = (p + 2 * (t0 + q) >= m - 2) ? v[q] : v[q] - x2 * wv[q];
      v[q] = x2;
    }
  }
}

// The two tdma passes for two right-hand sides with the same tables at once (the x and y
// projections of the pressure stage): the maps' multipliers depend on the tables only, so
// they are composed and scanned once; each right-hand side keeps its own offset.  Every
// value is formed as by tdma_forward / tdma_backward.
template <int SLICE>
__device__ __forceinline__ void tdma_forward2(double (&v)[SLICE], double (&u)[SLICE], const double (&of)[SLICE],
                                              const double (&rn)[SLICE], int lane, int t0, int cnt) {
#pragma clang fp contract(off)
  constexpr int H = SLICE / 2;
  double A1 = 1.0, B1 = 0.0, C1 = 0.0, A2 = 1.0, B2 = 0.0, C2 = 0.0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < cnt) {
      if (t0 + j == 0) {
        A1 = 0.0;
        B1 = v[j] * rn[j];
        C1 = u[j] * rn[j];
      } else {
        A1 = -(of[j] * rn[j]) * A1;
        B1 = (v[j] - B1 * of[j]) * rn[j];
        C1 = (u[j] - C1 * of[j]) * rn[j];
      }
    }
    const int q = H + j;
    if (q < cnt) {
      A2 = -(of[q] * rn[q]) * A2;
      B2 = (v[q] - B2 * of[q]) * rn[q];
      C2 = (u[q] - C2 * of[q]) * rn[q];
    }
  }
  double pa = A2 * A1, pb = A2 * B1 + B2, pc = A2 * C1 + C2;
  for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
    const double qa = __shfl_up(pa, o2), qb = __shfl_up(pb, o2), qc = __shfl_up(pc, o2);
    if (lane >= o2) {
      pb = pa * qb + pb;
      pc = pa * qc + pc;
      pa = pa * qa;
    }
  }
  double g1v = __shfl_up(pb, 1), h1v = __shfl_up(pc, 1);
  if (lane == 0) {
    g1v = 0.0;
    h1v = 0.0;
  }
  double g2v = A1 * g1v + B1, h2v = A1 * h1v + C1;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < cnt) {
      g1v = (t0 + j == 0) ? v[j] * rn[j] : (v[j] - g1v * of[j]) * rn[j];
      h1v = (t0 + j == 0) ? u[j] * rn[j] : (u[j] - h1v * of[j]) * rn[j];
      v[j] = g1v;
      u[j] = h1v;
    }
    const int q = H + j;
    if (q < cnt) {
      g2v = (v[q] - g2v * of[q]) * rn[q];
      h2v = (u[q] - h2v * of[q]) * rn[q];
      v[q] = g2v;
      u[q] = h2v;
    }
  }
}

template <int SLICE>
__device__ __forceinline__ void tdma_backward2(double (&v)[SLICE], double (&u)[SLICE], const double (&wv)[SLICE],
                                               int lane, int t0, int cnt, int p, int m) {
#pragma clang fp contract(off)
  constexpr int H = SLICE / 2;
  double A1 = 1.0, B1 = 0.0, C1 = 0.0, A2 = 1.0, B2 = 0.0, C2 = 0.0;  // L and U, top-down
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) {
      if (p + 2 * (t0 + j) >= m - 2) {
        A1 = 0.0;
        B1 = v[j];
        C1 = u[j];
      } else {
        A1 = -wv[j] * A1;
        B1 = v[j] - B1 * wv[j];
        C1 = u[j] - C1 * wv[j];
      }
    }
    const int q = H + j;
    if (q < cnt) {
      if (p + 2 * (t0 + q) >= m - 2) {
        A2 = 0.0;
        B2 = v[q];
        C2 = u[q];
      } else {
        A2 = -wv[q] * A2;
        B2 = v[q] - B2 * wv[q];
        C2 = u[q] - C2 * wv[q];
      }
    }
  }
  double pa = A1 * A2, pb = A1 * B2 + B1, pc = A1 * C2 + C1;
  for (int o2 = 1; o2 < WAVE; o2 <<= 1) {
    const double qa = __shfl_down(pa, o2), qb = __shfl_down(pb, o2), qc = __shfl_down(pc, o2);
    if (lane + o2 < WAVE) {
      pb = pa * qb + pb;
      pc = pa * qc + pc;
      pa = pa * qa;
    }
  }
  double x2 = __shfl_down(pb, 1), y2 = __shfl_down(pc, 1);
  if (lane == WAVE - 1) {
    x2 = 0.0;
    y2 = 0.0;
  }
  double x1 = A2 * x2 + B2, y1 = A2 * y2 + C2;
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) {
      const bool top = p + 2 * (t0 + j) >= m - 2;
      x1 = top ? v[j] : v[j] - x1 * wv[j];
      y1 = top ? u[j] : u[j] - y1 * wv[j];
      v[j] = x1;
      u[j] = y1;
    }
    const int q = H + j;
    if (q < cnt) {
      const bool top = p + 2 * (t0 + q) >= m - 2;
      x2 = top ? v[q] : v[q] - x2 * wv[q];
      y2 = top ? u[q] : u[q] - y2 * wv[q];
      v[q] = x2;
      u[q] = y2;
    }
  }
}

// suffix sums along the chain: v_t <- v_t + v_{t+1} + ... (the derivative recurrence
// d_i = g_i + d_{i+2} of ortho.rs:107-125, summed from the top as the reference does)
template <int SLICE>
__device__ __forceinline__ void suffix_sum(double (&v)[SLICE], int lane, int cnt) {
#pragma clang fp contract(off)
  constexpr int H = SLICE / 2;
  double t1 = 0.0, t2 = 0.0;
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) t1 = v[j] + t1;
    if (H + j < cnt) t2 = v[H + j] + t2;
  }
  double inc = t1 + t2;  // inclusive suffix sum over lanes
  for (int off = 1; off < WAVE; off <<= 1) {
    const double o = __shfl_down(inc, off);
    if (lane + off < WAVE) inc = inc + o;
  }
  double a2 = __shfl_down(inc, 1);
  if (lane == WAVE - 1) a2 = 0.0;
  double a1 = t2 + a2;
#pragma unroll
  for (int j = H - 1; j >= 0; --j) {
    if (j < cnt) {
      a1 = v[j] + a1;
      v[j] = a1;
    }
    if (H + j < cnt) {
      a2 = v[H + j] + a2;
      v[H + j] = a2;
    }
  }
}

}  // namespace chains
}  // namespace rbc
