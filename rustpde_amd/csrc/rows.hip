// Banded row solves of the tensor solvers (fdma_tensor.rs:195-234 with the per-row Fdma
// of fdma.rs:101-118), FP64, optionally fused with the preconditioner matvec that
// precedes them (MatVecFdma along y, matvec.rs:172-193).
//
// The offsets -2, 0, 2, 4 of the 4-diagonal systems couple only indices of equal parity,
// so a row is two independent chains, each solved by one wavefront; lane l owns chain
// positions [l per, (l + 1) per).  A pass over a chain is a linear recurrence:
// each lane composes the affine map of its slice, the maps are scanned across the
// wavefront with shuffles, and each lane replays its slice with the reference's
// arithmetic from its incoming state.
//
// Memory layout for coalescing: the right-hand side row is loaded with unit stride into
// a padded LDS buffer of the wavefront and the slices are read from there (results go
// back the same way).  Diagonals and band taps are stored on the host in slice order,
// [array][parity][q][lane] per diagonal set, so lane-parallel loads are contiguous.
// Absent entries are stored as zeros (band taps, off-diagonals beyond the chain end):
// adding their zero products leaves every sum as the reference forms it.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "ops.hpp"

namespace rbc {

namespace {

constexpr int WAVE = 64;
constexpr int WAVES = 4;  // wavefronts (rows) per block

__host__ __device__ __forceinline__ int pad(int j) { return j + (j >> 4); }

// LDS doubles per wavefront for rows of nr values (two guard zeros on each side)
inline int buf_len(int nr) { return pad(nr + 6) + 1; }

// Tables of up to four batch entries (blockIdx.y), or one shared by all entries.
struct RowTables {
  const double* band[4];
  const double* diag[4];
  int shared;
};

// One wavefront per (row, parity chain); the two chains of a row share its LDS buffer.
template <int SLICE>
__global__ __launch_bounds__(WAVE* WAVES) void fdma_rows_lds_kernel(const double* __restrict__ R, int ldr, int nr,
                                                                    double* __restrict__ X, int ld, int nrows, int m,
                                                                    int nsets, int per, int bl, RowTables tb, long sr,
                                                                    long sx) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int slot = w >> 1, p = w & 1;  // row slot of the block, chain parity
  const int ti = tb.shared ? 0 : (int)blockIdx.y;
  const double* __restrict__ bandp = tb.band[ti];
  const double* __restrict__ diagp = tb.diag[ti];
  R += sr * blockIdx.y;
  X += sx * blockIdx.y;
  double* buf = lds + slot * bl;
  // rows r and r + nsets share a diagonal set: one block per set when nrows = 2 nsets
  int r, set;
  bool valid;
  if (nrows == 2 * nsets) {
    set = blockIdx.x;
    r = set + slot * nsets;
    valid = true;
  } else {
    r = blockIdx.x * (WAVES / 2) + slot;
    valid = r < nrows;
    set = valid ? r % nsets : 0;
  }
  // (1) row -> LDS: element i at pad(i + 2); i in [-2, 0) and [nr, nr + 4) read as zero
  for (int j = p * WAVE + lane; j < nr + 6; j += 2 * WAVE) {
    const int i = j - 2;
    buf[pad(j)] = (valid && i >= 0 && i < nr) ? R[(size_t)r * ldr + i] : 0.0;
  }
  __syncthreads();
  const int len = p == 0 ? (m + 1) / 2 : m / 2;
  const int t0 = lane * per;
  const int cnt = max(0, min(per, len - t0));
  double v[SLICE];
  // The slice-ordered arrays carry SLICE * WAVE doubles of tail padding, so the loads of
  // all q < SLICE are in bounds and are issued unconditionally, ahead of their use.
  const size_t setbase = (size_t)set * (8 * (size_t)per * WAVE);
#define DG(a, q) diagp[setbase + ((size_t)((a) * 2 + p) * per + (q)) * WAVE + lane]
  if (bandp) {  // taps at offsets 0, -2, +2, +4, summed in that order
    const size_t tp = (size_t)2 * per * WAVE;
    double c[4][SLICE];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int q = 0; q < SLICE; ++q) c[k][q] = bandp[k * tp + ((size_t)p * per + q) * WAVE + lane];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      double val = 0.0;
      if (q < cnt) {
        const int j = p + 2 * (t0 + q) + 2;
        val = c[0][q] * buf[pad(j)];
        val = val + c[1][q] * buf[pad(j - 2)];
        val = val + c[2][q] * buf[pad(j + 2)];
        val = val + c[3][q] * buf[pad(j + 4)];
      }
      v[q] = val;
    }
  } else {
#pragma unroll
    for (int q = 0; q < SLICE; ++q) v[q] = q < cnt ? buf[pad(p + 2 * (t0 + q) + 2)] : 0.0;
  }
  // ---- forward elimination: y_t = x_t - y_{t-1} low[i-2]  (i = p + 2t >= 2)
  {
    double lo[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) lo[q] = DG(0, q);
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
  // ---- back substitution: x_t = (y_t - x_{t+1} up1[i] - x_{t+2} up2[i]) / dia[i]
  {
    double dg[SLICE], g1[SLICE], g2[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      dg[q] = DG(1, q);
      g1[q] = DG(2, q);
      g2[q] = DG(3, q);
    }
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
#undef DG
  // (3) slices -> LDS -> row, unit stride
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SLICE; ++q)
    if (q < cnt) buf[pad(p + 2 * (t0 + q) + 2)] = v[q];
  __syncthreads();
  if (valid)
    for (int i = p * WAVE + lane; i < m; i += 2 * WAVE) X[(size_t)r * ld + i] = buf[pad(i + 2)];
}

// from_ortho of a composite basis per row (composite_stencil.rs:250-276, linalg.rs:14-57):
// d_i = p_i diag_i + p_{i+2} low2_i, then g_i = (d_i - g_{i-2} off_{i-2}) / den_i and
// x_i = g_i - x_{i+2} w_i (x_i = g_i for the last two i), arithmetic as tdma_y.  One
// wavefront per (row, parity chain).  With `project` the result is subtracted from X as
// the velocity projection does (navier.rs:683-695): X = x (-1) + X.
template <int SLICE>
__global__ __launch_bounds__(WAVE* WAVES) void tdma_rows_lds_kernel(const double* __restrict__ R, int ldr,
                                                                    double* __restrict__ X, int ld, int rows, int m,
                                                                    int per, int bl, const double* __restrict__ tab,
                                                                    long sr, long sx, int project) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int slot = w >> 1, p = w & 1;
  R += sr * blockIdx.y;
  X += sx * blockIdx.y;
  double* buf = lds + slot * bl;
  const int r = blockIdx.x * (WAVES / 2) + slot, nr = m + 2;
  const bool valid = r < rows;
  for (int j = p * WAVE + lane; j < nr + 6; j += 2 * WAVE) {
    const int i = j - 2;
    buf[pad(j)] = (valid && i >= 0 && i < nr) ? R[(size_t)r * ldr + i] : 0.0;
  }
  __syncthreads();
  const int len = p == 0 ? (m + 1) / 2 : m / 2;
  const int t0 = lane * per;
  const int cnt = max(0, min(per, len - t0));
  double v[SLICE];
#define TB(a, q) tab[((size_t)((a) * 2 + p) * per + (q)) * WAVE + lane]
  {
    double dg[SLICE], l2[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      dg[q] = TB(0, q);
      l2[q] = TB(1, q);
    }
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      const int j = p + 2 * (t0 + q) + 2;
      v[q] = q < cnt ? buf[pad(j)] * dg[q] + buf[pad(j + 2)] * l2[q] : 0.0;
    }
  }
  {  // forward: g_i = (d_i - g_{i-2} off_{i-2}) / den_i  (i >= 2), g_i = d_i / den_i (i < 2)
    double of[SLICE], dn[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      of[q] = TB(2, q);
      dn[q] = TB(3, q);
    }
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
  {  // back: x_i = g_i for i >= m - 2, else x_i = g_i - x_{i+2} w_i
    double wv[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) wv[q] = TB(4, q);
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
#undef TB
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SLICE; ++q)
    if (q < cnt) buf[pad(p + 2 * (t0 + q) + 2)] = v[q];
  __syncthreads();
  if (valid)
    for (int i = p * WAVE + lane; i < m; i += 2 * WAVE) {
      double* o = X + (size_t)r * ld + i;
      const double x = buf[pad(i + 2)];
      *o = project ? x * (-1.0) + *o : x;
    }
}

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

int fdma_rows_lds_per(int m) { return ((m + 1) / 2 + WAVE - 1) / WAVE; }

bool fdma_rows_lds_supported(int m, int nr) {
  return m >= 5 && fdma_rows_lds_per(m) <= 16 && nr >= m &&
         (size_t)(WAVES / 2) * buf_len(nr) * sizeof(double) <= 65536;
}

static void launch_rows(hipStream_t s, const double* R, int ldr, int nr, double* X, int ld, int nrows, int m,
                        int nsets, const RowTables& tb, int batch, long sr, long sx) {
  if (!fdma_rows_lds_supported(m, nr)) throw std::invalid_argument("fdma_rows_lds: unsupported row length");
  if (nsets < 1 || nrows < 1 || batch < 1 || batch > 65535) throw std::invalid_argument("fdma_rows_lds: bad sizes");
  if (!tb.shared && batch > 4) throw std::invalid_argument("fdma_rows_lds: at most 4 entries with own tables");
  for (int k = 0; k < (tb.shared ? 1 : batch); ++k)
    if (!tb.band[k] && nr != m) throw std::invalid_argument("fdma_rows_lds: rhs length must equal m without a band");
  const int per = fdma_rows_lds_per(m), bl = buf_len(nr);
  const unsigned blocks = nrows == 2 * nsets ? (unsigned)nsets : (unsigned)((nrows + 1) / 2);
  const size_t lds = (size_t)(WAVES / 2) * bl * sizeof(double);
  const dim3 grid(blocks, batch), block(WAVE * WAVES);
  if (per <= 4)
    RBC_LAUNCH(fdma_rows_lds_kernel<4>, grid, block, lds, s, R, ldr, nr, X, ld, nrows, m, nsets, per, bl, tb,
                       sr, sx);
  else if (per <= 8)
    RBC_LAUNCH(fdma_rows_lds_kernel<8>, grid, block, lds, s, R, ldr, nr, X, ld, nrows, m, nsets, per, bl, tb,
                       sr, sx);
  else
    RBC_LAUNCH(fdma_rows_lds_kernel<16>, grid, block, lds, s, R, ldr, nr, X, ld, nrows, m, nsets, per, bl, tb,
                       sr, sx);
  check("fdma_rows_lds");
}

void fdma_rows_lds(hipStream_t s, const double* R, int ldr, int nr, double* X, int ld, int nrows, int m, int nsets,
                   const double* bandp, const double* diagp, int batch, long sr, long sx) {
  RowTables tb{};
  tb.band[0] = bandp;
  tb.diag[0] = diagp;
  tb.shared = 1;
  launch_rows(s, R, ldr, nr, X, ld, nrows, m, nsets, tb, batch, sr, sx);
}

void fdma_rows_lds_multi(hipStream_t s, const double* R, int ldr, int nr, double* X, int ld, int nrows, int m,
                         int nsets, const double* const* bandps, const double* const* diagps, int batch, long sr,
                         long sx) {
  if (batch < 1 || batch > 4) throw std::invalid_argument("fdma_rows_lds_multi: 1 to 4 entries");
  RowTables tb{};
  for (int k = 0; k < batch; ++k) {
    tb.band[k] = bandps[k];
    tb.diag[k] = diagps[k];
  }
  tb.shared = 0;
  launch_rows(s, R, ldr, nr, X, ld, nrows, m, nsets, tb, batch, sr, sx);
}

bool tdma_rows_lds_supported(int m) {
  return m >= 4 && fdma_rows_lds_per(m) <= 16 && (size_t)(WAVES / 2) * buf_len(m + 2) * sizeof(double) <= 65536;
}

void tdma_rows_lds(hipStream_t s, const double* R, int ldr, double* X, int ld, int rows, int m, const double* tabp,
                   int batch, long sr, long sx, bool project) {
  if (!tdma_rows_lds_supported(m)) throw std::invalid_argument("tdma_rows_lds: unsupported row length");
  if (rows < 1 || batch < 1 || batch > 65535) throw std::invalid_argument("tdma_rows_lds: bad sizes");
  const int per = fdma_rows_lds_per(m), bl = buf_len(m + 2), pj = project ? 1 : 0;
  const dim3 grid((unsigned)((rows + 1) / 2), batch), block(WAVE * WAVES);
  const size_t lds = (size_t)(WAVES / 2) * bl * sizeof(double);
  if (per <= 4)
    RBC_LAUNCH(tdma_rows_lds_kernel<4>, grid, block, lds, s, R, ldr, X, ld, rows, m, per, bl, tabp, sr, sx, pj);
  else if (per <= 8)
    RBC_LAUNCH(tdma_rows_lds_kernel<8>, grid, block, lds, s, R, ldr, X, ld, rows, m, per, bl, tabp, sr, sx, pj);
  else
    RBC_LAUNCH(tdma_rows_lds_kernel<16>, grid, block, lds, s, R, ldr, X, ld, rows, m, per, bl, tabp, sr, sx,
                       pj);
  check("tdma_rows_lds");
}

}  // namespace rbc
