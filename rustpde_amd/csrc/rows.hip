// Banded row solves of the tensor solvers (fdma_tensor.rs:195-234 with the per-row Fdma
// of fdma.rs:101-118), FP64, optionally fused with the preconditioner matvec that
// precedes them (MatVecFdma along y, matvec.rs:172-193): the row-solve kernel around the
// block body of rowsolve.hpp, and the from_ortho row kernel of the composite bases.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "chains.hpp"
#include "ops.hpp"
#include "rowsolve.hpp"

namespace rbc {

namespace {

using chains::buf_len;
using chains::pad;
using chains::WAVE;
using rowsolve::WAVES;

// Tables of up to four batch entries (blockIdx.y), or one shared by all entries.
struct RowTables {
  const double* band[4];
  const double* diag[4];
  int shared;
};

// One block per set (nrows = 2 nsets) or per two rows, blockIdx.y the batch entry.
// With pairg > 0 (one block per set, entries 0 and 1 on the same tables) the grid is 1-D:
// blocks [0, 16 pairg) hold entries 0 and 1 interleaved in groups of 16, block b of a group
// taking set 8 g + (b mod 8) of entry (b / 8) mod 2, so that the two blocks of a set are
// dispatched 8 apart and land on the same XCD (workgroups go round-robin over the 8 XCDs):
// the second reads the set's tables from that XCD's L2.  Blocks from 16 pairg on take the
// further entries, nsets blocks each.
template <int SLICE>
__global__ __launch_bounds__(WAVE* WAVES, SLICE <= 8 ? 4 : 2) void fdma_rows_lds_kernel(const double* __restrict__ R, int ldr, int nr,
                                                                    double* __restrict__ X, int ld, int nrows, int m,
                                                                    int nsets, int per, int bl, RowTables tb, long sr,
                                                                    long sx, int pairg) {
  extern __shared__ double lds[];
  int entry = blockIdx.y, bx = blockIdx.x;
  if (pairg > 0) {
    const int pb = 16 * pairg;
    if (bx < pb) {
      entry = (bx >> 3) & 1;
      bx = (bx >> 4) * 8 + (bx & 7);
      if (bx >= nsets) return;  // the whole block: no barrier reached
    } else {
      bx -= pb;
      entry = 2 + bx / nsets;
      bx -= (entry - 2) * nsets;
    }
  }
  const int ti = tb.shared ? 0 : entry;
  rowsolve::fdma_block<SLICE>(lds, R + sr * entry, ldr, nr, X + sx * entry, ld, nrows, m, nsets, per, bl, tb.band[ti],
                              tb.diag[ti], bx);
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
  constexpr int MAXQ = (2 * SLICE * WAVE + 8 + 2 * WAVE - 1) / (2 * WAVE);
  chains::stage_row<MAXQ>(buf, R + (size_t)r * ldr, valid, nr, p * WAVE + lane, 2 * WAVE);
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
    double of[SLICE], rn[SLICE];  // off_{i-2}, 1 / den_i
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      of[q] = TB(2, q);
      rn[q] = TB(3, q);
    }
    chains::tdma_forward(v, of, rn, lane, t0, cnt);
  }
  {  // back: x_i = g_i for i >= m - 2, else x_i = g_i - x_{i+2} w_i
    double wv[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) wv[q] = TB(4, q);
    chains::tdma_backward(v, wv, lane, t0, cnt, p, m);
  }
#undef TB
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SLICE; ++q)
    if (q < cnt) buf[pad(p + 2 * (t0 + q) + 2)] = v[q];
  __syncthreads();
  chains::unstage_row<MAXQ>(buf, X + (size_t)r * ld, valid, m, p * WAVE + lane, 2 * WAVE, project != 0);
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
  // entries 0 and 1 on the same tables, one block per set: interleaved 1-D grid
  const bool paired = !tb.shared && batch >= 2 && nrows == 2 * nsets && tb.band[0] == tb.band[1] &&
                      tb.diag[0] == tb.diag[1];
  const int pairg = paired ? (nsets + 7) / 8 : 0;
  const dim3 grid = paired ? dim3((unsigned)(16 * pairg + (batch - 2) * nsets), 1) : dim3(blocks, batch),
             block(WAVE * WAVES);
  if (per <= 4)
    RBC_LAUNCH(fdma_rows_lds_kernel<4>, grid, block, lds, s, R, ldr, nr, X, ld, nrows, m, nsets, per, bl, tb,
                       sr, sx, pairg);
  else if (per <= 8)
    RBC_LAUNCH(fdma_rows_lds_kernel<8>, grid, block, lds, s, R, ldr, nr, X, ld, nrows, m, nsets, per, bl, tb,
                       sr, sx, pairg);
  else
    RBC_LAUNCH(fdma_rows_lds_kernel<16>, grid, block, lds, s, R, ldr, nr, X, ld, nrows, m, nsets, per, bl, tb,
                       sr, sx, pairg);
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
