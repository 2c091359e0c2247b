// One workgroup of the banded row solve of the tensor solvers (fdma_tensor.rs:195-234 with
// the per-row Fdma of fdma.rs:101-118), FP64, optionally fused with the preconditioner
// matvec that precedes it (MatVecFdma along y, matvec.rs:172-193).  Shared by the row-solve
// kernel and by the pressure-stage launch, which solves a further field's rows in blocks of
// its own.
//
// The offsets -2, 0, 2, 4 of the 4-diagonal systems couple only indices of equal parity,
// so a row is two independent chains, each solved by one wavefront of the recurrences of
// chains.hpp.  The right-hand side row is loaded with unit stride into a padded LDS buffer
// and the slices are read from there (results go back the same way).  Diagonals and band
// taps are stored on the host in slice order, [array][parity][q][lane] per diagonal set,
// so lane-parallel loads are contiguous.  Absent entries are stored as zeros (band taps,
// off-diagonals beyond the chain end): adding their zero products leaves every sum as the
// reference forms it.
#pragma once
#include <hip/hip_runtime.h>

#include "chains.hpp"

namespace rbc {
namespace rowsolve {

using chains::pad;
using chains::WAVE;
constexpr int WAVES = 4;  // wavefronts per block: two row slots of two parity chains

// Block bx of a solve of nrows rows (R, X already offset to the batch entry): one wavefront
// per (row, parity chain), the two chains of a row sharing an LDS buffer of bl doubles at
// lds + slot * bl.  With nrows = 2 nsets block bx holds the rows bx and bx + nsets of set
// bx, else the rows 2 bx and 2 bx + 1.  Every thread of the block must call it.
template <int SLICE>
__device__ __forceinline__ void fdma_block(double* lds, const double* __restrict__ R, int ldr, int nr,
                                           double* __restrict__ X, int ld, int nrows, int m, int nsets, int per,
                                           int bl, const double* __restrict__ bandp,
                                           const double* __restrict__ diagp, int bx) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int slot = w >> 1, p = w & 1;  // row slot of the block, chain parity
  double* buf = lds + slot * bl;
  // rows r and r + nsets share a diagonal set: one block per set when nrows = 2 nsets
  int r, set;
  bool valid;
  if (nrows == 2 * nsets) {
    set = bx;
    r = set + slot * nsets;
    valid = true;
  } else {
    r = bx * (WAVES / 2) + slot;
    valid = r < nrows;
    set = valid ? r % nsets : 0;
  }
  // (1) row -> LDS: element i at pad(i + 2); i in [-2, 0) and [nr, nr + 4) read as zero
  constexpr int MAXQ = (2 * SLICE * WAVE + 8 + 2 * WAVE - 1) / (2 * WAVE);
  chains::stage_row<MAXQ>(buf, R + (size_t)r * ldr, valid, nr, p * WAVE + lane, 2 * WAVE);
  __syncthreads();
  const int len = p == 0 ? (m + 1) / 2 : m / 2;
  const int t0 = lane * per;
  const int cnt = max(0, min(per, len - t0));
  double v[SLICE];
  // The slice-ordered arrays carry SLICE * WAVE doubles of tail padding, so the loads of
  // all q < SLICE are in bounds and are issued unconditionally, ahead of their use.
  const size_t setbase = (size_t)set * (8 * (size_t)per * WAVE);
#define RBC_DG(a, q) diagp[setbase + ((size_t)((a) * 2 + p) * per + (q)) * WAVE + lane]
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
  // ---- forward elimination and back substitution (diagonal tables of this set)
  {
    double lo[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) lo[q] = RBC_DG(0, q);
    chains::fdma_forward(v, lo, lane, t0, cnt);
  }
  {
    double rd[SLICE], g1[SLICE], g2[SLICE];  // 1 / dia, up1, up2
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      rd[q] = RBC_DG(1, q);
      g1[q] = RBC_DG(2, q);
      g2[q] = RBC_DG(3, q);
    }
    chains::fdma_backward(v, rd, g1, g2, lane, t0, cnt, len);
  }
#undef RBC_DG
  // (3) slices -> LDS -> row, unit stride
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SLICE; ++q)
    if (q < cnt) buf[pad(p + 2 * (t0 + q) + 2)] = v[q];
  __syncthreads();
  chains::unstage_row<MAXQ>(buf, X + (size_t)r * ld, valid, m, p * WAVE + lane, 2 * WAVE, false);
}

}  // namespace rowsolve
}  // namespace rbc
