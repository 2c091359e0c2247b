// The pressure stage of the periodic step (navier.rs:683-721) fused per x-wavenumber, FP64.
//
// With a Fourier x basis every operation of the stage acts on the two spectral rows
// (real and imaginary plane) of one wavenumber k: the divergence
//     div = S_u(ik ux / sc) + d/dy S_u(uy)                          (navier.rs:698-703)
// the Poisson rows  P_y div -> pseu (Fdma of set k, fdma_tensor.rs:195-234), pseu(0,0) = 0,
// the projection    ux -= FO_u(S_p(ik pseu / sc)),  uy -= FO_u(d/dy S_p(pseu))
// and the update    pres = (pres - div nu) + S_p(pseu) / dt.
// One workgroup owns wavenumber k: four wavefronts, one per (plane, parity chain); the rows
// live in padded LDS buffers between the phases, so the stage reads ux and pres once (uy
// twice: its staged copy makes room for pseu) and writes ux, uy, pres, pseu once.  Every value is formed with the arithmetic of the unfused
// kernels it replaces (band stencils, the num-complex Fourier derivative, the derivative
// suffix sums, the Fdma and tdma recurrences of chains.hpp).
#include <hip/hip_runtime.h>

#include <cmath>
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

// Plane r of (a + ib)(0 + ik) / (sc + 0i) as num-complex forms it (fourier_diff, times = 1),
// from the other plane's value x: r = 0 gives (a 0 - x k) / sc (x = b), r = 1 gives
// (x k + b 0) / sc (x = a); the products with 0 only decide the sign of a zero result.
// isc = 1 / sc when sc is a power of two (else 0): (y sc + 0) / (sc sc) is then exactly y isc.
// P2: sc is known to be a power of two (isc != 0), which drops the division path.
template <bool P2>
__device__ __forceinline__ double fdiff_plane(double x, int r, double kk, double sc, double isc) {
#pragma clang fp contract(off)
  double v = r ? x * kk : 0.0 - x * kk;
  if (sc != 1.0) v = (P2 || isc != 0.0) ? v * isc : (v * sc) / (sc * sc + 0.0 * 0.0);
  return v;
}

// Stencil taps of a composite basis (composite_stencil.rs:117-162) at ortho index j, from
// the table coef (kind 0: d at coef[j], l at coef[n + j]) or formed in place for the
// Dirichlet (kind 1: d = 1, l = -1) and Neumann (kind 2: l = -(j-2)^2 / j^2) bases with
// the host's arithmetic, so the values are bitwise the table's.  m = n - 2 inputs.
__device__ __forceinline__ double tap_d(int kind, const double* __restrict__ coef, int j, int m) {
  return kind ? (j < m ? 1.0 : 0.0) : coef[j];
}
__device__ __forceinline__ double tap_l(int kind, const double* __restrict__ coef, int j, int n, int m) {
  if (kind == 0) return coef[n + j];
  if (j < 2 || j - 2 >= m) return 0.0;
  if (kind == 1) return -1.0;
  const long long i = j - 2;
  return -1.0 * (double)(i * i) / (double)((long long)j * j);
}

// RBC: the periodic Rayleigh-Benard bases (velocity Dirichlet, pseudo pressure Neumann)
// with a power-of-two x scale, fixed at compile time so that only their tap and scale code
// is emitted (the general kernel selects them at run time; the arithmetic is the same).
// Blocks from mx on solve the further field's rows (a.rows_*), which the stage does not touch:
// they fill the workgroup slots the mx stage blocks leave free.
template <int SLICE, bool RBC>
__global__ __launch_bounds__(256, 3) void pressure_rows_kernel(PressureArgs a) {
#pragma clang fp contract(off)
  const int su_kind = RBC ? 1 : a.su_kind, sp_kind = RBC ? 2 : a.sp_kind;
  extern __shared__ double lds[];
  if ((int)blockIdx.x >= a.mx) {  // the whole block
    rowsolve::fdma_block<SLICE>(lds, a.rows_R, a.rows_ldr, a.rows_nr, a.rows_X, a.rows_ld, 2 * a.rows_nsets,
                                a.rows_m, a.rows_nsets, a.rows_per, a.rows_bl, a.rows_band,
                                a.rows_diag, (int)blockIdx.x - a.mx);
    return;
  }
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
  const int r = w >> 1, p = w & 1;  // plane (0 real, 1 imaginary) and chain parity of this wavefront
  const int k = blockIdx.x, mx = a.mx, ny = a.ny, mu = a.mu, N = ny - 1, bl = a.bl;
  const double kk = (double)k;
  // LDS rows (index = plane): XA ux, later projection results; YB uy, later pseu; DV div
#define XA(pl) (lds + (pl) * bl)
#define YB(pl) (lds + (2 + (pl)) * bl)
#define DV(pl) (lds + (4 + (pl)) * bl)
#define ROWU(pl) ((size_t)((pl) ? mx + k : k) * mu)
#define ROWN(pl) ((size_t)((pl) ? mx + k : k) * ny)
  // stencil taps of the velocity (u) and pressure (p) bases
#define SU0(j) tap_d(su_kind, a.su, (j), mu)
#define SU1(j) tap_l(su_kind, a.su, (j), ny, mu)
#define SP0(j) tap_d(sp_kind, a.sp, (j), mu)
#define SP1(j) tap_l(sp_kind, a.sp, (j), ny, mu)

  // rows of up to 2 SLICE WAVE + 1 values: unrolled staging bounds (all loads before stores)
  constexpr int QW = (2 * SLICE * WAVE + 8 + WAVE - 1) / WAVE;  // per wavefront
  constexpr int QH = (2 * SLICE * WAVE + 8 + 127) / 128;       // per half block
  const int h = tid >> 7, th = tid & 127;                      // half block: plane h
  // (1) ux, uy rows -> LDS (element i at pad(i + 2), zero guards at i in [-2, 0), [mu, mu + 4)):
  //     wavefront w stages ux re, ux im, uy re, uy im; the zero guards of div around
  //     i in [0, ny) (read by the Poisson band) are set here
  chains::stage_row<QW>(lds + w * bl, (w < 2 ? a.UX : a.UY) + ROWU(w & 1), true, mu, lane, WAVE);
  if (tid < 12) {  // plane tid & 1; i = j - 2 in {-2, -1} (tid < 4) or [ny, ny + 4)
    const int j = tid < 4 ? (tid >> 1) : ny + 2 + ((tid - 4) >> 1);
    DV(tid & 1)[pad(j)] = 0.0;
  }
  chains::lds_barrier();
  // (2) div = S_u(ik ux / sc) + d/dy S_u(uy) for i < ny.  The derivative: g_i = 2 (i + 1) s_{i+1}
  //     summed from the top along the chains 1, 3, .., N - 1 and 2, 4, .., N, then
  //     d_0 = s_1 + d_2 / 2 (ortho.rs:107-125); each entry then adds the x part (taps i,
  //     i - 2 of plane r of ik ux / sc) as the unfused kernels add the two terms
  const int Cd = N / 2, perd = (Cd + WAVE - 1) / WAVE, t0d = lane * perd;
  const int cntd = max(0, min(perd, Cd - t0d));
  const int per = a.per, t0 = lane * per;
  // the Poisson phase's P_y taps and forward-elimination multipliers, issued here so that
  // their round trip overlaps the divergence (the P_y band of the pinv has no tap at -2:
  // the Rayleigh-Benard instantiation does not load it)
  const size_t tp = (size_t)2 * per * WAVE;
  const size_t sb = (size_t)k * (8 * (size_t)per * WAVE);
#define DG(arr, q) a.pdiag[sb + ((size_t)((arr) * 2 + p) * per + (q)) * WAVE + lane]
  double pc0[SLICE], pc1[SLICE], pc2[SLICE], pc3[SLICE], plo[SLICE];
  auto load_poisson = [&]() {
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      const size_t b = ((size_t)p * per + q) * WAVE + lane;
      pc0[q] = a.pybp[b];
      pc1[q] = RBC ? 0.0 : a.pybp[tp + b];
      pc2[q] = a.pybp[2 * tp + b];
      pc3[q] = a.pybp[3 * tp + b];
      plo[q] = DG(0, q);
    }
  };
  if constexpr (RBC) load_poisson();  // the general instantiation loads them in (3)
  {
    double v[SLICE];
    const double* Y = YB(r);
    const double* Xo = XA(1 - r);
    // x part of entry i: (ik ux / sc)_r at i and i - 2 with the stencil taps
    auto divx = [&](int i) {
      const double c0 = fdiff_plane<RBC>(Xo[pad(i + 2)], r, kk, a.kscale, a.isc);
      const double c2 = fdiff_plane<RBC>(Xo[pad(i)], r, kk, a.kscale, a.isc);
      return c0 * SU0(i) + c2 * SU1(i);
    };
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      v[q] = 0.0;
      if (q < cntd) {
        const int i = (p ? 1 : 2) + 2 * (t0d + q);
        if (i <= N - 1) {
          const double s = Y[pad(i + 3)] * SU0(i + 1) + Y[pad(i + 1)] * SU1(i + 1);
          v[q] = (2.0 * (double)(i + 1)) * s;
        }
      }
    }
    chains::suffix_sum(v, lane, cntd);
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      if (q < cntd) {
        const int i = (p ? 1 : 2) + 2 * (t0d + q);
        double d = v[q];
        if (a.inv_scy != 1.0) d = d * a.inv_scy;
        DV(r)[pad(i + 2)] = d + 1.0 * divx(i);
      }
    }
    if (p == 0 && lane == 0) {
      const double s1 = Y[pad(3)] * SU0(1) + Y[pad(1)] * SU1(1);
      double d = s1 + v[0] / 2.0;
      if (a.inv_scy != 1.0) d = d * a.inv_scy;
      DV(r)[pad(2)] = d + 1.0 * divx(0);
    }
  }
  chains::lds_barrier();
  // (3) pseu = Poisson rows of div (P_y taps, then the Fdma of set k); pseu(0,0) = 0
  const int lenu = p == 0 ? (mu + 1) / 2 : mu / 2;
  const int cnt = max(0, min(per, lenu - t0));
  if constexpr (!RBC) load_poisson();
  {
    double rd[SLICE], g1[SLICE], g2[SLICE];  // 1 / dia, up1, up2
    auto load_back = [&]() {
#pragma unroll
      for (int q = 0; q < SLICE; ++q) {
        rd[q] = DG(1, q);
        g1[q] = DG(2, q);
        g2[q] = DG(3, q);
      }
    };
    if constexpr (RBC) load_back();  // issued before the band and the forward pass
    double v[SLICE];
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      double val = 0.0;
      if (q < cnt) {
        const int j = p + 2 * (t0 + q) + 2;
        const double* D = DV(r);
        val = pc0[q] * D[pad(j)];
        if (!RBC) val = val + pc1[q] * D[pad(j - 2)];
        val = val + pc2[q] * D[pad(j + 2)];
        val = val + pc3[q] * D[pad(j + 4)];
      }
      v[q] = val;
    }
    chains::fdma_forward(v, plo, lane, t0, cnt);
    if constexpr (!RBC) load_back();
    chains::fdma_backward(v, rd, g1, g2, lane, t0, cnt, lenu);
#undef DG
    if (k == 0 && p == 0 && lane == 0) v[0] = 0.0;  // navier.rs:710-715
#pragma unroll
    for (int q = 0; q < SLICE; ++q)
      if (q < cnt) YB(r)[pad(p + 2 * (t0 + q) + 2)] = v[q];
  }
  chains::lds_barrier();
  // (4) pseu -> HBM; pres = (pres - div nu) + S_p(pseu) / dt (half block h: plane h); the new
  //     pres values stay in registers for (8)
  double pn[QH];
  {
    double pv[QH], c0[QH], c1[QH];
    double* prow = a.P0 + ROWN(h);
#pragma unroll
    for (int q = 0; q < QH; ++q) {
      const int i = th + q * 128;
      pv[q] = i < ny ? prow[i] : 0.0;
      c0[q] = i < ny ? SP0(i) : 0.0;
      c1[q] = i < ny ? SP1(i) : 0.0;
    }
    const double* Pp = YB(h);
#pragma unroll
    for (int q = 0; q < QH; ++q) {
      const int i = th + q * 128;
      pn[q] = 0.0;
      if (i >= ny) continue;
      if (i < mu) a.P1[ROWU(h) + i] = Pp[pad(i + 2)];
      const double w3 = Pp[pad(i + 2)] * c0[q] + Pp[pad(i)] * c1[q];
      const double t = pv[q] - DV(h)[pad(i + 2)] * a.nu;
      pn[q] = t + w3 * a.inv_dt;
      prow[i] = pn[q];
    }
  }
  chains::lds_barrier();
  // the from_ortho tables of the projections (1 / den, w), issued here so that their round
  // trip overlaps the derivative below
#define TB(arr, q) a.ttab[((size_t)((arr) * 2 + p) * per + (q)) * WAVE + lane]
  double trn[SLICE], twv[SLICE];
  auto load_tdma = [&]() {
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      trn[q] = TB(3, q);
      twv[q] = TB(4, q);
    }
  };
  if constexpr (RBC) load_tdma();  // the general instantiation loads them in (6)
  // (5) d/dy S_p(pseu) into DV (the chains of (2b); div is no longer needed)
  {
    double v[SLICE];
    const double* Pp = YB(r);
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      v[q] = 0.0;
      if (q < cntd) {
        const int i = (p ? 1 : 2) + 2 * (t0d + q);
        if (i <= N - 1) {
          const double s = Pp[pad(i + 3)] * SP0(i + 1) + Pp[pad(i + 1)] * SP1(i + 1);
          v[q] = (2.0 * (double)(i + 1)) * s;
        }
      }
    }
    chains::suffix_sum(v, lane, cntd);
#pragma unroll
    for (int q = 0; q < SLICE; ++q) {
      if (q < cntd) {
        const int i = (p ? 1 : 2) + 2 * (t0d + q);
        double d = v[q];
        if (a.inv_scy != 1.0) d = d * a.inv_scy;
        DV(r)[pad(i + 2)] = d;
      }
    }
    if (p == 0 && lane == 0) {
      const double s1 = Pp[pad(3)] * SP0(1) + Pp[pad(1)] * SP1(1);
      double d = s1 + v[0] / 2.0;
      if (a.inv_scy != 1.0) d = d * a.inv_scy;
      DV(r)[pad(2)] = d;
    }
  }
  chains::lds_barrier();
  // (6) both projections, from_ortho of g = S_p(ik pseu / sc) (ux) and of g = DV (uy):
  //     d_i = g_i diag_i + g_{i+2} low2_i, then the tdma recurrences with the tables shared
  //     by the two right-hand sides
  // the from_ortho tables of the velocity basis: diag, low2 and off_{i-2} = diag_i low2_{i-2}
  // are 1, -1, -1 for the Dirichlet basis (0 for off at the chain start)
#define TB_DG(q) (su_kind == 1 ? 1.0 : TB(0, q))
#define TB_L2(q) (su_kind == 1 ? -1.0 : TB(1, q))
#define TB_OF(q) (su_kind == 1 ? (t0 + (q) >= 1 ? -1.0 : 0.0) : TB(2, q))
  {
    double vx[SLICE], vy[SLICE];
    {
      double dg[SLICE], l2[SLICE];
#pragma unroll
      for (int q = 0; q < SLICE; ++q) {
        dg[q] = TB_DG(q);
        l2[q] = TB_L2(q);
      }
      // plane r of ik pseu / sc at the chain elements i0 - 2 + 2 j (j <= cnt + 1) and the
      // stencil taps at i0 + 2 j (j <= cnt), each formed once for the slice
      const int i0 = p + 2 * t0;
      const double* Xo = YB(1 - r);
      double cs[SLICE + 2], ta[SLICE + 1], tb[SLICE + 1];
#pragma unroll
      for (int j = 0; j < SLICE + 2; ++j)
        cs[j] = (cnt > 0 && j <= cnt + 1) ? fdiff_plane<RBC>(Xo[pad(i0 + 2 * j)], r, kk, a.kscale, a.isc) : 0.0;
#pragma unroll
      for (int j = 0; j < SLICE + 1; ++j) {
        ta[j] = (cnt > 0 && j <= cnt) ? SP0(i0 + 2 * j) : 0.0;
        tb[j] = (cnt > 0 && j <= cnt) ? SP1(i0 + 2 * j) : 0.0;
      }
#pragma unroll
      for (int q = 0; q < SLICE; ++q) {
        vx[q] = 0.0;
        vy[q] = 0.0;
        if (q < cnt) {
          const double g0 = cs[q + 1] * ta[q] + cs[q] * tb[q];
          const double g2 = cs[q + 2] * ta[q + 1] + cs[q + 1] * tb[q + 1];
          vx[q] = g0 * dg[q] + g2 * l2[q];
          const int j = p + 2 * (t0 + q) + 2;
          vy[q] = DV(r)[pad(j)] * dg[q] + DV(r)[pad(j + 2)] * l2[q];
        }
      }
    }
    {
      double of[SLICE];  // off_{i-2}; 1 / den_i and w_i were loaded before (5)
      if constexpr (!RBC) load_tdma();
#pragma unroll
      for (int q = 0; q < SLICE; ++q) of[q] = TB_OF(q);
      chains::tdma_forward2(vx, vy, of, trn, lane, t0, cnt);
      chains::tdma_backward2(vx, vy, twv, lane, t0, cnt, p, mu);
    }
    // ux = x (-1) + ux in place (XA still holds the staged ux); the uy projection replaces
    // the entries of DV that only this wavefront reads
#pragma unroll
    for (int q = 0; q < SLICE; ++q)
      if (q < cnt) {
        const int j = pad(p + 2 * (t0 + q) + 2);
        XA(r)[j] = vx[q] * (-1.0) + XA(r)[j];
        DV(r)[j] = vy[q];
      }
  }
#undef TB
#undef TB_DG
#undef TB_L2
#undef TB_OF
#undef SU0
#undef SU1
#undef SP0
#undef SP1
  chains::lds_barrier();
  // (7) the updated ux rows -> HBM; uy = x (-1) + uy (uy reloaded: its staged copy became pseu)
  chains::unstage_row<QH>(XA(h), a.UX + ROWU(h), true, mu, th, 128, false);
  chains::unstage_row<QH>(DV(h), a.UY + ROWU(h), true, mu, th, 128, true);
  if (a.dP) {
    // (8) d/dy of the updated pres (ortho.rs:107-125, the chains of (2) on the Chebyshev
    //     coefficients) for the next step's right-hand sides: pres rows (kept from (4)) ->
    //     YB (element i at pad(i + 2), zero guards as stage_row leaves them), derivative ->
    //     XA -> dP
    //     (the general instantiation reloads them: it has no registers to spare, and its
    //     barrier waits for the pres stores of (4))
    if constexpr (RBC) chains::lds_barrier();
    else __syncthreads();
    if constexpr (RBC) {
#pragma unroll
      for (int q = 0; q < QH; ++q) {
        const int i = th + q * 128;
        if (i < ny) YB(h)[pad(i + 2)] = pn[q];
      }
      if (th < 6) YB(h)[pad(th < 2 ? th : ny + th)] = 0.0;
    } else {
      chains::stage_row<QH>(YB(h), a.P0 + ROWN(h), true, ny, th, 128);
    }
    chains::lds_barrier();
    {
      double v[SLICE];
      const double* Y = YB(r);
#pragma unroll
      for (int q = 0; q < SLICE; ++q) {
        v[q] = 0.0;
        if (q < cntd) {
          const int i = (p ? 1 : 2) + 2 * (t0d + q);
          if (i <= N - 1) v[q] = (2.0 * (double)(i + 1)) * Y[pad(i + 3)];
        }
      }
      chains::suffix_sum(v, lane, cntd);
#pragma unroll
      for (int q = 0; q < SLICE; ++q) {
        if (q < cntd) {
          const int i = (p ? 1 : 2) + 2 * (t0d + q);
          double d = v[q];
          if (a.inv_scy != 1.0) d = d * a.inv_scy;
          XA(r)[pad(i + 2)] = d;
        }
      }
      if (p == 0 && lane == 0) {
        double d = Y[pad(3)] + v[0] / 2.0;
        if (a.inv_scy != 1.0) d = d * a.inv_scy;
        XA(r)[pad(2)] = d;
      }
    }
    chains::lds_barrier();
    chains::unstage_row<QH>(XA(h), a.dP + ROWN(h), true, ny, th, 128, false);
  }
#undef XA
#undef YB
#undef DV
#undef ROWU
#undef ROWN
}

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

bool pressure_rows_supported(int ny, int mu) {
  const int per = fdma_rows_lds_per(mu), perd = ((ny - 1) / 2 + WAVE - 1) / WAVE;
  return ny >= 9 && (ny & 1) == 1 && mu == ny - 2 && per <= 8 && perd <= 8 &&
         (size_t)6 * buf_len(ny) * sizeof(double) <= 65536;
}

void pressure_rows(hipStream_t s, const PressureArgs& args) {
  if (!pressure_rows_supported(args.ny, args.mu)) throw std::invalid_argument("pressure_rows: unsupported sizes");
  if (args.mx < 1) throw std::invalid_argument("pressure_rows: bad mx");
  PressureArgs a = args;
  a.per = fdma_rows_lds_per(a.mu);
  a.bl = buf_len(a.ny);
  int ex = 0;
  a.isc = (a.kscale > 0.0 && std::frexp(a.kscale, &ex) == 0.5) ? 1.0 / a.kscale : 0.0;
  const int perd = ((a.ny - 1) / 2 + WAVE - 1) / WAVE, sl = a.per > perd ? a.per : perd;
  const size_t lds = (size_t)6 * a.bl * sizeof(double);
  const bool rbc = a.su_kind == 1 && a.sp_kind == 2 && a.isc != 0.0;
  if (a.rows_nsets > 0) {  // the further rows: slices within the stage's SLICE, buffers within its LDS
    a.rows_per = fdma_rows_lds_per(a.rows_m);
    a.rows_bl = buf_len(a.rows_nr);
    if (!fdma_rows_lds_supported(a.rows_m, a.rows_nr) || (a.rows_per <= 4 ? 4 : 8) > (sl <= 4 ? 4 : 8) ||
        (size_t)(rowsolve::WAVES / 2) * a.rows_bl * sizeof(double) > lds || !a.rows_R || !a.rows_X || !a.rows_diag)
      throw std::invalid_argument("pressure_rows: unsupported further rows");
  } else {
    a.rows_nsets = 0;
  }
  const dim3 grid((unsigned)(a.mx + a.rows_nsets));
  if (sl <= 4) {
    if (rbc) RBC_LAUNCH((pressure_rows_kernel<4, true>), grid, dim3(256), lds, s, a);
    else RBC_LAUNCH((pressure_rows_kernel<4, false>), grid, dim3(256), lds, s, a);
  } else {
    if (rbc) RBC_LAUNCH((pressure_rows_kernel<8, true>), grid, dim3(256), lds, s, a);
    else RBC_LAUNCH((pressure_rows_kernel<8, false>), grid, dim3(256), lds, s, a);
  }
  check("pressure_rows");
}

}  // namespace rbc
