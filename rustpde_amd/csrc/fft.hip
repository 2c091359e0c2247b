// Batched real FFTs along axis 0 (the strided x axis) for the Fourier (fourier_r2c)
// direction of the periodic step, FP64.
//
// Two real columns a, b of length n (a power of two) go through one complex DFT of
// z = a + i b (radix-8 Stockham passes in LDS, fft_lds.hpp); their spectra follow from
// Z by A_k = (Z_k + conj Z_{n-k}) / 2, B_k = (Z_k - conj Z_{n-k}) / (2i).  The inverse
// builds Z_k = A_k + i B_k from the two half spectra (Hermitian extension) and runs the
// forward DFT on conj Z.  A workgroup owns C adjacent columns; its loads and stores are
// row segments of C doubles, and blocks are remapped so that neighbouring column groups
// share an XCD.  Spectral data are split planes (rows [0, n/2] real parts, rows
// [n/2 + 1, n + 1] imaginary parts).
//   r2c: X_k = sum_j x_j e^{-2 pi i jk/n}, k = 0..n/2   (r2c.rs:160-175, unnormalised)
//   c2r: x_j = 1/n sum_k X_k e^{+2 pi i jk/n}         (r2c.rs:211-226), Im X_0 and Im X_{n/2}
//        are ignored (half-spectrum convention); an optional (i k)^d / sc is applied to
//        X_k first (r2c.rs:88-99 with the num-complex division of space2.rs:256-262).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "fft_lds.hpp"
#include "ops.hpp"

namespace rbc {

namespace {

using fftlds::lds_dft;
using fftlds::lp;
using fftlds::lp_size;
using fftlds::xcd_remap;

// geometry for length L and LDS value size VB: PP column pairs per block, E = PP L complex
// values, T = E / 8.  Blocks of L >= 512 take 8 columns (row segments of 64 bytes in FP64,
// 32 in FP32); shorter lengths take 2048 / L column pairs.
template <int L, int VB = 8>
struct CGeo {
  static constexpr int PP = L >= 2048 ? 1 : (L >= 512 ? 4096 : 2048) / L;
  static constexpr int E = PP * L;
  static constexpr int T = E / 8;
  static constexpr int C = 2 * PP;  // columns per block
  static constexpr int N = L / 2;
  static constexpr int m = N + 1;  // half-spectrum length
};

// Tin: input type, V: arithmetic (and LDS) type, Tout: output type.  Columns j >= nzc are
// zero on input (the rows a y-dealiasing cut cleared): a block whose columns all lie there
// writes their zero spectra without loading or transforming.
template <int L, typename Tin, typename V, typename Tout>
__global__ __launch_bounds__((CGeo<L, sizeof(V)>::T)) void rfft_cols_kernel(const Tin* __restrict__ X, Tout* __restrict__ Y,
                                                               int cols, int cut, const V* __restrict__ tw,
                                                               long sx, long sy, int nzc) {
  using G = CGeo<L, sizeof(V)>;
  constexpr int C = G::C, T = G::T, m = G::m;
  constexpr int S = fftlds::spread_stride(L), NE = fftlds::spread_size(L, G::PP);  // transform p at p S
  __shared__ __attribute__((aligned(16))) V zmem[2 * NE];
  const fftlds::LdsZ<V, NE> z{zmem};
  X += sx * blockIdx.y;
  Y += sy * blockIdx.y;
  const int tid = threadIdx.x;
  const int j0 = xcd_remap(blockIdx.x, gridDim.x) * C;
  if (j0 >= nzc) {
    for (int e = tid; e < m * C; e += T) {
      const int c = e % C, k = e / C, j = j0 + c;
      if (j >= cols) continue;
      Y[(size_t)k * cols + j] = (Tout)0;
      Y[(size_t)(m + k) * cols + j] = (Tout)0;
    }
    return;
  }
  // load: row t, column c -> z_t of pair c / 2 (real part for even c, imaginary for odd)
  constexpr int QL = L * C / T;
  V v[QL];
#pragma unroll
  for (int q = 0; q < QL; ++q) {
    const int e = tid + q * T, c = e % C, t = e / C, j = j0 + c;
    v[q] = j < cols ? (V)X[(size_t)t * cols + j] : (V)0;
  }
#pragma unroll
  for (int q = 0; q < QL; ++q) {
    const int e = tid + q * T, c = e % C, t = e / C;
    z.c(c & 1, (c >> 1) * S + lp(t)) = v[q];
  }
  __syncthreads();
  lds_dft<L, T, G::E, S>(z, tw);
  // split: A_k = (Z_k + conj Z_{L-k}) / 2, B_k = (Z_k - conj Z_{L-k}) / (2i)
  for (int e = tid; e < m * C; e += T) {
    const int c = e % C, k = e / C, j = j0 + c;
    if (j >= cols) continue;
    const int p = (c >> 1) * S, k2 = (L - k) & (L - 1);
    V zr, zi, yr, yi;
    z.ld(p + lp(k), zr, zi);
    z.ld(p + lp(k2), yr, yi);
    const V half = (V)0.5;
    V xr, xi;
    if (c & 1) {
      xr = half * (zi + yi);
      xi = half * (yr - zr);
    } else {
      xr = half * (zr + yr);
      xi = half * (zi - yi);
    }
    if (k >= cut) {
      xr = 0;
      xi = 0;
    }
    Y[(size_t)k * cols + j] = (Tout)xr;
    Y[(size_t)(m + k) * cols + j] = (Tout)xi;
  }
}

// The work of one irfft block on columns j0 .. j0 + C - 1; FULL: all of them lie below cols
// (every block but the last), so the column guards drop out.
template <bool FULL, int L, typename Tin, typename V, typename Tout, int S, typename ZB>
__device__ __forceinline__ void irfft_block(const ZB& z, const Tin* __restrict__ X, Tout* __restrict__ Y, int cols,
                                            int times, double sc_d, const V* __restrict__ tw, int j0) {
#pragma clang fp contract(off)
  using G = CGeo<L, sizeof(V)>;
  constexpr int C = G::C, T = G::T, N = G::N, m = G::m;
  const V sc = (V)sc_d;
  // sc a power of two (and normal): the complex division by the real sc is a product
  const bool pow2 = sc_d > 0.0 && (__double_as_longlong(sc_d) & 0xFFFFFFFFFFFFFLL) == 0;
  const V isc = (V)(1.0 / sc_d);
  const int tid = threadIdx.x;
  // (1) half spectra with (i k)^times / sc applied; the two columns of a pair (lanes e and
  //     e ^ 1: C is even) exchange theirs by a lane swap and write conj Z of the pair:
  //     conj Z_k = (Ar - Bi) - i (Ai + Br), conj Z_{L-k} = (Ar + Bi) + i (Ai - Br)
  //     (Im A_0, Im A_N ignored: the half-spectrum convention)
  constexpr int QS = (m * C + T - 1) / T;
  V va[QS], vb[QS];
  // every load is issued, from element 0 of each plane where the entry lies outside (a
  // guarded load would become a branch that waits for its own round trip)
#pragma unroll
  for (int q = 0; q < QS; ++q) {
    const int e = tid + q * T, c = e % C, k = e / C, j = j0 + c;
    const bool ok = e < m * C && (FULL || j < cols);
    const size_t ia = ok ? (size_t)k * cols + j : 0;
    const Tin xa = X[ia], xb = X[ia + (size_t)m * cols];
    va[q] = ok ? (V)xa : (V)0;
    vb[q] = ok ? (V)xb : (V)0;
  }
#pragma unroll
  for (int q = 0; q < QS; ++q) {
    const int e = tid + q * T, c = e % C, k = e / C;
    V a = va[q], b = vb[q];
    const V kk = (V)k, zero = 0;
    for (int r = 0; r < times; ++r) {
      const V na = a * zero - b * kk, nb = a * kk + b * zero;
      a = na;
      b = nb;
    }
    if (sc != (V)1) {
      if (pow2) {  // (a sc + 0) / (sc sc) is exactly a / sc = a isc for a power-of-two sc
        a = a * isc;
        b = b * isc;
      } else {
        const V ns = sc * sc + zero * zero;
        const V na = (a * sc + b * zero) / ns, nb = (b * sc - a * zero) / ns;
        a = na;
        b = nb;
      }
    }
    if (k == 0 || k == N) b = 0;
    const V pa = __shfl_xor(a, 1), pb = __shfl_xor(b, 1);  // every lane takes part
    if (e >= m * C) continue;
    const bool odd = (c & 1) != 0;
    const V ar = odd ? pa : a, ai = odd ? pb : b, br = odd ? a : pa, bi = odd ? b : pb;
    const int zb = (c >> 1) * S;
    if (!odd) {
      z.st(zb + lp(k), ar - bi, -(ai + br));
    } else if (k > 0 && k < N) {
      z.st(zb + lp(L - k), ar + bi, ai - br);
    }
  }
  __syncthreads();
  lds_dft<L, T, G::E, S>(z, tw);
  // (3) a_j = Re W_j / L, b_j = -Im W_j / L with W = DFT(conj Z)
  const V inv = (V)(1.0 / (double)L);
  for (int e = tid; e < L * C; e += T) {
    const int c = e % C, t = e / C, j = j0 + c;
    if (!FULL && j >= cols) continue;
    const int p = (c >> 1) * S;
    Y[(size_t)t * cols + j] = (Tout)((c & 1) ? -z.i(p + lp(t)) * inv : z.r(p + lp(t)) * inv);
  }
}

// Batch entries b < split (all of them when split = 0) apply (i k)^times / sc and write Y;
// entries b >= split read input b - split again, apply (i k)^times1 / sc1 and write Y1 (the
// values and the x-derivatives of the same fields in one launch).
template <int L, typename Tin, typename V, typename Tout>
__global__ __launch_bounds__((CGeo<L, sizeof(V)>::T)) void irfft_cols_kernel(const Tin* __restrict__ X, Tout* __restrict__ Y,
                                                                int cols, int times, double sc_d,
                                                                const V* __restrict__ tw, long sx, long sy,
                                                                Tout* __restrict__ Y1, int split, int times1,
                                                                double sc1) {
  using G = CGeo<L, sizeof(V)>;
  constexpr int C = G::C;
  constexpr int S = fftlds::spread_stride(L), NE = fftlds::spread_size(L, G::PP);  // transform p at p S
  __shared__ __attribute__((aligned(16))) V zmem[2 * NE];
  const fftlds::LdsZ<V, NE> z{zmem};
  int b = blockIdx.y;
  if (split > 0 && b >= split) {
    b -= split;
    Y = Y1;
    times = times1;
    sc_d = sc1;
  }
  X += sx * b;
  Y += sy * b;
  const int j0 = xcd_remap(blockIdx.x, gridDim.x) * C;
  if (j0 + C <= cols) irfft_block<true, L, Tin, V, Tout, S>(z, X, Y, cols, times, sc_d, tw, j0);
  else irfft_block<false, L, Tin, V, Tout, S>(z, X, Y, cols, times, sc_d, tw, j0);
}

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void check_n(int n) {
  if (n < 8 || (n & (n - 1)) != 0) throw std::invalid_argument("fft: length must be a power of two >= 8");
  if (n > 2048) throw std::invalid_argument("fft: length above 2048 exceeds the LDS budget");
}

template <int L, typename Tin, typename V, typename Tout>
void launch_rfft(hipStream_t s, const Tin* X, Tout* Y, int cols, int cut, int batch, long sx, long sy, int nzc) {
  using G = CGeo<L, sizeof(V)>;
  const unsigned blocks = (unsigned)((cols + G::C - 1) / G::C);
  RBC_LAUNCH((rfft_cols_kernel<L, Tin, V, Tout>), dim3(blocks, batch), dim3(G::T), 0, s, X, Y, cols, cut,
             twiddles<V>(L), sx, sy, nzc < 0 ? cols : nzc);
}
template <int L, typename Tin, typename V, typename Tout>
void launch_irfft(hipStream_t s, const Tin* X, Tout* Y, int cols, int times, double sc, int batch, long sx, long sy,
                  Tout* Y1 = nullptr, int split = 0, int times1 = 0, double sc1 = 1.0) {
  using G = CGeo<L, sizeof(V)>;
  const unsigned blocks = (unsigned)((cols + G::C - 1) / G::C);
  RBC_LAUNCH((irfft_cols_kernel<L, Tin, V, Tout>), dim3(blocks, batch), dim3(G::T), 0, s, X, Y, cols, times, sc,
             twiddles<V>(L), sx, sy, Y1, split, times1, sc1);
}

}  // namespace

bool fft_supported(int n) { return n >= 8 && n <= 2048 && (n & (n - 1)) == 0; }

#define RBC_FFT_SWITCH(NN, CALL)                                            \
  switch (NN) {                                                             \
    case 8: CALL(8); break;                                                 \
    case 16: CALL(16); break;                                               \
    case 32: CALL(32); break;                                               \
    case 64: CALL(64); break;                                               \
    case 128: CALL(128); break;                                             \
    case 256: CALL(256); break;                                             \
    case 512: CALL(512); break;                                             \
    case 1024: CALL(1024); break;                                           \
    case 2048: CALL(2048); break;                                           \
    default: throw std::invalid_argument("fft: unsupported length");        \
  }

void rfft_cols(hipStream_t s, const double* X, double* Y, int n, int cols, int cut, int batch, long sx, long sy,
               int nzc) {
  check_n(n);
  if (batch < 1 || batch > 65535 || cols < 1) throw std::invalid_argument("rfft_cols: bad batch or cols");
  cut = std::min(cut, n / 2 + 1);
#define RBC_RFFT(LL) launch_rfft<LL, double, double, double>(s, X, Y, cols, cut, batch, sx, sy, nzc)
  RBC_FFT_SWITCH(n, RBC_RFFT)
#undef RBC_RFFT
  check("rfft_cols");
}

void irfft_cols(hipStream_t s, const double* X, double* Y, int n, int cols, int times, double sc, int batch, long sx,
                long sy) {
  check_n(n);
  if (batch < 1 || batch > 65535 || cols < 1) throw std::invalid_argument("irfft_cols: bad batch or cols");
#define RBC_IRFFT(LL) launch_irfft<LL, double, double, double>(s, X, Y, cols, times, sc, batch, sx, sy)
  RBC_FFT_SWITCH(n, RBC_IRFFT)
#undef RBC_IRFFT
  check("irfft_cols");
}

void irfft_cols_pair(hipStream_t s, const double* X, double* Y0, double* Y1, int n, int cols, int times1, double sc1,
                     int nin, long sx, long sy) {
  check_n(n);
  if (nin < 1 || 2 * nin > 65535 || cols < 1 || !Y1) throw std::invalid_argument("irfft_cols_pair: bad arguments");
#define RBC_IRFFTP(LL) \
  launch_irfft<LL, double, double, double>(s, X, Y0, cols, 0, 1.0, 2 * nin, sx, sy, Y1, nin, times1, sc1)
  RBC_FFT_SWITCH(n, RBC_IRFFTP)
#undef RBC_IRFFTP
  check("irfft_cols_pair");
}

void irfft_cols_pair_f32(hipStream_t s, const double* X, float* Y0, float* Y1, int n, int cols, int times1, double sc1,
                         int nin, long sx, long sy) {
  check_n(n);
  if (nin < 1 || 2 * nin > 65535 || cols < 1 || !Y1) throw std::invalid_argument("irfft_cols_pair_f32: bad arguments");
#define RBC_IRFFTP32(LL) \
  launch_irfft<LL, double, float, float>(s, X, Y0, cols, 0, 1.0, 2 * nin, sx, sy, Y1, nin, times1, sc1)
  RBC_FFT_SWITCH(n, RBC_IRFFTP32)
#undef RBC_IRFFTP32
  check("irfft_cols_pair_f32");
}

void rfft_cols_f32(hipStream_t s, const float* X, double* Y, int n, int cols, int cut, int batch, long sx, long sy) {
  check_n(n);
  if (batch < 1 || batch > 65535 || cols < 1) throw std::invalid_argument("rfft_cols_f32: bad batch or cols");
  cut = std::min(cut, n / 2 + 1);
#define RBC_RFFT32(LL) launch_rfft<LL, float, float, double>(s, X, Y, cols, cut, batch, sx, sy, -1)
  RBC_FFT_SWITCH(n, RBC_RFFT32)
#undef RBC_RFFT32
  check("rfft_cols_f32");
}

void rfft_cols_f32f(hipStream_t s, const float* X, float* Y, int n, int cols, int cut, int batch, long sx, long sy,
                    int nzc) {
  check_n(n);
  if (batch < 1 || batch > 65535 || cols < 1) throw std::invalid_argument("rfft_cols_f32f: bad batch or cols");
  cut = std::min(cut, n / 2 + 1);
#define RBC_RFFT32F(LL) launch_rfft<LL, float, float, float>(s, X, Y, cols, cut, batch, sx, sy, nzc)
  RBC_FFT_SWITCH(n, RBC_RFFT32F)
#undef RBC_RFFT32F
  check("rfft_cols_f32f");
}

void irfft_cols_f32(hipStream_t s, const double* X, float* Y, int n, int cols, int times, double sc, int batch,
                    long sx, long sy) {
  check_n(n);
  if (batch < 1 || batch > 65535 || cols < 1) throw std::invalid_argument("irfft_cols_f32: bad batch or cols");
#define RBC_IRFFT32(LL) launch_irfft<LL, double, float, float>(s, X, Y, cols, times, sc, batch, sx, sy)
  RBC_FFT_SWITCH(n, RBC_IRFFT32)
#undef RBC_IRFFT32
  check("irfft_cols_f32");
}

}  // namespace rbc
