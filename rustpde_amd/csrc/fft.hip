// Batched real FFTs along axis 0 (the strided x axis) for the Fourier (fourier_r2c)
// direction of the periodic step, FP64.
//
// A real transform of length n = 2N is done as one complex FFT of length N of
// z_t = x_{2t} + i x_{2t+1} (radix-2 decimation in frequency, in LDS, bit-reversed
// output) plus the usual split step.  Each workgroup owns C consecutive columns; loads
// and stores of a row segment are contiguous.  Spectral data are split planes
// (rows [0, N] real parts, rows [N+1, 2N+1] imaginary parts).
//   r2c: X_k = sum_j x_j e^{-2 pi i jk/n}, k = 0..N   (r2c.rs:160-175, unnormalised)
//   c2r: x_j = 1/n sum_k X_k e^{+2 pi i jk/n}         (r2c.rs:211-226), Im X_0 and Im X_N
//        are ignored (half-spectrum convention); an optional (i k)^d / sc is applied to
//        X_k first (r2c.rs:88-99 with the num-complex division of space2.rs:256-262).
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ops.hpp"

namespace rbc {

namespace {

constexpr int FFT_THREADS = 256;  // threads per workgroup
// columns per workgroup: LDS use 2 * C * (n/2 + 1) doubles stays <= 64 KB
inline int fft_cols_per_block(int n) { return n <= 1024 ? 4 : (n <= 2048 ? 2 : 1); }

struct Twiddles {
  double* stage = nullptr;  // e^{-2 pi i t/N}, t < N/2, interleaved (re, im)
  double* post = nullptr;   // e^{-2 pi i k/n}, k <= N, interleaved
};

std::mutex g_tw_mutex;
std::map<std::pair<int, int>, Twiddles> g_tw;  // (device, n) -> tables

const Twiddles& twiddles(int n) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  auto key = std::make_pair(dev, n);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  const int N = n / 2;
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<double> st(2 * std::max(N / 2, 1)), po(2 * (N + 1));
  for (int t = 0; t < N / 2; ++t) {
    long double a = -2.0L * pi * (long double)t / (long double)N;
    st[2 * t] = (double)cosl(a);
    st[2 * t + 1] = (double)sinl(a);
  }
  for (int k = 0; k <= N; ++k) {
    long double a = -2.0L * pi * (long double)k / (long double)n;
    po[2 * k] = (double)cosl(a);
    po[2 * k + 1] = (double)sinl(a);
  }
  Twiddles tw;
  if (hipMalloc(&tw.stage, st.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&tw.post, po.size() * sizeof(double)) != hipSuccess)
    throw std::runtime_error("fft: twiddle allocation failed");
  if (hipMemcpy(tw.stage, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(tw.post, po.data(), po.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("fft: twiddle upload failed");
  return g_tw.emplace(key, tw).first->second;
}

__device__ __forceinline__ int bitrev(int v, int bits) { return (int)(__brev((unsigned)v) >> (32 - bits)); }

// in-place radix-2 DIF FFT of C columns of length N held in LDS (re, im), sign -1 (fwd)
// or +1 (inverse, conjugate twiddles); output in bit-reversed order
template <int C>
__device__ void lds_fft(double* re, double* im, int N, int ldc, const double* __restrict__ tw, bool inverse) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x;
  for (int h = N / 2; h >= 1; h >>= 1) {
    const int stride = N / (2 * h);
    for (int e = tid; e < C * (N / 2); e += FFT_THREADS) {
      const int c = e % C, q = e / C;
      const int blk = q / h, i = q - blk * h;
      const int i0 = c * ldc + blk * 2 * h + i, i1 = i0 + h;
      const double wr = tw[2 * (i * stride)];
      const double wi = inverse ? -tw[2 * (i * stride) + 1] : tw[2 * (i * stride) + 1];
      const double ar = re[i0], ai = im[i0], br = re[i1], bi = im[i1];
      re[i0] = ar + br;
      im[i0] = ai + bi;
      const double dr = ar - br, di = ai - bi;
      re[i1] = dr * wr - di * wi;
      im[i1] = dr * wi + di * wr;
    }
    __syncthreads();
  }
}

template <int C>
__global__ __launch_bounds__(FFT_THREADS) void rfft_cols_kernel(const double* __restrict__ X, double* __restrict__ Y,
                                                                 int n, int cols, int cut,
                                                                 const double* __restrict__ tw,
                                                                 const double* __restrict__ post, long sx, long sy) {
#pragma clang fp contract(off)
  X += sx * blockIdx.y;
  Y += sy * blockIdx.y;
  extern __shared__ double lds[];
  const int N = n / 2, m = N + 1, bits = __ffs(N) - 1;
  double* re = lds;
  double* im = lds + C * N;
  const int j0 = blockIdx.x * C, tid = threadIdx.x;
  for (int e = tid; e < C * N; e += FFT_THREADS) {
    const int c = e % C, t = e / C, j = j0 + c;
    double a = 0.0, b = 0.0;
    if (j < cols) {
      a = X[(size_t)(2 * t) * cols + j];
      b = X[(size_t)(2 * t + 1) * cols + j];
    }
    re[c * N + t] = a;
    im[c * N + t] = b;
  }
  __syncthreads();
  lds_fft<C>(re, im, N, N, tw, false);
  for (int e = tid; e < C * m; e += FFT_THREADS) {
    const int c = e % C, k = e / C, j = j0 + c;
    if (j >= cols) continue;
    double xr = 0.0, xi = 0.0;
    if (k < cut) {
      const int p = c * N + bitrev(k % N, bits), q = c * N + bitrev((N - k) % N, bits);
      const double ar = re[p], ai = im[p], br = re[q], bi = -im[q];  // B = conj(Z_{N-k})
      const double er = 0.5 * (ar + br), ei = 0.5 * (ai + bi);
      const double dr = ar - br, di = ai - bi;
      const double orr = 0.5 * di, oi = -0.5 * dr;  // O = D / (2i)
      const double wr = post[2 * k], wi = post[2 * k + 1];
      xr = er + (orr * wr - oi * wi);
      xi = ei + (orr * wi + oi * wr);
    }
    Y[(size_t)k * cols + j] = xr;
    Y[(size_t)(m + k) * cols + j] = xi;
  }
}

template <int C>
__global__ __launch_bounds__(FFT_THREADS) void irfft_cols_kernel(const double* __restrict__ X, double* __restrict__ Y,
                                                                  int n, int cols, int times, double sc,
                                                                  const double* __restrict__ tw,
                                                                  const double* __restrict__ post, long sx, long sy) {
#pragma clang fp contract(off)
  X += sx * blockIdx.y;
  Y += sy * blockIdx.y;
  extern __shared__ double lds[];
  const int N = n / 2, m = N + 1, bits = __ffs(N) - 1, S = m;  // column stride S = N + 1
  double* re = lds;
  double* im = lds + C * S;
  const int j0 = blockIdx.x * C, tid = threadIdx.x;
  for (int e = tid; e < C * m; e += FFT_THREADS) {
    const int c = e % C, k = e / C, j = j0 + c;
    double a = 0.0, b = 0.0;
    if (j < cols) {
      a = X[(size_t)k * cols + j];
      b = X[(size_t)(m + k) * cols + j];
    }
    const double kk = (double)k;
    for (int q = 0; q < times; ++q) {
      const double na = a * 0.0 - b * kk, nb = a * kk + b * 0.0;
      a = na;
      b = nb;
    }
    if (sc != 1.0) {
      const double ns = sc * sc + 0.0 * 0.0;
      const double na = (a * sc + b * 0.0) / ns, nb = (b * sc - a * 0.0) / ns;
      a = na;
      b = nb;
    }
    if (k == 0 || k == N) b = 0.0;
    re[c * S + k] = a;
    im[c * S + k] = b;
  }
  __syncthreads();
  // Z_k = E_k + i O_k with E_k = (X_k + conj X_{N-k}) / 2, O_k = (X_k - conj X_{N-k}) e^{+2 pi i k/n} / 2.
  // One thread owns the pair (k, N-k) and overwrites both in place (k = 0 pairs with X_N).
  for (int e = tid; e < C * (N / 2 + 1); e += FFT_THREADS) {
    const int c = e % C, k = e / C, k2 = N - k;
    const double xr = re[c * S + k], xi = im[c * S + k];
    const double yr = re[c * S + k2], yi = im[c * S + k2];
    {  // Z_k from (X_k, X_{N-k})
      const double br = yr, bi = -yi;
      const double er = 0.5 * (xr + br), ei = 0.5 * (xi + bi);
      const double dr = 0.5 * (xr - br), di = 0.5 * (xi - bi);
      const double wr = post[2 * k], wi = -post[2 * k + 1];
      const double orr = dr * wr - di * wi, oi = dr * wi + di * wr;
      re[c * S + k] = er - oi;
      im[c * S + k] = ei + orr;
    }
    if (k2 != k && k2 < N) {  // Z_{N-k} from (X_{N-k}, X_k)
      const double br = xr, bi = -xi;
      const double er = 0.5 * (yr + br), ei = 0.5 * (yi + bi);
      const double dr = 0.5 * (yr - br), di = 0.5 * (yi - bi);
      const double wr = post[2 * k2], wi = -post[2 * k2 + 1];
      const double orr = dr * wr - di * wi, oi = dr * wi + di * wr;
      re[c * S + k2] = er - oi;
      im[c * S + k2] = ei + orr;
    }
  }
  __syncthreads();
  lds_fft<C>(re, im, N, S, tw, true);
  const double inv = 1.0 / (double)N;
  for (int e = tid; e < C * N; e += FFT_THREADS) {
    const int c = e % C, t = e / C, j = j0 + c;
    if (j >= cols) continue;
    const int p = c * S + bitrev(t, bits);
    Y[(size_t)(2 * t) * cols + j] = re[p] * inv;
    Y[(size_t)(2 * t + 1) * cols + j] = im[p] * inv;
  }
}

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void check_n(int n) {
  if (n < 8 || (n & (n - 1)) != 0) throw std::invalid_argument("fft: length must be a power of two >= 8");
  if (n > 4096) throw std::invalid_argument("fft: length above 4096 exceeds the LDS budget");
}

}  // namespace

bool fft_supported(int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; }

template <int C>
static void launch_rfft(hipStream_t s, const double* X, double* Y, int n, int cols, int cut, const Twiddles& tw,
                        int batch, long sx, long sy) {
  const size_t lds = sizeof(double) * 2 * C * (n / 2);
  const unsigned blocks = (unsigned)((cols + C - 1) / C);
  hipLaunchKernelGGL(rfft_cols_kernel<C>, dim3(blocks, batch), dim3(FFT_THREADS), lds, s, X, Y, n, cols, cut,
                     tw.stage, tw.post, sx, sy);
}
template <int C>
static void launch_irfft(hipStream_t s, const double* X, double* Y, int n, int cols, int times, double sc,
                         const Twiddles& tw, int batch, long sx, long sy) {
  const size_t lds = sizeof(double) * 2 * C * (n / 2 + 1);
  const unsigned blocks = (unsigned)((cols + C - 1) / C);
  hipLaunchKernelGGL(irfft_cols_kernel<C>, dim3(blocks, batch), dim3(FFT_THREADS), lds, s, X, Y, n, cols, times, sc,
                     tw.stage, tw.post, sx, sy);
}

void rfft_cols(hipStream_t s, const double* X, double* Y, int n, int cols, int cut, int batch, long sx, long sy) {
  check_n(n);
  if (batch < 1 || batch > 65535) throw std::invalid_argument("rfft_cols: bad batch");
  const Twiddles& tw = twiddles(n);
  cut = std::min(cut, n / 2 + 1);
  switch (fft_cols_per_block(n)) {
    case 4: launch_rfft<4>(s, X, Y, n, cols, cut, tw, batch, sx, sy); break;
    case 2: launch_rfft<2>(s, X, Y, n, cols, cut, tw, batch, sx, sy); break;
    default: launch_rfft<1>(s, X, Y, n, cols, cut, tw, batch, sx, sy); break;
  }
  check("rfft_cols");
}

void irfft_cols(hipStream_t s, const double* X, double* Y, int n, int cols, int times, double sc, int batch, long sx,
                long sy) {
  check_n(n);
  if (batch < 1 || batch > 65535) throw std::invalid_argument("irfft_cols: bad batch");
  const Twiddles& tw = twiddles(n);
  switch (fft_cols_per_block(n)) {
    case 4: launch_irfft<4>(s, X, Y, n, cols, times, sc, tw, batch, sx, sy); break;
    case 2: launch_irfft<2>(s, X, Y, n, cols, times, sc, tw, batch, sx, sy); break;
    default: launch_irfft<1>(s, X, Y, n, cols, times, sc, tw, batch, sx, sy); break;
  }
  check("irfft_cols");
}

}  // namespace rbc
