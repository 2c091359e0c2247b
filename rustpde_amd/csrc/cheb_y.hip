// Chebyshev transforms along axis 1 (the contiguous y axis) by FFT, FP64.
//
// For n = N + 1 Gauss-Lobatto points with N a power of two both directions are a DCT-I,
//     DCT1(u)_k = u_0 + (-1)^k u_N + 2 sum_{j=1}^{N-1} u_j cos(pi j k / N),
// i.e. the DFT of the even extension of u to length L = 2N.  Two rows a, b go through one
// complex DFT: the DFT of the even extension of a + i b has real part DCT1(a) and
// imaginary part DCT1(b), so no split step is needed.  The DFT runs in LDS as radix-8
// Stockham passes (plus one radix-4 or radix-2 pass), natural order in and out.
//   backward (ortho.rs:288-312): x_j = sum_k c_k T_k(x_j) = DCT1(u)_j with
//     u_k = (-1)^k c_k w_k, w_0 = w_N = 1, w_k = 1/2 otherwise (the reference's correction
//     vector, ortho.rs:52-58), exact in floating point.
//   forward (ortho.rs:235-258): c_k = (-1)^k h_k / N * DCT1(v)_k, h_0 = h_N = 1/2; rows
//     k >= cut are set to zero (the 2/3 dealias of the convective term).
// A composite basis is first reduced to Chebyshev coefficients by its stencil
// (composite_stencil.rs:207-229, taps as the banded kernels); one derivative
// (ortho.rs:107-125) is a suffix sum along each parity chain, done as a wavefront scan
// in LDS.  Backward results are written as out = x .* U + beta out (the convective
// product of conv_term.rs:35-41 fused into the transform).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ops.hpp"

namespace rbc {

namespace {

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// e^{-2 pi i t / L}, t < L, interleaved (re, im), per (device, L)
std::mutex g_mutex;
std::map<std::pair<int, int>, double*> g_tables;

const double* table(int L) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  auto key = std::make_pair(dev, L);
  auto it = g_tables.find(key);
  if (it != g_tables.end()) return it->second;
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<double> t(2 * (size_t)L);
  for (int i = 0; i < L; ++i) {
    const long double a = -2.0L * pi * (long double)i / (long double)L;
    t[2 * i] = (double)cosl(a);
    t[2 * i + 1] = (double)sinl(a);
  }
  double* d = nullptr;
  if (hipMalloc(&d, t.size() * sizeof(double)) != hipSuccess) throw std::runtime_error("cheb_y: table allocation failed");
  if (hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("cheb_y: table upload failed");
  g_tables.emplace(key, d);
  return d;
}

// block geometry for an FFT length L: E = P L complex values per block (P row pairs),
// T = E / 8 threads, so every pass gives each thread 8 values
template <int L>
struct Geo {
  static constexpr int E = L > 512 ? L : 512;
  static constexpr int P = E / L;
  static constexpr int T = E / 8;
  static constexpr int N = L / 2;
  static constexpr int n = N + 1;
};

constexpr double SQH = 0.70710678118654752440;  // sqrt(1/2)

// forward DFTs of 2, 4, 8 points (sign -1), in place, natural order
__device__ __forceinline__ void dft2(double* r, double* i) {
  const double ar = r[0], ai = i[0];
  r[0] = ar + r[1];
  i[0] = ai + i[1];
  r[1] = ar - r[1];
  i[1] = ai - i[1];
}
__device__ __forceinline__ void dft4(double& r0, double& i0, double& r1, double& i1, double& r2, double& i2, double& r3,
                                     double& i3) {
  const double t0r = r0 + r2, t0i = i0 + i2, t1r = r0 - r2, t1i = i0 - i2;
  const double t2r = r1 + r3, t2i = i1 + i3;
  const double t3r = i1 - i3, t3i = r3 - r1;  // (x1 - x3) (-i)
  r0 = t0r + t2r;
  i0 = t0i + t2i;
  r2 = t0r - t2r;
  i2 = t0i - t2i;
  r1 = t1r + t3r;
  i1 = t1i + t3i;
  r3 = t1r - t3r;
  i3 = t1i - t3i;
}
__device__ __forceinline__ void dft8(double* r, double* i) {
  dft4(r[0], i[0], r[2], i[2], r[4], i[4], r[6], i[6]);  // E_k in slots 0, 2, 4, 6
  dft4(r[1], i[1], r[3], i[3], r[5], i[5], r[7], i[7]);  // O_k in slots 1, 3, 5, 7
  // O_1 *= e^{-i pi/4}, O_2 *= -i, O_3 *= e^{-3 i pi/4}
  double o1r = SQH * (r[3] + i[3]), o1i = SQH * (i[3] - r[3]);
  double o2r = i[5], o2i = -r[5];
  double o3r = SQH * (i[7] - r[7]), o3i = -SQH * (r[7] + i[7]);
  const double e0r = r[0], e0i = i[0], e1r = r[2], e1i = i[2], e2r = r[4], e2i = i[4], e3r = r[6], e3i = i[6];
  const double o0r = r[1], o0i = i[1];
  r[0] = e0r + o0r; i[0] = e0i + o0i;
  r[4] = e0r - o0r; i[4] = e0i - o0i;
  r[1] = e1r + o1r; i[1] = e1i + o1i;
  r[5] = e1r - o1r; i[5] = e1i - o1i;
  r[2] = e2r + o2r; i[2] = e2i + o2i;
  r[6] = e2r - o2r; i[6] = e2i - o2i;
  r[3] = e3r + o3r; i[3] = e3i + o3i;
  r[7] = e3r - o3r; i[7] = e3i - o3i;
}

// one Stockham pass of radix R over the P transforms of length L held in (re, im)
template <int L, int R>
__device__ __forceinline__ void stockham_pass(double* re, double* im, int Ns, const double* __restrict__ tw) {
  using G = Geo<L>;
  constexpr int NB = L / R;            // butterflies per transform
  constexpr int PER = G::E / (R * G::T);  // butterflies per thread (8 / R)
  double xr[PER][R], xi[PER][R];
  int dst[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int g = threadIdx.x + q * G::T;
    const int p = g / NB, j = g - p * NB;
    const int k = j & (Ns - 1);
    const int base = p * L + j;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      xr[q][r] = re[base + r * NB];
      xi[q][r] = im[base + r * NB];
    }
    const int step = L / (Ns * R);  // twiddle e^{-2 pi i r k / (Ns R)} = table[r k step]
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const int t = r * k * step;
      const double wr = tw[2 * t], wi = tw[2 * t + 1];
      const double ar = xr[q][r], ai = xi[q][r];
      xr[q][r] = ar * wr - ai * wi;
      xi[q][r] = ar * wi + ai * wr;
    }
    if constexpr (R == 8) dft8(xr[q], xi[q]);
    else if constexpr (R == 4) dft4(xr[q][0], xi[q][0], xr[q][1], xi[q][1], xr[q][2], xi[q][2], xr[q][3], xi[q][3]);
    else dft2(xr[q], xi[q]);
    dst[q] = p * L + (j / Ns) * Ns * R + k;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      re[dst[q] + r * Ns] = xr[q][r];
      im[dst[q] + r * Ns] = xi[q][r];
    }
  __syncthreads();
}

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

template <int L>
__device__ __forceinline__ void lds_dft(double* re, double* im, const double* __restrict__ tw) {
  constexpr int NUM8 = ilog2(L) / 3, REM = ilog2(L) % 3;
  int Ns = 1;
#pragma unroll
  for (int s = 0; s < NUM8; ++s) {
    stockham_pass<L, 8>(re, im, Ns, tw);
    Ns *= 8;
  }
  if constexpr (REM == 2) stockham_pass<L, 4>(re, im, Ns, tw);
  else if constexpr (REM == 1) stockham_pass<L, 2>(re, im, Ns, tw);
}

// stencil value s_k of one row (taps at offsets 0, -2, +2, +4, summed in that order)
__device__ __forceinline__ double stencil_at(const double* __restrict__ row, int k, int mi, int n,
                                             const double* __restrict__ coef, const unsigned char* __restrict__ mask) {
#pragma clang fp contract(off)
  if (!coef) return k < mi ? row[k] : 0.0;
  const unsigned mk = mask[k];
  constexpr int off[4] = {0, -2, 2, 4};
  double acc = 0.0;
  bool first = true;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (!(mk & (1u << t))) continue;
    const double term = row[k + off[t]] * coef[t * n + k];
    acc = first ? term : acc + term;
    first = false;
  }
  return acc;
}

template <int L>
__global__ __launch_bounds__(Geo<L>::T) void cheb_y_bwd_kernel(
    const double* __restrict__ in, int ldi, int mi, double* __restrict__ out, int ldo, int rows,
    const double* __restrict__ coef, const unsigned char* __restrict__ mask, int times, double inv_sc,
    const double* __restrict__ U, int ldu, double beta, const double* __restrict__ tw, long si, long so, long su) {
  using G = Geo<L>;
  constexpr int N = G::N, n = G::n, P = G::P, T = G::T;
  __shared__ double re[G::E], im[G::E];
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  if (U) U += su * blockIdx.y;
  const int tid = threadIdx.x, row0 = 2 * P * blockIdx.x;
  // (1) Chebyshev coefficients of both rows, or g_k of the derivative recurrence:
  //     g_0 = s_1, g_k = 2 (k + 1) s_{k+1} (1 <= k <= n - 2), g_{n-1} = 0
  for (int e = tid; e < P * n; e += T) {
    const int p = e / n, k = e - p * n;
    const int ra = row0 + 2 * p, rb = ra + 1;
    const int kk = times ? k + 1 : k;
    double sa = 0.0, sb = 0.0;
    if (kk < n && (!times || k < n - 1)) {
      if (ra < rows) sa = stencil_at(in + (size_t)ra * ldi, kk, mi, n, coef, mask);
      if (rb < rows) sb = stencil_at(in + (size_t)rb * ldi, kk, mi, n, coef, mask);
      if (times && k > 0) {
        const double f = 2.0 * (double)(k + 1);
        sa = f * sa;
        sb = f * sb;
      }
    }
    re[p * L + k] = sa;
    im[p * L + k] = sb;
  }
  __syncthreads();
  if (times) {
    // (2) d_k = g_k + d_{k+2} for k >= 1 (suffix sums along the chains 1, 3, .., N-1 and
    //     2, 4, .., N), then d_0 = g_0 + d_2 / 2.  One wavefront per chain.
    constexpr int C = N / 2, SEG = (C + 63) / 64, NCH = 4 * P;
    const int lane = tid & 63, wave = tid >> 6;
    for (int ch = wave; ch < NCH; ch += T / 64) {
      const int p = ch >> 2, comp = (ch >> 1) & 1, par = ch & 1;
      double* z = (comp ? im : re) + p * L + (par ? 1 : 2);
      const int t0 = lane * SEG;
      double tot = 0.0;
#pragma unroll
      for (int q = SEG - 1; q >= 0; --q)
        if (t0 + q < C) tot = z[2 * (t0 + q)] + tot;
      double inc = tot;  // inclusive suffix sum over lanes
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_down(inc, off);
        if (lane + off < 64) inc = inc + o;
      }
      double acc = __shfl_down(inc, 1);
      if (lane == 63) acc = 0.0;
#pragma unroll
      for (int q = SEG - 1; q >= 0; --q) {
        const int t = t0 + q;
        if (t < C) {
          acc = z[2 * t] + acc;
          z[2 * t] = acc;
        }
      }
      if (par == 0 && lane == 0) {  // acc = d_2
        double* z0 = (comp ? im : re) + p * L;
        z0[0] = z0[0] + acc / 2.0;
      }
    }
    __syncthreads();
  }
  // (3) u_k = (-1)^k w_k c_k / sc^times and the even extension u_{L-k} = u_k
  for (int e = tid; e < P * n; e += T) {
    const int p = e / n, k = e - p * n;
    double a = re[p * L + k], b = im[p * L + k];
    if (times && inv_sc != 1.0) {
      a = a * inv_sc;
      b = b * inv_sc;
    }
    const double w = ((k & 1) ? -1.0 : 1.0) * ((k == 0 || k == N) ? 1.0 : 0.5);
    a = a * w;
    b = b * w;
    re[p * L + k] = a;
    im[p * L + k] = b;
    if (k > 0 && k < N) {
      re[p * L + L - k] = a;
      im[p * L + L - k] = b;
    }
  }
  __syncthreads();
  lds_dft<L>(re, im, tw);
  // (4) x_j = Re / Im Z_j, j <= N; out = x .* U + beta out
  for (int e = tid; e < P * n; e += T) {
    const int p = e / n, j = e - p * n;
    const int ra = row0 + 2 * p;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      double v = h ? im[p * L + j] : re[p * L + j];
      double* o = out + (size_t)r * ldo + j;
      if (U) v = v * U[(size_t)r * ldu + j];
      if (beta != 0.0) v = v + beta * (*o);
      *o = v;
    }
  }
}

template <int L>
__global__ __launch_bounds__(Geo<L>::T) void cheb_y_fwd_kernel(const double* __restrict__ in, int ldi,
                                                               double* __restrict__ out, int ldo, int rows, int cut,
                                                               const double* __restrict__ tw, long si, long so) {
  using G = Geo<L>;
  constexpr int N = G::N, n = G::n, P = G::P, T = G::T;
  __shared__ double re[G::E], im[G::E];
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  const int tid = threadIdx.x, row0 = 2 * P * blockIdx.x;
  for (int e = tid; e < P * n; e += T) {
    const int p = e / n, k = e - p * n;
    const int ra = row0 + 2 * p, rb = ra + 1;
    const double a = ra < rows ? in[(size_t)ra * ldi + k] : 0.0;
    const double b = rb < rows ? in[(size_t)rb * ldi + k] : 0.0;
    re[p * L + k] = a;
    im[p * L + k] = b;
    if (k > 0 && k < N) {
      re[p * L + L - k] = a;
      im[p * L + L - k] = b;
    }
  }
  __syncthreads();
  lds_dft<L>(re, im, tw);
  constexpr double invN = 1.0 / (double)N;  // exact: N is a power of two
  for (int e = tid; e < P * n; e += T) {
    const int p = e / n, k = e - p * n;
    const int ra = row0 + 2 * p;
    const double s = ((k & 1) ? -invN : invN) * ((k == 0 || k == N) ? 0.5 : 1.0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const double v = h ? im[p * L + k] : re[p * L + k];
      out[(size_t)r * ldo + k] = k < cut ? v * s : 0.0;
    }
  }
}

template <int L>
void launch_bwd(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, const double* coef,
                const unsigned char* mask, int times, double inv_sc, const double* U, int ldu, double beta, int batch,
                long si, long so, long su) {
  using G = Geo<L>;
  const unsigned blocks = (unsigned)((rows + 2 * G::P - 1) / (2 * G::P));
  hipLaunchKernelGGL(cheb_y_bwd_kernel<L>, dim3(blocks, batch), dim3(G::T), 0, s, in, ldi, mi, out, ldo, rows, coef,
                     mask, times, inv_sc, U, ldu, beta, table(L), si, so, su);
}

template <int L>
void launch_fwd(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, int cut, int batch, long si,
                long so) {
  using G = Geo<L>;
  const unsigned blocks = (unsigned)((rows + 2 * G::P - 1) / (2 * G::P));
  hipLaunchKernelGGL(cheb_y_fwd_kernel<L>, dim3(blocks, batch), dim3(G::T), 0, s, in, ldi, out, ldo, rows, cut,
                     table(L), si, so);
}

}  // namespace

bool cheb_y_supported(int n) {
  const int N = n - 1;
  return N >= 16 && N <= 2048 && (N & (N - 1)) == 0;
}

#define RBC_CHEB_Y_SWITCH(L2, CALL)                                   \
  switch (L2) {                                                       \
    case 32: CALL(32); break;                                         \
    case 64: CALL(64); break;                                         \
    case 128: CALL(128); break;                                       \
    case 256: CALL(256); break;                                       \
    case 512: CALL(512); break;                                       \
    case 1024: CALL(1024); break;                                     \
    case 2048: CALL(2048); break;                                     \
    case 4096: CALL(4096); break;                                     \
    default: throw std::invalid_argument("cheb_y: unsupported length"); \
  }

void cheb_y_backward(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                     const double* coef, const unsigned char* mask, int times, double inv_sc, const double* U, int ldu,
                     double beta, int batch, long si, long so, long su) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_backward: unsupported length");
  if (times < 0 || times > 1) throw std::invalid_argument("cheb_y_backward: times must be 0 or 1");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_backward: bad batch or rows");
  if (mi > n || (!coef && mi < 1)) throw std::invalid_argument("cheb_y_backward: bad input length");
#define RBC_BWD(LL) launch_bwd<LL>(s, in, ldi, mi, out, ldo, rows, coef, mask, times, inv_sc, U, ldu, beta, batch, si, so, su)
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_BWD)
#undef RBC_BWD
  check("cheb_y_backward");
}

void cheb_y_forward(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, int n, int cut, int batch,
                    long si, long so) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_forward: unsupported length");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_forward: bad batch or rows");
#define RBC_FWD(LL) launch_fwd<LL>(s, in, ldi, out, ldo, rows, cut, batch, si, so)
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_FWD)
#undef RBC_FWD
  check("cheb_y_forward");
}

}  // namespace rbc
