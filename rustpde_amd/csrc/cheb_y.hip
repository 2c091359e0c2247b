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
#include <type_traits>
#include <vector>

#include "bufrow.hpp"
#include "fft_lds.hpp"
#include "wave.hpp"
#include "ops.hpp"

namespace rbc {

namespace {

void check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

std::mutex g_mutex;
std::map<std::pair<int, int>, double*> g_tables;
std::map<std::pair<int, int>, float*> g_tables32;

// e^{-2 pi i t / L}, t < L, interleaved (re, im), rounded from long double
std::vector<double> table_values(int L) {
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<double> t(2 * (size_t)L);
  for (int i = 0; i < L; ++i) {
    const long double a = -2.0L * pi * (long double)i / (long double)L;
    t[2 * i] = (double)cosl(a);
    t[2 * i + 1] = (double)sinl(a);
  }
  return t;
}

}  // namespace

const double* fft_table(int L) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  auto key = std::make_pair(dev, L);
  auto it = g_tables.find(key);
  if (it != g_tables.end()) return it->second;
  const std::vector<double> t = table_values(L);
  double* d = nullptr;
  if (hipMalloc(&d, t.size() * sizeof(double)) != hipSuccess) throw std::runtime_error("cheb_y: table allocation failed");
  if (hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("cheb_y: table upload failed");
  g_tables.emplace(key, d);
  return d;
}

const float* fft_table_f32(int L) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(g_mutex);
  auto key = std::make_pair(dev, L);
  auto it = g_tables32.find(key);
  if (it != g_tables32.end()) return it->second;
  const std::vector<double> t = table_values(L);
  std::vector<float> f(t.size());
  for (size_t i = 0; i < t.size(); ++i) f[i] = (float)t[i];
  float* d = nullptr;
  if (hipMalloc(&d, f.size() * sizeof(float)) != hipSuccess) throw std::runtime_error("cheb_y: table allocation failed");
  if (hipMemcpy(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("cheb_y: table upload failed");
  g_tables32.emplace(key, d);
  return d;
}

namespace {
using fftlds::lds_dft;
using fftlds::lp;
using fftlds::lp_size;

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

// stencil value s_k = d_k c_k + l_k c_{k-2} of one row (composite_stencil.rs:207-229);
// d = coef[0, n), l = coef[n, 2n) hold zeros where the stencil has no entry, which leaves
// the sum of the present terms unchanged.  Without coef the row is Chebyshev already.
template <typename V>
struct StencilLoad {
  V c0, c2, d, l;
};
template <typename V, typename Tin>
__device__ __forceinline__ StencilLoad<V> stencil_load(const Tin* __restrict__ row, bool valid, int k, int mi, int n,
                                                       const double* __restrict__ coef) {
  StencilLoad<V> v{0, 0, 1, 0};
  if (k >= n) return v;
  if (valid && k < mi) v.c0 = (V)row[k];
  if (coef) {
    v.d = (V)coef[k];
    v.l = (V)coef[n + k];
    if (valid && k >= 2 && k - 2 < mi) v.c2 = (V)row[k - 2];
  }
  return v;
}
template <typename V>
__device__ __forceinline__ V stencil_value(const StencilLoad<V>& v, bool composite) {
#pragma clang fp contract(off)
  return composite ? v.d * v.c0 + v.l * v.c2 : v.c0;
}

// Tin: input type, V: arithmetic (and LDS) type, Tout: type of out and U
// Extra products added to one batch entry after the epilogue: out += E1 F1, then out += E2 F2
// (temp's u . grad(fieldbc) terms, navier.rs:547-550); E of the output type, F FP64, both
// [rows, ldu], shared by the batch.  The convective launch also takes F with its own row
// stride ldf (0: one row for all, a boundary field constant in x) and f1 null (a zero x
// gradient: the term is left out).
template <typename Tout>
struct ExtraProducts {
  const Tout* e1;
  const double* f1;
  const Tout* e2;
  const double* f2;
  int slot;        // batch entry that receives them (-1: none)
  long ldf = -1;   // row stride of f1, f2 (-1: ldu)
};

// DUAL: the two components of a complex transform are row r of `in` (no derivative) and
// row r of `in2` (one derivative, times = 1), and the epilogue forms the convective term
//     out_r = d/dy(in2)_r .* U2_r + 1 (x_r .* U_r)
// with the operation order of two separate launches (out = x .* U, then out = y .* U2 +
// beta out with beta = 1).  Only the pairing of rows inside the complex transform differs
// from those launches, which moves the results by rounding only.
template <int L, bool transform, bool DUAL, typename Tin, typename V, typename Tout>
__device__ __forceinline__ void cheb_y_bwd_body(const Tin* __restrict__ in, int ldi, int mi, Tout* __restrict__ out,
                                                int ldo, int rows, const double* __restrict__ coef, int times,
                                                double inv_sc_d, const Tout* __restrict__ U, int ldu, double beta_d,
                                                const V* __restrict__ tw, long si, long so, long su,
                                                ExtraProducts<Tout> xp, const Tin* __restrict__ in2 = nullptr,
                                                const Tout* __restrict__ U2 = nullptr) {
  using G = Geo<L>;
  constexpr int N = G::N, n = G::n, P = G::P, T = G::T;
  __shared__ __attribute__((aligned(16))) V zmem[2 * lp_size(G::E)];
  const fftlds::LdsZ<V, lp_size(G::E)> z{zmem};
  const V inv_sc = (V)inv_sc_d, beta = (V)beta_d;
  in += si * blockIdx.y;
  if constexpr (DUAL) in2 += si * blockIdx.y;
  out += so * blockIdx.y;
  if (U) U += su * blockIdx.y;
  // component a: times_a derivatives, component b: times_b
  const int times_a = DUAL ? 0 : times, times_b = DUAL ? 1 : times;
  const int tid = threadIdx.x, row0 = (DUAL ? 1 : 2) * P * blockIdx.x;
  // (1) Chebyshev coefficients of both rows, or g_k of the derivative recurrence:
  //     g_0 = s_1, g_k = 2 (k + 1) s_{k+1} (1 <= k <= n - 2), g_{n-1} = 0.
  //     All loads of a thread are issued before any is used.
  constexpr int Q = (P * n + T - 1) / T;
  StencilLoad<V> la[Q], lb[Q];
  if constexpr (P == 1) {
    // one row per component: raw buffers whose range checks give the zeros of the guarded
    // form (c_k for k >= mi, c_{k-2} for k < 2 or k - 2 >= min(mi, n - 2), the taps for
    // k >= n, where both coefficients are zero and the tap's value does not matter)
    const int ra = row0, rb = DUAL ? ra : ra + 1;
    const int m2 = mi < n - 2 ? mi : n - 2;
    const Tin* pa = in + (size_t)(ra < rows ? ra : 0) * ldi;
    const Tin* pb = (DUAL ? in2 : in) + (size_t)(rb < rows ? rb : 0) * ldi;
    const BufRow<Tin> a0(pa, ra < rows ? mi : 0), a2(pa, ra < rows && coef ? m2 : 0);
    const BufRow<Tin> b0(pb, rb < rows ? mi : 0), b2(pb, rb < rows && coef ? m2 : 0);
    const BufRow<double> cd(coef, coef ? n : 0), cl(coef ? coef + n : coef, coef ? n : 0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = tid + q * T;
      const int ka = times_a ? k + 1 : k, kb = times_b ? k + 1 : k;
      la[q] = StencilLoad<V>{(V)a0[ka], (V)a2[ka - 2], (V)cd[ka], (V)cl[ka]};
      lb[q] = StencilLoad<V>{(V)b0[kb], (V)b2[kb - 2], (V)cd[kb], (V)cl[kb]};
    }
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      const int p = e / n, k = e - p * n;
      const int ra = DUAL ? row0 + p : row0 + 2 * p, rb = DUAL ? ra : ra + 1;
      const int ka = (e < P * n) ? (times_a ? k + 1 : k) : n;  // k = n loads nothing
      const int kb = (e < P * n) ? (times_b ? k + 1 : k) : n;
      la[q] = stencil_load<V>(in + (size_t)ra * ldi, ra < rows, ka, mi, n, coef);
      lb[q] = stencil_load<V>((DUAL ? in2 : in) + (size_t)rb * ldi, rb < rows, kb, mi, n, coef);
    }
  }
  const bool comp = coef != nullptr;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + q * T;
    if (e >= P * n) continue;
    const int p = e / n, k = e - p * n;
    V sa = stencil_value(la[q], comp), sb = stencil_value(lb[q], comp);
    if (times_a) {
      if (k == n - 1) sa = 0;
      else if (k > 0) sa = (V)(2.0 * (double)(k + 1)) * sa;
    }
    if (times_b) {
      if (k == n - 1) sb = 0;
      else if (k > 0) sb = (V)(2.0 * (double)(k + 1)) * sb;
    }
    z.st(lp(p * L + k), sa, sb);
  }
  __syncthreads();
  if (times) {
    // (2) d_k = g_k + d_{k+2} for k >= 1 (suffix sums along the chains 1, 3, .., N-1 and
    //     2, 4, .., N), then d_0 = g_0 + d_2 / 2.  One wavefront per chain.
    //     DUAL: the chains of component b only.
    constexpr int C = N / 2, SEG = (C + 63) / 64, NCH = (DUAL ? 2 : 4) * P;
    const int lane = tid & 63, wave = tid >> 6;
    for (int ch = wave; ch < NCH; ch += T / 64) {
      const int p = DUAL ? ch >> 1 : ch >> 2, comp = DUAL ? 1 : (ch >> 1) & 1, par = ch & 1;
      const int zb = p * L + (par ? 1 : 2);  // chain element t at z.c(comp, lp(zb + 2 t))
      const int t0 = lane * SEG;
      V tot = 0;
#pragma unroll
      for (int q = SEG - 1; q >= 0; --q)
        if (t0 + q < C) tot = z.c(comp, lp(zb + 2 * (t0 + q))) + tot;
      V acc;  // the chain's sum above this lane's segment
      wave::suffix_sum(tot, lane, acc);
#pragma unroll
      for (int q = SEG - 1; q >= 0; --q) {
        const int t = t0 + q;
        if (t < C) {
          acc = z.c(comp, lp(zb + 2 * t)) + acc;
          z.c(comp, lp(zb + 2 * t)) = acc;
        }
      }
      if (par == 0 && lane == 0) {  // acc = d_2
        z.c(comp, lp(p * L)) = z.c(comp, lp(p * L)) + acc / (V)2;
      }
    }
    __syncthreads();
  }
  // (3) u_k = (-1)^k w_k c_k / sc^times and the even extension u_{L-k} = u_k
  for (int e = tid; e < P * n && transform; e += T) {
    const int p = e / n, k = e - p * n;
    V a, b;
    z.ld(lp(p * L + k), a, b);
    if (times_a && inv_sc != (V)1) a = a * inv_sc;
    if (times_b && inv_sc != (V)1) b = b * inv_sc;
    const V w = (V)(((k & 1) ? -1.0 : 1.0) * ((k == 0 || k == N) ? 1.0 : 0.5));
    a = a * w;
    b = b * w;
    z.st(lp(p * L + k), a, b);
    if (k > 0 && k < N) z.st(lp(p * L + L - k), a, b);
  }
  if constexpr (transform) {
    __syncthreads();
    lds_dft<L, T, G::E>(z, tw);
  }
  // (4) x_j = Re / Im Z_j, j <= N, or the coefficients d_j / sc without the transform;
  //     out = x .* U + beta out (loads first), plus the extra products on one entry
  const bool extra = xp.slot == (int)blockIdx.y;
  if constexpr (DUAL) {
    V ua[Q], ub[Q];
    if constexpr (P == 1) {
      const int r = row0 < rows ? row0 : 0, cnt = row0 < rows ? n : 0;
      const BufRow<Tout> bu(U + (size_t)r * ldu, cnt), bu2(U2 + (size_t)r * ldu, cnt);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        ua[q] = (V)bu[tid + q * T];
        ub[q] = (V)bu2[tid + q * T];
      }
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int e = tid + q * T;
        const int p = e / n, j = e - p * n;
        const int r = row0 + p;
        const bool ok = e < P * n && r < rows;
        ua[q] = ok ? (V)U[(size_t)r * ldu + j] : (V)0;
        ub[q] = ok ? (V)U2[(size_t)r * ldu + j] : (V)0;
      }
    }
    // the extra products of the entry, formed ahead with all their loads issued together
    // (from element 0 where an entry lies outside; the branches are uniform)
    V x1[Q], x2[Q];
    if (extra) {
      auto idx = [&](int q, size_t& o, size_t& of) {
        const int e = tid + q * T;
        const int p = e / n, j = e - p * n, r = row0 + p;
        const bool ok = e < P * n && r < rows;
        o = ok ? (size_t)r * ldu + j : 0;
        of = ok ? (xp.ldf < 0 ? o : (size_t)r * xp.ldf + j) : 0;
      };
      if (xp.f1) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          size_t o, of;
          idx(q, o, of);
          x1[q] = (V)xp.e1[o] * (V)xp.f1[of];
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        size_t o, of;
        idx(q, o, of);
        x2[q] = (V)xp.e2[o] * (V)xp.f2[of];
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      if (e >= P * n) continue;
      const int p = e / n, j = e - p * n;
      const int r = row0 + p;
      if (r >= rows) continue;
      V za, zb;
      z.ld(lp(p * L + j), za, zb);
      const V xa = (V)(Tout)(za * ua[q]);  // the value the first launch stores
      V v = zb * ub[q];
      v = v + (V)1 * xa;
      if (extra) {
        if (xp.f1) v = v + x1[q];
        v = v + x2[q];
      }
      out[(size_t)r * ldo + j] = (Tout)v;
    }
    return;
  }
  V ua[Q], ub[Q], oa[Q], ob[Q];
  if constexpr (P == 1) {
    // the values of lanes past the row or of rows past `rows` are never stored, and ua / oa
    // are used only with U / beta != 0: raw buffers stand for the guards
    const int ra = row0, rb = ra + 1;
    const int ca = ra < rows ? n : 0, cb = rb < rows ? n : 0;
    const size_t ia = (size_t)(ra < rows ? ra : 0), ib = (size_t)(rb < rows ? rb : 0);
    const BufRow<Tout> bua(U ? U + ia * ldu : U, U ? ca : 0), bub(U ? U + ib * ldu : U, U ? cb : 0);
    const BufRow<Tout> boa(out + ia * ldo, beta != (V)0 ? ca : 0), bob(out + ib * ldo, beta != (V)0 ? cb : 0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = tid + q * T;
      ua[q] = (V)bua[j];
      ub[q] = (V)bub[j];
      oa[q] = (V)boa[j];
      ob[q] = (V)bob[j];
    }
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      const int p = e / n, j = e - p * n;
      const int ra = row0 + 2 * p, rb = ra + 1;
      const bool ok = e < P * n;
      ua[q] = (U && ok && ra < rows) ? (V)U[(size_t)ra * ldu + j] : (V)1;
      ub[q] = (U && ok && rb < rows) ? (V)U[(size_t)rb * ldu + j] : (V)1;
      oa[q] = (beta != (V)0 && ok && ra < rows) ? (V)out[(size_t)ra * ldo + j] : (V)0;
      ob[q] = (beta != (V)0 && ok && rb < rows) ? (V)out[(size_t)rb * ldo + j] : (V)0;
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + q * T;
    if (e >= P * n) continue;
    const int p = e / n, j = e - p * n;
    const int ra = row0 + 2 * p, rb = ra + 1;
    V va, vb;
    z.ld(lp(p * L + j), va, vb);
    if (!transform && inv_sc != (V)1) {
      va = va * inv_sc;
      vb = vb * inv_sc;
    }
    if (U) {
      va = va * ua[q];
      vb = vb * ub[q];
    }
    if (beta != (V)0) {
      va = va + beta * oa[q];
      vb = vb + beta * ob[q];
    }
    if (extra) {
      if (ra < rows) {
        const size_t o = (size_t)ra * ldu + j;
        va = va + (V)xp.e1[o] * (V)xp.f1[o];
        va = va + (V)xp.e2[o] * (V)xp.f2[o];
      }
      if (rb < rows) {
        const size_t o = (size_t)rb * ldu + j;
        vb = vb + (V)xp.e1[o] * (V)xp.f1[o];
        vb = vb + (V)xp.e2[o] * (V)xp.f2[o];
      }
    }
    if (ra < rows) out[(size_t)ra * ldo + j] = (Tout)va;
    if (rb < rows) out[(size_t)rb * ldo + j] = (Tout)vb;
  }
}

template <int L, typename Tin, typename V, typename Tout>
__global__ __launch_bounds__(Geo<L>::T) void cheb_y_bwd_kernel(
    const Tin* __restrict__ in, int ldi, int mi, Tout* __restrict__ out, int ldo, int rows,
    const double* __restrict__ coef, int times, double inv_sc, const Tout* __restrict__ U, int ldu, double beta,
    const V* __restrict__ tw, long si, long so, long su, ExtraProducts<Tout> xp) {
  cheb_y_bwd_body<L, true, false, Tin, V, Tout>(in, ldi, mi, out, ldo, rows, coef, times, inv_sc, U, ldu, beta, tw, si,
                                                so, su, xp);
}
// the convective term of a batch of fields from their x-derivatives (dx) and values (val),
// both in x-physical / y-spectral form: out = d/dy(val) .* uy + dx .* ux (conv_term.rs:35-41)
template <int L, typename Tin, typename V, typename Tout>
__global__ __launch_bounds__(Geo<L>::T) void cheb_y_conv_kernel(const Tin* __restrict__ dx, const Tin* __restrict__ val,
                                                                int ldi, int mi, Tout* __restrict__ out, int ldo,
                                                                int rows, const double* __restrict__ coef,
                                                                double inv_sc, const Tout* __restrict__ ux,
                                                                const Tout* __restrict__ uy, int ldu,
                                                                const V* __restrict__ tw, long si, long so,
                                                                ExtraProducts<Tout> xp) {
  cheb_y_bwd_body<L, true, true, Tin, V, Tout>(dx, ldi, mi, out, ldo, rows, coef, 1, inv_sc, ux, ldu, 1.0, tw, si, so,
                                               0, xp, val, uy);
}
// The whole convective stage along y of the periodic step for P x-rows per block, all three
// fields (slots 0, 1, 2 = ux, uy, temp of dx / val, stride si) in the Dirichlet basis
// (composite_stencil.rs: s_k = c_k - c_{k-2}, taps d_k = 1 for k < m, l_k = -1 for
// 2 <= k < m + 2, formed in place with the table's values):
//   transform 0: (val_ux, val_uy) -> the physical velocity ux_r, uy_r, kept in LDS (the
//                values cheb_y_bwd_kernel stores for the convective launch);
//   transform f + 1: (dx_f, d/dy val_f) -> out_f = d/dy(val_f) .* uy + 1 (dx_f .* ux), plus
//                temp's u . grad(fieldbc) with the gradient arrays f1 (x; null where it is
//                zero) and f2 (y), row stride ldf (0 when the boundary field is constant in x);
//   then the three terms go through the dealiased forward transform (cheb_y_fwd_kernel's
//                epilogue, rows k >= cut zero) as the pairs (out_ux, out_uy) and
//                (out_temp, 0), and `out` receives the y-spectral rows.
// Each value is formed with the arithmetic of the launches it replaces (cheb_y_bwd_kernel,
// cheb_y_conv_kernel, cheb_y_fwd_kernel); only the pairing of rows inside the complex
// transforms differs, which moves the results by rounding only.  The physical velocity and
// the convective terms never leave the chip, and the rows of the next field are loaded
// while the current one is transformed.  FP32 (the mixed step), four blocks per CU.
template <int L, typename Tin, typename V, typename Tout>
__global__ __launch_bounds__(Geo<L>::T, 4) void cheb_y_conv3_kernel(
    const Tin* __restrict__ dx, const Tin* __restrict__ val, int ldi, int mi, Tout* __restrict__ out, int ldo, int rows,
    double inv_sc_d, const double* __restrict__ f1, const double* __restrict__ f2, long ldf, int cut,
    const V* __restrict__ tw, long si, long so) {
#pragma clang fp contract(off)
  using G = Geo<L>;
  constexpr int N = G::N, n = G::n, P = G::P, T = G::T;
  constexpr int Q = (P * n + T - 1) / T;
  constexpr double invN = 1.0 / (double)N;  // exact: N is a power of two
  __shared__ __attribute__((aligned(16))) V zmem[2 * lp_size(G::E)];
  const fftlds::LdsZ<V, lp_size(G::E)> z{zmem};
  __shared__ V uxs[P * n], uys[P * n];  // the physical velocity of the block's rows
  __shared__ V cs0[P * n], cs1[P * n];  // the convective terms of ux, then uy (temp reuses cs0)
  const V inv_sc = (V)inv_sc_d;
  const int tid = threadIdx.x, row0 = P * blockIdx.x;
  // raw coefficients c_kk, c_kk-2 of row r of one slot (zero outside the row)
  struct Raw {
    V c0[Q], c2[Q];
  };
  auto load = [&](const Tin* __restrict__ base, int shift, Raw& w) {
    if constexpr (P == 1) {
      // one row per block: a raw buffer over its mi = n - 2 values returns the zeros of
      // kk >= mi and of kk - 2 < 0 by its range check (kk >= n gives kk - 2 >= mi)
      const BufRow<Tin> row(base + (size_t)(row0 < rows ? row0 : 0) * ldi, row0 < rows ? mi : 0);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int kk = tid + q * T + shift;
        w.c0[q] = (V)row[kk];
        w.c2[q] = (V)row[kk - 2];
      }
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int e = tid + q * T;
        const int p = e / n, r = row0 + p, kk = e - p * n + shift;
        const bool ok = e < P * n && r < rows && kk < n;
        const Tin* row = base + (size_t)r * ldi;
        w.c0[q] = (ok && kk < mi) ? (V)row[kk] : (V)0;
        w.c2[q] = (ok && kk >= 2 && kk - 2 < mi) ? (V)row[kk - 2] : (V)0;
      }
    }
  };
  // the stencil value d c_kk + l c_kk-2 with the Dirichlet taps at kk (d = 1 below m = n - 2
  // and 0 above; l = -1 for 2 <= kk < n, 0 below): for kk >= n both rows load zeros, as the
  // reference's absent entries
  auto sval = [&](const Raw& w, int q, int kk) -> V {
    const V d = kk < n - 2 ? (V)1 : (V)0, l = (kk >= 2 && kk < n) ? (V)-1 : (V)0;
    return d * w.c0[q] + l * w.c2[q];
  };
  // even extension: element k of transform p at k and L - k
  auto put = [&](int q, V a, V b) {
    const int e = tid + q * T;
    if (e >= P * n) return;
    const int p = e / n, k = e - p * n;
    z.st(lp(p * L + k), a, b);
    if (k > 0 && k < N) z.st(lp(p * L + L - k), a, b);
  };
  // dealiased forward transform of the pair (cs0, cs1) into out slots fa (and fa + 1 with hasb)
  auto forward = [&](int fa, bool hasb) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      if (e < P * n) put(q, cs0[e], hasb ? cs1[e] : (V)0);
    }
    __syncthreads();
    lds_dft<L, T, G::E>(z, tw, fftlds::opaque_tid());
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      const int p = e / n, k = e - p * n, r = row0 + p;
      if (e >= P * n || r >= rows) continue;
      const V s = (V)(((k & 1) ? -invN : invN) * ((k == 0 || k == N) ? 0.5 : 1.0));
      V za, zb;
      z.ld(lp(p * L + k), za, zb);
      out[fa * so + (size_t)r * ldo + k] = (Tout)(k < cut ? za * s : (V)0);
      if (hasb) out[(fa + 1) * so + (size_t)r * ldo + k] = (Tout)(k < cut ? zb * s : (V)0);
    }
  };
  // ---- transform 0: the physical velocity (times = 0: u_k = (-1)^k w_k c_k); the rows of
  //      field 0 are loaded alongside
  Raw ra, rb;
  {
    Raw va, vb;
    load(val, 0, va);
    load(val + si, 0, vb);
    load(dx, 0, ra);
    load(val, 1, rb);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      const int k = e - (e / n) * n;
      const V w = (V)(((k & 1) ? -1.0 : 1.0) * ((k == 0 || k == N) ? 1.0 : 0.5));
      put(q, sval(va, q, k) * w, sval(vb, q, k) * w);
    }
  }
  __syncthreads();
  lds_dft<L, T, G::E>(z, tw, fftlds::opaque_tid());
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + q * T;
    const int p = e / n, j = e - p * n;
    if (e < P * n) {
      V za, zb;
      z.ld(lp(p * L + j), za, zb);
      uxs[e] = (V)(Tout)za;
      uys[e] = (V)(Tout)zb;
    }
  }
  // ---- transforms 1..3: the convective terms
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    __syncthreads();  // the previous transform's results are read
    // (1) component a: s_k of dx_f; component b: g_k of the derivative recurrence of val_f,
    //     g_0 = s_1, g_k = 2 (k + 1) s_{k+1}, g_{n-1} = 0
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      if (e >= P * n) continue;
      const int p = e / n, k = e - p * n;
      V b = sval(rb, q, k + 1);
      if (k == n - 1) b = 0;
      else if (k > 0) b = (V)(2.0 * (double)(k + 1)) * b;
      z.st(lp(p * L + k), sval(ra, q, k), b);
    }
    // the next field's rows (or temp's fieldbc gradient), in flight during this transform
    V g1[Q], g2[Q];
    if (f < 2) {
      load(dx + (f + 1) * si, 0, ra);
      load(val + (f + 1) * si, 1, rb);
    } else {
      // raw buffers over the gradient rows (an absent one has an empty range and reads as
      // zero), so that the loads are issued together instead of one guarded round trip each
      const int fcnt = ldf == 0 ? n : (int)((size_t)(rows - 1) * ldf + n);
      const BufRow<double> h1(f1, f1 ? fcnt : 0), h2(f2, f2 ? fcnt : 0);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int e = tid + q * T;
        const int p = e / n, j = e - p * n, r = row0 + p;
        const bool ok = e < P * n && r < rows;
        const int o = min(r, rows - 1) * (int)ldf + min(j, n - 1);  // always within the range
        const V a1 = (V)h1[o], a2 = (V)h2[o];
        g1[q] = ok ? a1 : (V)0;
        g2[q] = ok ? a2 : (V)0;
      }
    }
    __syncthreads();
    // (2) d_k = g_k + d_{k+2} along the chains 1, 3, .., N-1 and 2, 4, .., N of component b,
    //     then d_0 = g_0 + d_2 / 2; one wavefront per chain
    {
      constexpr int C = N / 2, SEG = (C + 63) / 64, NCH = 2 * P;
      const int lane = tid & 63, wave = tid >> 6;
      for (int ch = wave; ch < NCH; ch += T / 64) {
        const int p = ch >> 1, par = ch & 1;
        const int zb = p * L + (par ? 1 : 2);
        const int t0 = lane * SEG;
        V tot = 0;
#pragma unroll
        for (int q = SEG - 1; q >= 0; --q)
          if (t0 + q < C) tot = z.i(lp(zb + 2 * (t0 + q))) + tot;
        V acc;  // the chain's sum above this lane's segment
        wave::suffix_sum(tot, lane, acc);
#pragma unroll
        for (int q = SEG - 1; q >= 0; --q) {
          const int t = t0 + q;
          if (t < C) {
            acc = z.i(lp(zb + 2 * t)) + acc;
            z.i(lp(zb + 2 * t)) = acc;
          }
        }
        if (par == 0 && lane == 0) z.i(lp(p * L)) = z.i(lp(p * L)) + acc / (V)2;
      }
    }
    __syncthreads();
    // (3) u_k = (-1)^k w_k c_k (b: / sc) and the even extension
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      if (e >= P * n) continue;
      const int p = e / n, k = e - p * n;
      V a, b;
      z.ld(lp(p * L + k), a, b);
      if (inv_sc != (V)1) b = b * inv_sc;
      const V w = (V)(((k & 1) ? -1.0 : 1.0) * ((k == 0 || k == N) ? 1.0 : 0.5));
      put(q, a * w, b * w);
    }
    __syncthreads();
    lds_dft<L, T, G::E>(z, tw, fftlds::opaque_tid());
    // (4) out_f = d/dy(val_f) .* uy + 1 (dx_f .* ux) (+ temp's fieldbc terms)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      const int p = e / n, j = e - p * n, r = row0 + p;
      V v = 0;
      if (e < P * n && r < rows) {
        const V ux = uxs[e], uy = uys[e];
        V za, zb;
        z.ld(lp(p * L + j), za, zb);
        const V xa = (V)(Tout)(za * ux);
        v = zb * uy;
        v = v + (V)1 * xa;
        if (f == 2) {
          if (f1) v = v + ux * g1[q];
          if (f2) v = v + uy * g2[q];
        }
      }
      if (e < P * n) (f == 1 ? cs1 : cs0)[e] = (V)(Tout)v;
    }
    if (f == 1) forward(0, true);   // (out_ux, out_uy)
    if (f == 2) forward(2, false);  // (out_temp, 0)
  }
}

// the same without the transform: Chebyshev coefficients of the derivative
template <int L>
__global__ __launch_bounds__(Geo<L>::T) void cheb_y_diff_kernel(const double* __restrict__ in, int ldi, int mi,
                                                                double* __restrict__ out, int ldo, int rows,
                                                                const double* __restrict__ coef, double inv_sc,
                                                                double beta, long si, long so) {
  cheb_y_bwd_body<L, false, false, double, double, double>(in, ldi, mi, out, ldo, rows, coef, 1, inv_sc, nullptr, 0, beta,
                                                     nullptr, si, so, 0,
                                                     ExtraProducts<double>{nullptr, nullptr, nullptr, nullptr, -1});
}

template <int L, typename Tin, typename V, typename Tout>
__global__ __launch_bounds__(Geo<L>::T) void cheb_y_fwd_kernel(const Tin* __restrict__ in, int ldi,
                                                               Tout* __restrict__ out, int ldo, int rows, int cut,
                                                               const V* __restrict__ tw, long si, long so) {
  using G = Geo<L>;
  constexpr int N = G::N, n = G::n, P = G::P, T = G::T;
  __shared__ __attribute__((aligned(16))) V zmem[2 * lp_size(G::E)];
  const fftlds::LdsZ<V, lp_size(G::E)> z{zmem};
  in += si * blockIdx.y;
  out += so * blockIdx.y;
  const int tid = threadIdx.x, row0 = 2 * P * blockIdx.x;
  constexpr int Q = (P * n + T - 1) / T;
  V va[Q], vb[Q];
  if constexpr (P == 1) {
    const int ra = row0, rb = ra + 1;
    const BufRow<Tin> a(in + (size_t)(ra < rows ? ra : 0) * ldi, ra < rows ? n : 0);
    const BufRow<Tin> b(in + (size_t)(rb < rows ? rb : 0) * ldi, rb < rows ? n : 0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      va[q] = (V)a[tid + q * T];
      vb[q] = (V)b[tid + q * T];
    }
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int e = tid + q * T;
      const int p = e / n, k = e - p * n;
      const int ra = row0 + 2 * p, rb = ra + 1;
      const bool ok = e < P * n;
      va[q] = (ok && ra < rows) ? (V)in[(size_t)ra * ldi + k] : (V)0;
      vb[q] = (ok && rb < rows) ? (V)in[(size_t)rb * ldi + k] : (V)0;
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = tid + q * T;
    if (e >= P * n) continue;
    const int p = e / n, k = e - p * n;
    const V a = va[q], b = vb[q];
    z.st(lp(p * L + k), a, b);
    if (k > 0 && k < N) z.st(lp(p * L + L - k), a, b);
  }
  __syncthreads();
  lds_dft<L, T, G::E>(z, tw);
  constexpr double invN = 1.0 / (double)N;  // exact: N is a power of two
  for (int e = tid; e < P * n; e += T) {
    const int p = e / n, k = e - p * n;
    const int ra = row0 + 2 * p;
    const V s = (V)(((k & 1) ? -invN : invN) * ((k == 0 || k == N) ? 0.5 : 1.0));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = ra + h;
      if (r >= rows) continue;
      const V v = z.c(h, lp(p * L + k));
      out[(size_t)r * ldo + k] = (Tout)(k < cut ? v * s : (V)0);
    }
  }
}

template <int L, typename Tin, typename V, typename Tout>
void launch_bwd(hipStream_t s, const Tin* in, int ldi, int mi, Tout* out, int ldo, int rows, const double* coef,
                const unsigned char* mask, int times, double inv_sc, const Tout* U, int ldu, double beta, int batch,
                long si, long so, long su, bool transform,
                ExtraProducts<Tout> xp = ExtraProducts<Tout>{nullptr, nullptr, nullptr, nullptr, -1}) {
  using G = Geo<L>;
  const unsigned blocks = (unsigned)((rows + 2 * G::P - 1) / (2 * G::P));
  (void)mask;  // the zero coefficients of absent taps stand for the mask
  if (transform) {
    RBC_LAUNCH((cheb_y_bwd_kernel<L, Tin, V, Tout>), dim3(blocks, batch), dim3(G::T), 0, s, in, ldi, mi, out, ldo,
               rows, coef, times, inv_sc, U, ldu, beta, twiddles<V>(L), si, so, su, xp);
  } else {
    if constexpr (std::is_same<Tin, double>::value && std::is_same<Tout, double>::value)
      RBC_LAUNCH(cheb_y_diff_kernel<L>, dim3(blocks, batch), dim3(G::T), 0, s, in, ldi, mi, out, ldo, rows, coef,
                 inv_sc, beta, si, so);
    else
      throw std::invalid_argument("cheb_y_diff: FP64 only");
  }
}

template <int L, typename Tin, typename V, typename Tout>
void launch_conv(hipStream_t s, const Tin* dx, const Tin* val, int ldi, int mi, Tout* out, int ldo, int rows,
                 const double* coef, double inv_sc, const Tout* ux, const Tout* uy, int ldu, int batch, long si, long so,
                 ExtraProducts<Tout> xp) {
  using G = Geo<L>;
  const unsigned blocks = (unsigned)((rows + G::P - 1) / G::P);
  RBC_LAUNCH((cheb_y_conv_kernel<L, Tin, V, Tout>), dim3(blocks, batch), dim3(G::T), 0, s, dx, val, ldi, mi, out, ldo,
             rows, coef, inv_sc, ux, uy, ldu, twiddles<V>(L), si, so, xp);
}

template <int L, typename Tin, typename V, typename Tout>
void launch_conv3(hipStream_t s, const Tin* dx, const Tin* val, int ldi, int mi, Tout* out, int ldo, int rows,
                  double inv_sc, const double* f1, const double* f2, long ldf, int cut, long si, long so) {
  using G = Geo<L>;
  const unsigned blocks = (unsigned)((rows + G::P - 1) / G::P);
  RBC_LAUNCH((cheb_y_conv3_kernel<L, Tin, V, Tout>), dim3(blocks), dim3(G::T), 0, s, dx, val, ldi, mi, out, ldo, rows,
             inv_sc, f1, f2, ldf, cut, twiddles<V>(L), si, so);
}

template <int L, typename Tin, typename V, typename Tout>
void launch_fwd(hipStream_t s, const Tin* in, int ldi, Tout* out, int ldo, int rows, int cut, int batch, long si,
                long so) {
  using G = Geo<L>;
  const unsigned blocks = (unsigned)((rows + 2 * G::P - 1) / (2 * G::P));
  RBC_LAUNCH((cheb_y_fwd_kernel<L, Tin, V, Tout>), dim3(blocks, batch), dim3(G::T), 0, s, in, ldi, out, ldo, rows, cut,
             twiddles<V>(L), si, so);
}

}  // namespace

bool cheb_y_supported(int n) {  // the padded LDS buffers of L = 2 (n - 1) <= 2048 stay within 64 KB
  const int N = n - 1;
  return N >= 16 && N <= 1024 && (N & (N - 1)) == 0;
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
    default: throw std::invalid_argument("cheb_y: unsupported length"); \
  }

void cheb_y_backward(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                     const double* coef, const unsigned char* mask, int times, double inv_sc, const double* U, int ldu,
                     double beta, int batch, long si, long so, long su, const double* e1, const double* f1,
                     const double* e2, const double* f2, int xslot) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_backward: unsupported length");
  if (times < 0 || times > 1) throw std::invalid_argument("cheb_y_backward: times must be 0 or 1");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_backward: bad batch or rows");
  if (mi > n || (!coef && mi < 1)) throw std::invalid_argument("cheb_y_backward: bad input length");
#define RBC_BWD(LL)                                                                                          \
  launch_bwd<LL, double, double, double>(s, in, ldi, mi, out, ldo, rows, coef, mask, times, inv_sc, U, ldu, beta, \
                                         batch, si, so, su, true, ExtraProducts<double>{e1, f1, e2, f2, xslot})
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_BWD)
#undef RBC_BWD
  check("cheb_y_backward");
}

void cheb_y_conv(hipStream_t s, const double* dx, const double* val, int ldi, int mi, double* out, int ldo, int rows,
                 int n, const double* coef, double inv_sc, const double* ux, const double* uy, int ldu, int batch,
                 long si, long so, const double* e1, const double* f1, const double* e2, const double* f2, int xslot,
                 long ldf) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_conv: unsupported length");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_conv: bad batch or rows");
  if (mi > n || (!coef && mi < 1) || !ux || !uy) throw std::invalid_argument("cheb_y_conv: bad arguments");
#define RBC_CONV(LL)                                                                                            \
  launch_conv<LL, double, double, double>(s, dx, val, ldi, mi, out, ldo, rows, coef, inv_sc, ux, uy, ldu, batch, si, \
                                          so, ExtraProducts<double>{e1, f1, e2, f2, xslot, ldf})
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_CONV)
#undef RBC_CONV
  check("cheb_y_conv");
}

void cheb_y_conv_f32(hipStream_t s, const float* dx, const float* val, int ldi, int mi, float* out, int ldo, int rows,
                     int n, const double* coef, double inv_sc, const float* ux, const float* uy, int ldu, int batch,
                     long si, long so, const float* e1, const double* f1, const float* e2, const double* f2,
                     int xslot, long ldf) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_conv_f32: unsupported length");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_conv_f32: bad batch or rows");
  if (mi > n || (!coef && mi < 1) || !ux || !uy) throw std::invalid_argument("cheb_y_conv_f32: bad arguments");
#define RBC_CONV32(LL)                                                                                       \
  launch_conv<LL, float, float, float>(s, dx, val, ldi, mi, out, ldo, rows, coef, inv_sc, ux, uy, ldu, batch, si, \
                                       so, ExtraProducts<float>{e1, f1, e2, f2, xslot, ldf})
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_CONV32)
#undef RBC_CONV32
  check("cheb_y_conv_f32");
}

namespace {
void conv3_check(int n, int rows, int mi, const void* f2, long ldf, int cut, long si, long so, int ldo) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_conv3: unsupported length");
  if (rows < 1 || mi != n - 2) throw std::invalid_argument("cheb_y_conv3: bad rows or input length");
  if (ldf < 0 || (f2 == nullptr && ldf != 0)) throw std::invalid_argument("cheb_y_conv3: bad fieldbc stride");
  if (((double)(rows - 1) * (double)ldf + n) * 8.0 > 2147483647.0)  // the gradient rows' byte range, as an int
    throw std::invalid_argument("cheb_y_conv3: fieldbc rows too large");
  if (cut < 0 || cut > n || si < 0 || so < 0 || ldo < n) throw std::invalid_argument("cheb_y_conv3: bad output layout");
}
}  // namespace

void cheb_y_conv3_f32(hipStream_t s, const float* dx, const float* val, int ldi, int mi, float* out, int ldo, int rows,
                      int n, double inv_sc, const double* f1, const double* f2, long ldf, int cut, long si, long so) {
  conv3_check(n, rows, mi, f2, ldf, cut, si, so, ldo);
#define RBC_CONV3F(LL) \
  launch_conv3<LL, float, float, float>(s, dx, val, ldi, mi, out, ldo, rows, inv_sc, f1, f2, ldf, cut, si, so)
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_CONV3F)
#undef RBC_CONV3F
  check("cheb_y_conv3_f32");
}

void cheb_y_diff(hipStream_t s, const double* in, int ldi, int mi, double* out, int ldo, int rows, int n,
                 const double* coef, double inv_sc, double beta, int batch, long si, long so) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_diff: unsupported length");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_diff: bad batch or rows");
  if (mi > n || (!coef && mi < 1)) throw std::invalid_argument("cheb_y_diff: bad input length");
#define RBC_DIFF(LL)                                                                                             \
  launch_bwd<LL, double, double, double>(s, in, ldi, mi, out, ldo, rows, coef, nullptr, 1, inv_sc, nullptr, 0, beta, \
                                         batch, si, so, 0, false)
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_DIFF)
#undef RBC_DIFF
  check("cheb_y_diff");
}

void cheb_y_forward(hipStream_t s, const double* in, int ldi, double* out, int ldo, int rows, int n, int cut, int batch,
                    long si, long so) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_forward: unsupported length");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_forward: bad batch or rows");
#define RBC_FWD(LL) launch_fwd<LL, double, double, double>(s, in, ldi, out, ldo, rows, cut, batch, si, so)
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_FWD)
#undef RBC_FWD
  check("cheb_y_forward");
}

void cheb_y_backward_f32(hipStream_t s, const float* in, int ldi, int mi, float* out, int ldo, int rows, int n,
                         const double* coef, int times, double inv_sc, const float* U, int ldu, double beta, int batch,
                         long si, long so, long su, const float* e1, const double* f1, const float* e2,
                         const double* f2, int xslot) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_backward_f32: unsupported length");
  if (times < 0 || times > 1) throw std::invalid_argument("cheb_y_backward_f32: times must be 0 or 1");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_backward_f32: bad batch or rows");
  if (mi > n || (!coef && mi < 1)) throw std::invalid_argument("cheb_y_backward_f32: bad input length");
#define RBC_BWD32(LL)                                                                                              \
  launch_bwd<LL, float, float, float>(s, in, ldi, mi, out, ldo, rows, coef, nullptr, times, inv_sc, U, ldu, beta, \
                                      batch, si, so, su, true, ExtraProducts<float>{e1, f1, e2, f2, xslot})
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_BWD32)
#undef RBC_BWD32
  check("cheb_y_backward_f32");
}

void cheb_y_forward_f32(hipStream_t s, const float* in, int ldi, float* out, int ldo, int rows, int n, int cut,
                        int batch, long si, long so) {
  if (!cheb_y_supported(n)) throw std::invalid_argument("cheb_y_forward_f32: unsupported length");
  if (batch < 1 || batch > 65535 || rows < 1) throw std::invalid_argument("cheb_y_forward_f32: bad batch or rows");
#define RBC_FWD32(LL) launch_fwd<LL, float, float, float>(s, in, ldi, out, ldo, rows, cut, batch, si, so)
  RBC_CHEB_Y_SWITCH(2 * (n - 1), RBC_FWD32)
#undef RBC_FWD32
  check("cheb_y_forward_f32");
}

}  // namespace rbc
