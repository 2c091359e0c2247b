// In-LDS complex DFTs for the FFT-based kernels (FP64, or FP32 in the mixed-precision
// step; T is the arithmetic type): radix-8 Stockham passes plus one
// radix-4 or radix-2 pass, natural order in and out, forward sign (e^{-2 pi i jk/L}).
// A block holds E / L transforms of length L at stride L in one complex buffer (LdsZ); each of
// its T threads handles E / T = 8 values per pass.  Twiddles come from a table of
// e^{-2 pi i t / L}, t < L, interleaved (re, im), in the arithmetic type or FP64.
#pragma once
#include <hip/hip_runtime.h>

namespace rbc {
namespace fftlds {

// LDS index of element a of the transform buffers: three pad elements per 32.  The LDS
// serves a wavefront's access in half-waves of 32 lanes, and lanes whose element indices
// agree mod 32 conflict (b32 and b64 accesses alike, measured on MI355X).  With this
// padding the passes' reads (32 consecutive elements) and the writes of stride 8 are
// conflict-free, and the writes of the second radix-8 pass (64 u + k, u < 4, k < 8)
// conflict two-way; one pad per 16 conflicts two-way on every access.
__host__ __device__ constexpr int lp(int a) { return a + 3 * (a >> 5); }
// LDS elements of a buffer of E values
__host__ __device__ constexpr int lp_size(int E) { return lp(E) + 1; }

// A block's complex buffer of NE elements in LDS: (re, im) pairs (FP32: one 8-byte access
// per element, which moves twice the bytes of a 4-byte access in the same LDS time) or two
// arrays re[NE], im[NE] (FP64).  m points at 2 NE values, aligned to 16 bytes.
template <typename V>
struct alignas(2 * sizeof(V)) cpair {
  V r, i;
};
#ifndef RBC_LDS_PAIRS_FP64
#define RBC_LDS_PAIRS_FP64 0
#endif
template <typename V, int NE, bool PAIRS = (sizeof(V) == 4 || RBC_LDS_PAIRS_FP64)>
struct LdsZ {
  using value_type = V;
  V* m;
  __device__ __forceinline__ V& r(int e) const { return PAIRS ? m[2 * e] : m[e]; }
  __device__ __forceinline__ V& i(int e) const { return PAIRS ? m[2 * e + 1] : m[NE + e]; }
  __device__ __forceinline__ V& c(int comp, int e) const { return PAIRS ? m[2 * e + comp] : m[comp * NE + e]; }
  __device__ __forceinline__ void ld(int e, V& a, V& b) const {
    if constexpr (PAIRS) {
      const cpair<V> v = reinterpret_cast<const cpair<V>*>(m)[e];
      a = v.r;
      b = v.i;
    } else {
      a = m[e];
      b = m[NE + e];
    }
  }
  __device__ __forceinline__ void st(int e, V a, V b) const {
    if constexpr (PAIRS) {
      reinterpret_cast<cpair<V>*>(m)[e] = cpair<V>{a, b};
    } else {
      m[e] = a;
      m[NE + e] = b;
    }
  }
};

constexpr double SQH = 0.70710678118654752440;  // sqrt(1/2), rounded to T where used

// forward DFTs of 2, 4, 8 points (sign -1), in place, natural order
template <typename T>
__device__ __forceinline__ void dft2(T* r, T* i) {
  const T ar = r[0], ai = i[0];
  r[0] = ar + r[1];
  i[0] = ai + i[1];
  r[1] = ar - r[1];
  i[1] = ai - i[1];
}
template <typename T>
__device__ __forceinline__ void dft4(T& r0, T& i0, T& r1, T& i1, T& r2, T& i2, T& r3, T& i3) {
  const T t0r = r0 + r2, t0i = i0 + i2, t1r = r0 - r2, t1i = i0 - i2;
  const T t2r = r1 + r3, t2i = i1 + i3;
  const T t3r = i1 - i3, t3i = r3 - r1;  // (x1 - x3) (-i)
  r0 = t0r + t2r;
  i0 = t0i + t2i;
  r2 = t0r - t2r;
  i2 = t0i - t2i;
  r1 = t1r + t3r;
  i1 = t1i + t3i;
  r3 = t1r - t3r;
  i3 = t1i - t3i;
}
template <typename T>
__device__ __forceinline__ void dft8(T* r, T* i) {
  const T h = (T)SQH;
  dft4(r[0], i[0], r[2], i[2], r[4], i[4], r[6], i[6]);  // E_k in slots 0, 2, 4, 6
  dft4(r[1], i[1], r[3], i[3], r[5], i[5], r[7], i[7]);  // O_k in slots 1, 3, 5, 7
  // O_1 *= e^{-i pi/4}, O_2 *= -i, O_3 *= e^{-3 i pi/4}
  T o1r = h * (r[3] + i[3]), o1i = h * (i[3] - r[3]);
  T o2r = i[5], o2i = -r[5];
  T o3r = h * (i[7] - r[7]), o3i = -h * (r[7] + i[7]);
  const T e0r = r[0], e0i = i[0], e1r = r[2], e1i = i[2], e2r = r[4], e2i = i[4], e3r = r[6], e3i = i[6];
  const T o0r = r[1], o0i = i[1];
  r[0] = e0r + o0r; i[0] = e0i + o0i;
  r[4] = e0r - o0r; i[4] = e0i - o0i;
  r[1] = e1r + o1r; i[1] = e1i + o1i;
  r[5] = e1r - o1r; i[5] = e1i - o1i;
  r[2] = e2r + o2r; i[2] = e2i + o2i;
  r[6] = e2r - o2r; i[6] = e2i - o2i;
  r[3] = e3r + o3r; i[3] = e3i + o3i;
  r[7] = e3r - o3r; i[7] = e3i - o3i;
}

// one Stockham pass of radix R over the transforms of length L held in z (element t of
// transform p at p S + lp(t)), the pass after sub-transforms of length NS (NS = 1, 8, 64, ..).
// Every element a thread touches sits at its base index plus a constant, since
// lp(a + c) = lp(a) + lp(c) whenever (a mod 32) + (c mod 32) < 32: the reads take a = j
// (j < L / R) and c = r L / R; the writes a = (j / NS) NS R + k (k < NS) and c = r NS, and
// for NS = 1 (a mod 32 a multiple of R, r < R) and NS = 8 (a mod 32 = k < 8 when 8 R is a
// multiple of 32, c mod 32 = 8 r; R = 2: a mod 32 < 24, c < 16) that holds as it does for NS
// a multiple of 32.  So each thread forms two LDS indices per butterfly and the rest are
// immediate offsets.
template <int L, int R, int T, int E, int NS, int S, typename ZB, typename TW>
__device__ __forceinline__ void stockham_pass(const ZB& z, const TW* __restrict__ tw, int tid) {
  using V = typename ZB::value_type;
  static_assert(E % (R * T) == 0 && E % L == 0 && (L % 16 == 0 || L == 8), "block geometry");
  static_assert(NS == 1 || NS == 8 || NS % 32 == 0, "pass stride");
  constexpr int NB = L / R;          // butterflies per transform
  constexpr int PER = E / (R * T);   // butterflies per thread
  V xr[PER][R], xi[PER][R];
  int dst[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int g = tid + q * T;
    const int p = g / NB, j = g - p * NB;
    const int k = j & (NS - 1);
    const int src = p * S + lp(j);
#pragma unroll
    for (int r = 0; r < R; ++r) z.ld(src + lp(r * NB), xr[q][r], xi[q][r]);
    if constexpr (NS > 1) {
      // twiddles e^{-2 pi i r k / (NS R)} = table[r k L / (NS R)]: powers 1, 2, 4 from the
      // table, the others as products of two of them (the first pass has k = 0: none)
      const int t1 = k * (L / (NS * R));
      V wr[R], wi[R];
      wr[0] = 1;
      wi[0] = 0;
      if constexpr (R >= 2) { wr[1] = (V)tw[2 * t1]; wi[1] = (V)tw[2 * t1 + 1]; }
      if constexpr (R >= 4) {
        wr[2] = (V)tw[4 * t1]; wi[2] = (V)tw[4 * t1 + 1];
        wr[3] = wr[1] * wr[2] - wi[1] * wi[2]; wi[3] = wr[1] * wi[2] + wi[1] * wr[2];
      }
      if constexpr (R == 8) {
        wr[4] = (V)tw[8 * t1]; wi[4] = (V)tw[8 * t1 + 1];
        wr[5] = wr[1] * wr[4] - wi[1] * wi[4]; wi[5] = wr[1] * wi[4] + wi[1] * wr[4];
        wr[6] = wr[2] * wr[4] - wi[2] * wi[4]; wi[6] = wr[2] * wi[4] + wi[2] * wr[4];
        wr[7] = wr[3] * wr[4] - wi[3] * wi[4]; wi[7] = wr[3] * wi[4] + wi[3] * wr[4];
      }
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const V ar = xr[q][r], ai = xi[q][r];
        xr[q][r] = ar * wr[r] - ai * wi[r];
        xi[q][r] = ar * wi[r] + ai * wr[r];
      }
    }
    if constexpr (R == 8) dft8(xr[q], xi[q]);
    else if constexpr (R == 4) dft4(xr[q][0], xi[q][0], xr[q][1], xi[q][1], xr[q][2], xi[q][2], xr[q][3], xi[q][3]);
    else dft2(xr[q], xi[q]);
    dst[q] = p * S + lp((j / NS) * NS * R + k);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int r = 0; r < R; ++r) z.st(dst[q] + lp(r * NS), xr[q][r], xi[q][r]);
  __syncthreads();
}

// the passes from sub-transform length NS on: radix 8 while it divides, then one radix-4 or
// radix-2 pass
template <int L, int T, int E, int NS, int S, typename ZB, typename TW>
__device__ __forceinline__ void dft_passes(const ZB& z, const TW* __restrict__ tw, int tid) {
  if constexpr (NS * 8 <= L) {
    stockham_pass<L, 8, T, E, NS, S>(z, tw, tid);
    dft_passes<L, T, E, NS * 8, S>(z, tw, tid);
  } else if constexpr (NS * 4 == L) {
    stockham_pass<L, 4, T, E, NS, S>(z, tw, tid);
  } else if constexpr (NS * 2 == L) {
    stockham_pass<L, 2, T, E, NS, S>(z, tw, tid);
  }
}

// tid: the thread's index in the block (a caller that runs several transforms can pass a
// value the compiler cannot prove equal between them, so that each transform forms its
// LDS addresses afresh instead of holding one set live across the others)
// transform p of the block starts at element p S of the buffer (S >= lp(L - 1) + 1)
template <int L, int T, int E, int S = lp(L), typename ZB, typename TW>
__device__ __forceinline__ void lds_dft(const ZB& z, const TW* __restrict__ tw, int tid = threadIdx.x) {
  dft_passes<L, T, E, 1, S>(z, tw, tid);
}

// The stride between the transforms of a block whose threads also touch the same element
// of several transforms at once (the column kernels' loads and stores: lanes of one row run
// over the columns, i.e. over the transforms): S = 4 mod 32 puts element t of consecutive
// transforms on banks 4 apart (8 apart in 32-bit words for 8-byte pairs), where the plain
// stride lp(L) = 0 mod 32 would put them all on one bank.  Lengths below 256 (many short
// transforms per block) keep the plain stride, which costs no LDS.
constexpr int spread_stride(int L) { return L >= 256 ? lp(L) + ((4 - lp(L)) % 32 + 32) % 32 : lp(L); }
// LDS elements of PT transforms at stride S
constexpr int spread_size(int L, int PT) { return (PT - 1) * spread_stride(L) + lp(L - 1) + 1; }

// threadIdx.x behind an empty asm statement: a value the compiler cannot merge with other
// uses of the thread index
__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// block index remapped so that consecutive logical blocks run on the same XCD (blocks
// are dispatched round-robin over the 8 XCDs): neighbouring column groups then share
// that XCD's L2 lines
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int x = b & 7, s = b >> 3, q = nb >> 3, r = nb & 7;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + s;
}

}  // namespace fftlds

// e^{-2 pi i t / L}, t < L, interleaved, cached per (device, L) (defined in cheb_y.hip);
// the FP32 table holds the FP64 table's values rounded to FP32, so that the FP32 kernels
// read the twiddles they would convert
const double* fft_table(int L);
const float* fft_table_f32(int L);
// the table in the arithmetic type V of a kernel
template <typename V>
inline const V* twiddles(int L) {
  if constexpr (sizeof(V) == sizeof(float)) return fft_table_f32(L);
  else return fft_table(L);
}

}  // namespace rbc
