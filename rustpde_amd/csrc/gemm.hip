// GEMM on the CDNA4 matrix cores, FP64 or FP32: C = alpha (A B) [.* U] + beta C, row-major.
//
// A workgroup of 4 wavefronts computes a BM x BN tile (2 x 2 wavefronts, each a
// (BM/2) x (BN/2) sub-tile of 16 x 16 MFMA blocks).  K is staged through LDS in BK-deep
// slabs; the next slab is fetched into registers while the current one is consumed.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>

#include "fft_lds.hpp"
#include "ops.hpp"

namespace rbc {

namespace {

// The two matrix-core products: FP64 v_mfma_f64_16x16x4f64 (result register r of lane l:
// D[4 r + l / 16][l % 16]) and FP32 v_mfma_f32_16x16x4f32 (D[4 (l / 16) + r][l % 16]); both
// take A[l % 16][l / 16] and B[l / 16][l % 16] from lane l.
template <class V>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 fma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int lk) { return 4 * r + lk; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 fma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int lk) { return 4 * lk + r; }
};

// V: arithmetic type (A, U and the LDS tiles); B may be FP64 data read into an FP32 product
// (converted as it is staged); C is V or FP64.
// Batch entry z reads A at sA (z % amod), B at sB (z % bmod), U at sU (z % umod) (a modulus
// of 0: z itself) and writes C at sC z: operands shared by groups of entries (an operator
// per instance, used by several fields of it) need no copies.  Entries z >= split (when
// split > 0) take A2 / B2 (when given) in place of A / B: a second field with its own
// operators in the same launch.
template <int BM, int BN, int BK, class V, class TB, class TC>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(int M, int N, int K, const V* __restrict__ A, int lda, long sA,
                                                        const TB* __restrict__ B, int ldb, long sB, TC* __restrict__ C,
                                                        int ldc, long sC, double alpha, double beta,
                                                        const V* __restrict__ U, int ldu, long sU, int amod, int bmod,
                                                        int umod, const V* __restrict__ A2, const TB* __restrict__ B2,
                                                        int split, int zero00) {
  using MF = Mfma<V>;
  typedef typename MF::v4 v4;
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / 256, NB = BK * BN / 256;
  __shared__ V As[BK][BM + 1];
  __shared__ V Bs[BK][BN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm0 = (wid >> 1) * WM, wn0 = (wid & 1) * WN;
  // Tile (bx, by) of entry z from the remapped block index: consecutive logical blocks run
  // on one XCD, so the tiles that share an operand panel (the M tiles of one N panel and
  // the N tiles of one M panel of an entry) read it through the same L2.
  const int gx = gridDim.x, gy = gridDim.y;
  const int lg = fftlds::xcd_remap(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
  const int bx = lg % gx, by = (lg / gx) % gy;
  const int m0 = by * BM, n0 = bx * BN;
  const int lr = lane & 15, lk = lane >> 4;
  const int z = lg / (gx * gy);  // batch entry
  if (split > 0 && z >= split) {  // the second operator pair
    if (A2) A = A2;
    if (B2) B = B2;
  }
  A += sA * (amod ? z % amod : z);
  B += sB * (bmod ? z % bmod : z);
  C += sC * z;
  if (U) U += sU * (umod ? z % umod : z);
  v4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4{0, 0, 0, 0};
  V ra[NA], rb[NB];
  // B loads that convert are all issued, from element (0, 0) where the entry lies outside (a
  // guarded load followed by its conversion would wait for its own round trip)
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + q * 256, mm = e / BK, kk = e % BK, gm = m0 + mm, gk = k0 + kk;
      ra[q] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : (V)0;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + q * 256, kk = e / BN, nn = e % BN, gk = k0 + kk, gn = n0 + nn;
      const bool ok = gk < K && gn < N;
      if constexpr (sizeof(TB) != sizeof(V)) {
        const TB x = B[ok ? (size_t)gk * ldb + gn : 0];
        rb[q] = ok ? (V)x : (V)0;
      } else {
        rb[q] = ok ? (V)B[(size_t)gk * ldb + gn] : (V)0;
      }
    }
  };
  gload(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + q * 256;
      As[e % BK][e / BK] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + q * 256;
      Bs[e / BN][e % BN] = rb[q];
    }
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      V af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[kk + lk][wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[kk + lk][wn0 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::fma(af[i], bf[j], acc[i][j]);
    }
  }
  const V va = (V)alpha, vb = (V)beta;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    // the epilogue's operands U and C of this row of MFMA blocks, all loads issued before the
    // first is used (from element (0, 0) where an entry lies outside: a guarded load would wait
    // for its own round trip); the branches on U and beta are uniform
    V uv[FN][4], cv[FN][4];
    if (U) {
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm0 + i * 16 + MF::row(r, lk), gn = n0 + wn0 + j * 16 + lr;
          uv[j][r] = U[gm < M && gn < N ? (size_t)gm * ldu + gn : 0];
        }
    }
    if (beta != 0.0) {
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm0 + i * 16 + MF::row(r, lk), gn = n0 + wn0 + j * 16 + lr;
          cv[j][r] = (V)C[gm < M && gn < N ? (size_t)gm * ldc + gn : 0];
        }
    }
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm0 + i * 16 + MF::row(r, lk), gn = n0 + wn0 + j * 16 + lr;
        if (gm < M && gn < N) {
          V v = va * acc[i][j][r];
          if (U) v = v * uv[j][r];
          if (beta != 0.0) v = v + vb * cv[j][r];
          if (zero00 && gm == 0 && gn == 0) v = (V)0;
          C[(size_t)gm * ldc + gn] = (TC)v;
        }
      }
  }
}

// Tile shapes of the launch (BM x BN); the accumulation order over K is the same for all
// of them, so the choice never changes a result bit.
enum { TILE_64x64, TILE_32x32, TILE_64x32, TILE_32x64, NTILES };

template <int BM, int BN, int BK, class V, class TB, class TC>
void launch_tile(hipStream_t s, bool timed, int M, int N, int K, const V* A, int lda, long sA, const TB* B, int ldb,
                 long sB, TC* C, int ldc, long sC, int batch, double alpha, double beta, const V* U, int ldu, long sU,
                 const BatchMod& bm) {
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, batch);
  if (timed) {
    RBC_LAUNCH((gemm_mfma_kernel<BM, BN, BK, V, TB, TC>), grid, dim3(256), 0, s, M, N, K, A, lda, sA, B, ldb, sB, C,
               ldc, sC, alpha, beta, U, ldu, sU, bm.a, bm.b, bm.u, static_cast<const V*>(bm.a2),
               static_cast<const TB*>(bm.b2), bm.split, bm.zero00);
  } else {
    hipLaunchKernelGGL((gemm_mfma_kernel<BM, BN, BK, V, TB, TC>), grid, dim3(256), 0, s, M, N, K, A, lda, sA, B, ldb,
                       sB, C, ldc, sC, alpha, beta, U, ldu, sU, bm.a, bm.b, bm.u, static_cast<const V*>(bm.a2),
                       static_cast<const TB*>(bm.b2), bm.split, bm.zero00);
  }
}

template <class V, class TB, class TC>
void launch_variant(int v, hipStream_t s, bool timed, int M, int N, int K, const V* A, int lda, long sA, const TB* B,
                    int ldb, long sB, TC* C, int ldc, long sC, int batch, double alpha, double beta, const V* U,
                    int ldu, long sU, const BatchMod& bm) {
#define RBC_TILE(BM, BN, BK)                                                                                       \
  launch_tile<BM, BN, BK, V, TB, TC>(s, timed, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, alpha, beta, U, \
                                     ldu, sU, bm)
  switch (v) {
    case TILE_64x64: RBC_TILE(64, 64, 32); break;
    case TILE_32x32: RBC_TILE(32, 32, 32); break;
    case TILE_64x32: RBC_TILE(64, 32, 32); break;
    default: RBC_TILE(32, 64, 32); break;
  }
#undef RBC_TILE
}

// Autotuner: the first launch of a shape (device, M, N, K, batch, types) outside stream capture
// times every tile shape on the caller's operands, writing a scratch output, and keeps the
// fastest for that shape; under capture, or when timing fails, a shape without an entry
// gets the rule of thumb (64 x 64 tiles while they give four workgroups per CU, else
// 32 x 32).
std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int, int, int, int>, int> g_tune;

thread_local int g_no_tune = 0;

int tile_rule(int M, int N, int batch) {
  const long tiles64 = (long)((N + 63) / 64) * ((M + 63) / 64) * batch;
  return tiles64 >= 1024 ? TILE_64x64 : TILE_32x32;
}

template <class V, class TB, class TC>
int tile_for(hipStream_t s, int M, int N, int K, const V* A, int lda, long sA, const TB* B, int ldb, long sB, int ldc,
             long sC, int batch, const BatchMod& bm) {
  // RUSTPDE_AMD_GEMM_TILE = 0..3 forces one tile shape (for tests of each shape)
  if (const char* f = std::getenv("RUSTPDE_AMD_GEMM_TILE")) {
    const int v = std::atoi(f);
    if (v >= 0 && v < NTILES) return v;
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int types = (int)sizeof(V) + 16 * (int)sizeof(TB) + 256 * (int)sizeof(TC);
  const auto key = std::make_tuple(dev, M, N, K, batch, types);
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  if (g_no_tune) return tile_rule(M, N, batch);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipPeekAtLastError() != hipSuccess) return tile_rule(M, N, batch);  // leave a pending error to the caller
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return tile_rule(M, N, batch);
  // scratch output covering the launch's output footprint
  const size_t span = (size_t)(batch - 1) * (size_t)(sC > 0 ? sC : 0) + (size_t)(M - 1) * ldc + N;
  TC* scratch = nullptr;
  hipEvent_t ev[NTILES + 1] = {};
  int best = tile_rule(M, N, batch);
  bool ok = hipMalloc(&scratch, span * sizeof(TC)) == hipSuccess;
  for (int v = 0; ok && v <= NTILES; ++v) ok = hipEventCreate(&ev[v]) == hipSuccess;
  if (ok) {
    for (int v = 0; v < NTILES; ++v)  // warm-up of every variant
      launch_variant<V, TB, TC>(v, s, false, M, N, K, A, lda, sA, B, ldb, sB, scratch, ldc, sC, batch, 1.0, 0.0,
                                nullptr, 0, 0, bm);
    ok = hipEventRecord(ev[0], s) == hipSuccess;
    for (int v = 0; ok && v < NTILES; ++v) {
      for (int r = 0; r < 3; ++r)
        launch_variant<V, TB, TC>(v, s, false, M, N, K, A, lda, sA, B, ldb, sB, scratch, ldc, sC, batch, 1.0, 0.0,
                                  nullptr, 0, 0, bm);
      ok = hipEventRecord(ev[v + 1], s) == hipSuccess;
    }
    ok = ok && hipGetLastError() == hipSuccess && hipEventSynchronize(ev[NTILES]) == hipSuccess;
    float tbest = 0.0f;
    for (int v = 0; ok && v < NTILES; ++v) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[v], ev[v + 1]) != hipSuccess) {
        ok = false;
        break;
      }
      if (v == 0 || ms < tbest) {
        tbest = ms;
        best = v;
      }
    }
  }
  (void)hipGetLastError();
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  if (scratch) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(scratch);
  }
  if (!ok) best = tile_rule(M, N, batch);
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tune[key] = best;
  return best;
}

template <class V, class TB, class TC>
void gemm_typed(hipStream_t s, int M, int N, int K, const V* A, int lda, long sA, const TB* B, int ldb, long sB, TC* C,
                int ldc, long sC, int batch, double alpha, double beta, const V* U, int ldu, long sU,
                const BatchMod& bm) {
  if (bm.a < 0 || bm.b < 0 || bm.u < 0) throw std::runtime_error("gemm_mfma: negative batch modulus");
  if (M <= 0 || N <= 0 || batch <= 0) return;
  if (K <= 0) throw std::runtime_error("gemm_mfma: K must be positive");
  if (batch > 65535) throw std::runtime_error("gemm_mfma: batch exceeds the grid limit");
  const int v = tile_for<V, TB, TC>(s, M, N, K, A, lda, sA, B, ldb, sB, ldc, sC, batch, bm);
  launch_variant<V, TB, TC>(v, s, true, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, alpha, beta, U, ldu, sU,
                            bm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string("gemm_mfma: ") + hipGetErrorString(e));
}

}  // namespace

GemmNoTune::GemmNoTune() { ++g_no_tune; }
GemmNoTune::~GemmNoTune() { --g_no_tune; }

void gemm_mfma_batched(hipStream_t s, int M, int N, int K, const double* A, int lda, long sA, const double* B, int ldb,
                       long sB, double* C, int ldc, long sC, int batch, double alpha, double beta, const double* U,
                       int ldu, long sU, BatchMod bm) {
  gemm_typed<double, double, double>(s, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, alpha, beta, U, ldu, sU,
                                     bm);
}

void gemm_mfma_f32_batched(hipStream_t s, int M, int N, int K, const float* A, int lda, long sA, const void* B,
                           bool b_f64, int ldb, long sB, void* C, bool c_f64, int ldc, long sC, int batch, double alpha,
                           double beta, const float* U, int ldu, long sU, BatchMod bm) {
  if (b_f64 && c_f64)
    gemm_typed<float, double, double>(s, M, N, K, A, lda, sA, static_cast<const double*>(B), ldb, sB,
                                      static_cast<double*>(C), ldc, sC, batch, alpha, beta, U, ldu, sU, bm);
  else if (b_f64)
    gemm_typed<float, double, float>(s, M, N, K, A, lda, sA, static_cast<const double*>(B), ldb, sB,
                                     static_cast<float*>(C), ldc, sC, batch, alpha, beta, U, ldu, sU, bm);
  else if (c_f64)
    gemm_typed<float, float, double>(s, M, N, K, A, lda, sA, static_cast<const float*>(B), ldb, sB,
                                     static_cast<double*>(C), ldc, sC, batch, alpha, beta, U, ldu, sU, bm);
  else
    gemm_typed<float, float, float>(s, M, N, K, A, lda, sA, static_cast<const float*>(B), ldb, sB,
                                    static_cast<float*>(C), ldc, sC, batch, alpha, beta, U, ldu, sU, bm);
}

void gemm_mfma(hipStream_t s, int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C,
               int ldc, double alpha, double beta, const double* U, int ldu) {
  gemm_mfma_batched(s, M, N, K, A, lda, 0, B, ldb, 0, C, ldc, 0, 1, alpha, beta, U, ldu, 0);
}

}  // namespace rbc
