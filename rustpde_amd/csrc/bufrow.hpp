// Raw buffer access to one row of values.
#pragma once
#include <hip/hip_runtime.h>

namespace rbc {

// A raw buffer over one row of `count` values: element i reads as the value for 0 <= i <
// count and as 0 outside (the buffer's range check on the byte offset, which a negative i
// exceeds as an unsigned value), and a store outside is dropped, so guarded accesses take no
// branch or select.  The row must be the same for the whole wavefront: the descriptor lives
// in scalar registers.
template <typename T>
struct BufRow {
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ BufRow(const T* row, int count)
      : rs(__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(row), 0, count * (int)sizeof(T), 0x00020000)) {}
  __device__ __forceinline__ T operator[](int i) const {
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, i * 4, 0, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, i * 8, 0, 0));
  }
  __device__ __forceinline__ void store(int i, T v) const {
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), rs, i * 4, 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned int, v),
                                            rs, i * 8, 0, 0);
  }
};

// a pointer that is the same across the wavefront, as the compiler's scalar value
template <typename T>
__device__ __forceinline__ T* wave_uniform(T* p) {
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u), hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace rbc
