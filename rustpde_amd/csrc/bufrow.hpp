// Raw buffer access to one row of values.
#pragma once
#include <hip/hip_runtime.h>

namespace rbc {

// A raw buffer over one row of `count` values: element i reads as the value for 0 <= i <
// count and as 0 outside (the buffer's range check on the byte offset, which a negative i
// exceeds as an unsigned value), so guarded loads take no branch or select.  The row must be
// the same for the whole wavefront: the descriptor lives in scalar registers.  The word
// 0x00020000 is the descriptor's format word for raw (untyped, unswizzled) 32-bit access on
// the GFX9 family.
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
};

}  // namespace rbc
