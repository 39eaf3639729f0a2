// gr_plane.hpp -- device helpers shared by the fused passes.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#ifndef GR_STREAMING_HINTS
#define GR_STREAMING_HINTS 1
#endif

namespace gr {

// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>) in order, stopping
// after the first call that returns false: a loop unrolled by construction.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (void)(f(std::integral_constant<int, I>{}) && ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Plane accesses through buffer resources: the plane base and size live in a
// scalar descriptor and each lane supplies a 32-bit byte offset, so no per-lane
// 64-bit address arithmetic is spent on the dozen loads of every plane step.
// Word 3 is the raw-buffer format of CDNA (32-bit data, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
// NT: non-temporal load (the same cache-policy bit as bstore).
template <typename T, bool NT = false>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned off) {
  constexpr int aux = (NT && GR_STREAMING_HINTS) ? 2 : 0;
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, aux));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, aux));
}
// NT: non-temporal store (the CDNA cache-policy bit NT = 2).
template <bool NT, typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  constexpr int aux = (NT && GR_STREAMING_HINTS) ? 2 : 0;
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, aux);
  else
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                          (int)off, 0, aux);
}

// A coefficient or roc2 word of storage type C at element offset e of a plane
// resource, widened to T (bf16 widens exactly: its bits move to the top half).
// NT: non-temporal load (the same cache-policy bit as bstore).
template <typename T, typename C, bool NT>
__device__ __forceinline__ T bload_coef(__amdgpu_buffer_rsrc_t r, unsigned e) {
  constexpr int aux = (NT && GR_STREAMING_HINTS) ? 2 : 0;
  if constexpr (sizeof(C) == 2) {
    const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)(e * 2u), 0, aux);
    return static_cast<T>(__uint_as_float((unsigned)b << 16));
  } else if constexpr (sizeof(C) == 4) {
    return static_cast<T>(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(e * 4u), 0, aux)));
  } else {
    return static_cast<T>(__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(e * 8u), 0, aux)));
  }
}

}  // namespace gr
