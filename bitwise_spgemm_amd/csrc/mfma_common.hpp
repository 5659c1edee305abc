// mfma_common.hpp — helpers shared by the MFMA kernels (gfx950): 16-bit
// conversions, the 16x16x32 MFMA wrapper, the XCD-aware tile order, the
// direct-to-LDS copies (global_load_lds) with their wait, and the transposing
// LDS read.
#pragma once

#include "bws_internal.hpp"

namespace bws {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ uint32_t f32_to_bf16_rne(float x) {
  const __bf16 h = (__bf16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ uint32_t f32_to_f16_rne(float x) {
  const _Float16 h = (_Float16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ float bf16_to_f32(uint32_t u) { return __uint_as_float(u << 16); }

// PREC: 1 = BF16, 2 = FP16, 3 = BF16X2.  TB: B element type 0 = f32, 1 = bf16, 2 = f16.
template <int PREC>
__device__ __forceinline__ f32x4 mfma(const uint4& a, const uint4& b, const f32x4& c) {
  if constexpr (PREC == 2) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
}

__device__ __forceinline__ int64_t xcd_tile(int64_t g, int64_t per_xcd) {
  return (g & 7) * per_xcd + (g >> 3);
}

__device__ __forceinline__ void dma16(const void* g, uint32_t lds_addr) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(g), "s"(lds_addr)
               : "memory", "m0");
}
__device__ __forceinline__ void dma4(const void* g, uint32_t lds_addr) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dword %0, off"
               :
               : "v"(g), "s"(lds_addr)
               : "memory", "m0");
}
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
__device__ __forceinline__ uint32_t lds_u32(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
typedef short v4s __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 tr_read(uint32_t lds_addr) {
  const v4s v = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
      (v4s __attribute__((address_space(3)))*)(uintptr_t)lds_addr);
  return __builtin_bit_cast(uint2, v);
}

}  // namespace
}  // namespace bws
