// MFMA fragment helpers for gfx950 (CDNA4), wave64.
//
// v_mfma_f32_16x16x32_bf16   D[16x16] += A[16x32] * B[32x16]
//   A: lane l holds A[l%16][8*(l/16) .. 8*(l/16)+7]
//   B: lane l holds B[8*(l/16) .. +7][l%16]
//   D: lane l holds D[4*(l/16) .. +3][l%16]
// v_mfma_f32_16x16x16_bf16 (the _1k form): same with 4 values per lane,
//   A: A[l%16][4*(l/16) .. +3], B: B[4*(l/16) .. +3][l%16].
// The layouts are checked on the device by pvdb_mfma_selftest.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pvdb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16x16x32(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_f16: same operand layouts as the bf16 form.
__device__ __forceinline__ f32x4 mfma16x16x32h(const f16x8& a, const f16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma16x16x16(const s16x4& a, const s16x4& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Round-to-nearest-even float -> bf16 bits (finite inputs).
__device__ __forceinline__ uint16_t f2bf_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf_bits2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

}  // namespace pvdb
