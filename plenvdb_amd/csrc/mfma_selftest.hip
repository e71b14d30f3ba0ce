// Device check of the MFMA operand layouts documented in mfma_util.h: one wave
// multiplies a 16x32 by a 32x16 bf16 matrix (and a 16x16 by 16x16 one) using
// the assumed fragment layouts; the host compares with a reference product.
#include "mfma_util.h"
#include "pvdb_common.h"
#include "plenvdb_hip.h"

namespace {

__global__ void __launch_bounds__(64) k_mfma_selftest(const uint16_t* __restrict__ A,
                                                      const uint16_t* __restrict__ B,
                                                      float* __restrict__ D,
                                                      const uint16_t* __restrict__ A16,
                                                      const uint16_t* __restrict__ B16,
                                                      float* __restrict__ D16) {
    const int l = threadIdx.x;
    const int r = l & 15, g = l >> 4;
    pvdb::bf16x8 a, b;
    for (int j = 0; j < 8; ++j) {
        uint16_t av = A[r * 32 + 8 * g + j];
        uint16_t bv = B[(8 * g + j) * 16 + r];
        a[j] = __builtin_bit_cast(__bf16, av);
        b[j] = __builtin_bit_cast(__bf16, bv);
    }
    pvdb::f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = pvdb::mfma16x16x32(a, b, c);
    for (int i = 0; i < 4; ++i) D[(4 * g + i) * 16 + r] = c[i];
    pvdb::s16x4 a4, b4;
    for (int j = 0; j < 4; ++j) {
        a4[j] = (short)A16[r * 16 + 4 * g + j];
        b4[j] = (short)B16[(4 * g + j) * 16 + r];
    }
    pvdb::f32x4 c4 = {0.f, 0.f, 0.f, 0.f};
    c4 = pvdb::mfma16x16x16(a4, b4, c4);
    for (int i = 0; i < 4; ++i) D16[(4 * g + i) * 16 + r] = c4[i];
}

__global__ void __launch_bounds__(64) k_mfma_selftest_f16(const uint16_t* __restrict__ A,
                                                          const uint16_t* __restrict__ B,
                                                          float* __restrict__ D) {
    const int l = threadIdx.x;
    const int r = l & 15, g = l >> 4;
    pvdb::f16x8 a, b;
    for (int j = 0; j < 8; ++j) {
        a[j] = __builtin_bit_cast(_Float16, A[r * 32 + 8 * g + j]);
        b[j] = __builtin_bit_cast(_Float16, B[(8 * g + j) * 16 + r]);
    }
    pvdb::f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = pvdb::mfma16x16x32h(a, b, c);
    for (int i = 0; i < 4; ++i) D[(4 * g + i) * 16 + r] = c[i];
}

}  // namespace

extern "C" int pvdb_mfma_selftest_f16(const uint16_t* A, const uint16_t* B, float* D,
                                      void* stream) {
    PVDB_CHECK_ARG(A && B && D, "null pointer");
    hipLaunchKernelGGL(k_mfma_selftest_f16, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
    PVDB_LAUNCHED("k_mfma_selftest_f16");
    return PVDB_OK;
}

extern "C" int pvdb_mfma_selftest(const uint16_t* A, const uint16_t* B, float* D,
                                  const uint16_t* A16, const uint16_t* B16, float* D16,
                                  void* stream) {
    PVDB_CHECK_ARG(A && B && D && A16 && B16 && D16, "null pointer");
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D, A16,
                       B16, D16);
    PVDB_LAUNCHED("k_mfma_selftest");
    return PVDB_OK;
}
