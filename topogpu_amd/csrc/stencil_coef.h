// Storage of coarse-level stencil coefficients: FP32, or BF16 bits (round to
// nearest even of the FP32 value; a BF16 value widens to FP32 exactly).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

__device__ __forceinline__ void coef_store(float* p, float v) { *p = v; }
__device__ __forceinline__ void coef_store(uint16_t* p, float v) {
  uint32_t u = __builtin_bit_cast(uint32_t, v);
  u += 0x7fffu + ((u >> 16) & 1u);   // finite values only (sums of finite element entries)
  *p = (uint16_t)(u >> 16);
}
__device__ __forceinline__ float coef_load(const float* p) { return *p; }
__device__ __forceinline__ float coef_load(const uint16_t* p) {
  return __builtin_bit_cast(float, (uint32_t)*p << 16);
}

// a * b rounded to FP32 on its own, never fused into a later addition: the
// prescaled input of the fused smoothing start feeds both the products
// (explicit fmaf) and the epilogue's x = d0 + d, and every kernel must round
// it the same way.
__device__ __forceinline__ float mul_rounded(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}

}  // namespace tg
