// Per-node output stage shared by every operator kernel. The operator hands
// over A u at one owned node (three components), the node's free mask and
// the node's own (masked) input values; the epilogue writes the outputs and
// returns the node's contribution to the optional grid reduction.
#pragma once

#include "common.h"

namespace tg {

struct Ep32 {
  int mode;
  const float* rhs;
  const float* dinv;
  float* d;
  float* x_out;
  float* r;
  float a1, a2, a3;
  float* y;

  __device__ __forceinline__ double operator()(long long node, unsigned m, const float (&au)[3],
                                               const float (&u)[3]) const {
    const long long b = 3 * node;
    double part = 0.0;
    switch (mode) {
      case TG_EP_STORE:
#pragma unroll
        for (int c = 0; c < 3; ++c) y[b + c] = (m >> c & 1u) ? au[c] : 0.0f;
        break;
      case TG_EP_DOT:
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float v = (m >> c & 1u) ? au[c] : 0.0f;
          y[b + c] = v;
          part += (double)u[c] * (double)v;
        }
        break;
      case TG_EP_RESID:
#pragma unroll
        for (int c = 0; c < 3; ++c) y[b + c] = (m >> c & 1u) ? rhs[b + c] - au[c] : 0.0f;
        break;
      case TG_EP_RESID_NORM:
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float v = (m >> c & 1u) ? rhs[b + c] - au[c] : 0.0f;
          y[b + c] = v;
          part += (double)v * (double)v;
        }
        break;
      case TG_EP_CHEB:
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (m >> c & 1u) {
            const float t = dinv[b + c] * (rhs[b + c] - au[c]);
            const float dn = (a2 != 0.0f) ? a1 * t + a2 * d[b + c] : a1 * t;
            d[b + c] = dn;
            x_out[b + c] = u[c] + dn;
          } else {
            d[b + c] = 0.0f;
            x_out[b + c] = 0.0f;
          }
        }
        break;
      case TG_EP_CHEB_R:
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (m >> c & 1u) {
            const float rn = r[b + c] - au[c];
            const float dn = a1 * dinv[b + c] * rn + a2 * u[c];
            r[b + c] = rn;
            d[b + c] = dn;
            x_out[b + c] += a3 * u[c] + dn;
          } else {
            r[b + c] = 0.0f;
            d[b + c] = 0.0f;
            x_out[b + c] = 0.0f;
          }
        }
        break;
      case TG_EP_RESID_UPDATE:
#pragma unroll
        for (int c = 0; c < 3; ++c) r[b + c] = (m >> c & 1u) ? r[b + c] - au[c] : 0.0f;
        break;
      default:
        break;
    }
    return part;
  }
  __host__ __device__ bool reduces() const { return mode == TG_EP_DOT || mode == TG_EP_RESID_NORM; }
};

// FP64 level-0 operator output: y = A u or y = rhs - A u, optional FP32 copy
// and sum of squares.
struct Ep64 {
  const double* rhs;
  double* y;
  float* y32;
  bool norm;

  __device__ __forceinline__ double operator()(long long node, unsigned m, const double (&au)[3],
                                               const double (&)[3]) const {
    const long long b = 3 * node;
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      if (m >> c & 1u) v = rhs ? rhs[b + c] - au[c] : au[c];
      if (y) y[b + c] = v;
      if (y32) y32[b + c] = (float)v;
      part += v * v;
    }
    return part;
  }
  __host__ __device__ bool reduces() const { return norm; }
};

}  // namespace tg
