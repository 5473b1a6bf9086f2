// Per-node output stage shared by every operator kernel.
//
// The operator hands over A u at one owned node (three components), the
// node's free mask and the node's own (masked) input values. The epilogue is
// split in two phases so that kernels can issue every global load of all
// their owned nodes before the first store: load() fetches the vectors the
// mode needs (unconditionally, independent of the mask), store() combines,
// applies the mask and writes, returning the node's contribution to the
// optional grid reduction.
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

  struct In {
    float p[3], q[3], s[3];
  };

  __device__ __forceinline__ void load(long long node, In& in) const {
    const long long b = 3 * node;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      in.p[c] = in.q[c] = in.s[c] = 0.0f;
      switch (mode) {
        case TG_EP_RESID:
        case TG_EP_RESID_NORM:
          in.p[c] = rhs[b + c];
          break;
        case TG_EP_CHEB:
          in.p[c] = rhs[b + c];
          in.q[c] = dinv[b + c];
          if (a2 != 0.0f) in.s[c] = d[b + c];
          break;
        case TG_EP_CHEB_R:
          in.p[c] = r[b + c];
          in.q[c] = dinv[b + c];
          in.s[c] = x_out[b + c];
          break;
        case TG_EP_RESID_UPDATE:
          in.p[c] = r[b + c];
          break;
        default:
          break;
      }
    }
  }

  __device__ __forceinline__ double store(long long node, unsigned m, const float (&au)[3], const float (&u)[3],
                                          const In& in) const {
    const long long b = 3 * node;
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool f = (m >> c) & 1u;
      switch (mode) {
        case TG_EP_STORE:
          y[b + c] = f ? au[c] : 0.0f;
          break;
        case TG_EP_DOT: {
          const float v = f ? au[c] : 0.0f;
          y[b + c] = v;
          part += (double)u[c] * (double)v;
          break;
        }
        case TG_EP_RESID:
          y[b + c] = f ? in.p[c] - au[c] : 0.0f;
          break;
        case TG_EP_RESID_NORM: {
          const float v = f ? in.p[c] - au[c] : 0.0f;
          y[b + c] = v;
          part += (double)v * (double)v;
          break;
        }
        case TG_EP_CHEB: {
          const float t = in.q[c] * (in.p[c] - au[c]);
          const float dn = f ? ((a2 != 0.0f) ? a1 * t + a2 * in.s[c] : a1 * t) : 0.0f;
          d[b + c] = dn;
          x_out[b + c] = f ? u[c] + dn : 0.0f;
          break;
        }
        case TG_EP_CHEB_R: {
          const float rn = in.p[c] - au[c];
          const float dn = a1 * in.q[c] * rn + a2 * u[c];
          r[b + c] = f ? rn : 0.0f;
          d[b + c] = f ? dn : 0.0f;
          x_out[b + c] = f ? in.s[c] + a3 * u[c] + dn : 0.0f;
          break;
        }
        case TG_EP_RESID_UPDATE:
          r[b + c] = f ? in.p[c] - au[c] : 0.0f;
          break;
        default:
          break;
      }
    }
    return part;
  }

  __device__ __forceinline__ double operator()(long long node, unsigned m, const float (&au)[3],
                                               const float (&u)[3]) const {
    In in;
    load(node, in);
    return store(node, m, au, u, in);
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

  struct In {
    double p[3];
  };

  __device__ __forceinline__ void load(long long node, In& in) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) in.p[c] = rhs ? rhs[3 * node + c] : 0.0;
  }

  __device__ __forceinline__ double store(long long node, unsigned m, const double (&au)[3], const double (&)[3],
                                          const In& in) const {
    const long long b = 3 * node;
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      if (m >> c & 1u) v = rhs ? in.p[c] - au[c] : au[c];
      if (y) y[b + c] = v;
      if (y32) y32[b + c] = (float)v;
      part += v * v;
    }
    return part;
  }
  __host__ __device__ bool reduces() const { return norm; }
};

}  // namespace tg
