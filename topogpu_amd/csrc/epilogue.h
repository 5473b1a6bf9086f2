// Per-node output stage shared by every operator kernel.
//
// The update formulas are written with explicit fmaf: a multiply-add left to
// the compiler may be contracted either way round, and differently in
// different kernels; written out, every kernel that applies an epilogue
// rounds it the same way.
//
// The operator hands over A u at one owned node (three components), the
// node's free mask and the node's own (masked) input values. The epilogue is
// split in two phases so that kernels can issue every global load of all
// their owned nodes before the first store: load() fetches the vectors the
// mode needs (unconditionally, independent of the mask), store() combines,
// applies the mask and writes, returning the node's contribution to the
// optional grid reduction.
//
// The mode is a template parameter: each kernel is instantiated per mode, so
// the per-node code holds only that mode's loads, arithmetic and stores. The
// formulas are written once against an accessor that maps (vector, component)
// to memory; PtrAcc addresses plain global memory, the MFMA operator uses
// buffer descriptors with 32-bit offsets (hex_mfma.hip).
#pragma once

#include "common.h"

namespace tg {

// Runtime arguments of the FP32 epilogue (tg_ep32 plus the output y).
struct Ep32 {
  int mode;
  const float* rhs;
  const float* dinv;
  float* d;
  float* x_out;
  float* r;
  float a1, a2, a3;
  float* y;
};

enum EpVec { EV_RHS = 0, EV_DINV, EV_D, EV_X, EV_R, EV_Y, EV_COUNT };

// The epilogue of right-hand side r of a batch: every per-RHS vector offset
// by off = r * (3 * nodes); dinv belongs to the operator and is shared.
__host__ __device__ __forceinline__ Ep32 ep_slice(Ep32 e, long long off) {
  if (e.rhs) e.rhs += off;
  if (e.d) e.d += off;
  if (e.x_out) e.x_out += off;
  if (e.r) e.r += off;
  if (e.y) e.y += off;
  return e;
}

template <int MODE>
struct EpMode {
  struct In {
    float p[3], q[3], s[3];
  };
  static constexpr bool kReduces = MODE == TG_EP_DOT || MODE == TG_EP_RESID_NORM;
  // the operator's input is d0 = a3 * dinv * u, formed as it is read
  static constexpr bool kPrescale = MODE == TG_EP_CHEB_FIRST || MODE == TG_EP_RESID_FIRST;

  // One component c of a node: fetch the mode's inputs.
  template <class Acc>
  __device__ __forceinline__ static void load_one(const Ep32& e, const Acc& a, int c, float& p, float& q, float& s) {
    p = q = s = 0.0f;
    if constexpr (MODE == TG_EP_RESID || MODE == TG_EP_RESID_NORM) {
      p = a.template ld<EV_RHS>(c);
    } else if constexpr (MODE == TG_EP_CHEB) {
      p = a.template ld<EV_RHS>(c);
      q = a.template ld<EV_DINV>(c);
      if (e.a2 != 0.0f) s = a.template ld<EV_D>(c);
    } else if constexpr (MODE == TG_EP_CHEB_R || MODE == TG_EP_CHEB_LAST) {
      p = a.template ld<EV_R>(c);
      q = a.template ld<EV_DINV>(c);
      s = a.template ld<EV_X>(c);
    } else if constexpr (MODE == TG_EP_RESID_UPDATE) {
      p = a.template ld<EV_R>(c);
    } else if constexpr (MODE == TG_EP_CHEB_FIRST) {
      p = a.template ld<EV_RHS>(c);
      q = a.template ld<EV_DINV>(c);
    } else if constexpr (MODE == TG_EP_RESID_FIRST) {
      p = a.template ld<EV_RHS>(c);
    }
  }

  // One component: combine A u (au) with the inputs, apply the free bit f,
  // write; returns the component's share of the reduction.
  template <class Acc>
  __device__ __forceinline__ static double store_one(const Ep32& e, const Acc& a, int c, bool f, float au, float u,
                                                     float p, float q, float s) {
    double part = 0.0;
    if constexpr (MODE == TG_EP_STORE) {
      a.template st<EV_Y>(c, f ? au : 0.0f);
    } else if constexpr (MODE == TG_EP_DOT) {
      const float v = f ? au : 0.0f;
      a.template st<EV_Y>(c, v);
      part = (double)u * (double)v;
    } else if constexpr (MODE == TG_EP_RESID) {
      a.template st<EV_Y>(c, f ? p - au : 0.0f);
    } else if constexpr (MODE == TG_EP_RESID_NORM) {
      const float v = f ? p - au : 0.0f;
      a.template st<EV_Y>(c, v);
      part = (double)v * (double)v;
    } else if constexpr (MODE == TG_EP_CHEB) {
      const float t = q * (p - au);
      const float dn = f ? ((e.a2 != 0.0f) ? fmaf(e.a1, t, e.a2 * s) : e.a1 * t) : 0.0f;
      a.template st<EV_D>(c, dn);
      a.template st<EV_X>(c, f ? u + dn : 0.0f);
    } else if constexpr (MODE == TG_EP_CHEB_R) {
      const float rn = p - au;
      const float dn = fmaf(e.a1 * q, rn, e.a2 * u);
      a.template st<EV_R>(c, f ? rn : 0.0f);
      a.template st<EV_D>(c, f ? dn : 0.0f);
      a.template st<EV_X>(c, f ? fmaf(e.a3, u, s) + dn : 0.0f);
    } else if constexpr (MODE == TG_EP_RESID_UPDATE) {
      a.template st<EV_R>(c, f ? p - au : 0.0f);
    } else if constexpr (MODE == TG_EP_CHEB_FIRST) {
      // u is d0 here: CHEB_R's arithmetic from the start x = d0, r = rhs
      const float rn = p - au;
      const float dn = fmaf(e.a1 * q, rn, e.a2 * u);
      a.template st<EV_R>(c, f ? rn : 0.0f);
      a.template st<EV_D>(c, f ? dn : 0.0f);
      a.template st<EV_X>(c, f ? u + dn : 0.0f);
    } else if constexpr (MODE == TG_EP_RESID_FIRST) {
      a.template st<EV_R>(c, f ? p - au : 0.0f);
      a.template st<EV_X>(c, f ? u : 0.0f);
    } else if constexpr (MODE == TG_EP_CHEB_LAST) {
      const float rn = p - au;
      const float dn = fmaf(e.a1 * q, rn, e.a2 * u);
      a.template st<EV_X>(c, f ? fmaf(e.a3, u, s) + dn : 0.0f);
    }
    return part;
  }

  template <class Acc>
  __device__ __forceinline__ static void load(const Ep32& e, const Acc& a, In& in) {
#pragma unroll
    for (int c = 0; c < 3; ++c) load_one(e, a, c, in.p[c], in.q[c], in.s[c]);
  }

  template <class Acc>
  __device__ __forceinline__ static double store(const Ep32& e, const Acc& a, unsigned m, const float (&au)[3],
                                                 const float (&u)[3], const In& in) {
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) part += store_one(e, a, c, (m >> c) & 1u, au[c], u[c], in.p[c], in.q[c], in.s[c]);
    return part;
  }
};

// Global-memory accessor bound to one node (b = 3 * node).
struct PtrAcc {
  const Ep32& e;
  long long b;
  template <int V>
  __device__ __forceinline__ const float* ptr() const {
    if constexpr (V == EV_RHS) return e.rhs;
    else if constexpr (V == EV_DINV) return e.dinv;
    else if constexpr (V == EV_D) return e.d;
    else if constexpr (V == EV_X) return e.x_out;
    else if constexpr (V == EV_R) return e.r;
    else return e.y;
  }
  template <int V>
  __device__ __forceinline__ float ld(int c) const { return ptr<V>()[b + c]; }
  template <int V>
  __device__ __forceinline__ void st(int c, float v) const { const_cast<float*>(ptr<V>())[b + c] = v; }
};

// Epilogue with the mode fixed at compile time, node-indexed interface.
template <int MODE>
struct Ep32M : Ep32 {
  using Mode = EpMode<MODE>;
  using In = typename Mode::In;
  static constexpr int kMode = MODE;
  static constexpr bool kPrescale = Mode::kPrescale;
  __device__ __forceinline__ void load(long long node, In& in) const { Mode::load(*this, PtrAcc{*this, 3 * node}, in); }
  __device__ __forceinline__ double store(long long node, unsigned m, const float (&au)[3], const float (&u)[3],
                                          const In& in) const {
    return Mode::store(*this, PtrAcc{*this, 3 * node}, m, au, u, in);
  }
  __device__ __forceinline__ double operator()(long long node, unsigned m, const float (&au)[3],
                                               const float (&u)[3]) const {
    In in;
    load(node, in);
    return store(node, m, au, u, in);
  }
  __host__ __device__ static constexpr bool reduces() { return Mode::kReduces; }
  __device__ __forceinline__ Ep32M slice(long long off) const { return Ep32M{ep_slice(*this, off)}; }
};

// Call f(Ep32M<mode>{e}) for the runtime mode of e; returns f's status, or
// TG_ERR_ARG for an unknown mode.
template <class F>
int with_mode(const Ep32& e, F&& f) {
  switch (e.mode) {
    case TG_EP_STORE: return f(Ep32M<TG_EP_STORE>{e});
    case TG_EP_DOT: return f(Ep32M<TG_EP_DOT>{e});
    case TG_EP_RESID: return f(Ep32M<TG_EP_RESID>{e});
    case TG_EP_RESID_NORM: return f(Ep32M<TG_EP_RESID_NORM>{e});
    case TG_EP_CHEB: return f(Ep32M<TG_EP_CHEB>{e});
    case TG_EP_CHEB_R: return f(Ep32M<TG_EP_CHEB_R>{e});
    case TG_EP_RESID_UPDATE: return f(Ep32M<TG_EP_RESID_UPDATE>{e});
    case TG_EP_CHEB_FIRST: return f(Ep32M<TG_EP_CHEB_FIRST>{e});
    case TG_EP_RESID_FIRST: return f(Ep32M<TG_EP_RESID_FIRST>{e});
    case TG_EP_CHEB_LAST: return f(Ep32M<TG_EP_CHEB_LAST>{e});
    default: return set_error(TG_ERR_ARG, "unknown epilogue mode %d", e.mode);
  }
}

// FP64 level-0 operator output: y = A u or y = rhs - A u, optional FP32 copy
// and sum of squares.
struct Ep64 {
  const double* rhs;
  double* y;
  float* y32;
  bool norm;
  static constexpr bool kPrescale = false;

  __device__ __forceinline__ Ep64 slice(long long off) const {
    Ep64 e = *this;
    if (e.rhs) e.rhs += off;
    if (e.y) e.y += off;
    if (e.y32) e.y32 += off;
    return e;
  }

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
