// Density filter, projection, compliance sensitivity and OC update (FP64).
//
// The cone filter of pub_simp_solver.py:215-243 is a fixed stencil on the
// structured grid: weight rmin - d for centre distances d <= rmin - 1e-10,
// rmin for the element itself, rows normalised by Hs = W 1. H = Hs^-1 W and
// H^T = W Hs^-1 (W is symmetric), so both are one stencil pass with an input
// or an output scaling.
#include <math.h>

#include "common.h"
#include "reduce.h"

namespace tg {
namespace {

constexpr int kMaxOffs = 128;
struct Stencil {
  int m;
  int8_t off[kMaxOffs][3];
  double w[kMaxOffs];
};

int make_stencil(int m, const int32_t* offs, const double* w, Stencil& st) {
  TG_REQUIRE(m >= 1 && m <= kMaxOffs && offs && w, "stencil: need 1..%d offsets", kMaxOffs);
  st.m = m;
  for (int k = 0; k < m; ++k) {
    for (int d = 0; d < 3; ++d) {
      TG_REQUIRE(offs[3 * k + d] >= -127 && offs[3 * k + d] <= 127, "stencil offset out of range");
      st.off[k][d] = (int8_t)offs[3 * k + d];
    }
    st.w[k] = w[k];
  }
  return TG_OK;
}

__device__ __forceinline__ double stencil_at(const Grid& g, const Stencil& st, int i, int j, int k,
                                             const double* __restrict__ in, const double* __restrict__ in_scale) {
  double s = 0.0;
  for (int q = 0; q < st.m; ++q) {
    const int a = i + st.off[q][0], b = j + st.off[q][1], c = k + st.off[q][2];
    if (a < 0 || a >= g.nx || b < 0 || b >= g.ny || c < 0 || c >= g.nz) continue;
    const long long e = g.elem(a, b, c);
    const double v = in_scale ? in[e] * in_scale[e] : in[e];
    s = fma(st.w[q], v, s);
  }
  return s;
}

__global__ void k_stencil(Grid g, Stencil st, const double* __restrict__ in, const double* __restrict__ in_scale,
                          const double* __restrict__ out_scale, double* __restrict__ out) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.elems()) return;
  const int k = (int)(e % g.nz), j = (int)((e / g.nz) % g.ny), i = (int)(e / ((long long)g.nz * g.ny));
  const double s = stencil_at(g, st, i, j, k, in, in_scale);
  out[e] = out_scale ? s * out_scale[e] : s;
}

__device__ __forceinline__ double proj(double x, double beta, double eta) {
  const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
  return (tanh(beta * eta) + tanh(beta * (x - eta))) / den;
}

__global__ void k_density_chain(long long n, const double* __restrict__ rf, double beta, double eta, double penal,
                                double emin, double e0, double* rp, double* dh, double* E64, float* E32) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double x = rf[e];
  double p = x, d = 1.0;
  if (beta >= 1e-6) {
    const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
    const double th = tanh(beta * (x - eta));
    p = (tanh(beta * eta) + th) / den;
    d = beta * (1.0 - th * th) / den;
  }
  const double E = emin + (e0 - emin) * pow(p, penal);
  if (rp) rp[e] = p;
  if (dh) dh[e] = d;
  if (E64) E64[e] = E;
  if (E32) E32[e] = (float)E;
}

__global__ void k_sensitivity(Grid g, const double* __restrict__ U, const double* __restrict__ rp,
                              const double* __restrict__ dh, double penal, double erange, double* ce_out,
                              double* dcf) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.elems()) return;
  const int k = (int)(e % g.nz), j = (int)((e / g.nz) % g.ny), i = (int)(e / ((long long)g.nz * g.ny));
  double ue[24];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const double* p = U + 3 * g.node(i + corner_x(c), j + corner_y(c), k + corner_z(c));
    ue[3 * c] = p[0];
    ue[3 * c + 1] = p[1];
    ue[3 * c + 2] = p[2];
  }
  double ce = 0.0;
#pragma unroll
  for (int a = 0; a < 24; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 24; ++b) t = fma(c_ke64[a * 24 + b], ue[b], t);
    ce = fma(ue[a], t, ce);
  }
  if (ce_out) ce_out[e] = ce;
  const double x = rp[e];
  const double dcp = -penal * erange * pow(x, penal - 1.0) * ce;
  dcf[e] = dh ? dcp * dh[e] : dcp;
}

__device__ __forceinline__ double oc_candidate(double q, double lo, double hi, double lm) {
  // pub_simp_solver.py:326-328: clip to [rho-move, rho+move], then [1e-3, 1]
  const double x = q * (1.0 / sqrt(lm));   // same arithmetic as k_oc_pass
  return fmin(fmax(fmin(fmax(x, lo), hi), 1e-3), 1.0);
}

__global__ void k_oc_prepare(long long n, const double* __restrict__ rho, const double* __restrict__ dc,
                             const double* __restrict__ dv, double move, double* q, double* lo, double* hi,
                             double* state) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e == 0) {
    state[0] = 0.0;   // l1
    state[1] = 1e9;   // l2
    state[2] = 0.5e9; // lm
    state[3] = 0.0;   // done
    state[4] = 0.0;   // steps
    state[5] = 0.0;   // volume of last step
  }
  if (e >= n) return;
  const double dcs = fmin(dc[e], -1e-12), dvs = fmax(dv[e], 1e-12);
  q[e] = rho[e] * sqrt(-dcs / dvs);
  lo[e] = rho[e] - move;
  hi[e] = rho[e] + move;
}

// One OC pass covers OC_LEVELS bisection steps at once. The 2^L - 1 candidate
// multipliers are the midpoints of the binary tree below the current bracket
// [l1, l2], each computed as 0.5*(a+b) of its own bracket -- the same
// arithmetic the sequential bisection performs on whichever path it takes.
// The filtered/projected volume is monotone in lambda, so walking the tree
// with the evaluated volumes reproduces the sequential decisions exactly.
constexpr int OC_LEVELS = 4;
constexpr int OC_NPTS = (1 << OC_LEVELS) - 1;   // tree nodes in heap order

__device__ __forceinline__ void oc_tree(double l1, double l2, double (&lam)[OC_NPTS]) {
  double lo_[OC_NPTS], hi_[OC_NPTS];
  lo_[0] = l1;
  hi_[0] = l2;
#pragma unroll
  for (int t = 0; t < OC_NPTS; ++t) {
    lam[t] = 0.5 * (lo_[t] + hi_[t]);
    if (2 * t + 2 < OC_NPTS) {
      lo_[2 * t + 1] = lo_[t];  hi_[2 * t + 1] = lam[t];   // volume too small -> left
      lo_[2 * t + 2] = lam[t];  hi_[2 * t + 2] = hi_[t];   // volume too large -> right
    }
  }
}

__global__ void __launch_bounds__(256) k_oc_pass(Grid g, Stencil st, const double* __restrict__ hs_inv,
                                                const double* __restrict__ q, const double* __restrict__ lo,
                                                const double* __restrict__ hi, double beta, double eta,
                                                int use_h, const double* state, Workspace ws) {
  if (state[3] != 0.0) return;   // uniform: converged bracket
  double lam[OC_NPTS], rs[OC_NPTS];
  oc_tree(state[0], state[1], lam);
  // 1/sqrt(lambda) once per candidate: the element loop then multiplies
  // (an FP64 divide per neighbour and candidate would dominate the pass)
#pragma unroll
  for (int t = 0; t < OC_NPTS; ++t) rs[t] = 1.0 / sqrt(lam[t]);
  double v[OC_NPTS];
#pragma unroll
  for (int t = 0; t < OC_NPTS; ++t) v[t] = 0.0;
  const bool proj_on = use_h && beta > 1e-6;
  const long long n = g.elems();
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(e % g.nz), j = (int)((e / g.nz) % g.ny), i = (int)(e / ((long long)g.nz * g.ny));
    double s[OC_NPTS];
#pragma unroll
    for (int t = 0; t < OC_NPTS; ++t) s[t] = 0.0;
    for (int o = 0; o < st.m; ++o) {
      const int a = i + st.off[o][0], b = j + st.off[o][1], c = k + st.off[o][2];
      if (a < 0 || a >= g.nx || b < 0 || b >= g.ny || c < 0 || c >= g.nz) continue;
      const long long f = g.elem(a, b, c);
      const double qf = q[f], lf = lo[f], hf = hi[f], w = st.w[o];
#pragma unroll
      for (int t = 0; t < OC_NPTS; ++t)
        s[t] = fma(w, fmin(fmax(fmin(fmax(qf * rs[t], lf), hf), 1e-3), 1.0), s[t]);
    }
    const double hinv = hs_inv[e];
#pragma unroll
    for (int t = 0; t < OC_NPTS; ++t) {
      const double rf = s[t] * hinv;
      v[t] += proj_on ? proj(rf, beta, eta) : rf;
    }
  }
  block_partials<OC_NPTS>(v, ws);
}

// One block: the volumes of the pass's candidates, then up to OC_LEVELS
// bracket updates of pub_simp_solver.py:333-338, stopping when l2-l1 < 1e-9.
__global__ void __launch_bounds__(256) k_oc_walk(int nb, long long n, double volfrac, int max_steps, double* state,
                                                Workspace ws) {
  if (state[3] != 0.0) return;
  __shared__ double svol[OC_NPTS];
  for (int t = 0; t < OC_NPTS; ++t) {
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += ws.partials[(size_t)t * nb + b];
    block_sum<1>(v);
    if (threadIdx.x == 0) svol[t] = v[0];
  }
  if (threadIdx.x == 0) {
    double lam[OC_NPTS];
    oc_tree(state[0], state[1], lam);
    int node = 0;
    for (int lev = 0; lev < OC_LEVELS; ++lev) {
      if (state[4] >= max_steps) { state[3] = 1.0; break; }
      const double lm = lam[node];
      state[5] = svol[node];
      const bool right = svol[node] / (double)n > volfrac;
      if (right) state[0] = lm; else state[1] = lm;
      state[2] = lm;
      state[4] += 1.0;
      if (state[1] - state[0] < 1e-9 || state[4] >= max_steps) { state[3] = 1.0; break; }
      node = 2 * node + (right ? 2 : 1);
    }
  }
}

__global__ void __launch_bounds__(256) k_oc_finalize(long long n, const double* __restrict__ rho,
                                                    const double* __restrict__ q, const double* __restrict__ lo,
                                                    const double* __restrict__ hi, const double* state,
                                                    double* rho_new, Workspace ws) {
  const double lm = state[2];
  double mx = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const double r = oc_candidate(q[e], lo[e], hi[e], lm);
    rho_new[e] = r;
    mx = fmax(mx, fabs(r - rho[e]));
  }
  __shared__ double smax[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) b = fmax(b, smax[w]);
    ws.partials[blockIdx.x] = b;
  }
}

// One block: max over the per-block maxima (order-independent).
__global__ void __launch_bounds__(256) k_max_finalize(int nb, double* out, Workspace ws) {
  __shared__ double smax[4];
  double mx = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) mx = fmax(mx, ws.partials[b]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
}

int grid_blocks(long long n) {
  long long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace
}  // namespace tg

using namespace tg;

#define TG_NEED_INIT() TG_REQUIRE(::tg::ke_ready(), "tg_init has not been called on this device")

extern "C" int tg_stencil64(const tg_grid* gp, int32_t m, const int32_t* offs, const double* w, const double* in,
                            const double* in_scale, const double* out_scale, double* out, void* s) {
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(in && out && in != out, "stencil64: bad buffers");
  Stencil st;
  if (int r = make_stencil(m, offs, w, st)) return r;
  const Grid g = to_grid(gp);
  hipLaunchKernelGGL(k_stencil, dim3((unsigned)((g.elems() + 255) / 256)), dim3(256), 0, (hipStream_t)s, g, st, in,
                     in_scale, out_scale, out);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_density_chain(int64_t n, const double* rf, double beta, double eta, double penal, double emin,
                                double e0, double* rp, double* dh, double* E64, float* E32, void* s) {
  TG_REQUIRE(n > 0 && rf, "density_chain: bad argument");
  hipLaunchKernelGGL(k_density_chain, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, (long long)n,
                     rf, beta, eta, penal, emin, e0, rp, dh, E64, E32);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_sensitivity64(const tg_grid* gp, const double* U, const double* rp, const double* dh,
                                double penal, double erange, double* ce, double* dcf, void* s) {
  TG_NEED_INIT();
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(U && rp && dcf, "sensitivity64: NULL argument");
  const Grid g = to_grid(gp);
  hipLaunchKernelGGL(k_sensitivity, dim3((unsigned)((g.elems() + 255) / 256)), dim3(256), 0, (hipStream_t)s, g, U,
                     rp, dh, penal, erange, ce, dcf);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_oc_prepare(int64_t n, const double* rho, const double* dc, const double* dv, double move,
                             double* q, double* lo, double* hi, double* state, void* s) {
  TG_REQUIRE(n > 0 && rho && dc && dv && q && lo && hi && state, "oc_prepare: NULL argument");
  hipLaunchKernelGGL(k_oc_prepare, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, (long long)n,
                     rho, dc, dv, move, q, lo, hi, state);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_oc_bisect_step(const tg_grid* gp, int32_t m, const int32_t* offs, const double* w,
                                 const double* hs_inv, const double* q, const double* lo, const double* hi,
                                 double beta, double eta, double volfrac, int use_h, double* state, void* s) {
  TG_NEED_INIT();
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(hs_inv && q && lo && hi && state, "oc_bisect_step: NULL argument");
  Stencil st;
  if (int r = make_stencil(m, offs, w, st)) return r;
  const Grid g = to_grid(gp);
  const int nb = grid_blocks(g.elems());
  hipLaunchKernelGGL(k_oc_pass, dim3(nb), dim3(256), 0, (hipStream_t)s, g, st, hs_inv, q, lo, hi, beta, eta, use_h,
                     state, workspace());
  TG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_oc_walk, dim3(1), dim3(256), 0, (hipStream_t)s, nb, g.elems(), volfrac, 100, state,
                     workspace());
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_oc_finalize(int64_t n, const double* rho, const double* q, const double* lo, const double* hi,
                              const double* state, double* rho_new, double* change, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && rho && q && lo && hi && state && rho_new && change, "oc_finalize: NULL argument");
  const int nb = grid_blocks(n);
  hipLaunchKernelGGL(k_oc_finalize, dim3(nb), dim3(256), 0, (hipStream_t)s, (long long)n, rho, q, lo, hi, state,
                     rho_new, workspace());
  TG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max_finalize, dim3(1), dim3(256), 0, (hipStream_t)s, nb, change, workspace());
  TG_LAUNCH_CHECK();
  return TG_OK;
}
