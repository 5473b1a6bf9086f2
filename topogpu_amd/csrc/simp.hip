// Density filter, projection, compliance sensitivity and OC update (FP64).
//
// The cone filter of pub_simp_solver.py:215-243 is a fixed stencil on the
// structured grid: weight rmin - d for centre distances d <= rmin - 1e-10,
// rmin for the element itself, rows normalised by Hs = W 1. H = Hs^-1 W and
// H^T = W Hs^-1 (W is symmetric), so both are one stencil pass with an input
// or an output scaling.
#include <math.h>

#include <type_traits>

#include "common.h"
#include "reduce.h"

namespace tg {
namespace {

// rmin up to 4 (the largest radius a continuation action may set): 251
// offsets
constexpr int kMaxOffs = 256;
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

// FP64 tanh within about one ulp: for |y| < 0.625 the rational
// approximation y + y^3 P(y^2)/Q(y^2) of S. L. Moshier's Cephes library
// (tanh.c), else 1 - 2/(exp(2|y|) + 1) with the sign of y. The Heaviside
// projection evaluates it once per element and trial multiplier in the OC
// passes (beta = 1 keeps |y| <= 1/2, the rational branch); the generic
// library routine there cost most of the pass.
__device__ __forceinline__ double tanh_d(double y) {
  const double a = fabs(y);
  if (a < 0.625) {
    const double z = y * y;
    const double p = (-9.64399179425052238628e-1 * z - 9.92877231001918586564e1) * z - 1.61468768441708447952e3;
    const double q = ((z + 1.12811678491632931402e2) * z + 2.23548839060100448583e3) * z + 4.84406305325125486048e3;
    return y + y * z * p / q;
  }
  const double t = exp(-2.0 * a);
  return copysign((1.0 - t) / (1.0 + t), y);
}

__device__ __forceinline__ double proj(double x, double beta, double eta) {
  const double den = tanh_d(beta * eta) + tanh_d(beta * (1.0 - eta));
  return (tanh_d(beta * eta) + tanh_d(beta * (x - eta))) / den;
}

__global__ void k_density_chain(long long n, const double* __restrict__ rf, double beta, double eta, double penal,
                                double emin, double e0, double* rp, double* dh, double* E64, float* E32) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double x = rf[e];
  double p = x, d = 1.0;
  if (beta > 1e-6) {
    const double den = tanh_d(beta * eta) + tanh_d(beta * (1.0 - eta));
    const double th = tanh_d(beta * (x - eta));
    p = (tanh_d(beta * eta) + th) / den;
    d = beta * (1.0 - th * th) / den;
  }
  const double E = emin + (e0 - emin) * pow(p, penal);
  if (rp) rp[e] = p;
  if (dh) dh[e] = d;
  if (E64) E64[e] = E;
  if (E32) E32[e] = (float)E;
}

// Weighted sum over a batch of load cases (U holds nrhs contiguous slices of
// 3 * nodes; w = NULL weighs every case 1): ce = sum_r w_r u_r,e^T KE u_r,e.
// KE is staged in LDS and read there (every lane reads the same entry, a
// broadcast): held in scalar registers, its 576 doubles overflow them and
// the compiler spills them to vector-register lanes (2,300 extra
// instructions per element).
__global__ void __launch_bounds__(256) k_sensitivity(Grid g, int nrhs, const double* __restrict__ U,
                                                    const double* __restrict__ w, const double* __restrict__ rp,
                                                    const double* __restrict__ dh, double penal, double erange,
                                                    double* ce_out, double* dcf) {
  __shared__ double ske[576];
  for (int t = threadIdx.x; t < 576; t += blockDim.x) ske[t] = c_ke64[t];
  __syncthreads();
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.elems()) return;
  const int k = (int)(e % g.nz), j = (int)((e / g.nz) % g.ny), i = (int)(e / ((long long)g.nz * g.ny));
  const long long stride = 3 * g.nodes();
  double ce = 0.0;
  for (int r = 0; r < nrhs; ++r) {
    const double* Ur = U + r * stride;
    double ue[24];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const double* p = Ur + 3 * g.node(i + corner_x(c), j + corner_y(c), k + corner_z(c));
      ue[3 * c] = p[0];
      ue[3 * c + 1] = p[1];
      ue[3 * c + 2] = p[2];
    }
    double cr = 0.0;
    // rows node by node, the node loop not unrolled: its KE rows are then
    // read from LDS as they are needed instead of all 576 entries being
    // hoisted into registers; the node's three values are picked by selects
    // (an indexed register array would go to scratch memory)
#pragma unroll 1
    for (int na = 0; na < 8; ++na) {
      double un[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double v = ue[q];
#pragma unroll
        for (int c = 1; c < 8; ++c) v = (na == c) ? ue[3 * c + q] : v;
        un[q] = v;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double* krow = ske + (3 * na + q) * 24;
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < 24; ++b) t = fma(krow[b], ue[b], t);
        cr = fma(un[q], t, cr);
      }
    }
    ce = fma(w ? w[r] : 1.0, cr, ce);
  }
  if (ce_out) ce_out[e] = ce;
  const double x = rp[e];
  const double dcp = -penal * erange * pow(x, penal - 1.0) * ce;
  dcf[e] = dh ? dcp * dh[e] : dcp;
}

// ---- compliance in double-double precision (finite-difference checks)
// Error-free transformations: a + b = s + e exactly (TwoSum), a * b = p + e
// exactly (fma TwoProduct).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  e = fma(a, b, -p);
}
struct DD {
  double hi, lo;
};
__device__ __forceinline__ void dd_add(DD& x, double y) {
  double s, e;
  two_sum(x.hi, y, s, e);
  x.hi = s;
  x.lo += e;
}
__device__ __forceinline__ void dd_add(DD& x, const DD& y) {
  double s, e;
  two_sum(x.hi, y.hi, s, e);
  x.hi = s;
  x.lo += e + y.lo;
}

// Per block: F.u and u^T K u = sum_e E_e d_e^T KE d_e (d_e = u_e minus the
// element's first-node displacement; KE annihilates translations) with every
// product and sum carried in double-double. Partials: 4 doubles per block.
__global__ void __launch_bounds__(256) k_compliance_dd(Grid g, const double* __restrict__ E,
                                                      const double* __restrict__ U, const double* __restrict__ F,
                                                      Workspace ws) {
  DD fu{0.0, 0.0}, uku{0.0, 0.0};
  const long long nd = 3 * g.nodes(), ne = g.elems();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nd; i += stride) {
    double p, e;
    two_prod(F[i], U[i], p, e);
    dd_add(fu, p);
    fu.lo += e;
  }
  for (long long el = blockIdx.x * (long long)blockDim.x + threadIdx.x; el < ne; el += stride) {
    const int k = (int)(el % g.nz), j = (int)((el / g.nz) % g.ny), i = (int)(el / ((long long)g.nz * g.ny));
    // d = u_e - u_node0 exactly as d + dl (TwoSum): the quadratic form is
    // then d^T KE d + 2 dl^T KE d to second order in the rounding of d
    double d[24], dl[24];
    const double* u0 = U + 3 * g.node(i, j, k);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const double* pc = U + 3 * g.node(i + corner_x(c), j + corner_y(c), k + corner_z(c));
#pragma unroll
      for (int q = 0; q < 3; ++q) two_sum(pc[q], -u0[q], d[3 * c + q], dl[3 * c + q]);
    }
    DD w{0.0, 0.0};
    double corr = 0.0;
    for (int a = 3; a < 24; ++a)        // first-node rows and columns of d are zero
      for (int b = 3; b < 24; ++b) {
        const double kab = c_ke64[a * 24 + b];
        double p, e, p2, e2;
        two_prod(d[a], d[b], p, e);
        two_prod(p, kab, p2, e2);
        dd_add(w, p2);
        w.lo += e2 + e * kab;
        corr = fma(2.0 * dl[a] * kab, d[b], corr);
      }
    w.lo += corr;
    double p, e;
    two_prod(w.hi, E[el], p, e);
    dd_add(uku, p);
    uku.lo += e + w.lo * E[el];
  }
  // block reduction of the two pairs in fixed order
  __shared__ DD s_fu[256], s_uku[256];
  s_fu[threadIdx.x] = fu;
  s_uku[threadIdx.x] = uku;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      dd_add(s_fu[threadIdx.x], s_fu[threadIdx.x + off]);
      dd_add(s_uku[threadIdx.x], s_uku[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* pt = ws.partials + 4 * (size_t)blockIdx.x;
    pt[0] = s_fu[0].hi;
    pt[1] = s_fu[0].lo;
    pt[2] = s_uku[0].hi;
    pt[3] = s_uku[0].lo;
  }
}

__global__ void k_compliance_dd_finalize(int nb, double* out, Workspace ws) {
  if (threadIdx.x != 0) return;
  DD fu{0.0, 0.0}, uku{0.0, 0.0};
  for (int b = 0; b < nb; ++b) {
    const double* pt = ws.partials + 4 * (size_t)b;
    dd_add(fu, DD{pt[0], pt[1]});
    dd_add(uku, DD{pt[2], pt[3]});
  }
  double s, e;
  two_sum(fu.hi, fu.lo, s, e);
  out[0] = s;
  out[1] = e;
  two_sum(uku.hi, uku.lo, s, e);
  out[2] = s;
  out[3] = e;
}

// pub_simp_solver.py:326-328 clips to [rho-move, rho+move], then to
// [1e-3, 1]. For lo <= hi that is one clip to [lo', hi'] with lo', hi' the
// ends clipped to [1e-3, 1] (clips are monotone), which k_oc_prepare stores.
__device__ __forceinline__ double clip_bounds(double v) { return fmin(fmax(v, 1e-3), 1.0); }

__device__ __forceinline__ double oc_candidate(double q, double lo, double hi, double lm) {
  const double x = q * (1.0 / sqrt(lm));   // same arithmetic as k_oc_pass
  return fmin(fmax(x, lo), hi);
}

// OC bisection state (device doubles).
enum OcState {
  S_L1 = 0, S_L2, S_LM, S_DONE, S_STEPS, S_PASSES,
  S_A, S_VA,       // largest multiplier known to give volume above the target, and that volume
  S_B, S_VB,       // smallest multiplier known to give volume at or below the target
  S_SA, S_SB,      // as S_A / S_B, restricted to volumes clear of the target by the margin
  S_CAND,          // OC_NPTS multipliers of the next evaluation pass
  S_PREV = S_CAND + 15,   // the last bisection's result (kept across calls; 0: none)
  S_LEN = TG_OC_STATE_LEN
};
constexpr int OC_NPTS = 15;
static_assert(S_PREV >= S_CAND + OC_NPTS && S_PREV < S_LEN, "state layout");

__global__ void k_oc_prepare(long long n, const double* __restrict__ rho, const double* __restrict__ dc,
                             const double* __restrict__ dv, double move, double* q, double* lo, double* hi,
                             double* state) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e == 0) {
    const double prev = state[S_PREV];
    for (int k = 0; k < S_LEN; ++k)
      if (k != S_PREV) state[k] = 0.0;
    state[S_L1] = 0.0;
    state[S_L2] = 1e9;
    state[S_LM] = 0.5e9;
    state[S_A] = state[S_SA] = -1.0;      // none yet
    state[S_B] = state[S_SB] = HUGE_VAL;  // none yet
    if (prev > 1e-10 && prev < 1e9) {
      // first pass around the previous iteration's multiplier, which the
      // new crossing is usually near: it, twelve geometric neighbours from
      // e^(+-1e-4) to e^(+-1.5) times it, and the ends of the bisection's
      // range (a two-sided bracket whatever the crossing). Which multipliers
      // a pass evaluates changes only how many steps it decides, not the
      // decisions.
      const double d[6] = {1e-4, 1e-3, 1e-2, 0.05, 0.25, 1.5};
      state[S_CAND] = prev;
      for (int k = 0; k < 6; ++k) {
        state[S_CAND + 1 + 2 * k] = prev * exp(-d[k]);
        state[S_CAND + 2 + 2 * k] = prev * exp(d[k]);
      }
      state[S_CAND + 13] = 1e-10;
      state[S_CAND + 14] = 1e9;
    } else {
      // first pass: multipliers log-spaced over [1e-10, 1e9] (ratio 10^(19/14))
      for (int t = 0; t < OC_NPTS; ++t) state[S_CAND + t] = pow(10.0, -10.0 + 19.0 * t / (OC_NPTS - 1));
    }
  }
  if (e >= n) return;
  const double dcs = fmin(dc[e], -1e-12), dvs = fmax(dv[e], 1e-12);
  q[e] = rho[e] * sqrt(-dcs / dvs);
  lo[e] = clip_bounds(rho[e] - move);
  hi[e] = clip_bounds(rho[e] + move);
}

// Volume sum(heaviside(H x(lambda))) at the OC_NPTS multipliers of the state,
// one partial per block and multiplier.
__global__ void __launch_bounds__(256) k_oc_pass(Grid g, Stencil st, const double* __restrict__ hs_inv,
                                                const double* __restrict__ q, const double* __restrict__ lo,
                                                const double* __restrict__ hi, double beta, double eta,
                                                int use_h, const double* state, Workspace ws) {
  if (state[S_DONE] != 0.0) return;   // uniform: converged bracket
  // 1/sqrt(lambda) once per candidate: the element loop then multiplies
  // (an FP64 divide per neighbour and candidate would dominate the pass)
  double rs[OC_NPTS];
#pragma unroll
  for (int t = 0; t < OC_NPTS; ++t) rs[t] = 1.0 / sqrt(state[S_CAND + t]);
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
        s[t] = fma(w, fmin(fmax(qf * rs[t], lf), hf), s[t]);
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

// Tiled forms of k_stencil and k_oc_pass for radius-1 stencils whose offsets
// come as filter_stencil lists them: the element itself first, then the other
// offsets in increasing lexicographic (dx, dy, dz) order. A block owns a
// TX x TY x TZ element tile and stages the tile and its one-element halo in
// LDS (zero outside the grid, which adds w * 0 like the skipped neighbour of
// k_stencil); each thread then sums a z-run of T_RUN elements from LDS, one
// (dx, dy) column of T_RUN + 2 values at a time read as 16-byte words. Every
// output adds its terms in the stencil's order, so for finite inputs the
// result is that of k_stencil bit for bit.
constexpr int TX = 8, TY = 8, TZ = 16, T_THREADS = 256, T_RUN = TX * TY * TZ / T_THREADS;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, T_NH = HX * HY * HZ;
constexpr int T_HIT = (T_NH + T_THREADS - 1) / T_THREADS;
constexpr int OC_G = 3, OC_NG = OC_NPTS / OC_G;   // multipliers per block, blocks per tile
constexpr int kXcds = 8;
static_assert(OC_NPTS % OC_G == 0, "multiplier groups");
static_assert(TZ % T_RUN == 0 && T_RUN % 2 == 0 && HZ % 2 == 0, "z-runs of 16-byte aligned columns");

// w[(dx + 1) * 9 + (dy + 1) * 3 + dz + 1], zero where the stencil holds no
// offset; bit t of mask set where it does
struct Cube27 {
  double w[27];
  unsigned mask;
};

bool make_cube27(const Stencil& st, Cube27& c) {
  if (st.off[0][0] != 0 || st.off[0][1] != 0 || st.off[0][2] != 0) return false;
  c.mask = 1u << 13;
  for (int t = 0; t < 27; ++t) c.w[t] = 0.0;
  c.w[13] = st.w[0];
  int last = -1;
  for (int k = 1; k < st.m; ++k) {
    const int dx = st.off[k][0], dy = st.off[k][1], dz = st.off[k][2];
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return false;
    const int t = (dx + 1) * 9 + (dy + 1) * 3 + dz + 1;
    if (t <= last || t == 13) return false;
    last = t;
    c.w[t] = st.w[k];
    c.mask |= 1u << t;
  }
  return true;
}

struct Tile {
  int i0, j0, k0;
};

__device__ __forceinline__ Tile tile_origin(const Grid& g, int t) {
  const int ntz = (g.nz + TZ - 1) / TZ, nty = (g.ny + TY - 1) / TY;
  return {(t / (ntz * nty)) * TX, ((t / ntz) % nty) * TY, (t % ntz) * TZ};
}

// halo slot idx -> element index f of the grid; false outside it (the tiled
// kernels run on grids of fewer than 2^31 elements: 32-bit index arithmetic)
__device__ __forceinline__ bool halo_elem(const Grid& g, const Tile& tl, int idx, int& f) {
  const int hz = idx % HZ, hy = (idx / HZ) % HY, hx = idx / (HZ * HY);
  const int i = tl.i0 - 1 + hx, j = tl.j0 - 1 + hy, k = tl.k0 - 1 + hz;
  // unsigned compares test both ends; & rather than && keeps it branch-free
  const bool in = (idx < T_NH) & ((unsigned)i < (unsigned)g.nx) & ((unsigned)j < (unsigned)g.ny) &
                  ((unsigned)k < (unsigned)g.nz);
  f = in ? (i * g.ny + j) * g.nz + k : 0;
  return in;
}

__device__ __forceinline__ void load_col(const double* s, double (&x)[T_RUN + 2]) {
  const double2* p = reinterpret_cast<const double2*>(s);
#pragma unroll
  for (int u = 0; u < (T_RUN + 2) / 2; ++u) {
    const double2 v = p[u];
    x[2 * u] = v.x;
    x[2 * u + 1] = v.y;
  }
}

// acc[c][r] = the stencil sum for the thread's r-th element from slab c of
// sx, over the offsets of MASK (a compile-time set, so the sum is straight
// FMAs: a run-time mask would be if-converted into selects)
template <int G, unsigned MASK>
__device__ __forceinline__ void cube27_run(const Cube27& cw, const double (*sx)[T_NH], int tx, int ty, int tz,
                                           double (&acc)[G][T_RUN]) {
  const int base = tz * T_RUN;
#pragma unroll
  for (int c = 0; c < G; ++c) {
    double x[T_RUN + 2];
    load_col(&sx[c][((tx + 1) * HY + ty + 1) * HZ + base], x);
#pragma unroll
    for (int r = 0; r < T_RUN; ++r) acc[c][r] = cw.w[13] * x[r + 1];
  }
#pragma unroll
  for (int dx = 0; dx < 3; ++dx)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int t0 = dx * 9 + dy * 3;
      constexpr unsigned kAll = 7u;
      const unsigned m = (MASK >> t0) & (dx == 1 && dy == 1 ? 5u : kAll);
      if (m == 0) continue;
      const double w0 = cw.w[t0], w1 = cw.w[t0 + 1], w2 = cw.w[t0 + 2];
#pragma unroll
      for (int c = 0; c < G; ++c) {
        double x[T_RUN + 2];
        load_col(&sx[c][((tx + dx) * HY + ty + dy) * HZ + base], x);
#pragma unroll
        for (int r = 0; r < T_RUN; ++r) {
          if (m & 1u) acc[c][r] = fma(w0, x[r], acc[c][r]);
          if (m & 2u) acc[c][r] = fma(w1, x[r + 1], acc[c][r]);
          if (m & 4u) acc[c][r] = fma(w2, x[r + 2], acc[c][r]);
        }
      }
    }
}

// The offset sets of the radius-1 cone filter (filter_stencil): the element
// alone (rmin <= 1), with its 6 face neighbours (rmin <= sqrt 2), with the
// 12 edge neighbours too (rmin <= sqrt 3), the whole cube. Another set runs
// as the whole cube with zero weights, whose terms 0 * x leave a sum of
// finite values unchanged.
constexpr unsigned cube_mask(int max_d2) {
  unsigned m = 0;
  for (int t = 0; t < 27; ++t) {
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    if (dx * dx + dy * dy + dz * dz <= max_d2) m |= 1u << t;
  }
  return m;
}
constexpr unsigned kCube1 = cube_mask(0), kCube7 = cube_mask(1), kCube19 = cube_mask(2), kCube27 = cube_mask(3);

// fn(std::integral_constant<unsigned, MASK>) for the kernel instance of mask
template <class F>
void with_cube_mask(unsigned mask, F&& fn) {
  if (mask == kCube19) fn(std::integral_constant<unsigned, kCube19>());
  else if (mask == kCube7) fn(std::integral_constant<unsigned, kCube7>());
  else if (mask == kCube1) fn(std::integral_constant<unsigned, kCube1>());
  else fn(std::integral_constant<unsigned, kCube27>());
}

// The OC pass on tiles: OC_NPTS / OC_G blocks per tile, each evaluating
// OC_G of the pass's multipliers (it clips the tile's candidate densities for
// them into LDS once, instead of once per stencil neighbour). A tile's blocks
// are numbered kXcds apart, so that they run back to back on one XCD, where
// all but the first read the tile's inputs from that XCD's L2.
template <unsigned MASK>
__global__ void __launch_bounds__(T_THREADS) k_oc_pass_tiled(Grid g, Cube27 cw, int nb, const double* __restrict__ hs_inv, const double* __restrict__ q,
                const double* __restrict__ lo, const double* __restrict__ hi, double beta, double eta, int use_h,
                const double* state, Workspace ws) {
  if (state[S_DONE] != 0.0) return;   // uniform: converged bracket
  const int s = blockIdx.x / kXcds, grp = s % OC_NG, t = (s / OC_NG) * kXcds + blockIdx.x % kXcds;
  if (t >= nb) return;   // uniform: padding of the last round of kXcds tiles
  __shared__ __align__(16) double sx[OC_G][T_NH];
  const Tile tl = tile_origin(g, t);
  double rs[OC_G];
#pragma unroll
  for (int c = 0; c < OC_G; ++c) rs[c] = 1.0 / sqrt(state[S_CAND + grp * OC_G + c]);
#pragma unroll
  for (int it = 0; it < T_HIT; ++it) {
    const int idx = threadIdx.x + it * T_THREADS;
    int f;
    const bool in = halo_elem(g, tl, idx, f);
    const double hq = q[f], hl = lo[f], hh = hi[f];   // element 0 stands in outside the grid
    if (idx < T_NH)
#pragma unroll
      for (int c = 0; c < OC_G; ++c) sx[c][idx] = in ? fmin(fmax(hq * rs[c], hl), hh) : 0.0;
  }
  const int tz = threadIdx.x % (TZ / T_RUN), ty = (threadIdx.x / (TZ / T_RUN)) % TY,
            tx = threadIdx.x / ((TZ / T_RUN) * TY);
  const int oi = tl.i0 + tx, oj = tl.j0 + ty, ok0 = tl.k0 + tz * T_RUN;
  double hinv[T_RUN];
#pragma unroll
  for (int r = 0; r < T_RUN; ++r)
    hinv[r] = oi < g.nx && oj < g.ny && ok0 + r < g.nz ? hs_inv[g.elem(oi, oj, ok0 + r)] : 0.0;
  __syncthreads();
  const bool proj_on = use_h && beta > 1e-6;
  const double p0 = tanh_d(beta * eta), ipden = 1.0 / (p0 + tanh_d(beta * (1.0 - eta)));
  double vg[OC_G];
#pragma unroll
  for (int c = 0; c < OC_G; ++c) {
    // one multiplier at a time: its sums are projected before the next one's
    // LDS reads are issued (sched_barrier), which bounds the live registers
    double acc[1][T_RUN];
    cube27_run<1, MASK>(cw, &sx[c], tx, ty, tz, acc);
    vg[c] = 0.0;
#pragma unroll
    for (int r = 0; r < T_RUN; ++r)
      if (oi < g.nx && oj < g.ny && ok0 + r < g.nz) {
        const double rf = acc[0][r] * hinv[r];
        vg[c] += proj_on ? (p0 + tanh_d(beta * (rf - eta))) * ipden : rf;
      }
    __builtin_amdgcn_sched_barrier(0);
  }
  // partials in the layout of block_partials<OC_NPTS>
  block_sum<OC_G>(vg);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < OC_G; ++c) ws.partials[(size_t)(grp * OC_G + c) * nb + t] = vg[c];
}

template <unsigned MASK>
__global__ void __launch_bounds__(T_THREADS) k_stencil_tiled(Grid g, Cube27 cw, const double* __restrict__ in,
                                                            const double* __restrict__ in_scale,
                                                            const double* __restrict__ out_scale,
                                                            double* __restrict__ out) {
  __shared__ double sx[1][T_NH];
  const Tile tl = tile_origin(g, blockIdx.x);
#pragma unroll
  for (int it = 0; it < T_HIT; ++it) {
    const int idx = threadIdx.x + it * T_THREADS;
    int f;
    const bool in_g = halo_elem(g, tl, idx, f);
    const double x = in_scale ? in[f] * in_scale[f] : in[f];   // element 0 stands in outside the grid
    const double v = in_g ? x : 0.0;
    if (idx < T_NH) sx[0][idx] = v;
  }
  __syncthreads();
  const int tz = threadIdx.x % (TZ / T_RUN), ty = (threadIdx.x / (TZ / T_RUN)) % TY,
            tx = threadIdx.x / ((TZ / T_RUN) * TY);
  const int oi = tl.i0 + tx, oj = tl.j0 + ty, ok0 = tl.k0 + tz * T_RUN;
  double acc[1][T_RUN];
  cube27_run<1, MASK>(cw, sx, tx, ty, tz, acc);
#pragma unroll
  for (int r = 0; r < T_RUN; ++r) {
    if (oi < g.nx && oj < g.ny && ok0 + r < g.nz) {
      const long long e = g.elem(oi, oj, ok0 + r);
      out[e] = out_scale ? acc[0][r] * out_scale[e] : acc[0][r];
    }
  }
}

// out[c] = sum over blocks b of partials[c * nb + b] for c < nc, 16 lanes of
// a 256-thread block per candidate (16 candidates per round), each lane
// summing a fixed stride of blocks, then the lanes in a fixed tree.
__device__ void sum_candidate_partials(const double* partials, int nb, int nc, double* out) {
  const int c16 = threadIdx.x >> 4, l = threadIdx.x & 15;
  for (int c0 = 0; c0 < nc; c0 += 16) {
    const int c = c0 + c16;
    double v = 0.0;
    if (c < nc) {
      // the same additions in the same order, with eight loads in flight at
      // a time (the partials come from other XCDs' writes: each load is a
      // trip to memory)
      const double* pc = partials + (size_t)c * nb;
      int b = l;
      for (; b + 16 * 7 < nb; b += 16 * 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = pc[b + 16 * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += x[u];
      }
      for (; b < nb; b += 16) v += pc[b];
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 16);
    if (c < nc && l == 0) out[c] = v;
  }
  __syncthreads();
}

// One block: sums the pass's volumes, then
//  1. tightens the bracket [a, b] of the volume crossing;
//  2. advances the bisection of pub_simp_solver.py:333-338 (lm = (l1+l2)/2,
//     l1 = lm when volume/n > volfrac else l2 = lm, stop when l2-l1 < 1e-9
//     or after max_steps) through every step the evaluations decide: a
//     midpoint that was itself evaluated uses its own volume, exactly like
//     the sequential loop; a midpoint at or below an evaluated multiplier
//     whose volume is above the target (at or above one whose volume is at
//     or below it) takes that decision, which the monotonicity of the volume
//     in lambda implies. Only evaluations whose volume clears the target by
//     more than any rounding-level non-monotonicity could bridge are used
//     this way (sa, sb);
//  3. chooses the next pass's multipliers: the three midpoints of the next
//     two bisection levels (exact decisions, so every pass advances at least
//     two steps), one third-level midpoint, and a secant prediction of the
//     crossing in s = lambda^(-1/2) (the unclipped volume is linear in s)
//     with ten geometrically spaced neighbours, which shrink [a, b] around
//     the crossing so that whole runs of later midpoints are decided. Once
//     [a, b] is narrower than the bisection's final resolution, the pass
//     evaluates the 15 midpoints of the next four levels instead.
__global__ void __launch_bounds__(256) k_oc_plan(int nb, long long n, double volfrac, int max_steps, double* state,
                                                Workspace ws) {
  if (state[S_DONE] != 0.0) return;
  __shared__ double svol[OC_NPTS];
  __shared__ double scand[OC_NPTS];
  if (threadIdx.x < OC_NPTS) scand[threadIdx.x] = state[S_CAND + threadIdx.x];
  sum_candidate_partials(ws.partials, nb, OC_NPTS, svol);   // ends with a barrier
  // one wavefront walks the bisection: lane t < OC_NPTS holds candidate t in
  // registers, so the per-step lookup is a ballot rather than a serial scan
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const double dn = (double)n;
  constexpr double kMargin = 2e-14;   // on volume/n; a faithfully rounded projection moves it by ~1e-16
  const double my_cand = lane < OC_NPTS ? scand[lane] : -1.0;
  const int my_dec = lane < OC_NPTS ? (svol[lane] / dn > volfrac ? 1 : 0) : 0;
  double a = state[S_A], va = state[S_VA], b = state[S_B], vb = state[S_VB];
  double sa = state[S_SA], sb = state[S_SB];
  for (int t = 0; t < OC_NPTS; ++t) {
    const double f = svol[t] / dn, ct = scand[t];
    if (f > volfrac) {
      if (ct > a) { a = ct; va = svol[t]; }
      if (f - volfrac > kMargin && ct > sa) sa = ct;
    } else {
      if (ct < b) { b = ct; vb = svol[t]; }
      if (volfrac - f > kMargin && ct < sb) sb = ct;
    }
  }
  double l1 = state[S_L1], l2 = state[S_L2], lm_last = state[S_LM], steps = state[S_STEPS];
  bool done = false;
  while (!done) {
    const double lm = 0.5 * (l1 + l2);
    int dec = -1;   // 1: volume above target (l1 = lm), 0: at or below (l2 = lm)
    // the lowest-numbered candidate equal to the midpoint decides
    const unsigned long long hit = __ballot(lane < OC_NPTS && my_cand == lm);
    if (hit) dec = __shfl(my_dec, __ffsll((long long)hit) - 1, 64);
    if (dec < 0 && sa >= 0.0 && lm <= sa) dec = 1;
    if (dec < 0 && sb < HUGE_VAL && lm >= sb) dec = 0;
    if (dec < 0) break;
    if (dec) l1 = lm; else l2 = lm;
    lm_last = lm;
    steps += 1.0;
    if (l2 - l1 < 1e-9 || steps >= max_steps) done = true;
  }
  if (lane == 0) {
    state[S_L1] = l1; state[S_L2] = l2; state[S_LM] = lm_last; state[S_STEPS] = steps;
    state[S_A] = a; state[S_VA] = va; state[S_B] = b; state[S_VB] = vb;
    state[S_SA] = sa; state[S_SB] = sb;
  }
  if (lane != 0) return;
  state[S_PASSES] += 1.0;
  state[S_DONE] = done ? 1.0 : 0.0;
  if (done) {
    state[S_PREV] = lm_last;
    return;
  }

  // next pass
  double nc[OC_NPTS];
  if (a >= 0.0 && b < HUGE_VAL && b - a < 1e-9) {
    // crossing resolved beyond the final bracket width: the next four
    // levels of midpoints, in heap order
    double lo_[OC_NPTS], hi_[OC_NPTS];
    lo_[0] = l1;
    hi_[0] = l2;
    for (int t = 0; t < OC_NPTS; ++t) {
      nc[t] = 0.5 * (lo_[t] + hi_[t]);
      if (2 * t + 2 < OC_NPTS) {
        lo_[2 * t + 1] = lo_[t]; hi_[2 * t + 1] = nc[t];
        lo_[2 * t + 2] = nc[t];  hi_[2 * t + 2] = hi_[t];
      }
    }
    for (int t = 0; t < OC_NPTS; ++t) state[S_CAND + t] = nc[t];
    return;
  }
  const double m0 = 0.5 * (l1 + l2), mlo = 0.5 * (l1 + m0), mhi = 0.5 * (m0 + l2);
  nc[0] = m0; nc[1] = mlo; nc[2] = mhi;
  const double hb = fmin(b, l2);   // l2 >= b whenever b exists
  double P = m0;                   // without a two-sided bracket: follow the bisection
  double L = 1.0;                  // log-width of the bracket
  if (a > 0.0 && b < HUGE_VAL && a < hb) {
    const double sa = 1.0 / sqrt(a), sb = 1.0 / sqrt(b);
    double s = 0.5 * (sa + sb);
    if (va > vb) s = sb + (volfrac * dn - vb) * (sa - sb) / (va - vb);
    P = 1.0 / (s * s);
    if (!(P > a && P < b)) P = sqrt(a * b);
    L = log(b / a);
  }
  // third-level midpoint on the side the prediction points to
  nc[3] = (P > m0) ? ((P > mhi) ? 0.5 * (mhi + l2) : 0.5 * (m0 + mhi))
                   : ((P > mlo) ? 0.5 * (mlo + m0) : 0.5 * (l1 + mlo));
  nc[4] = P;
  double dlt = (L > 0.0 && L < 50.0) ? L : 1.0;
  for (int k = 0; k < 5; ++k) {
    dlt *= 0.1;
    nc[5 + 2 * k] = P * exp(-dlt);
    nc[6 + 2 * k] = P * exp(dlt);
  }
  for (int t = 0; t < OC_NPTS; ++t) state[S_CAND + t] = nc[t] > 0.0 ? nc[t] : m0;
}

__global__ void __launch_bounds__(256) k_oc_finalize(long long n, const double* __restrict__ rho,
                                                    const double* __restrict__ q, const double* __restrict__ lo,
                                                    const double* __restrict__ hi, const double* state,
                                                    double* rho_new, Workspace ws) {
  const double lm = state[S_LM];
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

// ---------------------------------------------------------------------------
// Mean-volume OC of the reference GPU loop (src/topogpu/solver.py:94-117):
//   cand(lam) = clip(rho sqrt(max(-min(dc, -1e-12) / max(lam, 1e-40), 0)),
//                    max(rho - move, rho_min), min(rho + move, 1));
//   hi = 1, lo = 0; while mean(cand(hi)) > V and hi < 1e40: lo = hi, hi *= 2;
//   80 times: mid = (lo + hi) / 2; mean(cand(mid)) > V ? lo = mid : hi = mid;
//             stop when |mean(cand(mid)) - V| < 1e-8; result cand(mid).
// Each pass evaluates the mean at up to OCM_MAXC multipliers at once: the
// next 64 doublings of hi, or the 63 midpoints of the next six bisection
// levels (every path the sequential loop can take, each midpoint computed by
// the same floating-point operations). A one-block planner then walks the
// sequential loop through the evaluated means, so the decisions and the
// final multiplier are the loop's own. In the passes the candidate is
// rho sqrt(-dc) / sqrt(lam) (one rounding apart from the loop's form, which
// moves a mean by ~1e-16 of itself); the final density uses the loop's
// exact expression.
constexpr int OCM_MAXC = 64, OCM_DEPTH = 6;
enum OcmState {
  M_PHASE,   // 0: doubling hi, 1: bisection, 2: done
  M_LO, M_HI,
  M_STEPS,   // bisection steps taken
  M_LAM,     // final multiplier
  M_KBASE,   // first doubling exponent of the pass
  M_NC,      // candidates in the pass
  M_DONE,
  M_CAND,
  M_LEN = M_CAND + OCM_MAXC
};
static_assert(M_LEN <= TG_OC_MEAN_STATE_LEN, "mean-OC state layout");

__device__ __forceinline__ double ocm_exact(double rho, double dc, double move, double rho_min, double lam) {
  const double dcs = fmin(dc, -1.0e-12);
  const double x = rho * sqrt(fmax(-dcs / fmax(lam, 1.0e-40), 0.0));
  return fmin(fmax(x, fmax(rho - move, rho_min)), fmin(rho + move, 1.0));
}

// the 63 midpoints of the bisection subtree below (lo, hi), heap order:
// node p has children 2p+1 (hi = mid) and 2p+2 (lo = mid)
__device__ void ocm_tree(double lo, double hi, double* cand) {
  double L[OCM_MAXC], H[OCM_MAXC];
  L[0] = lo;
  H[0] = hi;
  constexpr int nodes = (1 << OCM_DEPTH) - 1;
  for (int p = 0; p < nodes; ++p) {
    const double mid = 0.5 * (L[p] + H[p]);
    cand[p] = mid;
    if (2 * p + 2 < nodes) {
      L[2 * p + 1] = L[p];
      H[2 * p + 1] = mid;
      L[2 * p + 2] = mid;
      H[2 * p + 2] = H[p];
    }
  }
}

__global__ void k_ocm_prepare(double* state) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < M_LEN; ++k) state[k] = 0.0;
  state[M_HI] = 1.0;
  state[M_NC] = OCM_MAXC;
  for (int k = 0; k < OCM_MAXC; ++k) state[M_CAND + k] = ldexp(1.0, k);
}

// Per block and candidate: the sum of the candidate densities of the
// block's elements (grid-stride, fixed assignment), 16 candidates at a time.
__global__ void __launch_bounds__(256) k_ocm_pass(long long n, const double* __restrict__ rho,
                                                  const double* __restrict__ dc, double move, double rho_min,
                                                  const double* state, Workspace ws) {
  if (state[M_DONE] != 0.0) return;   // uniform
  const int nc = (int)state[M_NC], nb = gridDim.x;
  for (int c0 = 0; c0 < nc; c0 += 16) {
    double rs[16], acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      rs[c] = c0 + c < nc ? 1.0 / sqrt(fmax(state[M_CAND + c0 + c], 1.0e-40)) : 0.0;
      acc[c] = 0.0;
    }
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)nb * blockDim.x) {
      const double r = rho[e], q = r * sqrt(-fmin(dc[e], -1.0e-12));
      const double lo = fmax(r - move, rho_min), hi = fmin(r + move, 1.0);
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[c] += fmin(fmax(q * rs[c], lo), hi);
    }
    block_sum<16>(acc);
    if (threadIdx.x == 0)
#pragma unroll
      for (int c = 0; c < 16; ++c)
        if (c0 + c < nc) ws.partials[(size_t)(c0 + c) * nb + blockIdx.x] = acc[c];
  }
}

// One block: candidate means (partials summed in block order), then the
// sequential loop walked through them; writes the next candidates.
__global__ void __launch_bounds__(256) k_ocm_plan(int nb, long long n, double volfrac, double* state, Workspace ws) {
  if (state[M_DONE] != 0.0) return;
  __shared__ double mean[OCM_MAXC];
  const int nc = (int)state[M_NC];
  sum_candidate_partials(ws.partials, nb, nc, mean);
  if (threadIdx.x != 0) return;
  for (int c = 0; c < nc; ++c) mean[c] /= (double)n;
  if (state[M_PHASE] == 0.0) {
    // candidates are hi = 2^(kbase + c): the loop moves on while the mean
    // stays above the target and hi < 1e40
    const int kb = (int)state[M_KBASE];
    for (int c = 0; c < nc; ++c) {
      const double hi = state[M_CAND + c];
      if (mean[c] > volfrac && hi < 1.0e40) {
        state[M_LO] = hi;
        continue;
      }
      state[M_HI] = hi;
      state[M_PHASE] = 1.0;
      ocm_tree(state[M_LO], hi, state + M_CAND);
      state[M_NC] = (1 << OCM_DEPTH) - 1;
      return;
    }
    state[M_KBASE] = kb + nc;
    for (int c = 0; c < OCM_MAXC; ++c) state[M_CAND + c] = ldexp(1.0, kb + nc + c);
    state[M_NC] = OCM_MAXC;
    return;
  }
  // bisection: walk the subtree
  double lo = state[M_LO], hi = state[M_HI];
  int steps = (int)state[M_STEPS];
  int p = 0;
  for (int d = 0; d < OCM_DEPTH; ++d) {
    const double mid = state[M_CAND + p];
    const double m = mean[p];
    ++steps;
    if (m > volfrac) {
      lo = mid;
      p = 2 * p + 2;
    } else {
      hi = mid;
      p = 2 * p + 1;
    }
    if (fabs(m - volfrac) < 1.0e-8 || steps >= 80) {
      state[M_LAM] = mid;
      state[M_STEPS] = steps;
      state[M_DONE] = 1.0;
      state[M_PHASE] = 2.0;
      return;
    }
  }
  state[M_LO] = lo;
  state[M_HI] = hi;
  state[M_STEPS] = steps;
  ocm_tree(lo, hi, state + M_CAND);
}

__global__ void __launch_bounds__(256) k_ocm_finalize(long long n, const double* __restrict__ rho,
                                                      const double* __restrict__ dc, double move, double rho_min,
                                                      const double* state, double* rho_new, Workspace ws) {
  const double lam = state[M_LAM];
  double mx = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const double r = ocm_exact(rho[e], dc[e], move, rho_min, lam);
    rho_new[e] = r;
    mx = fmax(mx, fabs(r - rho[e]));
  }
  __shared__ double smax[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) ws.partials[blockIdx.x] = fmax(fmax(smax[0], smax[1]), fmax(smax[2], smax[3]));
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
  Cube27 cw;
  if (make_cube27(st, cw) && g.elems() < (1LL << 31)) {
    const long long tiles = (long long)((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY) * ((g.nz + TZ - 1) / TZ);
    TG_REQUIRE(tiles < (1LL << 31), "stencil64: grid too large");
    with_cube_mask(cw.mask, [&](auto mk) {
      hipLaunchKernelGGL(k_stencil_tiled<decltype(mk)::value>, dim3((unsigned)tiles), dim3(T_THREADS), 0,
                         (hipStream_t)s, g, cw, in, in_scale, out_scale, out);
    });
  } else {
    hipLaunchKernelGGL(k_stencil, dim3((unsigned)((g.elems() + 255) / 256)), dim3(256), 0, (hipStream_t)s, g, st,
                       in, in_scale, out_scale, out);
  }
  TG_LAUNCH_CHECK();
  return TG_OK;
}

// Passive elements after the OC update (apply_passive_mask,
// bc_generator.py:454-470): void elements to rho_min, solid ones to 1.
__global__ void k_passive(long long n, const int32_t* __restrict__ mask, double* __restrict__ rho, double rho_min) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int m = mask[e];
  if (m == 1) rho[e] = rho_min;
  else if (m == 2) rho[e] = 1.0;
}

extern "C" int tg_passive64(int64_t n, const int32_t* mask, double* rho, double rho_min, void* stream) {
  TG_REQUIRE(n > 0 && mask && rho, "passive64: bad argument");
  hipLaunchKernelGGL(k_passive, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, (long long)n,
                     mask, rho, rho_min);
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

extern "C" int tg_sensitivity64_b(const tg_grid* gp, int32_t nrhs, const double* U, const double* weights,
                                  const double* rp, const double* dh, double penal, double erange, double* ce,
                                  double* dcf, void* s) {
  TG_NEED_INIT();
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(U && rp && dcf && nrhs >= 1 && nrhs <= 64, "sensitivity64: bad argument");
  const Grid g = to_grid(gp);
  hipLaunchKernelGGL(k_sensitivity, dim3((unsigned)((g.elems() + 255) / 256)), dim3(256), 0, (hipStream_t)s, g, nrhs,
                     U, weights, rp, dh, penal, erange, ce, dcf);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_sensitivity64(const tg_grid* gp, const double* U, const double* rp, const double* dh,
                                double penal, double erange, double* ce, double* dcf, void* s) {
  return tg_sensitivity64_b(gp, 1, U, nullptr, rp, dh, penal, erange, ce, dcf, s);
}

extern "C" int tg_compliance_dd(const tg_grid* gp, const double* E, const double* U, const double* F, double* out,
                                void* s) {
  TG_NEED_INIT();
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(E && U && F && out, "compliance_dd: NULL argument");
  const Grid g = to_grid(gp);
  constexpr int nb = 1024;
  hipLaunchKernelGGL(k_compliance_dd, dim3(nb), dim3(256), 0, (hipStream_t)s, g, E, U, F, workspace());
  TG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_compliance_dd_finalize, dim3(1), dim3(64), 0, (hipStream_t)s, nb, out, workspace());
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
  Cube27 cw;
  int nb;
  if (make_cube27(st, cw) && g.elems() < (1LL << 31)) {
    const long long tiles = (long long)((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY) * ((g.nz + TZ - 1) / TZ);
    TG_REQUIRE(tiles * OC_NPTS <= kMaxPartials, "grid too large for the reduction workspace");
    nb = (int)tiles;
    const unsigned blocks = (unsigned)((nb + kXcds - 1) / kXcds) * kXcds * OC_NG;
    with_cube_mask(cw.mask, [&](auto mk) {
      hipLaunchKernelGGL(k_oc_pass_tiled<decltype(mk)::value>, dim3(blocks), dim3(T_THREADS), 0, (hipStream_t)s, g,
                         cw, nb, hs_inv, q, lo, hi, beta, eta, use_h, state, workspace());
    });
  } else {
    nb = grid_blocks(g.elems());
    hipLaunchKernelGGL(k_oc_pass, dim3(nb), dim3(256), 0, (hipStream_t)s, g, st, hs_inv, q, lo, hi, beta, eta,
                       use_h, state, workspace());
  }
  TG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_oc_plan, dim3(1), dim3(256), 0, (hipStream_t)s, nb, g.elems(), volfrac, 100, state,
                     workspace());
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_oc_mean_prepare(double* state, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(state, "oc_mean_prepare: NULL state");
  hipLaunchKernelGGL(k_ocm_prepare, dim3(1), dim3(64), 0, (hipStream_t)s, state);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_oc_mean_step(int64_t n, const double* rho, const double* dc, double volfrac, double move,
                               double rho_min, double* state, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && rho && dc && state, "oc_mean_step: NULL argument");
  const int nb = grid_blocks(n) < 1024 ? grid_blocks(n) : 1024;
  TG_REQUIRE((long long)nb * OCM_MAXC <= kMaxPartials, "oc_mean_step: workspace too small");
  hipLaunchKernelGGL(k_ocm_pass, dim3(nb), dim3(256), 0, (hipStream_t)s, (long long)n, rho, dc, move, rho_min, state,
                     workspace());
  TG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ocm_plan, dim3(1), dim3(256), 0, (hipStream_t)s, nb, (long long)n, volfrac, state,
                     workspace());
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_oc_mean_finalize(int64_t n, const double* rho, const double* dc, double move, double rho_min,
                                   const double* state, double* rho_new, double* change, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && rho && dc && state && rho_new && change, "oc_mean_finalize: NULL argument");
  const int nb = grid_blocks(n);
  hipLaunchKernelGGL(k_ocm_finalize, dim3(nb), dim3(256), 0, (hipStream_t)s, (long long)n, rho, dc, move, rho_min,
                     state, rho_new, workspace());
  TG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max_finalize, dim3(1), dim3(256), 0, (hipStream_t)s, nb, change, workspace());
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
