// Node-centric hex8 operators on the vector ALUs (FP64 / FP32).
//
// Each thread owns one node and gathers A u at that node from the (up to)
// eight elements around it:
//     (A u)_n = sum_{e containing n} W_e * KE[rows of n in e] . u_e
// This is the transpose-free ("gather") form of the element-by-element
// product: every output is written exactly once, so there are no atomics and
// results are bit-reproducible, and the flop count equals the scatter form
// (8 elements x 3 rows x 24 = 576 FMA per node vs 24 x 24 per element).
// A workgroup stages the node brick of its 4x4x16 node tile plus a one-node
// halo and the element weights in LDS; each thread then holds its 27-node
// neighbourhood in registers.
#include <type_traits>

#include "common.h"
#include "epilogue.h"
#include "reduce.h"

namespace tg {

int launch_hex_mfma(const Grid& g, const uint8_t* mask, const float* E, const float* u, const Ep32& ep,
                    double* dot, bool split, int nrhs, hipStream_t s);
int launch_mfma_probe(const uint16_t* a, const uint16_t* b, float* d, hipStream_t s);
template <typename C>
int launch_stencil(const Grid& g, const uint8_t* mask, const C* coef, const float* u, const Ep32& ep,
                   double* dot, int nrhs, hipStream_t s);
int launch_q8_mfma(const Grid& g, const uint8_t* mask, const float* Ef, const float* u, const Ep32& ep, double* dot,
                   bool split, int nrhs, hipStream_t s);

namespace {

constexpr int NTX = 4, NTY = 4, NTZ = 16, NT = NTX * NTY * NTZ;   // owned nodes
constexpr int BX = NTX + 2, BY = NTY + 2, BZ = NTZ + 2;           // node brick
constexpr int EX = NTX + 1, EY = NTY + 1, EZ = NTZ + 1;           // element brick

__device__ __forceinline__ void tile_origin(const Grid& g, int& i0, int& j0, int& k0) {
  const int ntz = (g.NZ() + NTZ - 1) / NTZ, nty = (g.NY() + NTY - 1) / NTY;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  k0 = (t % ntz) * NTZ;
  j0 = ((t / ntz) % nty) * NTY;
  i0 = (t / (ntz * nty)) * NTX;
}

int node_tiles(const Grid& g) {
  return ((g.NX() + NTX - 1) / NTX) * ((g.NY() + NTY - 1) / NTY) * ((g.NZ() + NTZ - 1) / NTZ);
}

template <typename T> __device__ __forceinline__ const T* ke_const();
template <> __device__ __forceinline__ const double* ke_const<double>() { return c_ke64; }
template <> __device__ __forceinline__ const float* ke_const<float>() { return c_ke32; }

// Element weights: E per element (one matrix KE).
template <typename T>
struct HexW {
  static constexpr int NS = 1;
  const T* E;
  __device__ T operator()(const Grid& g, int i, int j, int k, int) const { return E[g.elem(i, j, k)]; }
  __device__ const T* mat(int) const { return ke_const<T>(); }
};

// Implicit Galerkin level 1: coarse element (I,J,K) has matrix
// sum_s E_fine(2I+sx, 2J+sy, 2K+sz) * Q_s (requires even fine dims).
struct Q8W {
  static constexpr int NS = 8;
  const float* Ef;
  Grid fine;
  __device__ float operator()(const Grid&, int I, int J, int K, int s) const {
    return Ef[fine.elem(2 * I + (s >> 2 & 1), 2 * J + (s >> 1 & 1), 2 * K + (s & 1))];
  }
  __device__ const float* mat(int s) const { return c_q8_32 + 576 * s; }
};

// PRE: the brick holds (pa * dinv) * u instead of u (the Chebyshev start
// d0 of the fused first smoothing step, rounded as the separate start is).
template <typename T, bool PRE>
__device__ __forceinline__ void load_node_brick(const Grid& g, const uint8_t* __restrict__ mask,
                                                const T* __restrict__ u, int i0, int j0, int k0,
                                                T* su, const float* __restrict__ pd, float pa) {
  // All loads first (none depends on another), then the masked LDS stores.
  constexpr int NB = BX * BY * BZ, NIT = (NB + NT - 1) / NT;
  T v[NIT][3];
  unsigned vm[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = threadIdx.x + it * NT;
    const int bz = idx % BZ, by = (idx / BZ) % BY, bx = idx / (BZ * BY);
    const int i = i0 - 1 + bx, j = j0 - 1 + by, k = k0 - 1 + bz;
    const bool in = idx < NB && i >= 0 && i <= g.nx && j >= 0 && j <= g.ny && k >= 0 && k <= g.nz;
    const long long n = in ? g.node(i, j, k) : 0;
    vm[it] = in ? (unsigned)mask[n] : 0u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if constexpr (PRE) v[it][c] = in ? (pa * pd[3 * n + c]) * u[3 * n + c] : T(0);
      else v[it][c] = in ? u[3 * n + c] : T(0);
    }
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = threadIdx.x + it * NT;
    if (idx < NB)
#pragma unroll
      for (int c = 0; c < 3; ++c) su[3 * idx + c] = (vm[it] >> c & 1u) ? v[it][c] : T(0);
  }
}

// Generic node-centric operator with element weights W.
template <typename T, class W, class Ep>
__global__ void __launch_bounds__(NT) k_node_op(Grid g, const uint8_t* __restrict__ mask, W w,
                                               const T* __restrict__ u, Ep ep_all, double* dot,
                                               Workspace ws) {
  constexpr int NS = W::NS;
  __shared__ T su[BX * BY * BZ * 3];
  __shared__ T sw[EX * EY * EZ * NS];
  // right-hand side blockIdx.y of a batch
  const long long roff = (long long)blockIdx.y * 3 * g.nodes();
  u += roff;
  const Ep ep = ep_all.slice(roff);
  int i0, j0, k0;
  tile_origin(g, i0, j0, k0);
  if constexpr (Ep::kPrescale) load_node_brick<T, true>(g, mask, u, i0, j0, k0, su, ep.dinv, ep.a3);
  else load_node_brick<T, false>(g, mask, u, i0, j0, k0, su, nullptr, 0.0f);
  for (int idx = threadIdx.x; idx < EX * EY * EZ; idx += blockDim.x) {
    const int ez = idx % EZ, ey = (idx / EZ) % EY, ex = idx / (EZ * EY);
    const int i = i0 - 1 + ex, j = j0 - 1 + ey, k = k0 - 1 + ez;
    const bool in = i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
#pragma unroll
    for (int s = 0; s < NS; ++s) sw[idx * NS + s] = in ? w(g, i, j, k, s) : T(0);
  }
  __syncthreads();

  const int tz = threadIdx.x % NTZ, ty = (threadIdx.x / NTZ) % NTY, tx = threadIdx.x / (NTZ * NTY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  const bool own = i <= g.nx && j <= g.ny && k <= g.nz;
  // epilogue inputs are fetched now so their latency hides behind the FMAs
  const long long n_own = own ? g.node(i, j, k) : 0;
  const unsigned m_own = own ? (unsigned)mask[n_own] : 0u;
  typename Ep::In ein;
  if (own) ep.load(n_own, ein);
  double part = 0.0;
  if (own && sizeof(T) == 4) {
    // FP32: all 27 neighbours in registers (81 values), each read from LDS once
    T nb[27][3];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
      const int dx = d / 9, dy = (d / 3) % 3, dz = d % 3;
      const int idx = ((tx + dx) * BY + (ty + dy)) * BZ + tz + dz;
#pragma unroll
      for (int c = 0; c < 3; ++c) nb[d][c] = su[3 * idx + c];
    }
    T acc[3] = {0, 0, 0};
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = corner_x(o), oy = corner_y(o), oz = corner_z(o);
      const int eidx = ((tx - ox + 1) * EY + (ty - oy + 1)) * EZ + (tz - oz + 1);
#pragma unroll 1
      for (int s = 0; s < NS; ++s) {
        const T we = sw[eidx * NS + s];
        const T* K = w.mat(s) + 72 * o;     // rows 3*o .. 3*o+2
        T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int nbi = ((1 - ox + corner_x(c)) * 3 + (1 - oy + corner_y(c))) * 3 + (1 - oz + corner_z(c));
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const T uv = nb[nbi][q];
            t0 = fma(K[0 * 24 + 3 * c + q], uv, t0);
            t1 = fma(K[1 * 24 + 3 * c + q], uv, t1);
            t2 = fma(K[2 * 24 + 3 * c + q], uv, t2);
          }
        }
        acc[0] = fma(we, t0, acc[0]);
        acc[1] = fma(we, t1, acc[1]);
        acc[2] = fma(we, t2, acc[2]);
      }
    }
    const T own_u[3] = {nb[13][0], nb[13][1], nb[13][2]};
    part = ep.store(n_own, m_own, acc, own_u, ein);
  } else if (own) {
    // FP64: one element at a time, its 8 corner displacements read from LDS
    // per element (24 values in registers instead of 81 doubles), which
    // leaves room for several waves per SIMD to hide the FMA and LDS
    // latencies (2.7x faster than the 27-neighbour form at 160x80x80).
    T acc[3] = {0, 0, 0};
#pragma unroll 1
    for (int o = 0; o < 8; ++o) {
      const int ox = corner_x(o), oy = corner_y(o), oz = corner_z(o);
      const int eidx = ((tx - ox + 1) * EY + (ty - oy + 1)) * EZ + (tz - oz + 1);
      const int ebase = ((tx - ox + 1) * BY + (ty - oy + 1)) * BZ + (tz - oz + 1);   // element's corner 0
      T ue[8][3];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int idx = ebase + (corner_x(c) * BY + corner_y(c)) * BZ + corner_z(c);
#pragma unroll
        for (int q = 0; q < 3; ++q) ue[c][q] = su[3 * idx + q];
      }
#pragma unroll 1
      for (int s = 0; s < NS; ++s) {
        const T we = sw[eidx * NS + s];
        const T* K = w.mat(s) + 72 * o;     // rows 3*o .. 3*o+2
        T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const T uv = ue[c][q];
            t0 = fma(K[0 * 24 + 3 * c + q], uv, t0);
            t1 = fma(K[1 * 24 + 3 * c + q], uv, t1);
            t2 = fma(K[2 * 24 + 3 * c + q], uv, t2);
          }
        acc[0] = fma(we, t0, acc[0]);
        acc[1] = fma(we, t1, acc[1]);
        acc[2] = fma(we, t2, acc[2]);
      }
    }
    const int cidx = ((tx + 1) * BY + (ty + 1)) * BZ + tz + 1;
    const T own_u[3] = {su[3 * cidx], su[3 * cidx + 1], su[3 * cidx + 2]};
    part = ep.store(n_own, m_own, acc, own_u, ein);
  }
  if (ep.reduces()) {
    double v[1] = {part};
    block_partials<1>(v, ws);
  }
}

// Stored element matrices (576 FP32 per element, row-major).
template <class Ep>
__global__ void __launch_bounds__(NT) k_node_stored(Grid g, const uint8_t* __restrict__ mask,
                                                   const float* __restrict__ mats,
                                                   const float* __restrict__ u, Ep ep_all, double* dot,
                                                   Workspace ws) {
  __shared__ float su[BX * BY * BZ * 3];
  const long long roff = (long long)blockIdx.y * 3 * g.nodes();
  u += roff;
  const Ep ep = ep_all.slice(roff);
  int i0, j0, k0;
  tile_origin(g, i0, j0, k0);
  if constexpr (Ep::kPrescale) load_node_brick<float, true>(g, mask, u, i0, j0, k0, su, ep.dinv, ep.a3);
  else load_node_brick<float, false>(g, mask, u, i0, j0, k0, su, nullptr, 0.0f);
  __syncthreads();
  const int tz = threadIdx.x % NTZ, ty = (threadIdx.x / NTZ) % NTY, tx = threadIdx.x / (NTZ * NTY);
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  const bool own = i <= g.nx && j <= g.ny && k <= g.nz;
  double part = 0.0;
  if (own) {
    float nb[27][3];
#pragma unroll
    for (int d = 0; d < 27; ++d) {
      const int dx = d / 9, dy = (d / 3) % 3, dz = d % 3;
      const int idx = ((tx + dx) * BY + (ty + dy)) * BZ + tz + dz;
#pragma unroll
      for (int c = 0; c < 3; ++c) nb[d][c] = su[3 * idx + c];
    }
    float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = corner_x(o), oy = corner_y(o), oz = corner_z(o);
      const int ei = i - ox, ej = j - oy, ek = k - oz;
      if (ei < 0 || ei >= g.nx || ej < 0 || ej >= g.ny || ek < 0 || ek >= g.nz) continue;
      const float* K = mats + 576 * g.elem(ei, ej, ek) + 72 * o;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int nbi = ((1 - ox + corner_x(c)) * 3 + (1 - oy + corner_y(c))) * 3 + (1 - oz + corner_z(c));
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const float uv = nb[nbi][q];
          acc[0] = fmaf(K[0 * 24 + 3 * c + q], uv, acc[0]);
          acc[1] = fmaf(K[1 * 24 + 3 * c + q], uv, acc[1]);
          acc[2] = fmaf(K[2 * 24 + 3 * c + q], uv, acc[2]);
        }
      }
    }
    const long long n = g.node(i, j, k);
    const float own_u[3] = {nb[13][0], nb[13][1], nb[13][2]};
    part = ep(n, mask[n], acc, own_u);
  }
  if (ep.reduces()) {
    double v[1] = {part};
    block_partials<1>(v, ws);
  }
}

// ---- diagonals: sum over the elements at a node of W * K[r][r]
__device__ __forceinline__ void write_dinv(long long n, unsigned m, const double (&d)[3], double* d64,
                                           float* d32) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = (!(m >> c & 1u) || fabs(d[c]) <= 1e-12) ? 1.0 : 1.0 / d[c];
    if (d64) d64[3 * n + c] = v;
    if (d32) d32[3 * n + c] = (float)v;
  }
}

// KIND 0: E per element (KE); 1: fine E (Q8 sums); 2: stored FP64 matrices.
// Each element's weights are loaded once (all eight children's for KIND 1)
// and the loops are unrolled, so a node's loads are all in flight together;
// every d[r] adds its terms in (corner, child) order.
template <int KIND>
__global__ void __launch_bounds__(256) k_diag(Grid g, const uint8_t* __restrict__ mask, const double* __restrict__ w,
                                             Grid fine, double* d64, float* d32) {
  const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;   // nodes < 2^32 (checked by the launcher)
  if (n >= (unsigned)g.nodes()) return;
  const unsigned t = n / (unsigned)g.NZ();
  const int k = (int)(n - t * (unsigned)g.NZ()), j = (int)(t % (unsigned)g.NY()), i = (int)(t / (unsigned)g.NY());
  double d[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ei = i - corner_x(o), ej = j - corner_y(o), ek = k - corner_z(o);
    if (ei < 0 || ei >= g.nx || ej < 0 || ej >= g.ny || ek < 0 || ek >= g.nz) continue;
    if constexpr (KIND == 0) {
      const double we = w[g.elem(ei, ej, ek)];
#pragma unroll
      for (int r = 0; r < 3; ++r) d[r] += we * c_ke64[(3 * o + r) * 25];
    } else if constexpr (KIND == 1) {
      double ws[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) ws[s] = w[fine.elem(2 * ei + (s >> 2 & 1), 2 * ej + (s >> 1 & 1), 2 * ek + (s & 1))];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 8; ++s) d[r] += ws[s] * (double)c_q8_32[576 * s + (3 * o + r) * 25];
    } else {
      const double* m = w + 576 * g.elem(ei, ej, ek);
#pragma unroll
      for (int r = 0; r < 3; ++r) d[r] += m[(3 * o + r) * 25];
    }
  }
  write_dinv(n, mask[n], d, d64, d32);
}

template <typename T, class W, class Ep>
int launch_node(const Grid& g, const uint8_t* mask, W w, const T* u, const Ep& ep, double* dot, int nrhs,
                hipStream_t s) {
  const int nb = node_tiles(g);
  TG_REQUIRE((long long)nb * nrhs <= kMaxPartials, "grid too large for the reduction workspace");
  hipLaunchKernelGGL((k_node_op<T, W, Ep>), dim3(nb, nrhs), dim3(NT), 0, s, g, mask, w, u, ep, dot, workspace());
  TG_LAUNCH_CHECK();
  return ep.reduces() ? finalize_batched(1, nb, nrhs, dot, nullptr, s) : TG_OK;
}

// Scratch of the allocation-free FP64 residual: the per-block partials of
// its reduction, in a file-scope device array (one copy per GPU, held until
// the process ends) instead of the workspace tg_init allocates. Host code
// never uses the array's name as an address: tg_init resolves each device's
// copy with hipGetSymbolAddress into g_scratch[dev], and the kernels receive
// that pointer.
constexpr int kStaticPartials = 1 << 16;   // 512 KiB: grids up to 65,536 node tiles
static __device__ __align__(256) double scratch[kStaticPartials];
constexpr int kScratchDevices = 64;
double* g_scratch[kScratchDevices] = {};

}  // namespace

int init_static_scratch(int dev) {
  TG_REQUIRE(dev >= 0 && dev < kScratchDevices, "device %d out of range", dev);
  if (g_scratch[dev]) return TG_OK;
  void* ptr = nullptr;
  TG_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  TG_REQUIRE(ptr != nullptr, "static scratch has no device address");
  g_scratch[dev] = static_cast<double*>(ptr);
  return TG_OK;
}

}  // namespace tg

using namespace tg;

#define TG_NEED_INIT() TG_REQUIRE(::tg::ke_ready(), "tg_init has not been called on this device")

extern "C" int tg_apply32_b(const tg_op32* op, int32_t nrhs, const float* u, float* y, const tg_ep32* ep,
                            void* stream) {
  TG_NEED_INIT();
  TG_REQUIRE(op && ep && u, "tg_apply32: NULL operator, epilogue or input");
  TG_REQUIRE(nrhs >= 1 && nrhs <= 64, "tg_apply32: nrhs must be 1..64");
  if (int rc = check_grid(&op->grid)) return rc;
  TG_REQUIRE(op->mask, "tg_apply32: mask is NULL");
  const Grid g = to_grid(&op->grid);
  Ep32 e{ep->mode, ep->rhs, ep->dinv, ep->d, ep->x, ep->r, ep->a1, ep->a2, ep->a3, y};
  switch (ep->mode) {
    case TG_EP_STORE: TG_REQUIRE(y, "y is NULL"); break;
    case TG_EP_DOT: case TG_EP_RESID_NORM:
      TG_REQUIRE(y && ep->dot, "y or dot is NULL");
      TG_REQUIRE(ep->mode == TG_EP_DOT || ep->rhs, "rhs is NULL");
      break;
    case TG_EP_RESID: TG_REQUIRE(y && ep->rhs, "y or rhs is NULL"); break;
    case TG_EP_CHEB:
      TG_REQUIRE(ep->rhs && ep->dinv && ep->d && ep->x, "CHEB needs rhs, dinv, d, x");
      TG_REQUIRE(ep->x != u, "CHEB x must differ from the input");
      break;
    case TG_EP_CHEB_R:
      TG_REQUIRE(ep->r && ep->dinv && ep->d && ep->x, "CHEB_R needs r, dinv, d, x");
      TG_REQUIRE(ep->d != u && ep->r != u && ep->x != u, "CHEB_R outputs must differ from the input");
      break;
    case TG_EP_RESID_UPDATE:
      TG_REQUIRE(ep->r && ep->r != u, "RESID_UPDATE needs r distinct from the input");
      break;
    case TG_EP_CHEB_FIRST:
      TG_REQUIRE(ep->rhs && ep->dinv && ep->d && ep->x && ep->r, "CHEB_FIRST needs rhs, dinv, d, x, r");
      TG_REQUIRE(ep->d != u && ep->r != u && ep->x != u && ep->d != ep->rhs && ep->r != ep->rhs &&
                     ep->x != ep->rhs && ep->d != ep->x && ep->d != ep->r && ep->x != ep->r,
                 "CHEB_FIRST outputs must be distinct and differ from the input and rhs");
      break;
    case TG_EP_RESID_FIRST:
      TG_REQUIRE(ep->rhs && ep->dinv && ep->x && ep->r, "RESID_FIRST needs rhs, dinv, x, r");
      TG_REQUIRE(ep->r != u && ep->x != u && ep->r != ep->rhs && ep->x != ep->rhs && ep->x != ep->r,
                 "RESID_FIRST outputs must be distinct and differ from the input and rhs");
      break;
    case TG_EP_CHEB_LAST:
      TG_REQUIRE(ep->r && ep->dinv && ep->x, "CHEB_LAST needs r, dinv, x");
      TG_REQUIRE(ep->x != u && ep->x != ep->r, "CHEB_LAST x must differ from the input and r");
      break;
    default: return set_error(TG_ERR_ARG, "unknown epilogue mode %d", ep->mode);
  }
  hipStream_t s = (hipStream_t)stream;
  switch (op->kind) {
    case TG_OP_HEX_F32:
      TG_REQUIRE(op->weights, "weights NULL");
      return with_mode(e, [&](const auto& em) {
        return launch_node<float>(g, op->mask, HexW<float>{op->weights}, u, em, ep->dot, nrhs, s);
      });
    case TG_OP_HEX_BF16:
    case TG_OP_HEX_BF16X3:
      TG_REQUIRE(op->weights, "weights NULL");
      return launch_hex_mfma(g, op->mask, op->weights, u, e, ep->dot, op->kind == TG_OP_HEX_BF16X3, nrhs, s);
    case TG_OP_Q8_F32: {
      TG_REQUIRE(op->weights, "weights NULL");
      const Grid fine{2 * g.nx, 2 * g.ny, 2 * g.nz};
      return with_mode(e, [&](const auto& em) {
        return launch_node<float>(g, op->mask, Q8W{op->weights, fine}, u, em, ep->dot, nrhs, s);
      });
    }
    case TG_OP_STORED_F32: {
      TG_REQUIRE(op->elem_mats, "elem_mats NULL");
      const int nb = node_tiles(g);
      return with_mode(e, [&](const auto& em) {
        using E = std::decay_t<decltype(em)>;
        TG_REQUIRE((long long)nb * nrhs <= kMaxPartials, "grid too large for the reduction workspace");
        hipLaunchKernelGGL((k_node_stored<E>), dim3(nb, nrhs), dim3(NT), 0, s, g, op->mask, op->elem_mats, u, em,
                           ep->dot, workspace());
        TG_LAUNCH_CHECK();
        return E::reduces() ? finalize_batched(1, nb, nrhs, ep->dot, nullptr, s) : TG_OK;
      });
    }
    case TG_OP_Q8_BF16:
    case TG_OP_Q8_BF16X3:
      TG_REQUIRE(op->weights, "weights NULL");
      return launch_q8_mfma(g, op->mask, op->weights, u, e, ep->dot, op->kind == TG_OP_Q8_BF16X3, nrhs, s);
    case TG_OP_STENCIL_F32:
      TG_REQUIRE(op->stencil, "stencil NULL");
      return launch_stencil(g, op->mask, static_cast<const float*>(op->stencil), u, e, ep->dot, nrhs, s);
    case TG_OP_STENCIL_BF16:
      TG_REQUIRE(op->stencil, "stencil NULL");
      return launch_stencil(g, op->mask, static_cast<const uint16_t*>(op->stencil), u, e, ep->dot, nrhs, s);
    default:
      return set_error(TG_ERR_ARG, "unknown operator kind %d", op->kind);
  }
}

extern "C" int tg_apply32(const tg_op32* op, const float* u, float* y, const tg_ep32* ep, void* stream) {
  return tg_apply32_b(op, 1, u, y, ep, stream);
}

extern "C" int tg_apply64_b(const tg_grid* gp, const uint8_t* mask, const double* E, int32_t nrhs, const double* u,
                            const double* rhs, double* y, float* y32, double* sumsq, void* stream) {
  TG_NEED_INIT();
  if (int rc = check_grid(gp)) return rc;
  TG_REQUIRE(mask && E && u && (y || y32), "tg_apply64: NULL argument");
  TG_REQUIRE(nrhs >= 1 && nrhs <= 64, "tg_apply64: nrhs must be 1..64");
  const Grid g = to_grid(gp);
  Ep64 e{rhs, y, y32, sumsq != nullptr};
  return launch_node<double>(g, mask, HexW<double>{E}, u, e, sumsq, nrhs, (hipStream_t)stream);
}
extern "C" int tg_apply64(const tg_grid* gp, const uint8_t* mask, const double* E, const double* u,
                          const double* rhs, double* y, float* y32, double* sumsq, void* stream) {
  return tg_apply64_b(gp, mask, E, 1, u, rhs, y, y32, sumsq, stream);
}

extern "C" int tg_apply64_noalloc(const tg_grid* gp, const uint8_t* mask, const double* E, const double* u,
                                  const double* rhs, double* y, float* y32, double* sumsq, void* stream) {
  TG_NEED_INIT();
  if (int rc = check_grid(gp)) return rc;
  TG_REQUIRE(mask && E && u && (y || y32), "tg_apply64_noalloc: NULL argument");
  int dev = -1;
  TG_HIP_CHECK(hipGetDevice(&dev));
  TG_REQUIRE(dev >= 0 && dev < kScratchDevices && g_scratch[dev], "tg_apply64_noalloc: no scratch on device %d",
             dev);
  const Grid g = to_grid(gp);
  const int nb = node_tiles(g);
  TG_REQUIRE(nb <= kStaticPartials,
             "tg_apply64_noalloc: %d node tiles exceed the static scratch of %d partials (grids up to %d nodes)",
             nb, kStaticPartials, kStaticPartials * NT);
  Ep64 e{rhs, y, y32, sumsq != nullptr};
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL((k_node_op<double, HexW<double>, Ep64>), dim3(nb, 1), dim3(NT), 0, s, g, mask, HexW<double>{E},
                     u, e, sumsq, Workspace{g_scratch[dev], nullptr});
  TG_LAUNCH_CHECK();
  return e.reduces() ? finalize_batched_at(g_scratch[dev], (size_t)kStaticPartials, 1, nb, 1, sumsq, nullptr, s)
                     : TG_OK;
}

extern "C" int tg_mfma_probe(const uint16_t* a, const uint16_t* b, float* d, void* stream) {
  TG_REQUIRE(a && b && d, "mfma_probe: NULL argument");
  return launch_mfma_probe(a, b, d, (hipStream_t)stream);
}

static int launch_diag(const tg_grid* gp, const uint8_t* mask, int kind, const double* w, Grid fine,
                       double* d64, float* d32, void* stream) {
  TG_NEED_INIT();
  if (int rc = check_grid(gp)) return rc;
  TG_REQUIRE(mask && w && (d64 || d32), "diag: NULL argument");
  const Grid g = to_grid(gp);
  const long long n = g.nodes();
  TG_REQUIRE(n < (1LL << 32) - 256, "diag: grid of %lld nodes exceeds 32-bit node indices", n);
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  const hipStream_t st = (hipStream_t)stream;
  if (kind == 0) hipLaunchKernelGGL(k_diag<0>, grid, block, 0, st, g, mask, w, fine, d64, d32);
  else if (kind == 1) hipLaunchKernelGGL(k_diag<1>, grid, block, 0, st, g, mask, w, fine, d64, d32);
  else hipLaunchKernelGGL(k_diag<2>, grid, block, 0, st, g, mask, w, fine, d64, d32);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_diag_inv_hex(const tg_grid* g, const uint8_t* mask, const double* E, double* d64,
                               float* d32, void* s) {
  return launch_diag(g, mask, 0, E, Grid{0, 0, 0}, d64, d32, s);
}
extern "C" int tg_diag_inv_q8(const tg_grid* g, const uint8_t* mask, const double* fineE, float* d32,
                              void* s) {
  if (int rc = check_grid(g)) return rc;
  return launch_diag(g, mask, 1, fineE, Grid{2 * g->nelx, 2 * g->nely, 2 * g->nelz}, nullptr, d32, s);
}
extern "C" int tg_diag_inv_stored(const tg_grid* g, const uint8_t* mask, const double* mats, float* d32,
                                  void* s) {
  return launch_diag(g, mask, 2, mats, Grid{0, 0, 0}, nullptr, d32, s);
}
