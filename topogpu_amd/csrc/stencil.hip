// Coarse Galerkin levels as assembled 27-point block stencils.
//
// A coarse level's operator M K M (K assembled from its element matrices,
// M the free mask) is stored per node and row component as the 27 x 3
// coefficients that couple it to the three components of its 27 neighbours
// (offsets -1..1 in x, y, z; neighbour index k = (dx+1)*9 + (dy+1)*3 + dz+1).
// Layout is structure-of-arrays, coef[(c*81 + 3*k + d) * nodes + node], so a
// thread per (node, row) reads its 81 coefficients with consecutive threads
// on consecutive addresses. Against stored element matrices (576 values per
// element, each element read by its 8 nodes) the stencil is 243 values per
// node: less than half the bytes, read once, and no per-element gather.
// Entries of fixed rows or fixed columns and entries reaching outside the grid
// are zero, so the product needs no bounds logic beyond clamping the
// neighbour index into the vector (a zero coefficient times a valid value).
#include <type_traits>

#include "common.h"
#include "epilogue.h"
#include "reduce.h"
#include "stencil_coef.h"

namespace tg {
namespace {

constexpr int SNT = 256;

// One thread per (node, row component): sum over the elements containing the
// node of their 3 x 24 row block, in FP64, then masked and rounded to FP32.
template <typename C>
__global__ void __launch_bounds__(SNT) k_assemble_stencil(Grid g, const uint8_t* __restrict__ mask,
                                                         const double* __restrict__ mats, C* __restrict__ coef) {
  const long long n = blockIdx.x * (long long)SNT + threadIdx.x;
  const int c = blockIdx.y;
  const long long nn = g.nodes();
  if (n >= nn) return;
  const int k = (int)(n % g.NZ()), j = (int)((n / g.NZ()) % g.NY()), i = (int)(n / ((long long)g.NZ() * g.NY()));
  double acc[27][3];
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q][0] = acc[q][1] = acc[q][2] = 0.0;
  const unsigned mrow = mask[n];
  if (mrow >> c & 1u) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {   // the node is corner a of element (i,j,k) - corner(a)
      const int ei = i - corner_x(a), ej = j - corner_y(a), ek = k - corner_z(a);
      if (ei < 0 || ei >= g.nx || ej < 0 || ej >= g.ny || ek < 0 || ek >= g.nz) continue;
      const double* K = mats + 576 * g.elem(ei, ej, ek) + (3 * a + c) * 24;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int q = (corner_x(b) - corner_x(a) + 1) * 9 + (corner_y(b) - corner_y(a) + 1) * 3 +
                      (corner_z(b) - corner_z(a) + 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[q][d] += K[3 * b + d];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int dx = q / 9 - 1, dy = (q / 3) % 3 - 1, dz = q % 3 - 1;
    const int a = i + dx, b = j + dy, e = k + dz;
    const bool in = a >= 0 && a < g.NX() && b >= 0 && b < g.NY() && e >= 0 && e < g.NZ();
    const unsigned mcol = in ? mask[g.node(a, b, e)] : 0u;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      coef_store(&coef[(long long)(c * 81 + 3 * q + d) * nn + n], (mcol >> d & 1u) ? (float)acc[q][d] : 0.0f);
  }
}

// One workgroup = SAP nodes x 3 waves; wave w sums the 9 neighbours in the
// x-plane dx = w - 1 (27 coefficient and 27 value loads per lane, all in
// flight at once: the coarse levels are small and bound by load latency, and
// a lane with all 81 products would need several dependent rounds of loads).
// Wave 2 adds the three plane sums in the order dx = -1, 0, 1 and applies the
// epilogue.
constexpr int SAP = 64;

template <typename C, int MODE>
__global__ void __launch_bounds__(3 * SAP) k_stencil_apply(Grid g, const uint8_t* __restrict__ mask,
                                                          const C* __restrict__ coef, const float* __restrict__ u,
                                                          Ep32 ep_all, Workspace ws) {
  using Mode = EpMode<MODE>;
  __shared__ float plane_sum[2][SAP];
  const long long nn = g.nodes();
  // right-hand side blockIdx.z of a batch
  u += (long long)blockIdx.z * 3 * nn;
  const Ep32 ep = ep_slice(ep_all, (long long)blockIdx.z * 3 * nn);
  const int lane = threadIdx.x % SAP;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / SAP);
  const long long n = blockIdx.x * (long long)SAP + lane;
  const int c = blockIdx.y;
  const bool own = n < nn;
  const long long nc = own ? n : nn - 1;
  // the operator's input (prescaled modes: the Chebyshev start a3 * dinv * u)
  auto in_at = [&](long long i) -> float {
    if constexpr (Mode::kPrescale) return mul_rounded(ep.a3 * ep.dinv[i], u[i]);
    else return u[i];
  };
  float p = 0.0f, q = 0.0f, s = 0.0f, uo = 0.0f;
  const PtrAcc acc{ep, 3 * nc};
  if (w == 2) {
    Mode::load_one(ep, acc, c, p, q, s);
    uo = in_at(3 * nc + c);
  }
  const long long sz = g.NZ(), sy = (long long)g.NZ() * g.NY();
  const C* cf = coef + (long long)(c * 81 + 27 * w) * nn + nc;
  const long long base = nc + (w - 1) * sy;
  float au = 0.0f;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    long long nb = base + (j / 3 - 1) * sz + (j % 3 - 1);
    nb = nb < 0 ? 0 : (nb >= nn ? nn - 1 : nb);
#pragma unroll
    for (int d = 0; d < 3; ++d) au = fmaf(coef_load(cf + (long long)(3 * j + d) * nn), in_at(3 * nb + d), au);
  }
  if (w < 2) plane_sum[w][lane] = au;
  __syncthreads();
  double part = 0.0;
  if (w == 2 && own) {
    au = (plane_sum[0][lane] + plane_sum[1][lane]) + au;
    const bool f = mask[n] >> c & 1u;
    part = Mode::store_one(ep, acc, c, f, au, f ? uo : 0.0f, p, q, s);
  }
  if constexpr (Mode::kReduces) {
    double v[1] = {part};
    block_partials<1>(v, ws);
  }
}

}  // namespace

template <typename C>
int launch_stencil(const Grid& g, const uint8_t* mask, const C* coef, const float* u, const Ep32& ep,
                   double* dot, int nrhs, hipStream_t s) {
  const long long nn = g.nodes();
  const long long nbx = (nn + SAP - 1) / SAP;
  TG_REQUIRE(nbx * 3 * nrhs <= kMaxPartials, "grid too large for the reduction workspace");
  return with_mode(ep, [&](const auto& em) {
    constexpr int M = std::decay_t<decltype(em)>::kMode;
    hipLaunchKernelGGL((k_stencil_apply<C, M>), dim3((unsigned)nbx, 3, nrhs), dim3(3 * SAP), 0, s, g, mask, coef, u,
                       ep, workspace());
    TG_LAUNCH_CHECK();
    return EpMode<M>::kReduces ? finalize_batched(1, (int)(nbx * 3), nrhs, dot, nullptr, s) : TG_OK;
  });
}

template int launch_stencil<float>(const Grid&, const uint8_t*, const float*, const float*, const Ep32&, double*,
                                   int, hipStream_t);
template int launch_stencil<uint16_t>(const Grid&, const uint8_t*, const uint16_t*, const float*, const Ep32&,
                                      double*, int, hipStream_t);

template <typename C>
int assemble_stencil(const tg_grid* gp, const uint8_t* mask, const double* mats, C* coef, void* stream) {
  if (int rc = check_grid(gp)) return rc;
  TG_REQUIRE(mask && mats && coef, "assemble_stencil: NULL argument");
  const Grid g = to_grid(gp);
  const long long nn = g.nodes();
  hipLaunchKernelGGL((k_assemble_stencil<C>), dim3((unsigned)((nn + SNT - 1) / SNT), 3), dim3(SNT), 0,
                     (hipStream_t)stream, g, mask, mats, coef);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" int tg_assemble_stencil(const tg_grid* gp, const uint8_t* mask, const double* mats, float* coef,
                                   void* stream) {
  return assemble_stencil(gp, mask, mats, coef, stream);
}
extern "C" int tg_assemble_stencil_bf16(const tg_grid* gp, const uint8_t* mask, const double* mats, uint16_t* coef,
                                        void* stream) {
  return assemble_stencil(gp, mask, mats, coef, stream);
}
