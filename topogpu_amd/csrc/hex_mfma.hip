// Level-0 hex8 operator with the element GEMM on BF16 MFMA (gfx950).
//
// A workgroup owns a tile of 4 (x) x 8 (y) node lines and 15 z-nodes
// [k0, k0+15). It computes every element touching those nodes: the 5 x 9
// element columns around the lines, each as a z-pencil of 16 elements
// [k0-1, k0+15) held by one wavefront. For a pencil the 16 element products
//     F(32 x 16) = KE'(32 x 32) * U(32 x 16)
// are two v_mfma_f32_16x16x32_bf16 (rows/columns in node-pad order, see
// mfma_layout.h), FP32 accumulate, then scaled by E in FP32. Before the BF16
// cast each element's displacement is taken relative to its first node:
// KE annihilates rigid translations, so this is exact in exact arithmetic and
// makes the rounding error proportional to the element's deformation rather
// than to its absolute displacement. The optional split mode adds the
// hi*lo and lo*hi products of a two-term BF16 expansion (3 MFMA per tile).
//
// Reduction without atomics: lane (g, j) ends with the forces of element j on
// its corner-g node of the lower z face (tile 0) and of the upper z face
// (tile 1). Node z = k0-1+j sums tile 0 of element j and tile 1 of element
// j-1 -- a one-lane DPP row shift. The four element columns meeting at a
// node line are then summed from LDS slots in fixed corner order. Halo
// elements are recomputed by the neighbouring tile instead of exchanged:
// MFMA throughput is far above what the memory traffic of this kernel needs.
#include "common.h"
#include "epilogue.h"
#include "mfma_layout.h"
#include "reduce.h"

namespace tg {

extern __constant__ uint16_t c_ke_bf16[2 * 2 * 64 * 8];

namespace {

constexpr int MTX = 4, MTY = 8, MTZ = 15;                 // owned node tile
constexpr int PX = MTX + 1, PY = MTY + 1, NPEN = PX * PY;   // element columns
constexpr int UBX = MTX + 2, UBY = MTY + 2, UBZ = MTZ + 2;  // node brick
constexpr int NTHR = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float row_shift_up(float x) {
  // lane l <- lane l-1 within each row of 16 lanes (DPP row_shr:1)
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x111, 0xf,
                                                               0xf, true));
}

__device__ __forceinline__ bf16x8 load_a(int hilo, int tile, int lane) {
  const uint4 w = *reinterpret_cast<const uint4*>(&c_ke_bf16[((hilo * 2 + tile) * 64 + lane) * 8]);
  return __builtin_bit_cast(bf16x8, w);
}

template <bool SPLIT, class Ep>
__global__ void __launch_bounds__(NTHR) k_hex_mfma(Grid g, const uint8_t* __restrict__ mask,
                                                  const float* __restrict__ E, const float* __restrict__ u,
                                                  Ep ep, double* dot, Workspace ws) {
  __shared__ float4 su[UBX * UBY * UBZ];
  // [line][corner][z][3]: consecutive lanes touch consecutive 3-float
  // records (odd stride in 4-byte banks), so writes and reads are
  // free of bank conflicts.
  __shared__ float slot[MTX * MTY][4][MTZ][3];

  const int ntz = (g.NZ() + MTZ - 1) / MTZ, nty = (g.NY() + MTY - 1) / MTY;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int k0 = (t % ntz) * MTZ;
  const int j0 = ((t / ntz) % nty) * MTY;
  const int i0 = (t / (ntz * nty)) * MTX;

  // Node brick: every load is issued before any is used (no load depends on
  // another), so a workgroup pays one memory round trip, not one per node.
  constexpr int NB = UBX * UBY * UBZ, NIT = (NB + NTHR - 1) / NTHR;
  float vx[NIT], vy[NIT], vz[NIT];
  unsigned vm[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = threadIdx.x + it * NTHR;
    const int bz = idx % UBZ, by = (idx / UBZ) % UBY, bx = idx / (UBZ * UBY);
    const int i = i0 - 1 + bx, j = j0 - 1 + by, k = k0 - 1 + bz;
    const bool in = idx < NB && i >= 0 && i <= g.nx && j >= 0 && j <= g.ny && k >= 0 && k <= g.nz;
    const long long n = in ? g.node(i, j, k) : 0;
    vm[it] = in ? (unsigned)mask[n] : 0u;
    vx[it] = in ? u[3 * n] : 0.f;
    vy[it] = in ? u[3 * n + 1] : 0.f;
    vz[it] = in ? u[3 * n + 2] : 0.f;
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = threadIdx.x + it * NTHR;
    if (idx < NB)
      su[idx] = make_float4((vm[it] & 1u) ? vx[it] : 0.f, (vm[it] & 2u) ? vy[it] : 0.f,
                            (vm[it] & 4u) ? vz[it] : 0.f, 0.f);
  }

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int je = mfma_b_col(lane);   // element within the pencil
  const int grp = lane >> 4;         // B: nodes 2grp, 2grp+1 ; D: corner grp
  const int c0 = 2 * grp, c1 = 2 * grp + 1;
  // Element moduli of this wave's pencils, fetched before the barrier so no
  // global-memory latency remains inside the pencil loop.
  constexpr int PPW = (NPEN + NTHR / 64 - 1) / (NTHR / 64);   // pencils per wave
  float ev[PPW];
#pragma unroll
  for (int q = 0; q < PPW; ++q) {
    const int p = wave + q * (NTHR / 64);
    const int px = p / PY, py = p % PY;
    const int ci = i0 - 1 + px, cj = j0 - 1 + py, ck = k0 - 1 + je;
    const bool ein = p < NPEN && ci >= 0 && ci < g.nx && cj >= 0 && cj < g.ny && ck >= 0 && ck < g.nz;
    ev[q] = ein ? E[g.elem(ci, cj, ck)] : 0.f;
  }
  const bf16x8 a_hi0 = load_a(0, 0, lane), a_hi1 = load_a(0, 1, lane);
  bf16x8 a_lo0 = a_hi0, a_lo1 = a_hi1;
  if (SPLIT) {
    a_lo0 = load_a(1, 0, lane);
    a_lo1 = load_a(1, 1, lane);
  }
  __syncthreads();

#pragma unroll
  for (int q = 0; q < PPW; ++q) {
    const int p = wave + q * (NTHR / 64);
    if (p >= NPEN) break;
    const int px = p / PY, py = p % PY;
    const float e = ev[q];
    const float4 u0 = su[(px * UBY + py) * UBZ + je];
    const float4 ua = su[((px + corner_x(c0)) * UBY + py + corner_y(c0)) * UBZ + je + corner_z(c0)];
    const float4 ub = su[((px + corner_x(c1)) * UBY + py + corner_y(c1)) * UBZ + je + corner_z(c1)];
    const floatx8 f = {ua.x - u0.x, ua.y - u0.y, ua.z - u0.z, 0.f,
                       ub.x - u0.x, ub.y - u0.y, ub.z - u0.z, 0.f};
    const bf16x8 b_hi = __builtin_convertvector(f, bf16x8);
    floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi0, b_hi, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi1, b_hi, d1, 0, 0, 0);
    if (SPLIT) {
      const bf16x8 b_lo = __builtin_convertvector(f - __builtin_convertvector(b_hi, floatx8), bf16x8);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi0, b_lo, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_hi1, b_lo, d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo0, b_hi, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_lo1, b_hi, d1, 0, 0, 0);
    }
    float s[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) s[v] = e * d0[v] + row_shift_up(e * d1[v]);
    const int ti = px - 1 + corner_x(grp), tj = py - 1 + corner_y(grp);
    if (je >= 1 && ti >= 0 && ti < MTX && tj >= 0 && tj < MTY) {
#pragma unroll
      for (int v = 0; v < 3; ++v) slot[ti * MTY + tj][grp][je - 1][v] = s[v];
    }
  }
  __syncthreads();

  // Output: all epilogue loads of this thread's nodes first, then the stores.
  constexpr int NOUT = MTX * MTY * MTZ, OIT = (NOUT + NTHR - 1) / NTHR;
  typename Ep::In in[OIT];
  long long nn[OIT];
  unsigned mo[OIT];
  bool ok[OIT];
#pragma unroll
  for (int it = 0; it < OIT; ++it) {
    const int idx = threadIdx.x + it * NTHR;
    const int z = idx % MTZ, line = idx / MTZ, ti = line / MTY, tj = line % MTY;
    const int i = i0 + ti, j = j0 + tj, k = k0 + z;
    ok[it] = idx < NOUT && i <= g.nx && j <= g.ny && k <= g.nz;
    nn[it] = ok[it] ? g.node(i, j, k) : 0;
    mo[it] = ok[it] ? (unsigned)mask[nn[it]] : 0u;
    if (ok[it]) ep.load(nn[it], in[it]);
  }
  double part = 0.0;
#pragma unroll
  for (int it = 0; it < OIT; ++it) {
    if (!ok[it]) continue;
    const int idx = threadIdx.x + it * NTHR;
    const int z = idx % MTZ, line = idx / MTZ, ti = line / MTY, tj = line % MTY;
    float au[3];
#pragma unroll
    for (int v = 0; v < 3; ++v)
      au[v] = ((slot[line][0][z][v] + slot[line][1][z][v]) + slot[line][2][z][v]) + slot[line][3][z][v];
    const float4 uo = su[((ti + 1) * UBY + tj + 1) * UBZ + z + 1];
    const float own[3] = {uo.x, uo.y, uo.z};
    part += ep.store(nn[it], mo[it], au, own, in[it]);
  }
  if (ep.reduces()) {
    double v[1] = {part};
    block_partials<1>(v, ws);
  }
}

// One wavefront, one MFMA on caller-packed operands: d[lane] = D fragment.
__global__ void __launch_bounds__(64) k_mfma_probe(const uint4* __restrict__ a, const uint4* __restrict__ b,
                                                  float4* __restrict__ d) {
  const int lane = threadIdx.x;
  const floatx4 z = {0.f, 0.f, 0.f, 0.f};
  const floatx4 r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[lane]),
                                                            __builtin_bit_cast(bf16x8, b[lane]), z, 0, 0, 0);
  d[lane] = make_float4(r[0], r[1], r[2], r[3]);
}

}  // namespace

int launch_mfma_probe(const uint16_t* a, const uint16_t* b, float* d, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, reinterpret_cast<const uint4*>(a),
                     reinterpret_cast<const uint4*>(b), reinterpret_cast<float4*>(d));
  TG_LAUNCH_CHECK();
  return TG_OK;
}

int launch_hex_mfma(const Grid& g, const uint8_t* mask, const float* E, const float* u, const Ep32& ep,
                    double* dot, bool split, hipStream_t s) {
  const int nb = ((g.NX() + MTX - 1) / MTX) * ((g.NY() + MTY - 1) / MTY) * ((g.NZ() + MTZ - 1) / MTZ);
  TG_REQUIRE(nb <= kMaxPartials, "grid too large for the reduction workspace");
  if (split)
    hipLaunchKernelGGL((k_hex_mfma<true, Ep32>), dim3(nb), dim3(NTHR), 0, s, g, mask, E, u, ep, dot,
                       workspace());
  else
    hipLaunchKernelGGL((k_hex_mfma<false, Ep32>), dim3(nb), dim3(NTHR), 0, s, g, mask, E, u, ep, dot,
                       workspace());
  TG_LAUNCH_CHECK();
  return ep.reduces() ? finalize_sums(1, nb, dot, nullptr, s) : TG_OK;
}

}  // namespace tg
