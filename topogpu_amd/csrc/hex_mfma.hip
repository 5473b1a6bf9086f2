// Level-0 hex8 operator with the element GEMM on BF16 MFMA (gfx950).
//
// A workgroup of NW wavefronts owns a tile of MTX = NW-1 (x) by MTY (y) node
// lines and MTZ = 15 z-nodes [k0, k0+15). It computes every element touching
// those nodes: NW x (MTY+1) element columns, each a z-pencil of 16 elements
// [k0-1, k0+15). Wavefront w takes the column at x offset w and walks its
// MTY+1 pencils along y. For one pencil the 16 element products
//     F(32 x 16) = KE'(32 x 32) * U(32 x 16)
// are two v_mfma_f32_16x16x32_bf16 (rows/columns in node-pad order, see
// mfma_layout.h), FP32 accumulate, then scaled by E in FP32. Before the BF16
// cast each element's displacement is taken relative to its first node:
// KE annihilates rigid translations, so this is exact in exact arithmetic and
// makes the rounding error proportional to the element's deformation rather
// than to its absolute displacement. The optional split mode adds the
// hi*lo and lo*hi products of a two-term BF16 expansion (3 MFMA per tile).
//
// Reduction without atomics:
//  * z: lane (g, j) ends with the forces of element j on its corner-g node of
//    the lower z face (tile 0) and of the upper face (tile 1). Node
//    z = k0-1+j sums tile 0 of element j and tile 1 of element j-1 -- a
//    one-lane DPP row shift.
//  * y: consecutive pencils use the two output layouts of mfma_layout.h, so
//    the y-line shared by pencils py and py+1 sits in the same lane half of
//    both: one lane-wise add completes that line's share from this column.
//  * x: the two element columns meeting at a node line (wavefronts w and
//    w+1) each leave their share in an LDS slot; the output stage adds the
//    two in fixed order.
// Halo elements are recomputed by the neighbouring tile instead of exchanged.
//
// Addressing: each workgroup addresses its x-slab of the vectors through
// buffer descriptors based at the slab start, so every offset is 32-bit.
// Nodes and elements outside the grid get an offset past the descriptor's
// range: the hardware returns zero for such loads and drops such stores,
// which replaces all bounds branches.
#include <type_traits>

#include "common.h"
#include "epilogue.h"
#include "mfma_layout.h"
#include "reduce.h"

namespace tg {

extern __constant__ uint16_t c_ke_bf16[2 * 2 * 2 * 64 * 8];

namespace {

constexpr int NW = 8;                                       // wavefronts = element columns in x
constexpr int NTHR = 64 * NW;
constexpr int MTX = NW - 1, MTY = 8, MTZ = 15;              // owned node tile
constexpr int PY = MTY + 1;                                 // pencils per wavefront
constexpr int UBX = MTX + 2, UBY = MTY + 2, UBZ = MTZ + 2;  // node brick
constexpr int SZ = 16;                                      // z records per slot side (15 used)
constexpr uint32_t kOOB = 0x80000000u;                      // offset past every descriptor
constexpr uint32_t kMaxRange = 0x7fffffffu;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, unsigned long long bytes) {
  const int n = (int)(bytes < kMaxRange ? bytes : kMaxRange);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, n, 0x00020000);
}

__device__ __forceinline__ float row_shift_up(float x) {
  // lane l <- lane l-1 within each row of 16 lanes (DPP row_shr:1)
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x111, 0xf,
                                                               0xf, true));
}

__device__ __forceinline__ bf16x8 load_a(int layout, int hilo, int tile, int lane) {
  const uint4 w = *reinterpret_cast<const uint4*>(&c_ke_bf16[(((layout * 2 + hilo) * 2 + tile) * 64 + lane) * 8]);
  return __builtin_bit_cast(bf16x8, w);
}

// Buffer accessor for the epilogue: one descriptor per vector, based at the
// workgroup's slab; b = byte offset of the node's first component.
struct BufAcc {
  rsrc_t rhs, dinv, d, x, r, y;
  uint32_t b;
  template <int V>
  __device__ __forceinline__ rsrc_t rs() const {
    if constexpr (V == EV_RHS) return rhs;
    else if constexpr (V == EV_DINV) return dinv;
    else if constexpr (V == EV_D) return d;
    else if constexpr (V == EV_X) return x;
    else if constexpr (V == EV_R) return r;
    else return y;
  }
  template <int V>
  __device__ __forceinline__ float ld(int c) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs<V>(), b + 4 * c, 0, 0));
  }
  template <int V>
  __device__ __forceinline__ void st(int c, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rs<V>(), b + 4 * c, 0, 0);
  }
};

struct TileShape {
  int ntz, nty;   // tiles along z and y
};

template <bool SPLIT, int MODE>
__global__ void __launch_bounds__(NTHR) k_hex_mfma(Grid g, TileShape ts, const uint8_t* __restrict__ mask,
                                                  const float* __restrict__ E, const float* __restrict__ u,
                                                  Ep32 ep, Workspace ws) {
  using Mode = EpMode<MODE>;
  __shared__ float4 su[UBX * UBY * UBZ];
  // [x line][y line][side][z][3]: side 0 is the share of the element column
  // on the -x side of the line, side 1 the +x side.
  __shared__ float slot[MTX * MTY * 2 * SZ * 3];

  const int NXn = g.nx + 1, NYn = g.ny + 1, NZn = g.nz + 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int tz = t % ts.ntz, ty = (t / ts.ntz) % ts.nty, tx = t / (ts.ntz * ts.nty);
  const int i0 = tx * MTX, j0 = ty * MTY, k0 = tz * MTZ;

  // Slab of x-planes this workgroup touches, as descriptors with 32-bit
  // offsets relative to the slab start.
  const int is = i0 > 0 ? i0 - 1 : 0;
  const long long nslab = (long long)is * NYn * NZn, eslab = (long long)is * g.ny * g.nz;
  const long long nn = (long long)NXn * NYn * NZn, ne = (long long)g.nx * g.ny * g.nz;
  const rsrc_t rs_m = make_rsrc(mask + nslab, (unsigned long long)(nn - nslab));
  const rsrc_t rs_u = make_rsrc(u + 3 * nslab, 12ull * (nn - nslab));
  const rsrc_t rs_e = make_rsrc(E + eslab, 4ull * (ne - eslab));
  const unsigned long long vbytes = 12ull * (nn - nslab);
  BufAcc acc;
  {
    const float* z = u;   // unused vectors: any valid base, never accessed
    acc.rhs = make_rsrc((ep.rhs ? ep.rhs : z) + 3 * nslab, ep.rhs ? vbytes : 0);
    acc.dinv = make_rsrc((ep.dinv ? ep.dinv : z) + 3 * nslab, ep.dinv ? vbytes : 0);
    acc.d = make_rsrc((ep.d ? ep.d : z) + 3 * nslab, ep.d ? vbytes : 0);
    acc.x = make_rsrc((ep.x_out ? ep.x_out : z) + 3 * nslab, ep.x_out ? vbytes : 0);
    acc.r = make_rsrc((ep.r ? ep.r : z) + 3 * nslab, ep.r ? vbytes : 0);
    acc.y = make_rsrc((ep.y ? ep.y : z) + 3 * nslab, ep.y ? vbytes : 0);
  }

  // ---- node brick: every load issued before any is used
  constexpr int NB = UBX * UBY * UBZ, NIT = (NB + NTHR - 1) / NTHR;
  uint32_t bw[NIT][3];
  uint32_t bm[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = threadIdx.x + it * NTHR;
    const int bz = idx % UBZ, line = idx / UBZ, by = line % UBY, bx = line / UBY;
    const int i = i0 - 1 + bx, j = j0 - 1 + by, k = k0 - 1 + bz;
    const bool in = idx < NB && (unsigned)i < (unsigned)NXn && (unsigned)j < (unsigned)NYn &&
                    (unsigned)k < (unsigned)NZn;
    const uint32_t rel = (uint32_t)(((i - is) * NYn + j) * NZn + k);
    const auto w3 = __builtin_amdgcn_raw_buffer_load_b96(rs_u, in ? 12u * rel : kOOB, 0, 0);
    bm[it] = __builtin_amdgcn_raw_buffer_load_b8(rs_m, in ? rel : kOOB, 0, 0);
    bw[it][0] = w3[0];
    bw[it][1] = w3[1];
    bw[it][2] = w3[2];
  }

  // ---- owned output nodes: offsets and epilogue inputs, fetched now so
  // their latency hides behind the pencil work
  constexpr int NOUT = MTX * MTY * MTZ, OIT = (NOUT + NTHR - 1) / NTHR;
  uint32_t ob[OIT];
  uint32_t om[OIT];
  typename Mode::In oin[OIT];
#pragma unroll
  for (int it = 0; it < OIT; ++it) {
    const int idx = threadIdx.x + it * NTHR;
    const int z = idx % MTZ, line = idx / MTZ, ti = line / MTY, tj = line % MTY;
    const int i = i0 + ti, j = j0 + tj, k = k0 + z;
    const bool own = idx < NOUT && i < NXn && j < NYn && k < NZn;
    const uint32_t rel = (uint32_t)(((i - is) * NYn + j) * NZn + k);
    ob[it] = own ? 12u * rel : kOOB;
    om[it] = __builtin_amdgcn_raw_buffer_load_b8(rs_m, own ? rel : kOOB, 0, 0);
    acc.b = ob[it];
    Mode::load(ep, acc, oin[it]);
  }

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int je = lane & 15;     // element within the pencil (B and D column)
  const int grp = lane >> 4;    // B: input corners 2grp, 2grp+1 ; D: output group

  // Element moduli of this wavefront's pencils.
  float ev[PY];
  {
    const int ci = i0 - 1 + w, ck = k0 - 1 + je;
    const bool okx = (unsigned)ci < (unsigned)g.nx && (unsigned)ck < (unsigned)g.nz;
    const uint32_t ebase = (uint32_t)(((ci - is) * g.ny + (j0 - 1)) * g.nz + ck);
#pragma unroll
    for (int py = 0; py < PY; ++py) {
      const bool ok = okx && (unsigned)(j0 - 1 + py) < (unsigned)g.ny;
      ev[py] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                             rs_e, ok ? 4u * (ebase + py * g.nz) : kOOB, 0, 0));
    }
  }
  bf16x8 ah[2][2], al[2][2];
#pragma unroll
  for (int ly = 0; ly < 2; ++ly)
#pragma unroll
    for (int tl = 0; tl < 2; ++tl) {
      ah[ly][tl] = load_a(ly, 0, tl, lane);
      al[ly][tl] = SPLIT ? load_a(ly, 1, tl, lane) : ah[ly][tl];
    }

#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = threadIdx.x + it * NTHR;
    if (idx < NB)
      su[idx] = make_float4((bm[it] & 1u) ? __builtin_bit_cast(float, bw[it][0]) : 0.f,
                            (bm[it] & 2u) ? __builtin_bit_cast(float, bw[it][1]) : 0.f,
                            (bm[it] & 4u) ? __builtin_bit_cast(float, bw[it][2]) : 0.f, 0.f);
  }
  __syncthreads();

  // ---- pencils of column w
  const int cA = 2 * grp, cB = 2 * grp + 1;
  const int offA = (corner_x(cA) * UBY + corner_y(cA)) * UBZ + corner_z(cA);
  const int offB = (corner_x(cB) * UBY + corner_y(cB)) * UBZ + corner_z(cB);
  const int ubase = w * UBY * UBZ + je;
  // slot writes: lane (g, je) of the pair (py-1, py) writes x-line w-1+cx
  const int cx = mfma_out_cx(grp), ti = w - 1 + cx;
  const bool wr = je >= 1 && (unsigned)ti < (unsigned)MTX;
  const bool wr_lo = wr && lane < 32, wr_hi = wr && lane >= 32;
  const int sbase = ((ti * MTY) * 2 + (1 - cx)) * SZ * 3 + (je - 1) * 3;   // used only when wr

  float rp[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int py = 0; py < PY; ++py) {
    const int ly = py & 1;
    const float4 u0 = su[ubase + py * UBZ];
    const float4 ua = su[ubase + py * UBZ + offA];
    const float4 ub = su[ubase + py * UBZ + offB];
    const floatx8 f = {ua.x - u0.x, ua.y - u0.y, ua.z - u0.z, 0.f, ub.x - u0.x, ub.y - u0.y, ub.z - u0.z, 0.f};
    const bf16x8 bh = __builtin_convertvector(f, bf16x8);
    floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ly][0], bh, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ly][1], bh, d1, 0, 0, 0);
    if constexpr (SPLIT) {
      const bf16x8 bl = __builtin_convertvector(f - __builtin_convertvector(bh, floatx8), bf16x8);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ly][0], bl, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ly][1], bl, d1, 0, 0, 0);
      d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ly][0], bh, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ly][1], bh, d1, 0, 0, 0);
    }
    const float e = ev[py];
    float rc[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) rc[v] = e * d0[v] + row_shift_up(e * d1[v]);
    if (py > 0) {
      // y-line py-1: lower lane half when py-1 is even, upper when odd
      const bool on = ((py - 1) & 1) ? wr_hi : wr_lo;
      if (on) {
        float* s = &slot[sbase + (py - 1) * 2 * SZ * 3];
#pragma unroll
        for (int v = 0; v < 3; ++v) s[v] = rp[v] + rc[v];
      }
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) rp[v] = rc[v];
  }
  __syncthreads();

  // ---- output: sum the two column shares, apply the epilogue
  double part = 0.0;
#pragma unroll
  for (int it = 0; it < OIT; ++it) {
    int idx = threadIdx.x + it * NTHR;
    idx = idx < NOUT ? idx : NOUT - 1;   // surplus threads read a valid slot; their stores are dropped
    const int z = idx % MTZ, line = idx / MTZ, tli = line / MTY, tlj = line % MTY;
    const float* s0 = &slot[(line * 2) * SZ * 3 + z * 3];
    float au[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) au[v] = s0[v] + s0[SZ * 3 + v];
    const float4 uo = su[((tli + 1) * UBY + tlj + 1) * UBZ + z + 1];
    const float own[3] = {uo.x, uo.y, uo.z};
    acc.b = ob[it];
    part += Mode::store(ep, acc, om[it], au, own, oin[it]);
  }
  if constexpr (Mode::kReduces) {
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
  const TileShape ts{(g.NZ() + MTZ - 1) / MTZ, (g.NY() + MTY - 1) / MTY};
  const long long nb = (long long)((g.NX() + MTX - 1) / MTX) * ts.nty * ts.ntz;
  TG_REQUIRE(nb <= kMaxPartials, "grid too large for the reduction workspace");
  // every slab offset must stay below 2^31 bytes
  TG_REQUIRE((long long)(MTX + 2) * g.NY() * g.NZ() * 12 < (long long)kMaxRange,
             "grid cross-section %d x %d nodes too large for 32-bit slab offsets", g.NY(), g.NZ());
  return with_mode(ep, [&](const auto& em) {
    constexpr int M = std::decay_t<decltype(em)>::kMode;
    if (split)
      hipLaunchKernelGGL((k_hex_mfma<true, M>), dim3((unsigned)nb), dim3(NTHR), 0, s, g, ts, mask, E, u, ep,
                         workspace());
    else
      hipLaunchKernelGGL((k_hex_mfma<false, M>), dim3((unsigned)nb), dim3(NTHR), 0, s, g, ts, mask, E, u, ep,
                         workspace());
    TG_LAUNCH_CHECK();
    return EpMode<M>::kReduces ? finalize_sums(1, (int)nb, dot, nullptr, s) : TG_OK;
  });
}

}  // namespace tg
