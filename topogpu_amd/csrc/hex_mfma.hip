// Hex8 element operators with the element GEMM on BF16 MFMA (gfx950):
//   k_hex_mfma : level 0, element matrix E_e * KE;
//   k_q8_mfma  : Galerkin level 1 of a fine grid with even dims, element
//                matrix sum_s E_child(s) * Q_s over the eight fine children.
//
// A workgroup of NW wavefronts owns a tile of MTX = NW-1 (x) by MTY (y) node
// lines and MTZ = 15 z-nodes [k0, k0+15). It computes every element touching
// those nodes: NW x (MTY+1) element columns, each a z-pencil of 16 elements
// [k0-1, k0+15). Wavefront w takes the column at x offset w and walks its
// MTY+1 pencils along y. For one pencil the 16 element products
//     F(32 x 16) = A(32 x 32) * U(32 x 16)
// are two v_mfma_f32_16x16x32_bf16 per matrix A (rows/columns in node-pad
// order, see mfma_layout.h), FP32 accumulate, scaled by the element weight in
// FP32. Before the BF16 cast each element's displacement is taken relative to
// its first node: KE and every Q_s annihilate rigid translations, so this is
// exact in exact arithmetic and makes the rounding error proportional to the
// element's deformation rather than to its absolute displacement. The split
// mode adds the hi*lo and lo*hi products of a two-term BF16 expansion
// (3 MFMA per tile), which brings the product to FP32-level accuracy.
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
extern __device__ uint16_t c_q8_bf16[8 * 2 * 2 * 2 * 64 * 8];

namespace {

constexpr uint32_t kMaxRange = 0x7fffffffu;


// Offset o when valid, else an offset past every descriptor: valid offsets
// stay below 2^31 (descriptor ranges are at most 2^31 - 1 bytes), so setting
// bit 31 is enough. Written without a select so that the compiler keeps the
// loads branch-free; a branch around a load makes it wait for every load in
// flight before reusing the branch's registers.
__device__ __forceinline__ uint32_t oob_unless(uint32_t o, bool valid) { return o | ((uint32_t)!valid << 31); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;

template <int NW_>
struct TileT {
  static constexpr int NW = NW_;                  // wavefronts = element columns in x
  static constexpr int NTHR = 64 * NW;
  static constexpr int MTX = NW - 1, MTY = 8, MTZ = 15;   // owned node tile
  static constexpr int PY = MTY + 1;              // pencils per wavefront
  static constexpr int UBX = MTX + 2, UBY = MTY + 2, UBZ = MTZ + 2;   // node brick
  static constexpr int SZ = 16;                   // z records per slot side (15 used)
  static constexpr int NB = UBX * UBY * UBZ;
  // Brick loads: thread t < LPR * UBZ takes z = t % UBZ of brick line t / UBZ
  // and, in round r, the same (y, z) BXR x-planes further on. LPR is a
  // multiple of UBY, so a round advances x only and the LDS index of a
  // thread's node is t + r * LPR * UBZ: no per-round index arithmetic.
  static constexpr int LPR = ((NTHR / UBZ) / UBY) * UBY, BXR = LPR / UBY;
  static constexpr int NIT = (UBX + BXR - 1) / BXR;   // load rounds
  static constexpr int BACT = LPR * UBZ;               // threads with brick work
  static constexpr int NBP = NIT * BACT + (NTHR - BACT);   // + a dump slot per idle thread
  // Output nodes, the same way: thread t < OLPR * MTZ takes z = t % MTZ of
  // owned line t / MTZ and, in round r, the line OXR x-lines further on.
  static constexpr int OLPR = ((NTHR / MTZ) / MTY) * MTY, OXR = OLPR / MTY;
  static constexpr int OIT = (MTX + OXR - 1) / OXR;    // output rounds
  static constexpr int OACT = OLPR * MTZ;
  static constexpr int SLOT = MTX * MTY * 2 * SZ * 3;
  static_assert(LPR >= UBY && OLPR >= MTY, "tile too large for the workgroup");
};
using Tile0 = TileT<8>;   // level 0
using Tile1 = TileT<4>;   // level 1 (8x fewer elements: smaller tiles keep the CUs busy)

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, unsigned long long bytes) {
  const int n = (int)(bytes < kMaxRange ? bytes : kMaxRange);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, n, 0x00020000);
}

__device__ __forceinline__ float row_shift_up(float x) {
  // lane l <- lane l-1 within each row of 16 lanes (DPP row_shr:1)
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x111, 0xf,
                                                               0xf, true));
}

__device__ __forceinline__ bf16x8 load_ke(int layout, int hilo, int tile, int lane) {
  const uint4 w = *reinterpret_cast<const uint4*>(&c_ke_bf16[(((layout * 2 + hilo) * 2 + tile) * 64 + lane) * 8]);
  return __builtin_bit_cast(bf16x8, w);
}

__device__ __forceinline__ bf16x8 load_q8(int s, int layout, int hilo, int tile, int lane) {
  const uint4 w =
      *reinterpret_cast<const uint4*>(&c_q8_bf16[((((s * 2 + layout) * 2 + hilo) * 2 + tile) * 64 + lane) * 8]);
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
  int ntz, nty;     // tiles along z and y
  uint32_t mz, my;  // ceil(2^32 / ntz), ceil(2^32 / nty) (0 for a divisor of 1)
};

// q = t / d for the tile counts: umulhi(t, ceil(2^32 / d)) is exact while
// t * d < 2^32, which the launcher checks.
__device__ __forceinline__ uint32_t div_tiles(uint32_t t, uint32_t m) { return m ? __umulhi(t, m) : t; }

// Workgroup geometry and node-vector descriptors.
struct Block {
  int i0, j0, k0;    // owned tile origin (nodes)
  int is;            // first x-plane of the slab
  int NXn, NYn, NZn; // node dims
  rsrc_t rs_m, rs_u;
  BufAcc acc;
};

// Geometry and descriptors of tile t; u and ep are those of one right-hand
// side.
template <class T>
__device__ __forceinline__ Block tile_block(const Grid& g, TileShape ts, int t, const uint8_t* mask, const float* u,
                                            const Ep32& ep) {
  Block b;
  b.NXn = g.nx + 1;
  b.NYn = g.ny + 1;
  b.NZn = g.nz + 1;
  const uint32_t q1 = div_tiles((uint32_t)t, ts.mz), q2 = div_tiles(q1, ts.my);
  const int tz = t - (int)q1 * ts.ntz, ty = (int)q1 - (int)q2 * ts.nty, tx = (int)q2;
  b.i0 = tx * T::MTX;
  b.j0 = ty * T::MTY;
  b.k0 = tz * T::MTZ;
  b.is = b.i0 > 0 ? b.i0 - 1 : 0;
  const long long nslab = (long long)b.is * b.NYn * b.NZn;
  const long long nn = (long long)b.NXn * b.NYn * b.NZn;
  b.rs_m = make_rsrc(mask + nslab, (unsigned long long)(nn - nslab));
  b.rs_u = make_rsrc(u + 3 * nslab, 12ull * (nn - nslab));
  const unsigned long long vbytes = 12ull * (nn - nslab);
  const float* z = u;   // unused vectors: any valid base, never accessed
  b.acc.rhs = make_rsrc((ep.rhs ? ep.rhs : z) + 3 * nslab, ep.rhs ? vbytes : 0);
  b.acc.dinv = make_rsrc((ep.dinv ? ep.dinv : z) + 3 * nslab, ep.dinv ? vbytes : 0);
  b.acc.d = make_rsrc((ep.d ? ep.d : z) + 3 * nslab, ep.d ? vbytes : 0);
  b.acc.x = make_rsrc((ep.x_out ? ep.x_out : z) + 3 * nslab, ep.x_out ? vbytes : 0);
  b.acc.r = make_rsrc((ep.r ? ep.r : z) + 3 * nslab, ep.r ? vbytes : 0);
  b.acc.y = make_rsrc((ep.y ? ep.y : z) + 3 * nslab, ep.y ? vbytes : 0);
  return b;
}

// One tile per workgroup. Right-hand side blockIdx.y of a batch: every
// per-RHS vector is offset by blockIdx.y * 3 * nodes (dinv is the
// operator's and shared).
template <class T>
__device__ __forceinline__ Block block_setup(const Grid& g, TileShape ts, const uint8_t* mask, const float* u,
                                             const Ep32& ep_all) {
  const long long roff = (long long)blockIdx.y * 3 * g.nodes();
  return tile_block<T>(g, ts, xcd_remap(blockIdx.x, gridDim.x), mask, u + roff, ep_slice(ep_all, roff));
}

// Node brick loads: every load issued before any is used. PRE also loads
// dinv at the same nodes (bd) for the prescaled input of the fused first
// smoothing step.
template <class T, bool PRE>
__device__ __forceinline__ void brick_load(const Block& b, uint32_t (&bw)[T::NIT][3], uint32_t (&bm)[T::NIT],
                                           uint32_t (&bd)[T::NIT][3]) {
  const int t = threadIdx.x, bz = t % T::UBZ, l0 = t / T::UBZ, bx0 = l0 / T::UBY, by = l0 % T::UBY;
  const int i0 = b.i0 - 1 + bx0, j = b.j0 - 1 + by, k = b.k0 - 1 + bz;
  const bool ok = (t < T::BACT) & ((unsigned)j < (unsigned)b.NYn) & ((unsigned)k < (unsigned)b.NZn);
  const uint32_t rel0 = (uint32_t)(((i0 - b.is) * b.NYn + j) * b.NZn + k);
  const uint32_t plane = (uint32_t)(T::BXR * b.NYn * b.NZn);
#pragma unroll
  for (int it = 0; it < T::NIT; ++it) {
    const bool in = ok & ((unsigned)(i0 + it * T::BXR) < (unsigned)b.NXn);
    const uint32_t rel = rel0 + it * plane;
    const auto w3 = __builtin_amdgcn_raw_buffer_load_b96(b.rs_u, oob_unless(12u * rel, in), 0, 0);
    bm[it] = __builtin_amdgcn_raw_buffer_load_b8(b.rs_m, oob_unless(rel, in), 0, 0);
    bw[it][0] = w3[0];
    bw[it][1] = w3[1];
    bw[it][2] = w3[2];
    if constexpr (PRE) {
      const auto d3 = __builtin_amdgcn_raw_buffer_load_b96(b.acc.dinv, oob_unless(12u * rel, in), 0, 0);
      bd[it][0] = d3[0];
      bd[it][1] = d3[1];
      bd[it][2] = d3[2];
    }
  }
}

// Masked node values into LDS (fixed DOFs and nodes outside the grid read
// as zero). Every thread stores unconditionally -- idle threads into a dump
// slot of their own past the brick -- because a conditional store would let
// the compiler sink the matching loads into the branch and wait there.
// PRE: the stored value is (pa * dinv) * u, rounded as the separate
// Chebyshev start rounds it.
template <class T, bool PRE>
__device__ __forceinline__ void brick_store(float4* su, const uint32_t (&bw)[T::NIT][3], const uint32_t (&bm)[T::NIT],
                                            const uint32_t (&bd)[T::NIT][3], float pa) {
  const int t = threadIdx.x;
  const bool act = t < T::BACT;
#pragma unroll
  for (int it = 0; it < T::NIT; ++it) {
    const int idx = act ? t + it * T::BACT : T::NIT * T::BACT + (t - T::BACT);
    float v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[c] = __builtin_bit_cast(float, bw[it][c]);
      if constexpr (PRE) v[c] = (pa * __builtin_bit_cast(float, bd[it][c])) * v[c];
    }
    su[idx] = make_float4((bm[it] & 1u) ? v[0] : 0.f, (bm[it] & 2u) ? v[1] : 0.f, (bm[it] & 4u) ? v[2] : 0.f, 0.f);
  }
}

// Owned output nodes: offsets, masks and the epilogue inputs, fetched early
// so their latency hides behind the pencil work.
template <class T, int MODE>
__device__ __forceinline__ void out_prefetch(Block& b, const Ep32& ep, uint32_t (&ob)[T::OIT], uint32_t (&om)[T::OIT],
                                             typename EpMode<MODE>::In (&oin)[T::OIT]) {
  const int t = threadIdx.x, z = t % T::MTZ, l0 = t / T::MTZ, ti0 = l0 / T::MTY, tj = l0 % T::MTY;
  const int i0 = b.i0 + ti0, j = b.j0 + tj, k = b.k0 + z;
  const bool ok = (t < T::OACT) & (j < b.NYn) & (k < b.NZn);
  const uint32_t rel0 = (uint32_t)(((i0 - b.is) * b.NYn + j) * b.NZn + k);
  const uint32_t plane = (uint32_t)(T::OXR * b.NYn * b.NZn);
#pragma unroll
  for (int it = 0; it < T::OIT; ++it) {
    const bool own = ok & (ti0 + it * T::OXR < T::MTX) & (i0 + it * T::OXR < b.NXn);
    const uint32_t rel = rel0 + it * plane;
    ob[it] = oob_unless(12u * rel, own);
    om[it] = __builtin_amdgcn_raw_buffer_load_b8(b.rs_m, oob_unless(rel, own), 0, 0);
    b.acc.b = ob[it];
    EpMode<MODE>::load(ep, b.acc, oin[it]);
  }
}

// Output stage: sum the two column shares of every owned node, apply the
// epilogue; returns this thread's share of the grid reduction.
template <class T, int MODE>
__device__ __forceinline__ double out_store(Block& b, const Ep32& ep, const float* slot, const float4* su,
                                            const uint32_t (&ob)[T::OIT], const uint32_t (&om)[T::OIT],
                                            const typename EpMode<MODE>::In (&oin)[T::OIT]) {
  // idle threads and rounds past the tile read a valid slot; their stores are dropped
  const int t = threadIdx.x < T::OACT ? threadIdx.x : 0;
  const int z = t % T::MTZ, l0 = t / T::MTZ, ti0 = l0 / T::MTY, tj = l0 % T::MTY;
  double part = 0.0;
#pragma unroll
  for (int it = 0; it < T::OIT; ++it) {
    const int tli = ti0 + it * T::OXR < T::MTX ? ti0 + it * T::OXR : T::MTX - 1;
    const int line = tli * T::MTY + tj;
    const float* s0 = &slot[(line * 2) * T::SZ * 3 + z * 3];
    float au[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) au[v] = s0[v] + s0[T::SZ * 3 + v];
    const float4 uo = su[((tli + 1) * T::UBY + tj + 1) * T::UBZ + z + 1];
    const float own[3] = {uo.x, uo.y, uo.z};
    b.acc.b = ob[it];
    part += EpMode<MODE>::store(ep, b.acc, om[it], au, own, oin[it]);
  }
  return part;
}

// Per-lane pencil addressing of wavefront w.
template <class T>
struct PencilLane {
  int lane, je, grp;
  int offA, offB, ubase;   // brick offsets of the lane's corners / element 0
  bool wr_lo, wr_hi;       // slot write of a pair (py-1, py) with py-1 even / odd
  int sbase;               // slot index of the lane's record (when it writes)

  __device__ __forceinline__ PencilLane(int w) {
    lane = threadIdx.x & 63;
    je = lane & 15;     // element within the pencil (B and D column)
    grp = lane >> 4;    // B: input corners 2grp, 2grp+1 ; D: output group
    const int cA = 2 * grp, cB = 2 * grp + 1;
    offA = (corner_x(cA) * T::UBY + corner_y(cA)) * T::UBZ + corner_z(cA);
    offB = (corner_x(cB) * T::UBY + corner_y(cB)) * T::UBZ + corner_z(cB);
    ubase = w * T::UBY * T::UBZ + je;
    // lane (g, je) of the pair (py-1, py) writes x-line w-1+cx
    const int cx = mfma_out_cx(grp), ti = w - 1 + cx;
    const bool wr = je >= 1 && (unsigned)ti < (unsigned)T::MTX;
    wr_lo = wr && lane < 32;
    wr_hi = wr && lane >= 32;
    sbase = ((ti * T::MTY) * 2 + (1 - cx)) * T::SZ * 3 + (je - 1) * 3;
  }

  // Relative displacements of pencil py's elements at this lane's two corners.
  __device__ __forceinline__ floatx8 rel_disp(const float4* su, int py) const {
    const float4 u0 = su[ubase + py * T::UBZ];
    const float4 ua = su[ubase + py * T::UBZ + offA];
    const float4 ub = su[ubase + py * T::UBZ + offB];
    const floatx8 f = {ua.x - u0.x, ua.y - u0.y, ua.z - u0.z, 0.f, ub.x - u0.x, ub.y - u0.y, ub.z - u0.z, 0.f};
    return f;
  }

  // Pair (py-1, py) complete: write the y-line py-1 share of this column.
  __device__ __forceinline__ void pair_store(float* slot, int py, const float (&rp)[3], const float (&rc)[3]) const {
    // y-line py-1: lower lane half when py-1 is even, upper when odd
    const bool on = ((py - 1) & 1) ? wr_hi : wr_lo;
    if (on) {
      float* s = &slot[sbase + (py - 1) * 2 * T::SZ * 3];
#pragma unroll
      for (int v = 0; v < 3; ++v) s[v] = rp[v] + rc[v];
    }
  }
};

template <bool SPLIT>
__device__ __forceinline__ void mma2(const bf16x8 (&ah)[2], const bf16x8 (&al)[2], const bf16x8& bh,
                                     const bf16x8& bl, floatx4& d0, floatx4& d1) {
  d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0], bh, d0, 0, 0, 0);
  d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1], bh, d1, 0, 0, 0);
  if constexpr (SPLIT) {
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0], bl, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1], bl, d1, 0, 0, 0);
    d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0], bh, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[1], bh, d1, 0, 0, 0);
  }
}

// Element moduli of wavefront w's pencils (lane element je) in tile b.
template <class T>
__device__ __forceinline__ void e_load(const Grid& g, const Block& b, const float* E, int w, int je,
                                       float (&ev)[T::PY]) {
  const long long eslab = (long long)b.is * g.ny * g.nz, ne = (long long)g.nx * g.ny * g.nz;
  const rsrc_t rs_e = make_rsrc(E + eslab, 4ull * (ne - eslab));
  const int ci = b.i0 - 1 + w, ck = b.k0 - 1 + je;
  const bool okx = ((unsigned)ci < (unsigned)g.nx) & ((unsigned)ck < (unsigned)g.nz);
  const uint32_t ebase = (uint32_t)(((ci - b.is) * g.ny + (b.j0 - 1)) * g.nz + ck);
#pragma unroll
  for (int py = 0; py < T::PY; ++py) {
    const bool ok = okx & ((unsigned)(b.j0 - 1 + py) < (unsigned)g.ny);
    ev[py] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                           rs_e, oob_unless(4u * (ebase + py * g.nz), ok), 0, 0));
  }
}

// KE fragments of both output layouts (hi part; lo part in split mode).
template <bool SPLIT>
__device__ __forceinline__ void ke_frags(int lane, bf16x8 (&ah)[2][2], bf16x8 (&al)[2][2]) {
#pragma unroll
  for (int ly = 0; ly < 2; ++ly)
#pragma unroll
    for (int tl = 0; tl < 2; ++tl) {
      ah[ly][tl] = load_ke(ly, 0, tl, lane);
      al[ly][tl] = SPLIT ? load_ke(ly, 1, tl, lane) : ah[ly][tl];
    }
}

// The wavefront's PY pencils: element products from the brick in LDS, z and
// y reductions in registers, x shares into the LDS slots.
template <bool SPLIT, class T>
__device__ __forceinline__ void pencils0(const float4* su, float* slot, const PencilLane<T>& pl,
                                         const bf16x8 (&ah)[2][2], const bf16x8 (&al)[2][2], const float (&ev)[T::PY]) {
  float rp[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int py = 0; py < T::PY; ++py) {
    const int ly = py & 1;
    const floatx8 f = pl.rel_disp(su, py);
    const bf16x8 bh = __builtin_convertvector(f, bf16x8);
    bf16x8 bl = bh;
    if constexpr (SPLIT) bl = __builtin_convertvector(f - __builtin_convertvector(bh, floatx8), bf16x8);
    floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    mma2<SPLIT>(ah[ly], al[ly], bh, bl, d0, d1);
    const float e = ev[py];
    float rc[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) rc[v] = e * d0[v] + row_shift_up(e * d1[v]);
    if (py > 0) pl.pair_store(slot, py, rp, rc);
#pragma unroll
    for (int v = 0; v < 3; ++v) rp[v] = rc[v];
  }
}

// ---------------------------------------------------------------- level 0
// Minimum waves per SIMD asked of the compiler for the PCG product (hi/lo
// split, DOT epilogue), which then keeps it within 512 / 6 = 80 VGPRs
// without spilling: six waves are three 8-wave workgroups per CU, so the
// 1,518 tiles of 160x80x80 run in two rounds instead of three (24.6 against
// 26.4 us). The single-pass instantiations fit 80 VGPRs by themselves; the
// other split ones would spill.
template <bool SPLIT, int MODE>
constexpr int hex_min_waves() {
  return (SPLIT && MODE == TG_EP_DOT) ? 6 : 1;
}

template <bool SPLIT, int MODE>
__global__ void __launch_bounds__(Tile0::NTHR, (hex_min_waves<SPLIT, MODE>())) k_hex_mfma(Grid g, TileShape ts, const uint8_t* __restrict__ mask,
                                                         const float* __restrict__ E, const float* __restrict__ u,
                                                         Ep32 ep, Workspace ws) {
  using T = Tile0;
  using Mode = EpMode<MODE>;
  __shared__ float4 su[T::NBP];
  // [x line][y line][side][z][3]: side 0 is the share of the element column
  // on the -x side of the line, side 1 the +x side.
  __shared__ float slot[T::SLOT];

  Block b = block_setup<T>(g, ts, mask, u, ep);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const PencilLane<T> pl(w);
  // loads in the order they are needed (they complete in issue order)
  bf16x8 ah[2][2], al[2][2];
  ke_frags<SPLIT>(pl.lane, ah, al);
  uint32_t bw[T::NIT][3], bm[T::NIT], bd[T::NIT][3];
  brick_load<T, Mode::kPrescale>(b, bw, bm, bd);
  float ev[T::PY];
  e_load<T>(g, b, E, w, pl.je, ev);
  uint32_t ob[T::OIT], om[T::OIT];
  typename Mode::In oin[T::OIT];
  out_prefetch<T, MODE>(b, ep, ob, om, oin);
  brick_store<T, Mode::kPrescale>(su, bw, bm, bd, ep.a3);
  __syncthreads();
  pencils0<SPLIT, T>(su, slot, pl, ah, al, ev);
  __syncthreads();

  const double part = out_store<T, MODE>(b, ep, slot, su, ob, om, oin);
  if constexpr (Mode::kReduces) {
    double v[1] = {part};
    block_partials<1>(v, ws);
  }
}

// ---------------------------------------------------------------- level 1
// Coarse element (I,J,K) of grid g has matrix sum_s E_f(2I+sx, 2J+sy, 2K+sz)
// Q_s (s = 4 sx + 2 sy + sz) with E_f on the fine grid 2g. Per pencil the
// eight products Q_s U share the B operand U; each is scaled by its child's
// modulus per element (column) and accumulated in FP32. The child matrices
// are streamed per s (8 fragments per lane) and applied to half of the
// wavefront's pencils at a time, which bounds the registers held.
template <bool SPLIT, int MODE>
__global__ void __launch_bounds__(Tile1::NTHR) k_q8_mfma(Grid g, TileShape ts, const uint8_t* __restrict__ mask,
                                                        const float* __restrict__ Ef, const float* __restrict__ u,
                                                        Ep32 ep, Workspace ws) {
  using T = Tile1;
  using Mode = EpMode<MODE>;
  constexpr int H0 = (T::PY + 1) / 2;   // pencils of the first half
  __shared__ float4 su[T::NBP];
  __shared__ float slot[T::SLOT];

  Block b = block_setup<T>(g, ts, mask, u, ep);
  const int fny = 2 * g.ny, fnz = 2 * g.nz;
  const long long fslab = 2ll * b.is * fny * fnz, fne = 8ll * g.nx * g.ny * g.nz;
  const rsrc_t rs_e = make_rsrc(Ef + fslab, 4ull * (fne - fslab));

  uint32_t bw[T::NIT][3], bm[T::NIT], bd[T::NIT][3];
  brick_load<T, Mode::kPrescale>(b, bw, bm, bd);
  uint32_t ob[T::OIT], om[T::OIT];
  typename Mode::In oin[T::OIT];
  out_prefetch<T, MODE>(b, ep, ob, om, oin);

  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const PencilLane<T> pl(w);
  // fine offset of child (0,0,0) of the lane's element in pencil py = 0
  const int ci = b.i0 - 1 + w, ck = b.k0 - 1 + pl.je;
  const bool okx = ((unsigned)ci < (unsigned)g.nx) & ((unsigned)ck < (unsigned)g.nz);
  const uint32_t fbase = (uint32_t)(((2 * (ci - b.is)) * fny + 2 * (b.j0 - 1)) * fnz + 2 * ck);
  auto child_e = [&](int s, int py) -> float {
    const bool ok = okx & ((unsigned)(b.j0 - 1 + py) < (unsigned)g.ny);
    const uint32_t off = fbase + (uint32_t)(((s >> 2) * fny + 2 * py + ((s >> 1) & 1)) * fnz + (s & 1));
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs_e, oob_unless(4u * off, ok), 0, 0));
  };
  brick_store<T, Mode::kPrescale>(su, bw, bm, bd, ep.a3);
  __syncthreads();

  float rp[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    constexpr int NPH = H0;   // slots per half (the second half uses PY - H0 of them)
    const int p0 = half ? H0 : 0, np = half ? T::PY - H0 : H0;
    bf16x8 bh[NPH], bl[NPH];
    floatx4 a0[NPH], a1[NPH];
#pragma unroll
    for (int q = 0; q < NPH; ++q) {
      if (q < np) {
        const floatx8 f = pl.rel_disp(su, p0 + q);
        bh[q] = __builtin_convertvector(f, bf16x8);
        bl[q] = bh[q];
        if constexpr (SPLIT) bl[q] = __builtin_convertvector(f - __builtin_convertvector(bh[q], floatx8), bf16x8);
      }
      a0[q] = floatx4{0.f, 0.f, 0.f, 0.f};
      a1[q] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
    float ec[NPH];
#pragma unroll
    for (int q = 0; q < NPH; ++q) ec[q] = q < np ? child_e(0, p0 + q) : 0.f;
#pragma unroll 1
    for (int s = 0; s < 8; ++s) {   // not unrolled: one child's operands live at a time
      bf16x8 ah[2][2], al[2][2];
#pragma unroll
      for (int ly = 0; ly < 2; ++ly)
#pragma unroll
        for (int tl = 0; tl < 2; ++tl) {
          ah[ly][tl] = load_q8(s, ly, 0, tl, pl.lane);
          al[ly][tl] = SPLIT ? load_q8(s, ly, 1, tl, pl.lane) : ah[ly][tl];
        }
      float en[NPH];
#pragma unroll
      for (int q = 0; q < NPH; ++q) en[q] = (s < 7 && q < np) ? child_e(s + 1, p0 + q) : 0.f;
#pragma unroll
      for (int q = 0; q < NPH; ++q) {
        if (q < np) {
          const int ly = (p0 + q) & 1;
          floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
          mma2<SPLIT>(ah[ly], al[ly], bh[q], bl[q], d0, d1);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            a0[q][v] = fmaf(ec[q], d0[v], a0[q][v]);
            a1[q][v] = fmaf(ec[q], d1[v], a1[q][v]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < NPH; ++q) ec[q] = en[q];
    }
#pragma unroll
    for (int q = 0; q < NPH; ++q) {
      if (q < np) {
        const int py = p0 + q;
        float rc[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) rc[v] = a0[q][v] + row_shift_up(a1[q][v]);
        if (py > 0) pl.pair_store(slot, py, rp, rc);
#pragma unroll
        for (int v = 0; v < 3; ++v) rp[v] = rc[v];
      }
    }
  }
  __syncthreads();

  const double part = out_store<T, MODE>(b, ep, slot, su, ob, om, oin);
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

template <class T>
int tile_launch_shape(const Grid& g, TileShape& ts, long long& nb) {
  const int ntz = (g.NZ() + T::MTZ - 1) / T::MTZ, nty = (g.NY() + T::MTY - 1) / T::MTY;
  auto magic = [](int d) { return d == 1 ? 0u : (uint32_t)(((1ull << 32) + (unsigned)d - 1) / (unsigned)d); };
  ts = TileShape{ntz, nty, magic(ntz), magic(nty)};
  nb = (long long)((g.NX() + T::MTX - 1) / T::MTX) * ts.nty * ts.ntz;
  TG_REQUIRE(nb <= kMaxPartials, "grid too large for the reduction workspace");
  // the tile index divisions (div_tiles) are exact while index * divisor < 2^32
  TG_REQUIRE(nb * (long long)(ntz > nty ? ntz : nty) < (1LL << 32), "too many tiles");
  // every slab offset must stay below 2^31 bytes
  TG_REQUIRE((long long)(T::MTX + 2) * g.NY() * g.NZ() * 12 < (long long)kMaxRange,
             "grid cross-section %d x %d nodes too large for 32-bit slab offsets", g.NY(), g.NZ());
  return TG_OK;
}

}  // namespace

int launch_mfma_probe(const uint16_t* a, const uint16_t* b, float* d, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, reinterpret_cast<const uint4*>(a),
                     reinterpret_cast<const uint4*>(b), reinterpret_cast<float4*>(d));
  TG_LAUNCH_CHECK();
  return TG_OK;
}

int launch_hex_mfma(const Grid& g, const uint8_t* mask, const float* E, const float* u, const Ep32& ep,
                    double* dot, bool split, int nrhs, hipStream_t s) {
  TileShape ts;
  long long nb;
  if (int rc = tile_launch_shape<Tile0>(g, ts, nb)) return rc;
  TG_REQUIRE(nb * nrhs <= kMaxPartials, "grid too large for the reduction workspace");
  return with_mode(ep, [&](const auto& em) {
    constexpr int M = std::decay_t<decltype(em)>::kMode;
    if (split)
      hipLaunchKernelGGL((k_hex_mfma<true, M>), dim3((unsigned)nb, nrhs), dim3(Tile0::NTHR), 0, s, g, ts, mask, E,
                         u, ep, workspace());
    else
      hipLaunchKernelGGL((k_hex_mfma<false, M>), dim3((unsigned)nb, nrhs), dim3(Tile0::NTHR), 0, s, g, ts, mask, E,
                         u, ep, workspace());
    TG_LAUNCH_CHECK();
    return EpMode<M>::kReduces ? finalize_batched(1, (int)nb, nrhs, dot, nullptr, s) : TG_OK;
  });
}

int launch_q8_mfma(const Grid& g, const uint8_t* mask, const float* Ef, const float* u, const Ep32& ep, double* dot,
                   bool split, int nrhs, hipStream_t s) {
  TileShape ts;
  long long nb;
  if (int rc = tile_launch_shape<Tile1>(g, ts, nb)) return rc;
  TG_REQUIRE(nb * nrhs <= kMaxPartials, "grid too large for the reduction workspace");
  // fine-grid slab offsets: NW element columns = 2 NW fine x-planes
  TG_REQUIRE((long long)(2 * Tile1::NW) * (2 * g.ny) * (2 * g.nz) * 4 < (long long)kMaxRange,
             "grid cross-section too large for 32-bit slab offsets");
  return with_mode(ep, [&](const auto& em) {
    constexpr int M = std::decay_t<decltype(em)>::kMode;
    if (split)
      hipLaunchKernelGGL((k_q8_mfma<true, M>), dim3((unsigned)nb, nrhs), dim3(Tile1::NTHR), 0, s, g, ts, mask, Ef, u,
                         ep, workspace());
    else
      hipLaunchKernelGGL((k_q8_mfma<false, M>), dim3((unsigned)nb, nrhs), dim3(Tile1::NTHR), 0, s, g, ts, mask, Ef,
                         u, ep, workspace());
    TG_LAUNCH_CHECK();
    return EpMode<M>::kReduces ? finalize_batched(1, (int)nb, nrhs, dot, nullptr, s) : TG_OK;
  });
}

}  // namespace tg
