// Geometric multigrid on the structured hex grid: trilinear transfers,
// Galerkin coarse element matrices and the dense coarsest solve.
//
// Coarse grid per axis: n_c = max(1, n_f / 2). Prolongation per axis (fine
// node i, coarse node I): even i takes weight 1 from I = i/2; odd i takes 1/2
// from i/2 and i/2 + 1; parents past n_c are dropped and each fine row is
// rescaled to sum 1 (the reference's row-normalised trilinear P,
// solver_v2.py:62-149). P is the Kronecker product of the three axes, so the
// normalisation factorises per axis.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.h"
#include "dense.h"
#include "reduce.h"

namespace tg {
namespace {

struct AxisW {
  int lo, hi;      // coarse parents (hi == lo when single)
  double wlo, whi;
};

__host__ __device__ __forceinline__ AxisW axis_weights(int i, int nc) {
  AxisW a;
  a.lo = i >> 1;
  if ((i & 1) == 0) {
    a.hi = a.lo; a.wlo = 1.0; a.whi = 0.0;
  } else if (a.lo + 1 <= nc) {
    a.hi = a.lo + 1; a.wlo = 0.5; a.whi = 0.5;
  } else {
    a.hi = a.lo; a.wlo = 1.0; a.whi = 0.0;   // lone parent after normalisation
  }
  return a;
}

// Weight of coarse node I for fine node i along one axis (0 if not a parent).
__host__ __device__ __forceinline__ double axis_w(int i, int I, int nc) {
  const AxisW a = axis_weights(i, nc);
  double w = 0.0;
  if (a.lo == I) w += a.wlo;
  if (a.hi == I && a.hi != a.lo) w += a.whi;
  return w;
}

// Transfers of a batch of right-hand sides: blockIdx.y selects the RHS, whose
// vectors are contiguous slices of 3 * nodes of their grid.
// The transfer weights are 1, 1/2 (and their products), exact in FP32.
// Both kernels evaluate every candidate term, with weight 0 and a clamped
// (valid) index for a non-parent, so all loads are independent and issue
// back to back; adding 0 * x leaves a finite sum unchanged, so the result is
// that of summing the parents alone, in the same order.
struct Par3 {
  int idx[3];
  float w[3];
};

// Fine nodes 2I-1, 2I, 2I+1 along one axis (nf fine elements, nc coarse)
// and their prolongation weights from coarse node I.
__device__ __forceinline__ Par3 restrict_axis(int I, int nf, int nc) {
  Par3 p;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int i = 2 * I + d - 1;
    const bool in = (i >= 0) & (i <= nf);
    p.idx[d] = in ? i : (i < 0 ? 0 : nf);
    p.w[d] = in ? (float)axis_w(i, I, nc) : 0.0f;
  }
  return p;
}

__device__ __forceinline__ float3 ld3(const float* p) { return make_float3(p[0], p[1], p[2]); }

__global__ void __launch_bounds__(256) k_prolong(Grid f, Grid c, const uint8_t* __restrict__ mask_f,
                                                 const float* __restrict__ ec, float* __restrict__ xf, int accumulate) {
  ec += (long long)blockIdx.y * 3 * c.nodes();
  xf += (long long)blockIdx.y * 3 * f.nodes();
  const unsigned n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (unsigned)f.nodes()) return;
  const unsigned t = n / (unsigned)f.NZ();
  const int k = (int)(n - t * (unsigned)f.NZ()), j = (int)(t % (unsigned)f.NY()), i = (int)(t / (unsigned)f.NY());
  const AxisW ax = axis_weights(i, c.nx), ay = axis_weights(j, c.ny), az = axis_weights(k, c.nz);
  // the weights are 0, 1/2 and 1 and their products: exact in FP32, so the
  // FP32 products equal the rounded FP64 ones
  const float wx[2] = {(float)ax.wlo, (float)ax.whi}, wy[2] = {(float)ay.wlo, (float)ay.whi},
              wz[2] = {(float)az.wlo, (float)az.whi};
  float3 v[8];
  float w[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int I = (a & 4) ? ax.hi : ax.lo, J = (a & 2) ? ay.hi : ay.lo, K = (a & 1) ? az.hi : az.lo;
    w[a] = wx[(a >> 2) & 1] * wy[(a >> 1) & 1] * wz[a & 1];
    v[a] = ld3(ec + 3 * ((I * c.NY() + J) * c.NZ() + K));
  }
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    acc[0] = fmaf(w[a], v[a].x, acc[0]);
    acc[1] = fmaf(w[a], v[a].y, acc[1]);
    acc[2] = fmaf(w[a], v[a].z, acc[2]);
  }
  const unsigned m = mask_f[n];
  float* o = xf + 3 * (size_t)n;
  const float3 base = accumulate ? ld3(o) : make_float3(0.f, 0.f, 0.f);
  o[0] = (m & 1u) ? base.x + acc[0] : base.x;
  o[1] = (m & 2u) ? base.y + acc[1] : base.y;
  o[2] = (m & 4u) ? base.z + acc[2] : base.z;
}

__global__ void __launch_bounds__(256) k_restrict(Grid f, Grid c, const uint8_t* __restrict__ mask_c,
                                                  const float* __restrict__ rf, float* __restrict__ rc) {
  rf += (long long)blockIdx.y * 3 * f.nodes();
  rc += (long long)blockIdx.y * 3 * c.nodes();
  const unsigned N = blockIdx.x * blockDim.x + threadIdx.x;
  if (N >= (unsigned)c.nodes()) return;
  const unsigned t = N / (unsigned)c.NZ();
  const int K = (int)(N - t * (unsigned)c.NZ()), J = (int)(t % (unsigned)c.NY()), I = (int)(t / (unsigned)c.NY());
  const Par3 px = restrict_axis(I, f.nx, c.nx), py = restrict_axis(J, f.ny, c.ny), pz = restrict_axis(K, f.nz, c.nz);
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float wxy = px.w[a] * py.w[b];
      const float* line = rf + 3 * ((px.idx[a] * f.NY() + py.idx[b]) * f.NZ());
      float3 v[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = ld3(line + 3 * pz.idx[d]);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float w = wxy * pz.w[d];
        acc[0] = fmaf(w, v[d].x, acc[0]);
        acc[1] = fmaf(w, v[d].y, acc[1]);
        acc[2] = fmaf(w, v[d].z, acc[2]);
      }
    }
  const unsigned m = mask_c[N];
  float* o = rc + 3 * (size_t)N;
  o[0] = (m & 1u) ? acc[0] : 0.f;
  o[1] = (m & 2u) ? acc[1] : 0.f;
  o[2] = (m & 4u) ? acc[2] : 0.f;
}

// One block per coarse element: K_C = sum_f (M_f P_f)^T K_f (M_f P_f) over
// its children f in (x, y, z) loop order (two per axis, three at the end of
// an odd fine extent). Up to kGalChildren children per round: their
// matrices, local prolongations and masks are loaded together (one memory
// round trip instead of one per child), each thread turns one row of one
// child's K into the same row of T = K (M P) in place, and every coarse entry
// adds the children's (M P)^T T in child order -- the sums and their order
// of a loop over one child at a time.
constexpr int kGalChildren = 8;

__global__ void __launch_bounds__(256) k_galerkin(Grid f, Grid c, const uint8_t* __restrict__ mask_f, int use_mask,
                                                 const double* __restrict__ fE, const double* __restrict__ fmats,
                                                 double* __restrict__ out64, float* __restrict__ out32) {
  __shared__ double sK[kGalChildren][576];    // K, then T, of each child of the round
  __shared__ double sP[kGalChildren][8][8];   // local P: fine corner x coarse corner
  __shared__ double sM[kGalChildren][24];
  const long long C = blockIdx.x;
  const int CK = (int)(C % c.nz), CJ = (int)((C / c.nz) % c.ny), CI = (int)(C / ((long long)c.nz * c.ny));
  // children ranges per axis
  const int fx0 = 2 * CI, fx1 = (CI == c.nx - 1) ? f.nx - 1 : 2 * CI + 1;
  const int fy0 = 2 * CJ, fy1 = (CJ == c.ny - 1) ? f.ny - 1 : 2 * CJ + 1;
  const int fz0 = 2 * CK, fz1 = (CK == c.nz - 1) ? f.nz - 1 : 2 * CK + 1;
  const int cny = fy1 - fy0 + 1, cnz = fz1 - fz0 + 1;
  const int nch = (fx1 - fx0 + 1) * cny * cnz;
  auto child = [&](int id, int& fi, int& fj, int& fk) {
    fk = fz0 + id % cnz;
    fj = fy0 + (id / cnz) % cny;
    fi = fx0 + id / (cnz * cny);
  };
  double acc[3] = {0.0, 0.0, 0.0};
  for (int c0 = 0; c0 < nch; c0 += kGalChildren) {
    const int cn = nch - c0 < kGalChildren ? nch - c0 : kGalChildren;
    for (int e = threadIdx.x; e < cn * 576; e += blockDim.x) {
      const int ch = e / 576, t = e - ch * 576;
      int fi, fj, fk;
      child(c0 + ch, fi, fj, fk);
      const long long fe = f.elem(fi, fj, fk);
      sK[ch][t] = fmats ? fmats[576 * fe + t] : fE[fe] * c_ke64[t];
    }
    for (int e = threadIdx.x; e < cn * 64; e += blockDim.x) {
      const int ch = e >> 6, a = (e >> 3) & 7, b = e & 7;   // fine corner a, coarse corner b
      int fi, fj, fk;
      child(c0 + ch, fi, fj, fk);
      sP[ch][a][b] = axis_w(fi + corner_x(a), CI + corner_x(b), c.nx) *
                     axis_w(fj + corner_y(a), CJ + corner_y(b), c.ny) *
                     axis_w(fk + corner_z(a), CK + corner_z(b), c.nz);
    }
    for (int e = threadIdx.x; e < cn * 24; e += blockDim.x) {
      const int ch = e / 24, l = e - ch * 24, a = l / 3, q = l % 3;
      double mv = 1.0;
      if (use_mask) {
        int fi, fj, fk;
        child(c0 + ch, fi, fj, fk);
        const long long n = f.node(fi + corner_x(a), fj + corner_y(a), fk + corner_z(a));
        mv = (mask_f[n] >> q & 1u) ? 1.0 : 0.0;
      }
      sM[ch][l] = mv;
    }
    __syncthreads();
    // T[a][B] = sum_b K[a][b] (MP)[b][B], one row of one child per thread
    if (threadIdx.x < cn * 24) {
      const int ch = threadIdx.x / 24, a = threadIdx.x % 24;
      double kr[24];
#pragma unroll
      for (int b = 0; b < 24; ++b) kr[b] = sK[ch][a * 24 + b];
#pragma unroll 1
      for (int Bn = 0; Bn < 8; ++Bn)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          double s = 0.0;
#pragma unroll
          for (int nb = 0; nb < 8; ++nb) s += kr[3 * nb + q] * sM[ch][3 * nb + q] * sP[ch][nb][Bn];
          sK[ch][a * 24 + 3 * Bn + q] = s;
        }
    }
    __syncthreads();
    // K_C[A][B] += sum_a (MP)[a][A] T[a][B], child by child
    for (int r = 0; r < 3; ++r) {
      const int t = threadIdx.x + r * 256;
      if (t < 576) {
        const int A = t / 24, B = t % 24, q = A % 3;
        for (int ch = 0; ch < cn; ++ch) {
          double s = 0.0;
#pragma unroll
          for (int na = 0; na < 8; ++na) s += sP[ch][na][A / 3] * sM[ch][3 * na + q] * sK[ch][(3 * na + q) * 24 + B];
          acc[r] += s;
        }
      }
    }
    __syncthreads();
  }
  for (int r = 0; r < 3; ++r) {
    const int t = threadIdx.x + r * 256;
    if (t < 576) {
      out64[576 * C + t] = acc[r];
      if (out32) out32[576 * C + t] = (float)acc[r];
    }
  }
}

// Row-wise dense assembly: thread per DOF of the grid. The row's (at most
// 27 x 3) nonzeros are summed in registers over the elements in corner order
// and stored once each (A is zero beforehand).
__global__ void __launch_bounds__(256) k_assemble_dense(Grid g, const double* __restrict__ mats,
                                                        const int* __restrict__ fidx, int nf, double* __restrict__ A) {
  const long long d = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (d >= 3 * g.nodes()) return;
  const int row = fidx[d];
  if (row < 0) return;
  const long long n = d / 3;
  const int r = (int)(d % 3);
  const int k = (int)(n % g.NZ()), j = (int)((n / g.NZ()) % g.NY()), i = (int)(n / ((long long)g.NZ() * g.NY()));
  double acc[27][3];
#pragma unroll
  for (int q = 0; q < 27; ++q) acc[q][0] = acc[q][1] = acc[q][2] = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {   // the node is corner a of element (i,j,k) - corner(a)
    const int ei = i - corner_x(a), ej = j - corner_y(a), ek = k - corner_z(a);
    if (ei < 0 || ei >= g.nx || ej < 0 || ej >= g.ny || ek < 0 || ek >= g.nz) continue;
    const double* K = mats + 576 * g.elem(ei, ej, ek) + (3 * a + r) * 24;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int q = (corner_x(b) - corner_x(a) + 1) * 9 + (corner_y(b) - corner_y(a) + 1) * 3 +
                    (corner_z(b) - corner_z(a) + 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[q][c] += K[3 * b + c];
    }
  }
  double* Ar = A + (long long)row * nf;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int a = i + q / 9 - 1, b = j + (q / 3) % 3 - 1, e = k + q % 3 - 1;
    if (a < 0 || a >= g.NX() || b < 0 || b >= g.NY() || e < 0 || e >= g.NZ()) continue;
    const long long m = g.node(a, b, e);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int col = fidx[3 * m + c];
      if (col >= 0) Ar[col] = acc[q][c];
    }
  }
}

// A (n x n, upper triangle of the row-major array valid) -> full symmetric.
__global__ void k_set_int(int* p, int v) { *p = v; }

// Matrix blockIdx.y of a batch at stride sa.
__global__ void k_symmetrize_from_upper(double* A, int n, long long sa) {
  A += blockIdx.y * sa;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)n * n) return;
  const int r = (int)(t / n), c = (int)(t % n);
  if (r > c) A[t] = A[(long long)c * n + r];
}

// One wavefront per row: ec[free[i]] = sum_j Ainv[i][j] rc[free[j]]; RHS
// blockIdx.y of a batch uses the slices at blockIdx.y * n_full. Blocks past
// the n / 4 row blocks write the fixed entries of ec as zero, one thread per
// entry (an entry is fixed when a binary search of the ascending free list
// misses it), so no entry is written twice and no fill launch is needed.
__global__ void k_coarsest_apply(const double* __restrict__ Ainv, int n, const int* __restrict__ fl,
                                 long long n_full, const float* __restrict__ rc, float* __restrict__ ec) {
  rc += blockIdx.y * n_full;
  ec += blockIdx.y * n_full;
  const int row_blocks = (n + 3) / 4;
  if ((int)blockIdx.x >= row_blocks) {
    const long long e = (long long)(blockIdx.x - row_blocks) * blockDim.x + threadIdx.x;
    if (e >= n_full) return;
    int lo = 0, hi = n;   // first free index >= e
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (fl[mid] < e) lo = mid + 1;
      else hi = mid;
    }
    if (lo == n || fl[lo] != e) ec[e] = 0.0f;
    return;
  }
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* a = Ainv + (long long)row * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s = fma(a[j], (double)rc[fl[j]], s);
  s = wave_sum(s);
  if (lane == 0) ec[fl[row]] = (float)s;
}

// Inverse of a symmetric positive definite n x n matrix (row-major) by
// Gauss-Jordan elimination without pivoting. The pivots of Gauss-Jordan on
// an SPD matrix are the pivots of its Cholesky factorisation squared, so
// they stay positive; a pivot <= 0 (or NaN) reports "not SPD" through info.
//
// Small matrices (n <= 64), and the diagonal blocks of the blocked form, are
// inverted by one workgroup with the matrix in registers: thread (tr, tc) =
// (tid / 32, tid % 32) owns the entries (tr + 32a, tc + 32b), a, b < NB.
// Step k publishes row k and column k through LDS and every thread updates
// its own entries, so each of the n steps costs two barriers and a few FMAs
// per thread, with no global-memory round trip.
//
// Larger matrices use the blocked form with BK-wide pivot blocks: for each
// diagonal block k, with P = inverse of the current A_kk,
//   A_kj <- P A_kj,  A_ij <- A_ij - A_ik A_kj (new A_kj),  A_ik <- -A_ik P,
//   A_kk <- P      (i, j outside block k),
// which leaves the inverse in place after the last block. The updates are
// tiled FP64 kernels over TS x TS tiles; each block step is four launches.
// The pivot block stays narrow: the panels are multiplied by an explicit
// block inverse, whose rounding on an ill-conditioned matrix grows with the
// block width (at condition 1e8, 64-wide blocks lose ~1e-5 relative
// accuracy, 16-wide blocks stay at the ~1e-9 of scalar elimination).
constexpr int GJ_THREADS = 1024;
constexpr int BK = 16, TS = 64;

// Invert the b x b matrix at A (leading dimension lda) into P (leading
// dimension ldp; P may be A). info: first failing pivot + base + 1; with
// zero_info, thread 0 clears *info first. Block x of a batch works on the
// matrices at A + x * sa and P + x * sp.
template <int NB>
__global__ void __launch_bounds__(GJ_THREADS) k_spd_inverse_gj(const double* A, int lda, int n, double* P, int ldp,
                                                             int* info, int info_base, int zero_info, long long sa,
                                                             long long sp) {
  A += blockIdx.x * sa;
  P += blockIdx.x * sp;
  __shared__ double prow[32 * NB];
  __shared__ double pcol[32 * NB];
  const int tid = threadIdx.x, tr = tid >> 5, tc = tid & 31;
  if (zero_info && tid == 0) *info = 0;
  double a[NB][NB];
#pragma unroll
  for (int x = 0; x < NB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      const int i = tr + 32 * x, j = tc + 32 * y;
      a[x][y] = (i < n && j < n) ? A[(long long)i * lda + j] : 0.0;
    }
  for (int k = 0; k < n; ++k) {
    const int kq = k >> 5, kr = k & 31;
    // column kq / row kq of the register tile, selected without dynamic
    // register indexing (which would move the tile to scratch memory)
    if (tc == kr)
#pragma unroll
      for (int x = 0; x < NB; ++x) {
        double v = 0.0;
#pragma unroll
        for (int y = 0; y < NB; ++y) v = (y == kq) ? a[x][y] : v;
        pcol[tr + 32 * x] = v;
      }
    if (tr == kr)
#pragma unroll
      for (int y = 0; y < NB; ++y) {
        double v = 0.0;
#pragma unroll
        for (int x = 0; x < NB; ++x) v = (x == kq) ? a[x][y] : v;
        prow[tc + 32 * y] = v;
      }
    __syncthreads();
    const double p = pcol[k];   // uniform
    if (!(p > 0.0)) {
      if (tid == 0) *info = info_base + k + 1;
      return;
    }
    const double ip = 1.0 / p;
    double pr[NB], pc[NB];
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      const int j = tc + 32 * y;
      pr[y] = (j == k ? 1.0 : prow[j]) * ip;
    }
#pragma unroll
    for (int x = 0; x < NB; ++x) pc[x] = pcol[tr + 32 * x];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y) {
        const int i = tr + 32 * x, j = tc + 32 * y;
        a[x][y] = (i == k) ? pr[y] : (j == k ? 0.0 : a[x][y]) - pc[x] * pr[y];
      }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < NB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      const int i = tr + 32 * x, j = tc + 32 * y;
      if (i < n && j < n) P[(long long)i * ldp + j] = a[x][y];
    }
}

// A_kj <- P A_kj: rows [k0, k0 + kw) of column tile blockIdx.x (columns
// inside the pivot block are left alone). In this and the next two kernels
// blockIdx.z selects the matrix A + z * sa of a batch and its pivot inverse
// P + z * BK * BK.
__global__ void __launch_bounds__(256) k_gj_row_panel(double* A, int n, int k0, int kw, const double* P, long long sa) {
  A += blockIdx.z * sa;
  P += blockIdx.z * BK * BK;
  __shared__ double sp[BK][BK + 1], sb[BK][TS + 1];
  const int j0 = blockIdx.x * TS, jw = min(TS, n - j0);
  for (int e = threadIdx.x; e < BK * BK; e += blockDim.x) {
    const int r = e / BK, c = e % BK;
    sp[r][c] = (r < kw && c < kw) ? P[r * BK + c] : 0.0;
  }
  for (int e = threadIdx.x; e < BK * TS; e += blockDim.x) {
    const int r = e / TS, c = e % TS;
    sb[r][c] = (r < kw && c < jw) ? A[(long long)(k0 + r) * n + j0 + c] : 0.0;
  }
  __syncthreads();
  const int c = threadIdx.x % TS, rg = threadIdx.x / TS;   // 4 row groups
  const int j = j0 + c;
  const bool colok = c < jw && (j < k0 || j >= k0 + kw);
#pragma unroll
  for (int q = 0; q < BK / 4; ++q) {
    const int r = rg + 4 * q;
    double acc = 0.0;
    for (int kk = 0; kk < kw; ++kk) acc = fma(sp[r][kk], sb[kk][c], acc);
    if (colok && r < kw) A[(long long)(k0 + r) * n + j] = acc;
  }
}

// A_ij -= A_ik A_kj on tile (blockIdx.y, blockIdx.x), rows and columns
// inside the pivot block excluded. Thread (tr, tc) owns rows tr + 16a and
// columns tc + 16b of the tile, a, b < 4.
__global__ void __launch_bounds__(256) k_gj_trailing(double* A, int n, int k0, int kw, long long stride) {
  A += blockIdx.z * stride;
  __shared__ double sa[TS][BK + 1], sb[BK][TS + 1];
  const int i0 = blockIdx.y * TS, iw = min(TS, n - i0), j0 = blockIdx.x * TS, jw = min(TS, n - j0);
  for (int e = threadIdx.x; e < TS * BK; e += blockDim.x) {
    const int r = e / BK, c = e % BK;
    sa[r][c] = (r < iw && c < kw) ? A[(long long)(i0 + r) * n + k0 + c] : 0.0;
  }
  for (int e = threadIdx.x; e < BK * TS; e += blockDim.x) {
    const int r = e / TS, c = e % TS;
    sb[r][c] = (r < kw && c < jw) ? A[(long long)(k0 + r) * n + j0 + c] : 0.0;
  }
  __syncthreads();
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4] = {};
  for (int kk = 0; kk < kw; ++kk) {
    double l[4], r[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) l[a] = sa[tr + 16 * a][kk];
#pragma unroll
    for (int b = 0; b < 4; ++b) r[b] = sb[kk][tc + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(l[a], r[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + tr + 16 * a;
    const bool rowok = tr + 16 * a < iw && (i < k0 || i >= k0 + kw);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tc + 16 * b;
      if (rowok && tc + 16 * b < jw && (j < k0 || j >= k0 + kw)) A[(long long)i * n + j] -= acc[a][b];
    }
  }
}

// A_ik <- -A_ik P for the rows of row tile blockIdx.x outside the pivot
// block; the pivot block's rows get A_kk <- P.
__global__ void __launch_bounds__(256) k_gj_col_panel(double* A, int n, int k0, int kw, const double* P,
                                                      long long stride) {
  A += blockIdx.z * stride;
  P += blockIdx.z * BK * BK;
  __shared__ double sa[TS][BK + 1], sp[BK][BK + 1];
  const int i0 = blockIdx.x * TS, iw = min(TS, n - i0);
  for (int e = threadIdx.x; e < BK * BK; e += blockDim.x) {
    const int r = e / BK, c = e % BK;
    sp[r][c] = (r < kw && c < kw) ? P[r * BK + c] : 0.0;
  }
  for (int e = threadIdx.x; e < TS * BK; e += blockDim.x) {
    const int r = e / BK, c = e % BK;
    sa[r][c] = (r < iw && c < kw) ? A[(long long)(i0 + r) * n + k0 + c] : 0.0;
  }
  __syncthreads();
  const int r = threadIdx.x % TS, cg = threadIdx.x / TS;   // 4 column groups
  const int i = i0 + r;
  if (r >= iw) return;
  const bool inblock = i >= k0 && i < k0 + kw;
#pragma unroll
  for (int q = 0; q < BK / 4; ++q) {
    const int c = cg + 4 * q;
    if (c >= kw) continue;
    double acc = 0.0;
    for (int kk = 0; kk < kw; ++kk) acc = fma(sa[r][kk], sp[kk][c], acc);
    A[(long long)i * n + k0 + c] = inblock ? sp[i - k0][c] : -acc;
  }
}

// The blocked Gauss-Jordan above in one workgroup with the whole matrix in
// LDS (n <= kLdsInvMaxN): each BK-wide pivot block is inverted in registers
// (threads (r, c) < BK, the scalar steps of k_spd_inverse_gj), then the row
// panel, the trailing update and the column panel are formed from LDS with
// the same sums in the same order as k_gj_row_panel, k_gj_trailing and
// k_gj_col_panel. One launch replaces four per pivot block; the upper
// triangle is mirrored before the write-back, as k_symmetrize_from_upper
// does.
constexpr int kLdsInvMaxN = 140;   // 140^2 doubles = 153 KiB of LDS
constexpr int kLdsInvNB = (kLdsInvMaxN + 31) / 32;

__global__ void __launch_bounds__(GJ_THREADS) k_spd_inverse_lds(double* __restrict__ A, int n, int* info) {
  __shared__ double sA[kLdsInvMaxN * kLdsInvMaxN];
  __shared__ double sP[BK * BK];
  __shared__ double prow[BK], pcol[BK];
  const int tid = threadIdx.x;
  const int nn = n * n;
  // several loads of the matrix in flight per thread (an unrolled loop)
#pragma unroll 6
  for (int e = tid; e < nn; e += GJ_THREADS) sA[e] = A[e];
  if (tid == 0) *info = 0;
  __syncthreads();
  const int pr_r = tid >> 4, pr_c = tid & 15;    // pivot-block entry of threads < BK * BK
  const int ti = tid >> 5, tj = tid & 31;          // trailing-update layout
  for (int k0 = 0; k0 < n; k0 += BK) {
    const int kw = n - k0 < BK ? n - k0 : BK;
    // P = inverse of the pivot block (scalar Gauss-Jordan, as k_spd_inverse_gj)
    const bool pin = tid < BK * BK && pr_r < kw && pr_c < kw;
    double a = pin ? sA[(k0 + pr_r) * n + k0 + pr_c] : 0.0;
    for (int k = 0; k < kw; ++k) {
      if (pin && pr_c == k) pcol[pr_r] = a;
      if (pin && pr_r == k) prow[pr_c] = a;
      __syncthreads();
      const double p = pcol[k];   // uniform
      if (!(p > 0.0)) {
        if (tid == 0) *info = k0 + k + 1;
        return;
      }
      const double ip = 1.0 / p;
      if (pin) {
        const double prj = (pr_c == k ? 1.0 : prow[pr_c]) * ip;
        const double pci = pcol[pr_r];
        a = (pr_r == k) ? prj : (pr_c == k ? 0.0 : a) - pci * prj;
      }
      __syncthreads();
    }
    if (pin) sP[pr_r * BK + pr_c] = a;
    __syncthreads();
    // row panel: A_kj <- P A_kj (k_gj_row_panel's sums)
    constexpr int RPW = (BK * kLdsInvMaxN + GJ_THREADS - 1) / GJ_THREADS;
    double rp[RPW];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int e = tid + q * GJ_THREADS, r = e / n, j = e - r * n;
      double acc = 0.0;
      if (r < kw)
        for (int kk = 0; kk < kw; ++kk) acc = fma(sP[r * BK + kk], sA[(k0 + kk) * n + j], acc);
      rp[q] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
      const int e = tid + q * GJ_THREADS, r = e / n, j = e - r * n;
      if (r < kw && (j < k0 || j >= k0 + kw)) sA[(k0 + r) * n + j] = rp[q];
    }
    __syncthreads();
    // trailing update: A_ij -= A_ik A_kj (k_gj_trailing's sums)
    double acc[kLdsInvNB][kLdsInvNB] = {};
    for (int kk = 0; kk < kw; ++kk) {
      double l[kLdsInvNB], r[kLdsInvNB];
#pragma unroll
      for (int x = 0; x < kLdsInvNB; ++x) {
        const int i = ti + 32 * x;
        l[x] = i < n ? sA[i * n + k0 + kk] : 0.0;
      }
#pragma unroll
      for (int y = 0; y < kLdsInvNB; ++y) {
        const int j = tj + 32 * y;
        r[y] = j < n ? sA[(k0 + kk) * n + j] : 0.0;
      }
#pragma unroll
      for (int x = 0; x < kLdsInvNB; ++x)
#pragma unroll
        for (int y = 0; y < kLdsInvNB; ++y) acc[x][y] = fma(l[x], r[y], acc[x][y]);
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < kLdsInvNB; ++x) {
      const int i = ti + 32 * x;
#pragma unroll
      for (int y = 0; y < kLdsInvNB; ++y) {
        const int j = tj + 32 * y;
        if (i < n && j < n && (i < k0 || i >= k0 + kw) && (j < k0 || j >= k0 + kw)) sA[i * n + j] -= acc[x][y];
      }
    }
    // column panel from the old A_ik, which the trailing update leaves alone
    // (k_gj_col_panel's sums); written once every thread has formed its own
    constexpr int CPW = (kLdsInvMaxN * BK + GJ_THREADS - 1) / GJ_THREADS;
    double cp[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const int e = tid + q * GJ_THREADS, i = e / BK, c = e - i * BK;
      double v = 0.0;
      if (i < n && c < kw) {
        if (i >= k0 && i < k0 + kw) {
          v = sP[(i - k0) * BK + c];
        } else {
          for (int kk = 0; kk < kw; ++kk) v = fma(sA[i * n + k0 + kk], sP[kk * BK + c], v);
          v = -v;
        }
      }
      cp[q] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const int e = tid + q * GJ_THREADS, i = e / BK, c = e - i * BK;
      if (i < n && c < kw) sA[i * n + k0 + c] = cp[q];
    }
    __syncthreads();
  }
#pragma unroll 6
  for (int e = tid; e < nn; e += GJ_THREADS) {
    const int r = e / n, c = e - r * n;
    A[e] = r > c ? sA[c * n + r] : sA[e];
  }
}

}  // namespace

struct SolverHandles {
  rocblas_handle h = nullptr;
};
SolverHandles g_handles[64];

int get_handle(hipStream_t s, rocblas_handle* out) {
  int dev = 0;
  TG_HIP_CHECK(hipGetDevice(&dev));
  TG_REQUIRE(dev >= 0 && dev < 64, "device id out of range");
  if (!g_handles[dev].h && rocblas_create_handle(&g_handles[dev].h) != rocblas_status_success)
    return set_error(TG_ERR_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(g_handles[dev].h, s) != rocblas_status_success)
    return set_error(TG_ERR_HIP, "rocblas_set_stream failed");
  *out = g_handles[dev].h;
  return TG_OK;
}

namespace {

// Constant Galerkin tables per device (FP64, row-major 576 per matrix):
// Q[t] = P_t^T KE P_t for the 8 children t of a coarse element, and
// R[8s+t] = P_s^T Q[t] P_s for the 64 grandchildren two levels down.
// P_s is the local trilinear prolongation of child s = 4 sx + 2 sy + sz.
struct GTables {
  double* q8 = nullptr;
  double* r64 = nullptr;
};
GTables g_tables[64];

void local_prolongation(int s, double P[8][8]) {
  static const double w1d[3][2] = {{1.0, 0.0}, {0.5, 0.5}, {0.0, 1.0}};
  const int sx = s >> 2 & 1, sy = s >> 1 & 1, sz = s & 1;
  for (int c = 0; c < 8; ++c)
    for (int C = 0; C < 8; ++C)
      P[c][C] = w1d[sx + corner_x(c)][corner_x(C)] * w1d[sy + corner_y(c)][corner_y(C)] *
                w1d[sz + corner_z(c)][corner_z(C)];
}

// out = P^T K P for a 24x24 K and the 8x8 node prolongation P (kron I3).
void ptkp(const double P[8][8], const double* K, double* out) {
  for (int A = 0; A < 24; ++A)
    for (int B = 0; B < 24; ++B) {
      double acc = 0.0;
      for (int a = 0; a < 8; ++a) {
        const double pa = P[a][A / 3];
        if (pa == 0.0) continue;
        for (int b = 0; b < 8; ++b) {
          const double pb = P[b][B / 3];
          if (pb == 0.0) continue;
          acc += pa * K[(3 * a + A % 3) * 24 + 3 * b + B % 3] * pb;
        }
      }
      out[A * 24 + B] = acc;
    }
}

int ensure_tables(GTables** out) {
  int dev = 0;
  TG_HIP_CHECK(hipGetDevice(&dev));
  TG_REQUIRE(dev >= 0 && dev < 64, "device id out of range");
  GTables& t = g_tables[dev];
  if (!t.q8) {
    static double q8[8 * 576], r64[64 * 576];
    double P[8][8];
    for (int s = 0; s < 8; ++s) {
      local_prolongation(s, P);
      ptkp(P, ke_host(), q8 + 576 * s);
    }
    for (int s = 0; s < 8; ++s) {
      local_prolongation(s, P);
      for (int c = 0; c < 8; ++c) ptkp(P, q8 + 576 * c, r64 + 576 * (8 * s + c));
    }
    TG_HIP_CHECK(hipMalloc((void**)&t.q8, sizeof(q8)));
    TG_HIP_CHECK(hipMalloc((void**)&t.r64, sizeof(r64)));
    TG_HIP_CHECK(hipMemcpy(t.q8, q8, sizeof(q8), hipMemcpyHostToDevice));
    TG_HIP_CHECK(hipMemcpy(t.r64, r64, sizeof(r64), hipMemcpyHostToDevice));
  }
  *out = &t;
  return TG_OK;
}

// Ew[C][j] = E of descendant j of coarse element C, `down` levels below:
// j = t (down = 1) or 8 s + t (down = 2), positions as in the tables above.
__global__ void k_gather_desc(Grid c, Grid f, int down, const double* __restrict__ fE, double* __restrict__ Ew) {
  const int m = down == 1 ? 8 : 64;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= c.elems() * m) return;
  const long long C = t / m;
  const int j = (int)(t % m);
  const int K = (int)(C % c.nz), J = (int)((C / c.nz) % c.ny), I = (int)(C / ((long long)c.nz * c.ny));
  int fi, fj, fk;
  if (down == 1) {
    fi = 2 * I + (j >> 2 & 1);
    fj = 2 * J + (j >> 1 & 1);
    fk = 2 * K + (j & 1);
  } else {
    const int s = j >> 3, u = j & 7;
    fi = 4 * I + 2 * (s >> 2 & 1) + (u >> 2 & 1);
    fj = 4 * J + 2 * (s >> 1 & 1) + (u >> 1 & 1);
    fk = 4 * K + 2 * (s & 1) + (u & 1);
  }
  Ew[t] = fE[f.elem(fi, fj, fk)];
}

__global__ void k_to_f32(long long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

}  // namespace
}  // namespace tg

using namespace tg;

#define TG_NEED_INIT() TG_REQUIRE(::tg::ke_ready(), "tg_init has not been called on this device")

static int check_pair(const tg_grid* fine, const tg_grid* coarse) {
  if (int rc = check_grid(fine)) return rc;
  if (int rc = check_grid(coarse)) return rc;
  TG_REQUIRE(coarse->nelx == (fine->nelx / 2 > 1 ? fine->nelx / 2 : 1) &&
                 coarse->nely == (fine->nely / 2 > 1 ? fine->nely / 2 : 1) &&
                 coarse->nelz == (fine->nelz / 2 > 1 ? fine->nelz / 2 : 1),
             "coarse grid must be max(1, n/2) of the fine grid");
  // the transfer kernels index vectors with 32-bit arithmetic
  TG_REQUIRE(3LL * to_grid(fine).nodes() < (1LL << 31), "grid too large for the transfer kernels");
  return TG_OK;
}

extern "C" int tg_restrict32_b(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_c, int32_t nrhs,
                               const float* rf, float* rc, void* s) {
  if (int r = check_pair(fine, coarse)) return r;
  TG_REQUIRE(mask_c && rf && rc && nrhs >= 1 && nrhs <= 64, "restrict: bad argument");
  const Grid f = to_grid(fine), c = to_grid(coarse);
  hipLaunchKernelGGL(k_restrict, dim3((unsigned)((c.nodes() + 255) / 256), nrhs), dim3(256), 0, (hipStream_t)s, f,
                     c, mask_c, rf, rc);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_restrict32(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_c, const float* rf,
                             float* rc, void* s) {
  return tg_restrict32_b(fine, coarse, mask_c, 1, rf, rc, s);
}

extern "C" int tg_prolong32_b(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_f, int32_t nrhs,
                              const float* ec, float* xf, int accumulate, void* s) {
  if (int r = check_pair(fine, coarse)) return r;
  TG_REQUIRE(mask_f && ec && xf && nrhs >= 1 && nrhs <= 64, "prolong: bad argument");
  const Grid f = to_grid(fine), c = to_grid(coarse);
  hipLaunchKernelGGL(k_prolong, dim3((unsigned)((f.nodes() + 255) / 256), nrhs), dim3(256), 0, (hipStream_t)s, f,
                     c, mask_f, ec, xf, accumulate);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_prolong32(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_f, const float* ec,
                            float* xf, int accumulate, void* s) {
  return tg_prolong32_b(fine, coarse, mask_f, 1, ec, xf, accumulate, s);
}

extern "C" int tg_galerkin_elems(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_f, int use_mask,
                                 const double* fine_E, const double* fine_mats, double* out64, float* out32,
                                 void* s) {
  TG_NEED_INIT();
  if (int r = check_pair(fine, coarse)) return r;
  TG_REQUIRE(out64 && (fine_E || fine_mats) && (!use_mask || mask_f), "galerkin: NULL argument");
  const Grid f = to_grid(fine), c = to_grid(coarse);
  TG_REQUIRE(c.elems() < (1LL << 31), "too many coarse elements");
  hipLaunchKernelGGL(k_galerkin, dim3((unsigned)c.elems()), dim3(256), 0, (hipStream_t)s, f, c, mask_f, use_mask,
                     fine_E, fine_mats, out64, out32);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_assemble_dense(const tg_grid* gp, const double* mats, const int32_t* fidx, int32_t nf, double* A,
                                 void* s) {
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(mats && fidx && A && nf > 0, "assemble_dense: bad argument");
  const Grid g = to_grid(gp);
  if (int r = tg_fill64((int64_t)nf * nf, A, 0.0, s)) return r;
  const long long nd = 3 * g.nodes();
  hipLaunchKernelGGL(k_assemble_dense, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, (hipStream_t)s, g, mats,
                     fidx, nf, A);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

int dense_inverse_blocked(double* A, int n, hipStream_t st);
int dense_inverse_cholesky(double* A, int n, hipStream_t st);
// Above kBlockedMaxN the 16-wide pivot steps of the blocked inverse (each a
// pass over the whole matrix) cost more than rocSOLVER's Cholesky
// factorisation and inverse (potrf + potri, 45 ms at n = 7,260 on MI355X).
// Up to kRegisterMaxN one workgroup holds the whole matrix in registers (NB x
// NB FP64 entries per thread, NB = ceil(n / 32)). Each pivot step costs every
// thread O(NB^2) selects and FMAs whatever the step, so beyond NB = 2 the
// blocked form is faster: at n = 135, 286 us for NB = 5 against 255 us for
// the blocked form in four launches per pivot block, and 160 us for the
// same blocked steps in one workgroup with the matrix in LDS
// (k_spd_inverse_lds, up to kLdsInvMaxN).
constexpr int kRegisterMaxN = 64;

// The one-workgroup inverses (n <= kLdsInvMaxN), queued on st without a host
// wait: A inverted in place and left symmetric, *info = 0 or the first
// failing pivot + 1 (each kernel clears it first).
int queue_one_workgroup_inverse(double* A, int n, int* info, hipStream_t st) {
  const dim3 g1(1), b1(GJ_THREADS);
  if (n <= kRegisterMaxN) {
    // the matrix in registers: NB x NB entries per thread
    if (n <= 32)
      hipLaunchKernelGGL((k_spd_inverse_gj<1>), g1, b1, 0, st, A, n, n, A, n, info, 0, 1, 0LL, 0LL);
    else
      hipLaunchKernelGGL((k_spd_inverse_gj<2>), g1, b1, 0, st, A, n, n, A, n, info, 0, 1, 0LL, 0LL);
    TG_LAUNCH_CHECK();
    const long long nn = (long long)n * n;   // exact symmetry for the V-cycle
    hipLaunchKernelGGL(k_symmetrize_from_upper, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, A, n, 0LL);
  } else {
    // the matrix in LDS, mirrored from its upper triangle on the way out
    hipLaunchKernelGGL(k_spd_inverse_lds, g1, b1, 0, st, A, n, info);
  }
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_dense_spd_inverse(double* A, int32_t n, void* s) {
  TG_REQUIRE(A && n > 0, "dense_spd_inverse: bad argument");
  TG_REQUIRE(::tg::ke_ready(), "tg_init has not been called on this device");
  const hipStream_t st = (hipStream_t)s;
  if (n > kLdsInvMaxN) return n <= kBlockedMaxN ? dense_inverse_blocked(A, n, st) : dense_inverse_cholesky(A, n, st);
  int* info_d = workspace().flags;
  if (int r = queue_one_workgroup_inverse(A, n, info_d, st)) return r;
  int info = 0;
  TG_HIP_CHECK(hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, st));
  TG_HIP_CHECK(hipStreamSynchronize(st));
  if (info != 0) return set_error(TG_ERR_NOT_SPD, "dense inverse: matrix not SPD (pivot %d)", info);
  return TG_OK;
}

extern "C" int tg_dense_spd_inverse_async(double* A, int32_t n, int32_t* info, void* s) {
  TG_REQUIRE(A && info && n > 0, "dense_spd_inverse_async: bad argument");
  TG_REQUIRE(n <= kLdsInvMaxN, "dense_spd_inverse_async: n = %d above %d", n, kLdsInvMaxN);
  return queue_one_workgroup_inverse(A, n, info, (hipStream_t)s);
}

// Blocked Gauss-Jordan with BK-wide pivot blocks on a batch of matrices at
// stride sa (the same pivot steps on every matrix, one launch per kernel and
// step); P (the inverted pivot blocks, BK x BK per matrix) lives in the
// reduction workspace, which no other kernel uses meanwhile on this stream.
// Leaves each inverse symmetric (mirrored from its upper triangle); a pivot
// <= 0 sets info (cleared by the caller).
namespace tg {
int spd_inverse_batched(double* A, int n, int batch, long long sa, hipStream_t st) {
  TG_REQUIRE(n >= 1 && n <= kBlockedMaxN && batch >= 1 && (long long)batch * BK * BK <= kMaxPartials,
             "batched inverse: bad shape");
  int* info_d = workspace().flags;
  double* P = workspace().partials;
  const unsigned nt = (unsigned)((n + TS - 1) / TS), nb = (unsigned)batch;
  for (int k0 = 0; k0 < n; k0 += BK) {
    const int kw = n - k0 < BK ? n - k0 : BK;
    hipLaunchKernelGGL((k_spd_inverse_gj<1>), dim3(nb), dim3(GJ_THREADS), 0, st, A + (long long)k0 * n + k0, n, kw,
                       P, BK, info_d, k0, 0, sa, (long long)BK * BK);
    TG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_row_panel, dim3(nt, 1, nb), dim3(256), 0, st, A, n, k0, kw, P, sa);
    TG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_trailing, dim3(nt, nt, nb), dim3(256), 0, st, A, n, k0, kw, sa);
    TG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_col_panel, dim3(nt, 1, nb), dim3(256), 0, st, A, n, k0, kw, P, sa);
    TG_LAUNCH_CHECK();
  }
  const long long nn = (long long)n * n;   // exact symmetry
  hipLaunchKernelGGL(k_symmetrize_from_upper, dim3((unsigned)((nn + 255) / 256), nb), dim3(256), 0, st, A, n, sa);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
}  // namespace tg

int dense_inverse_blocked(double* A, int n, hipStream_t st) {
  int* info_d = workspace().flags;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, info_d, 0);
  TG_LAUNCH_CHECK();
  if (int r = spd_inverse_batched(A, n, 1, 0, st)) return r;
  int info = 0;
  TG_HIP_CHECK(hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, st));
  TG_HIP_CHECK(hipStreamSynchronize(st));
  if (info != 0) return set_error(TG_ERR_NOT_SPD, "dense inverse: matrix not SPD (pivot %d)", info);
  return TG_OK;
}

// rocSOLVER Cholesky factorisation and inverse. A row-major symmetric A is
// its own column-major transpose: work on the column-major lower triangle
// (the row-major upper triangle), then mirror it.
int dense_inverse_cholesky(double* A, int n, hipStream_t st) {
  rocblas_handle h = nullptr;
  if (int r = get_handle(st, &h)) return r;
  rocblas_int* dinfo = reinterpret_cast<rocblas_int*>(workspace().flags);
  rocblas_int info = 0;
  rocblas_status rs = rocsolver_dpotrf(h, rocblas_fill_lower, n, A, n, dinfo);
  if (rs == rocblas_status_success) {
    TG_HIP_CHECK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
    TG_HIP_CHECK(hipStreamSynchronize(st));
  }
  if (rs == rocblas_status_success && info == 0) {
    rs = rocsolver_dpotri(h, rocblas_fill_lower, n, A, n, dinfo);
    if (rs == rocblas_status_success) {
      TG_HIP_CHECK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
      TG_HIP_CHECK(hipStreamSynchronize(st));
    }
  }
  if (rs != rocblas_status_success) return set_error(TG_ERR_HIP, "rocsolver potrf/potri status %d", (int)rs);
  if (info != 0) return set_error(TG_ERR_NOT_SPD, "dense inverse: matrix not SPD (Cholesky info=%d)", (int)info);
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(k_symmetrize_from_upper, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, A, n, 0LL);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_coarsest_apply_b(const double* Ainv, int32_t n, const int32_t* fl, int64_t n_full, int32_t nrhs,
                                   const float* rc, float* ec, void* s) {
  TG_REQUIRE(Ainv && fl && rc && ec && n > 0 && n_full >= n && nrhs >= 1 && nrhs <= 64,
             "coarsest_apply: bad argument");
  const long long blocks = (n + 3) / 4 + (n_full + 255) / 256;
  TG_REQUIRE(blocks < (1LL << 31), "coarsest_apply: too large");
  hipLaunchKernelGGL(k_coarsest_apply, dim3((unsigned)blocks, nrhs), dim3(256), 0, (hipStream_t)s, Ainv, n, fl,
                     (long long)n_full, rc, ec);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_coarsest_apply(const double* Ainv, int32_t n, const int32_t* fl, int64_t n_full, const float* rc,
                                 float* ec, void* s) {
  return tg_coarsest_apply_b(Ainv, n, fl, n_full, 1, rc, ec, s);
}

extern "C" int tg_galerkin_from_fine(const tg_grid* fine, int down, const double* fine_E, double* work,
                                     double* out64, float* out32, void* s) {
  TG_NEED_INIT();
  if (int r = check_grid(fine)) return r;
  TG_REQUIRE(down == 1 || down == 2, "levels_down must be 1 or 2");
  TG_REQUIRE(fine_E && work && out64, "galerkin_from_fine: NULL argument");
  const int f2 = 1 << down;
  TG_REQUIRE(fine->nelx % f2 == 0 && fine->nely % f2 == 0 && fine->nelz % f2 == 0,
             "galerkin_from_fine: fine dims must be divisible by %d", f2);
  const Grid f = to_grid(fine);
  const Grid c{f.nx / f2, f.ny / f2, f.nz / f2};
  GTables* tab = nullptr;
  if (int r = ensure_tables(&tab)) return r;
  const int m = down == 1 ? 8 : 64;
  const long long nw = c.elems() * m;
  hipLaunchKernelGGL(k_gather_desc, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, (hipStream_t)s, c, f, down,
                     fine_E, work);
  TG_LAUNCH_CHECK();
  // Row-major K (nc x 576) = Ew (nc x m) * T (m x 576); in column-major terms
  // K^T = T^T * Ew^T with the row-major arrays read as their transposes.
  rocblas_handle h = nullptr;
  if (int r = get_handle((hipStream_t)s, &h)) return r;
  const double one = 1.0, zero = 0.0;
  TG_REQUIRE(c.elems() < (1LL << 31), "too many coarse elements");
  const rocblas_status st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, 576, (rocblas_int)c.elems(),
                                          m, &one, down == 1 ? tab->q8 : tab->r64, 576, work, m, &zero, out64, 576);
  if (st != rocblas_status_success) return set_error(TG_ERR_HIP, "rocblas_dgemm status %d", (int)st);
  if (out32) {
    const long long n = 576 * c.elems();
    const long long nb = (n + 255) / 256 > 4096 ? 4096 : (n + 255) / 256;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)s, n, out64, out32);
    TG_LAUNCH_CHECK();
  }
  return TG_OK;
}
