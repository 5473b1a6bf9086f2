// Coarsest level as a block-tridiagonal system over x-planes, factored by
// block cyclic reduction (FP64).
//
// With the node order (i, j, k) -> (i*NY + j)*NZ + k, the DOFs of x-plane i
// are one contiguous range of m = 3*NY*NZ, and the 27-point operator couples
// plane i only to planes i-1, i, i+1:
//     A_i x_i + C_i x_{i+1} + C_{i-1}^T x_{i-1} = b_i.
// Fixed DOFs keep an identity row and column (their right-hand side is zero,
// so their solution is zero), which leaves every block SPD.
//
// One reduction level eliminates the odd positions of the active plane list
// (stride s = 2^level) in favour of the even ones. For odd plane o between
// even planes L and R, with U = C(L,o), V = C(o,R), Ai = A_o^-1:
//     Y = Ai U^T, X = Ai V (and the transposes YT = U Ai, XT = V^T Ai),
//     A_L -= U Y,  A_R -= V^T X,  C(L,R) = -U X.
// The last active plane (plane 0) is inverted. This is an exact block
// Cholesky factorisation in odd-even plane order: the solve it gives is the
// dense Cholesky solve of multigrid_v4.py:781-801, 977-978 up to rounding.
// Setup per SIMP iteration is a handful of batched inverses of m x m plane
// blocks and batched DGEMMs per level (log2 of the plane count levels),
// instead of the n^3 dense factorisation and inverse of the whole coarsest
// matrix; a solve is one GEMV kernel per level and direction:
//   forward  b_L -= Y^T b_o, b_R -= X^T b_o   (even planes, level by level)
//   final    x_0 = A_0^-1 b_0
//   backward x_o = Ai b_o - Y x_L - X x_R     (odd planes, top level first)
// Every block is column-major m x m; symmetric blocks are read either way.
#include "common.h"
#include "dense.h"
#include "reduce.h"

namespace tg {
namespace {

constexpr int kMaxBcrLevels = 32;

struct BcrLevel {
  int s, K, n_odd, n_x;              // stride, active planes, odd positions, odd ones with a right neighbour
  long long C, X, XT, Y, YT, Cnext;  // offsets (doubles) into the work buffer
};

struct BcrPlan {
  int P, m, nlev;
  long long mm, A, total;
  BcrLevel lv[kMaxBcrLevels];
};

int make_plan(const Grid& g, BcrPlan& p) {
  p.P = g.NX();
  p.m = 3 * g.NY() * g.NZ();
  TG_REQUIRE(p.m <= kBlockedMaxN, "bcr: plane of %d DOFs exceeds %d", p.m, kBlockedMaxN);
  p.mm = (long long)p.m * p.m;
  long long off = 0;
  p.A = off;
  off += p.P * p.mm;
  long long C = off;
  off += (p.P - 1) * p.mm;
  int s = 1, K = p.P, l = 0;
  while (K > 1) {
    TG_REQUIRE(l < kMaxBcrLevels, "bcr: too many levels");
    BcrLevel& v = p.lv[l];
    v.s = s;
    v.K = K;
    v.n_odd = K / 2;
    v.n_x = (K - 1) / 2;
    v.C = C;
    v.X = off;
    off += v.n_x * p.mm;
    v.XT = off;
    off += v.n_x * p.mm;
    v.Y = off;
    off += v.n_odd * p.mm;
    v.YT = off;
    off += v.n_odd * p.mm;
    v.Cnext = off;
    const int Kn = (K + 1) / 2;
    off += (Kn - 1) * p.mm;
    C = v.Cnext;
    s *= 2;
    K = Kn;
    ++l;
  }
  p.nlev = l;
  p.total = off;
  return TG_OK;
}

// Block assembly from the element matrices (576 FP64 per element, row-major
// 24 x 24): one thread per DOF row writes its entries of A_i (same plane) and
// of C_i (plane i+1); entries reaching plane i-1 belong to C_{i-1}, written
// by that plane's rows. A and C are zero on entry.
__global__ void __launch_bounds__(256) k_bcr_assemble(Grid g, const uint8_t* __restrict__ mask,
                                                      const double* __restrict__ mats, int m, double* A,
                                                      double* C) {
  const long long d = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (d >= 3 * g.nodes()) return;
  const long long n = d / 3;
  const int r = (int)(d % 3);
  const int k = (int)(n % g.NZ()), j = (int)((n / g.NZ()) % g.NY()), i = (int)(n / ((long long)g.NZ() * g.NY()));
  const long long mm = (long long)m * m;
  const int dl = (int)(d - (long long)i * m);
  if (!(mask[n] >> r & 1u)) {
    A[i * mm + dl + (long long)dl * m] = 1.0;
    return;
  }
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
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int dx = q / 9 - 1, b = j + (q / 3) % 3 - 1, e = k + q % 3 - 1;
    if (dx < 0 || i + dx >= g.NX() || b < 0 || b >= g.NY() || e < 0 || e >= g.NZ()) continue;
    const unsigned mc = mask[g.node(i + dx, b, e)];
    double* blk = dx == 0 ? A + i * mm : C + i * mm;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!(mc >> c & 1u)) continue;
      const int dlc = 3 * (b * g.NZ() + e) + c;
      blk[dl + (long long)dlc * m] = acc[q][c];
    }
  }
}

constexpr int BW = 4;   // wavefronts (= block rows) per workgroup of the solve kernels

// Forward step of one level: even position t (plane 2ts) takes
// b_e -= Y_t^T b_{o_t} + X_{t-1}^T b_{o_{t-1}}, one wavefront per row. The
// first level reads b from the FP32 right-hand side; later ones from w.
template <bool FIRST>
__global__ void __launch_bounds__(64 * BW) k_bcr_forward(int m, int s, int n_odd, int n_even, long long nfull,
                                                         const float* __restrict__ rc, double* vec,
                                                         const double* __restrict__ Y,
                                                         const double* __restrict__ X) {
  const int rb = (m + BW - 1) / BW;
  const int t = blockIdx.x / rb, r = (blockIdx.x % rb) * BW + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (t >= n_even || r >= m) return;   // uniform per wavefront
  rc += blockIdx.y * nfull;
  double* w = vec + blockIdx.y * 2 * nfull;
  const long long mm = (long long)m * m;
  auto src = [&](long long i) -> double { return FIRST ? (double)rc[i] : w[i]; };
  double acc = 0.0;
  if (t < n_odd) {
    const long long o = (long long)(2 * t + 1) * s * m;
    const double* col = Y + t * mm + (long long)r * m;
    for (int c = lane; c < m; c += 64) acc = fma(col[c], src(o + c), acc);
  }
  if (t >= 1) {   // odd t-1 has this plane as its right neighbour
    const long long o = (long long)(2 * t - 1) * s * m;
    const double* col = X + (t - 1) * mm + (long long)r * m;
    for (int c = lane; c < m; c += 64) acc = fma(col[c], src(o + c), acc);
  }
  acc = wave_sum(acc);
  const long long e = (long long)2 * t * s * m + r;
  if (lane == 0) w[e] = src(e) - acc;
}

// Plane 0: x_0 = A_0^-1 b_0 (b staged in LDS), also written as FP32 to ec.
template <bool FIRST>
__global__ void __launch_bounds__(1024) k_bcr_final(int m, long long nfull, const float* __restrict__ rc, double* vec,
                                                   const double* __restrict__ Ai, float* __restrict__ ec) {
  __shared__ double b[kBlockedMaxN];
  rc += blockIdx.y * nfull;
  ec += blockIdx.y * nfull;
  const double* w = vec + blockIdx.y * 2 * nfull;
  double* x = vec + blockIdx.y * 2 * nfull + nfull;
  for (int i = threadIdx.x; i < m; i += blockDim.x) b[i] = FIRST ? (double)rc[i] : w[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int r = wv; r < m; r += nw) {
    const double* row = Ai + (long long)r * m;   // symmetric: row r = column r
    double acc = 0.0;
    for (int c = lane; c < m; c += 64) acc = fma(row[c], b[c], acc);
    acc = wave_sum(acc);
    if (lane == 0) {
      x[r] = acc;
      ec[r] = (float)acc;
    }
  }
}

// Backward step of one level: odd position u (plane (2u+1)s) solves
// x_o = Ai_u b_o - (Y_u x_L + X_u x_R), one wavefront per row.
template <bool FIRST>
__global__ void __launch_bounds__(64 * BW) k_bcr_backward(int m, int s, int n_odd, int n_x, long long nfull,
                                                          const float* __restrict__ rc, double* vec,
                                                          const double* __restrict__ Ai, long long ai_stride,
                                                          const double* __restrict__ YT,
                                                          const double* __restrict__ XT, float* __restrict__ ec) {
  const int rb = (m + BW - 1) / BW;
  const int u = blockIdx.x / rb, r = (blockIdx.x % rb) * BW + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (u >= n_odd || r >= m) return;   // uniform per wavefront
  rc += blockIdx.y * nfull;
  ec += blockIdx.y * nfull;
  const double* w = vec + blockIdx.y * 2 * nfull;
  double* x = vec + blockIdx.y * 2 * nfull + nfull;
  const long long mm = (long long)m * m;
  const long long o = (long long)(2 * u + 1) * s * m, L = (long long)2 * u * s * m, R = (long long)(2 * u + 2) * s * m;
  const double* arow = Ai + u * ai_stride + (long long)r * m;
  double a = 0.0, c2 = 0.0;
  for (int c = lane; c < m; c += 64) a = fma(arow[c], FIRST ? (double)rc[o + c] : w[o + c], a);
  const double* yrow = YT + u * mm + (long long)r * m;
  for (int c = lane; c < m; c += 64) c2 = fma(yrow[c], x[L + c], c2);
  if (u < n_x) {
    const double* xrow = XT + u * mm + (long long)r * m;
    for (int c = lane; c < m; c += 64) c2 = fma(xrow[c], x[R + c], c2);
  }
  a = wave_sum(a);
  c2 = wave_sum(c2);
  if (lane == 0) {
    const double v = a - c2;
    x[o + r] = v;
    ec[o + r] = (float)v;
  }
}

}  // namespace
}  // namespace tg

using namespace tg;

#define TG_NEED_INIT() TG_REQUIRE(::tg::ke_ready(), "tg_init has not been called on this device")

extern "C" int tg_bcr_size(const tg_grid* gp, int64_t* n_work) {
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(n_work, "bcr_size: NULL argument");
  BcrPlan p;
  if (int r = make_plan(to_grid(gp), p)) return r;
  *n_work = p.total;
  return TG_OK;
}

extern "C" int tg_bcr_setup(const tg_grid* gp, const uint8_t* mask, const double* mats64, double* work, void* s) {
  TG_NEED_INIT();
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(mask && mats64 && work, "bcr_setup: NULL argument");
  const Grid g = to_grid(gp);
  BcrPlan p;
  if (int r = make_plan(g, p)) return r;
  const hipStream_t st = (hipStream_t)s;
  const int m = p.m;
  const long long mm = p.mm;
  double* A = work + p.A;
  if (int r = tg_fill64((2LL * p.P - 1) * mm, A, 0.0, s)) return r;
  const long long nd = 3 * g.nodes();
  hipLaunchKernelGGL(k_bcr_assemble, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, g, mask, mats64, m, A,
                     A + p.P * mm);   // C_0 follows the P diagonal blocks
  TG_LAUNCH_CHECK();
  int* info_d = workspace().flags;
  TG_HIP_CHECK(hipMemsetAsync(info_d, 0, sizeof(int), st));
  rocblas_handle h = nullptr;
  if (int r = get_handle(st, &h)) return r;
  const double one = 1.0, mone = -1.0, zero = 0.0;
  const auto N = rocblas_operation_none, T = rocblas_operation_transpose;
  auto gemm = [&](rocblas_operation ta, rocblas_operation tb, const double* alpha, const double* Am, long long sa,
                  const double* Bm, long long sb, const double* beta, double* Cm, long long sc, int batch) -> int {
    if (batch <= 0) return TG_OK;
    const rocblas_status rs = rocblas_dgemm_strided_batched(h, ta, tb, m, m, m, alpha, Am, m, sa, Bm, m, sb, beta,
                                                            Cm, m, sc, batch);
    if (rs != rocblas_status_success) return set_error(TG_ERR_HIP, "bcr: rocblas_dgemm_strided_batched status %d", (int)rs);
    return TG_OK;
  };
  for (int l = 0; l < p.nlev; ++l) {
    const BcrLevel& v = p.lv[l];
    double* Ai = A + v.s * mm;
    const long long sai = 2LL * v.s * mm;
    const double* U = work + v.C;
    const double* V = work + v.C + mm;
    if (int r = spd_inverse_batched(Ai, m, v.n_odd, sai, st)) return r;
    if (int r = gemm(N, T, &one, Ai, sai, U, 2 * mm, &zero, work + v.Y, mm, v.n_odd)) return r;
    if (int r = gemm(N, N, &one, U, 2 * mm, Ai, sai, &zero, work + v.YT, mm, v.n_odd)) return r;
    if (int r = gemm(N, N, &one, Ai, sai, V, 2 * mm, &zero, work + v.X, mm, v.n_x)) return r;
    if (int r = gemm(T, N, &one, V, 2 * mm, Ai, sai, &zero, work + v.XT, mm, v.n_x)) return r;
    if (int r = gemm(N, N, &mone, U, 2 * mm, work + v.Y, mm, &one, A, sai, v.n_odd)) return r;
    if (int r = gemm(T, N, &mone, V, 2 * mm, work + v.X, mm, &one, A + 2LL * v.s * mm, sai, v.n_x)) return r;
    if (int r = gemm(N, N, &mone, U, 2 * mm, work + v.X, mm, &zero, work + v.Cnext, mm, v.n_x)) return r;
  }
  if (int r = spd_inverse_batched(A, m, 1, 0, st)) return r;
  int info = 0;
  TG_HIP_CHECK(hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, st));
  TG_HIP_CHECK(hipStreamSynchronize(st));
  if (info != 0) return set_error(TG_ERR_NOT_SPD, "bcr: plane block not SPD (pivot %d)", info);
  return TG_OK;
}

extern "C" int tg_bcr_apply_b(const tg_grid* gp, const double* work, int32_t nrhs, const float* rc, float* ec,
                              double* vec, void* s) {
  if (int r = check_grid(gp)) return r;
  TG_REQUIRE(work && rc && ec && vec && nrhs >= 1 && nrhs <= 64, "bcr_apply: bad argument");
  const Grid g = to_grid(gp);
  BcrPlan p;
  if (int r = make_plan(g, p)) return r;
  const hipStream_t st = (hipStream_t)s;
  const int m = p.m;
  const long long mm = p.mm, nfull = (long long)p.P * m;
  const unsigned rb = (unsigned)((m + BW - 1) / BW);
  for (int l = 0; l < p.nlev; ++l) {
    const BcrLevel& v = p.lv[l];
    const int n_even = (v.K + 1) / 2;
    const dim3 grid(rb * n_even, nrhs);
    if (l == 0)
      hipLaunchKernelGGL(k_bcr_forward<true>, grid, dim3(64 * BW), 0, st, m, v.s, v.n_odd, n_even, nfull, rc, vec,
                         work + v.Y, work + v.X);
    else
      hipLaunchKernelGGL(k_bcr_forward<false>, grid, dim3(64 * BW), 0, st, m, v.s, v.n_odd, n_even, nfull, rc, vec,
                         work + v.Y, work + v.X);
    TG_LAUNCH_CHECK();
  }
  if (p.nlev == 0)
    hipLaunchKernelGGL(k_bcr_final<true>, dim3(1, nrhs), dim3(1024), 0, st, m, nfull, rc, vec, work + p.A, ec);
  else
    hipLaunchKernelGGL(k_bcr_final<false>, dim3(1, nrhs), dim3(1024), 0, st, m, nfull, rc, vec, work + p.A, ec);
  TG_LAUNCH_CHECK();
  for (int l = p.nlev - 1; l >= 0; --l) {
    const BcrLevel& v = p.lv[l];
    const dim3 grid(rb * v.n_odd, nrhs);
    const double* Ai = work + p.A + v.s * mm;
    const long long sai = 2LL * v.s * mm;
    if (l == 0)
      hipLaunchKernelGGL(k_bcr_backward<true>, grid, dim3(64 * BW), 0, st, m, v.s, v.n_odd, v.n_x, nfull, rc, vec,
                         Ai, sai, work + v.YT, work + v.XT, ec);
    else
      hipLaunchKernelGGL(k_bcr_backward<false>, grid, dim3(64 * BW), 0, st, m, v.s, v.n_odd, v.n_x, nfull, rc, vec,
                         Ai, sai, work + v.YT, work + v.XT, ec);
    TG_LAUNCH_CHECK();
  }
  return TG_OK;
}
