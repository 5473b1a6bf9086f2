// Library state, constant tables and Krylov vector kernels.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "common.h"
#include "mfma_layout.h"
#include "reduce.h"

namespace tg {

__constant__ double c_ke64[576];
__constant__ float c_ke32[576];
__constant__ float c_q8_32[8 * 576];
// BF16 bit patterns of the unit stiffness in MFMA A-operand order, see
// mfma_layout.h: [output layout][hi/lo][M tile][lane][8].
__constant__ uint16_t c_ke_bf16[2 * 2 * 2 * 64 * 8];
// The eight Galerkin child matrices Q_s in the same A-operand order:
// [child s][output layout][hi/lo][M tile][lane][8] (global memory: 128 KiB).
__device__ uint16_t c_q8_bf16[8 * 2 * 2 * 2 * 64 * 8];

namespace {
thread_local char g_err[512] = "";
constexpr int kMaxDevices = 64;
bool g_ready[kMaxDevices] = {};
Workspace g_ws[kMaxDevices] = {};

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return -1;
  return d;
}

uint16_t bf16_bits_rne(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);   // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_to_float(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
}  // namespace

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

bool ke_ready() {
  const int d = current_device();
  return d >= 0 && d < kMaxDevices && g_ready[d];
}

double g_ke_host[576];
const double* ke_host() { return g_ke_host; }

Workspace workspace() {
  const int d = current_device();
  return (d >= 0 && d < kMaxDevices) ? g_ws[d] : Workspace{nullptr, nullptr};
}

int check_grid(const tg_grid* g) {
  TG_REQUIRE(g != nullptr, "grid is NULL");
  TG_REQUIRE(g->nelx >= 1 && g->nely >= 1 && g->nelz >= 1, "grid dims must be >= 1 (got %d %d %d)",
             g->nelx, g->nely, g->nelz);
  const double nodes = (double)(g->nelx + 1) * (g->nely + 1) * (g->nelz + 1);
  TG_REQUIRE(3.0 * nodes < 2147483647.0 * 64.0, "grid too large");
  return TG_OK;
}

// ---------------------------------------------------------------- kernels

namespace {

constexpr int kVecBlock = 256;

int vec_blocks(long long n) {
  long long b = (n + kVecBlock * 4 - 1) / (kVecBlock * 4);
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}
#ifndef TG_RED_BLOCKS
#define TG_RED_BLOCKS 1024
#endif
// Reducing kernels: fewer, fuller blocks, so the one-block finalize sums
// fewer partials.
int red_blocks(long long n) {
  const int b = vec_blocks(n);
  return b < TG_RED_BLOCKS ? b : TG_RED_BLOCKS;
}

__global__ void k_copy32(long long n, float* __restrict__ dst, const float* __restrict__ src) {
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill32(long long n, float* x, float v) {
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}
__global__ void k_fill64(long long n, double* x, double v) {
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}

// Batched vector kernels: blockIdx.y selects the right-hand side; each RHS
// is a contiguous slice of n entries (slice r starts at r * n). Per-RHS
// scalars are arrays indexed by r. The grid-stride loops are unrolled (a
// thread has about four elements) so that their loads are in flight
// together.
// The FP32 vector kernels below give each thread groups of four consecutive
// entries counted from the start of the slice (groups t, t + S, ... for
// thread t of S), each group loaded as one float4 when every vector of the
// slice starts on a 16-byte boundary -- four times the bytes in flight per
// load of one float -- and as four floats otherwise (a right-hand side at an
// odd offset); the last group of n % 4 entries goes entry by entry. The
// entries of a thread and their order thus depend only on n and the launch
// shape, not on alignment: a slice of a batch gives the bits of the same
// slice alone.
__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

__device__ __forceinline__ float4 ld4(const float* p, long long g, bool vec) {
  if (vec) return reinterpret_cast<const float4*>(p)[g];
  const float* q = p + 4 * g;
  return make_float4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ void st4(float* p, long long g, bool vec, float4 v) {
  if (vec) {
    reinterpret_cast<float4*>(p)[g] = v;
  } else {
    float* q = p + 4 * g;
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
    q[3] = v.w;
  }
}

__global__ void __launch_bounds__(kVecBlock) k_dot32(long long n, const float* a, const float* b,
                                                    const float* c, Workspace ws) {
  const long long o = blockIdx.y * n;
  a += o;
  b += o;
  if (c) c += o;
  double v[2] = {0.0, 0.0};
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
  const bool vec = aligned16(a) && aligned16(b) && (!c || aligned16(c));
  const long long n4 = n >> 2;
  #pragma unroll 2
  for (long long g = t0; g < n4; g += stride) {
    const float4 x = ld4(a, g, vec), y = ld4(b, g, vec);
    const double ax = x.x, ay = x.y, az = x.z, aw = x.w;
    v[0] += ax * (double)y.x;
    v[0] += ay * (double)y.y;
    v[0] += az * (double)y.z;
    v[0] += aw * (double)y.w;
    if (c) {
      const float4 z = ld4(c, g, vec);
      v[1] += ax * (double)z.x;
      v[1] += ay * (double)z.y;
      v[1] += az * (double)z.z;
      v[1] += aw * (double)z.w;
    }
  }
  if (n4 % stride == t0)   // the thread whose next group is the partial one
    for (long long i = 4 * n4; i < n; ++i) {
      const double ai = a[i];
      v[0] += ai * (double)b[i];
      if (c) v[1] += ai * (double)c[i];
    }
  block_partials<2>(v, ws);
}

__global__ void __launch_bounds__(kVecBlock) k_dot64(long long n, const double* a, const double* b,
                                                    Workspace ws) {
  const long long o = blockIdx.y * n;
  a += o;
  b += o;
  double v[1] = {0.0};
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    v[0] = fma(a[i], b[i], v[0]);
  block_partials<1>(v, ws);
}

// Block r: out_k[r] = sum_b partials[k * nseg * seg + r * seg + b] in block
// order (the layout block_partials<K> writes for a grid of nseg slices of seg
// blocks each).
// Each thread adds its blocks b = tid, tid + 256, ... in that order; the
// loads go out eight at a time (the partials were just written by other
// XCDs, so every load is a trip to memory), padded with zeros past seg.
__device__ __forceinline__ double strided_sum(const double* pk, int seg) {
  double v = 0.0;
  for (int b0 = threadIdx.x; b0 < seg; b0 += 8 * blockDim.x) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = b0 + u * (int)blockDim.x;
      x[u] = b < seg ? pk[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) v += x[u];
  }
  return v;
}

__global__ void __launch_bounds__(256) k_finalize(int K, int seg, const double* partials, double* out0, double* out1) {
  const int r = blockIdx.x, nseg = gridDim.x;
  const double* p0 = partials + (size_t)r * seg;
  double v[2] = {strided_sum(p0, seg), K > 1 ? strided_sum(p0 + (size_t)nseg * seg, seg) : 0.0};
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    if (out0) out0[r] = v[0];
    if (K > 1 && out1) out1[r] = v[1];
  }
}

__global__ void __launch_bounds__(kVecBlock) k_pcg_xr(long long n, float* x, float* r, const float* p,
                                                     const float* q, const double* rz,
                                                     const double* pq, Workspace ws) {
  const int k = blockIdx.y;
  const long long o = k * n;
  x += o;
  r += o;
  p += o;
  q += o;
  const double den = pq[k];
  const float alpha = (float)(den != 0.0 ? rz[k] / den : 0.0);
  double v[1] = {0.0};
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
  const bool vec = aligned16(x) && aligned16(r) && aligned16(p) && aligned16(q);
  const long long n4 = n >> 2;
  #pragma unroll 2
  for (long long g = t0; g < n4; g += stride) {
    const float4 pi = ld4(p, g, vec), qi = ld4(q, g, vec), xi = ld4(x, g, vec), r0 = ld4(r, g, vec);
    st4(x, g, vec, make_float4(fmaf(alpha, pi.x, xi.x), fmaf(alpha, pi.y, xi.y), fmaf(alpha, pi.z, xi.z),
                               fmaf(alpha, pi.w, xi.w)));
    const float4 ri = make_float4(fmaf(-alpha, qi.x, r0.x), fmaf(-alpha, qi.y, r0.y), fmaf(-alpha, qi.z, r0.z),
                                  fmaf(-alpha, qi.w, r0.w));
    st4(r, g, vec, ri);
    v[0] += (double)ri.x * (double)ri.x;
    v[0] += (double)ri.y * (double)ri.y;
    v[0] += (double)ri.z * (double)ri.z;
    v[0] += (double)ri.w * (double)ri.w;
  }
  if (n4 % stride == t0)
    for (long long i = 4 * n4; i < n; ++i) {
      x[i] = fmaf(alpha, p[i], x[i]);
      const float ri = fmaf(-alpha, q[i], r[i]);
      r[i] = ri;
      v[0] += (double)ri * (double)ri;
    }
  block_partials<1>(v, ws);
}

__global__ void k_pcg_dir(long long n, float* p, const float* z, const double* zq, const double* pq) {
  const int k = blockIdx.y;
  const long long o = k * n;
  p += o;
  z += o;
  const double den = pq[k];
  const float beta = (float)(den != 0.0 ? -zq[k] / den : 0.0);
  const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
  const bool vec = aligned16(p) && aligned16(z);
  const long long n4 = n >> 2;
  #pragma unroll 2
  for (long long g = t0; g < n4; g += stride) {
    const float4 pi = ld4(p, g, vec), zi = ld4(z, g, vec);
    st4(p, g, vec, make_float4(fmaf(beta, pi.x, zi.x), fmaf(beta, pi.y, zi.y), fmaf(beta, pi.z, zi.z),
                               fmaf(beta, pi.w, zi.w)));
  }
  if (n4 % stride == t0)
    for (long long i = 4 * n4; i < n; ++i) p[i] = fmaf(beta, p[i], z[i]);
}

// dinv is shared by all right-hand sides (one operator).
__global__ void k_cheb_init(long long n, const float* rhs, const float* dinv, float a, float* d, float* x, float* r) {
  const long long o = blockIdx.y * n;
  rhs += o;
  d += o;
  x += o;
  if (r) r += o;
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = a * dinv[i] * rhs[i];
    d[i] = v;
    x[i] = v;
    if (r) r[i] = rhs[i];
  }
}

// REDUCE = false: the steps before the last, whose v.Dv nobody reads.
template <bool REDUCE>
__global__ void __launch_bounds__(kVecBlock) k_power_step(long long n, float* v, const float* kv, const float* dinv,
                                                         Workspace ws) {
  double acc[1] = {0.0};
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (REDUCE) {
      const double vi = v[i];
      acc[0] += vi * vi / (double)dinv[i];
    }
    v[i] = dinv[i] * kv[i];
  }
  if (REDUCE) block_partials<1>(acc, ws);
}

__global__ void k_axpy_f32_to_f64(long long n, double* x, const float* d) {
  #pragma unroll 4
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    x[i] += (double)d[i];
}

// y = (accumulate ? y : 0) + a * x, a = scale * (alpha ? *alpha : 1), x read
// from x64 or, when x64 is NULL, from x32. alpha is a device scalar written
// by an earlier kernel on the stream (a Gram-Schmidt coefficient).
__global__ void k_axpy64(long long n, double* __restrict__ y, const double* __restrict__ x64,
                         const float* __restrict__ x32, const double* __restrict__ alpha, double scale,
                         int accumulate) {
  const double a = alpha ? scale * *alpha : scale;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double xv = x64 ? x64[i] : (double)x32[i];
    y[i] = accumulate ? y[i] + a * xv : a * xv;
  }
}

// y = x / sqrt(*ss) (0 when *ss <= 0), y32 = float(y).
__global__ void k_normalize64(long long n, const double* __restrict__ x, const double* __restrict__ ss,
                              double* __restrict__ y, float* __restrict__ y32) {
  const double v = *ss;
  const double h = v > 0.0 ? sqrt(v) : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double q = h > 0.0 ? x[i] / h : 0.0;
    if (y) y[i] = q;
    if (y32) y32[i] = (float)q;
  }
}

}  // namespace
}  // namespace tg

using namespace tg;

// ---------------------------------------------------------------- C ABI

extern "C" int tg_axpy64(int64_t n, double* y, const double* x64, const float* x32, const double* alpha,
                         double scale, int32_t accumulate, void* s) {
  TG_REQUIRE(n > 0 && y && (x64 || x32), "axpy64: bad argument");
  hipLaunchKernelGGL(k_axpy64, dim3(vec_blocks(n)), dim3(kVecBlock), 0, (hipStream_t)s, (long long)n, y, x64, x32,
                     alpha, scale, (int)accumulate);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_normalize64(int64_t n, const double* x, const double* ss, double* y, float* y32, void* s) {
  TG_REQUIRE(n > 0 && x && ss && (y || y32), "normalize64: bad argument");
  hipLaunchKernelGGL(k_normalize64, dim3(vec_blocks(n)), dim3(kVecBlock), 0, (hipStream_t)s, (long long)n, x, ss, y,
                     y32);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" const char* tg_last_error(void) { return g_err; }
extern "C" int tg_version(void) { return 1; }

extern "C" int tg_mapped_pointer(const void* host, void** dev) {
  TG_REQUIRE(host && dev, "bad mapped_pointer args");
  void* d = nullptr;
  const hipError_t e = hipHostGetDevicePointer(&d, const_cast<void*>(host), 0);
  if (e != hipSuccess || !d) {
    (void)hipGetLastError();
    return ::tg::set_error(TG_ERR_ARG, "not pinned host memory mapped into the device: %s", hipGetErrorString(e));
  }
  *dev = d;
  return TG_OK;
}

extern "C" int tg_init(const double* ke) {
  TG_REQUIRE(ke != nullptr, "ke_unit_host is NULL");
  const int dev = current_device();
  TG_REQUIRE(dev >= 0 && dev < kMaxDevices, "no current HIP device");
  float ke32[576];
  for (int i = 0; i < 576; ++i) ke32[i] = (float)ke[i];

  // Galerkin child matrices Q_s = P_s^T KE P_s, child s = (sx,sy,sz) at
  // index 4*sx + 2*sy + sz. Fine local position t = s + corner offset in
  // {0,1,2} takes weights (1,0), (1/2,1/2), (0,1) from coarse nodes (0,1).
  static const double w1d[3][2] = {{1.0, 0.0}, {0.5, 0.5}, {0.0, 1.0}};
  float q8[8 * 576];
  for (int s = 0; s < 8; ++s) {
    const int sx = s >> 2 & 1, sy = s >> 1 & 1, sz = s & 1;
    double P[8][8];
    for (int c = 0; c < 8; ++c)
      for (int C = 0; C < 8; ++C)
        P[c][C] = w1d[sx + corner_x(c)][corner_x(C)] * w1d[sy + corner_y(c)][corner_y(C)] *
                  w1d[sz + corner_z(c)][corner_z(C)];
    for (int A = 0; A < 24; ++A)
      for (int B = 0; B < 24; ++B) {
        double acc = 0.0;
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) {
            const double pa = P[a][A / 3], pb = P[b][B / 3];
            if (pa == 0.0 || pb == 0.0) continue;
            acc += pa * ke[(3 * a + A % 3) * 24 + 3 * b + B % 3] * pb;
          }
        q8[s * 576 + A * 24 + B] = (float)acc;
      }
  }

  // KE in MFMA A-operand order with hi/lo BF16 split. Columns (input) are
  // node-pad order over the reference corners; rows (output) follow the two
  // alternating corner layouts of mfma_out_cx / mfma_out_cy.
  static uint16_t kb[2 * 2 * 2 * 64 * 8];
  for (int ly = 0; ly < 2; ++ly)
    for (int t = 0; t < 2; ++t)
      for (int l = 0; l < 64; ++l)
        for (int v = 0; v < 8; ++v) {
          const int row = mfma_a_row(t, l) & 15, col = mfma_a_col(l, v);
          const int g = row >> 2, cr = row & 3, nc = col >> 2, cc = col & 3;
          const int nr = corner_of(mfma_out_cx(g), mfma_out_cy(ly, g), t);
          const float x = (cr < 3 && cc < 3) ? ke32[(3 * nr + cr) * 24 + 3 * nc + cc] : 0.0f;
          const uint16_t hi = bf16_bits_rne(x);
          const uint16_t lo = bf16_bits_rne(x - bf16_to_float(hi));
          kb[(((ly * 2 + 0) * 2 + t) * 64 + l) * 8 + v] = hi;
          kb[(((ly * 2 + 1) * 2 + t) * 64 + l) * 8 + v] = lo;
        }

  static uint16_t qb[8 * 2 * 2 * 2 * 64 * 8];
  for (int sc = 0; sc < 8; ++sc)
    for (int ly = 0; ly < 2; ++ly)
      for (int t = 0; t < 2; ++t)
        for (int l = 0; l < 64; ++l)
          for (int v = 0; v < 8; ++v) {
            const int row = mfma_a_row(t, l) & 15, col = mfma_a_col(l, v);
            const int g = row >> 2, cr = row & 3, nc = col >> 2, cc = col & 3;
            const int nr = corner_of(mfma_out_cx(g), mfma_out_cy(ly, g), t);
            const float x = (cr < 3 && cc < 3) ? q8[sc * 576 + (3 * nr + cr) * 24 + 3 * nc + cc] : 0.0f;
            const uint16_t hi = bf16_bits_rne(x);
            const uint16_t lo = bf16_bits_rne(x - bf16_to_float(hi));
            qb[((((sc * 2 + ly) * 2 + 0) * 2 + t) * 64 + l) * 8 + v] = hi;
            qb[((((sc * 2 + ly) * 2 + 1) * 2 + t) * 64 + l) * 8 + v] = lo;
          }
  TG_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_q8_bf16), qb, sizeof(qb)));

  for (int i = 0; i < 576; ++i) g_ke_host[i] = ke[i];
  TG_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_ke64), ke, sizeof(double) * 576));
  TG_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_ke32), ke32, sizeof(ke32)));
  TG_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_q8_32), q8, sizeof(q8)));
  TG_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_ke_bf16), kb, sizeof(kb)));
  if (!g_ws[dev].partials) {
    void* p = nullptr;
    TG_HIP_CHECK(hipMalloc(&p, sizeof(double) * kMaxPartials + 256));
    g_ws[dev].partials = (double*)p;
    g_ws[dev].flags = (int*)((char*)p + sizeof(double) * kMaxPartials);
    TG_HIP_CHECK(hipMemset(g_ws[dev].flags, 0, 256));
  }
  if (int rc = init_static_scratch(dev)) return rc;
  TG_HIP_CHECK(hipDeviceSynchronize());
  g_ready[dev] = true;
  return TG_OK;
}

#define TG_STREAM(s) ((hipStream_t)(s))
#define TG_NEED_INIT() TG_REQUIRE(::tg::ke_ready(), "tg_init has not been called on this device")

extern "C" int tg_fill32(int64_t n, float* x, float v, void* s) {
  TG_REQUIRE(n >= 0 && (n == 0 || x), "bad fill32 args");
  if (n == 0) return TG_OK;
  hipLaunchKernelGGL(k_fill32, dim3(vec_blocks(n)), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, x, v);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_graph_begin(void* s) {
  TG_REQUIRE(s, "tg_graph_begin: capture needs a created stream, not the null stream");
  TG_HIP_CHECK(hipStreamBeginCapture(TG_STREAM(s), hipStreamCaptureModeThreadLocal));
  return TG_OK;
}
extern "C" int tg_graph_end(void* s, void** exec) {
  TG_REQUIRE(s && exec, "tg_graph_end: NULL argument");
  *exec = nullptr;
  hipGraph_t g = nullptr;
  TG_HIP_CHECK(hipStreamEndCapture(TG_STREAM(s), &g));
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (rc != hipSuccess) return set_error(TG_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(rc));
  *exec = e;
  return TG_OK;
}
extern "C" int tg_graph_launch(void* exec, void* s) {
  TG_REQUIRE(exec, "tg_graph_launch: NULL graph");
  TG_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)exec, TG_STREAM(s)));
  return TG_OK;
}
extern "C" int tg_graph_destroy(void* exec) {
  if (exec) TG_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)exec));
  return TG_OK;
}
extern "C" int tg_copy32(int64_t n, float* dst, const float* src, void* s) {
  TG_REQUIRE(n >= 0 && (n == 0 || (dst && src)), "bad copy32 args");
  if (n == 0) return TG_OK;
  hipLaunchKernelGGL(k_copy32, dim3(vec_blocks(n)), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, dst, src);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_fill64(int64_t n, double* x, double v, void* s) {
  TG_REQUIRE(n >= 0 && (n == 0 || x), "bad fill64 args");
  if (n == 0) return TG_OK;
  hipLaunchKernelGGL(k_fill64, dim3(vec_blocks(n)), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, x, v);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_dot32_b(int64_t n, int32_t nrhs, const float* a, const float* b, const float* c, double* out,
                          double* out2, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && nrhs >= 1 && a && b && out && (c == nullptr) == (out2 == nullptr), "bad dot32 args");
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot32, dim3(nb, nrhs), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, a, b, c,
                     workspace());
  TG_LAUNCH_CHECK();
  return finalize_batched(c ? 2 : 1, nb, nrhs, out, out2, TG_STREAM(s));
}
extern "C" int tg_dot32(int64_t n, const float* a, const float* b, const float* c, double* out,
                        double* out2, void* s) {
  return tg_dot32_b(n, 1, a, b, c, out, out2, s);
}
extern "C" int tg_dot64_b(int64_t n, int32_t nrhs, const double* a, const double* b, double* out, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && nrhs >= 1 && a && b && out, "bad dot64 args");
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot64, dim3(nb, nrhs), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, a, b, workspace());
  TG_LAUNCH_CHECK();
  return finalize_batched(1, nb, nrhs, out, nullptr, TG_STREAM(s));
}
extern "C" int tg_dot64(int64_t n, const double* a, const double* b, double* out, void* s) {
  return tg_dot64_b(n, 1, a, b, out, s);
}
extern "C" int tg_pcg_xr_b(int64_t n, int32_t nrhs, float* x, float* r, const float* p, const float* q,
                           const double* rz, const double* pq, double* rr, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && nrhs >= 1 && x && r && p && q && rz && pq && rr, "bad pcg_xr args");
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_pcg_xr, dim3(nb, nrhs), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, x, r, p, q, rz, pq,
                     workspace());
  TG_LAUNCH_CHECK();
  return finalize_batched(1, nb, nrhs, rr, nullptr, TG_STREAM(s));
}
extern "C" int tg_pcg_xr(int64_t n, float* x, float* r, const float* p, const float* q,
                         const double* rz, const double* pq, double* rr, void* s) {
  return tg_pcg_xr_b(n, 1, x, r, p, q, rz, pq, rr, s);
}
extern "C" int tg_pcg_dir_b(int64_t n, int32_t nrhs, float* p, const float* z, const double* zq, const double* pq,
                            void* s) {
  TG_REQUIRE(n > 0 && nrhs >= 1 && p && z && zq && pq, "bad pcg_dir args");
  hipLaunchKernelGGL(k_pcg_dir, dim3(vec_blocks(n), nrhs), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, p, z,
                     zq, pq);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_pcg_dir(int64_t n, float* p, const float* z, const double* zq, const double* pq,
                          void* s) {
  return tg_pcg_dir_b(n, 1, p, z, zq, pq, s);
}
extern "C" int tg_axpy_f32_to_f64(int64_t n, double* x, const float* d, void* s) {
  TG_REQUIRE(n > 0 && x && d, "bad axpy args");
  hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3(vec_blocks(n)), dim3(kVecBlock), 0, TG_STREAM(s),
                     (long long)n, x, d);
  TG_LAUNCH_CHECK();
  return TG_OK;
}

extern "C" int tg_cheb_init32_b(int64_t n, int32_t nrhs, const float* rhs, const float* dinv, float a, float* d,
                                float* x, float* r, void* s) {
  TG_REQUIRE(n > 0 && nrhs >= 1 && rhs && dinv && d && x && r != rhs, "bad cheb_init args");
  hipLaunchKernelGGL(k_cheb_init, dim3(vec_blocks(n), nrhs), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, rhs,
                     dinv, a, d, x, r);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
extern "C" int tg_cheb_init32(int64_t n, const float* rhs, const float* dinv, float a, float* d, float* x, float* r,
                              void* s) {
  return tg_cheb_init32_b(n, 1, rhs, dinv, a, d, x, r, s);
}
extern "C" int tg_power_step32(int64_t n, float* v, const float* kv, const float* dinv, double* vdv, void* s) {
  TG_NEED_INIT();
  TG_REQUIRE(n > 0 && v && kv && dinv, "bad power_step args");
  const int nb = red_blocks(n);
  if (!vdv) {
    hipLaunchKernelGGL(k_power_step<false>, dim3(nb), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, v, kv, dinv,
                       workspace());
    TG_LAUNCH_CHECK();
    return TG_OK;
  }
  hipLaunchKernelGGL(k_power_step<true>, dim3(nb), dim3(kVecBlock), 0, TG_STREAM(s), (long long)n, v, kv, dinv,
                     workspace());
  TG_LAUNCH_CHECK();
  return finalize_sums(1, nb, vdv, nullptr, TG_STREAM(s));
}

namespace tg {
int finalize_batched_at(const double* partials, size_t capacity, int K, int seg, int nseg, double* out0,
                        double* out1, hipStream_t s) {
  TG_REQUIRE(partials && K >= 1 && K <= 2 && seg >= 1 && nseg >= 1 && (size_t)K * seg * nseg <= capacity,
             "bad reduction shape");
  hipLaunchKernelGGL(k_finalize, dim3(nseg), dim3(256), 0, s, K, seg, partials, out0, out1);
  TG_LAUNCH_CHECK();
  return TG_OK;
}
int finalize_batched(int K, int seg, int nseg, double* out0, double* out1, hipStream_t s) {
  return finalize_batched_at(workspace().partials, (size_t)kMaxPartials, K, seg, nseg, out0, out1, s);
}
int finalize_sums(int K, int nb, double* out0, double* out1, hipStream_t s) {
  return finalize_batched(K, nb, 1, out0, out1, s);
}
}  // namespace tg
