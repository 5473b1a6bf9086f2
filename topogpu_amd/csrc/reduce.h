// Deterministic FP64 reductions across a grid without a second launch.
//
// Each block reduces its values in a fixed tree, writes one partial per
// output to the library workspace and takes a ticket; the block holding the
// last ticket sums all partials in block order and writes the result. The
// summation order depends only on the launch shape, so results are
// bit-reproducible run to run (unlike atomic accumulation).
#pragma once

#include "common.h"

namespace tg {

struct Workspace {
  double* partials;     // >= kMaxPartials doubles
  unsigned* ticket;     // one counter, returned to 0 by the last block
};
constexpr int kMaxPartials = 1 << 19;   // 4 MiB of FP64 partials
Workspace workspace();                  // per current device, valid after tg_init

// Sum over the 64 lanes of a wavefront.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of K values; result valid in thread 0. blockDim.x <= 1024.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  __shared__ double s_red[16][K];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) s_red[wid][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += s_red[w][k];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Grid-wide sum of K per-thread values into out[k] (device pointers; an
// out[k] may be nullptr). Must be reached by every thread of every block.
// Returns true in the block that performed the final sum (after out[] is
// written), which can then run follow-up single-block logic.
template <int K>
__device__ bool grid_sum(double (&v)[K], double* const (&out)[K], Workspace ws) {
  __shared__ bool s_last;
  block_sum<K>(v);
  const int nb = gridDim.x * gridDim.y * gridDim.z;
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) ws.partials[(size_t)k * nb + bid] = v[k];
    __threadfence();
    const unsigned t = atomicAdd(ws.ticket, 1u);
    s_last = (t == (unsigned)(nb - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      s += ws.partials[(size_t)k * nb + b];
    acc[k] = s;
  }
  block_sum<K>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (out[k]) out[k][0] = acc[k];
    *ws.ticket = 0u;
    __threadfence();
  }
  __syncthreads();
  return true;
}

}  // namespace tg
