// Deterministic FP64 reductions across a grid.
//
// A reducing kernel writes one partial per block (fixed tree inside the
// block) to the library workspace; a one-block finalize kernel launched
// right after it on the same stream sums the partials in block order. The
// kernel boundary makes the partials visible, so no device-scope fence is
// needed inside the hot kernel (on MI355X such a fence writes back the
// XCD's L2 and costs more than the reduction itself). The summation order
// depends only on the launch shape: results are bit-reproducible.
#pragma once

#include "common.h"

namespace tg {

struct Workspace {
  double* partials;     // >= kMaxPartials doubles
  int* flags;           // 64 device status words (flags[0]: dense inverse info)
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

// Write this block's K partial sums to ws.partials[k * nblocks + block].
// Must be reached by every thread of every block.
template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], Workspace ws) {
  block_sum<K>(v);
  if (threadIdx.x == 0) {
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int k = 0; k < K; ++k) ws.partials[(size_t)k * nb + bid] = v[k];
  }
}

// Host side: launch the one-block finalize that sums K x nblocks partials
// into out[0..K) (entries may be nullptr). Defined in core.hip.
int finalize_sums(int K, int nblocks, double* out0, double* out1, hipStream_t s);
// As finalize_sums for a grid of nseg slices of seg blocks each (one slice
// per right-hand side): out_k[r] sums slice r. One block per slice.
int finalize_batched(int K, int seg, int nseg, double* out0, double* out1, hipStream_t s);
// The same over partials at a given device address (capacity doubles).
int finalize_batched_at(const double* partials, size_t capacity, int K, int seg, int nseg, double* out0,
                        double* out1, hipStream_t s);
// Resolve the device address of the static scratch of the allocation-free
// entry point for device dev (defined in hex_vec.hip; called by tg_init).
int init_static_scratch(int dev);

}  // namespace tg
