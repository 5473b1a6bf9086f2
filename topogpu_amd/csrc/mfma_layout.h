// Operand layout of v_mfma_f32_16x16x32_bf16 on gfx950 (wave64):
//   D(16x16, f32) = A(16x32, bf16) * B(32x16, bf16) + C
//   A[i][k]: lane = i + 16*(k/8), element k%8 of the lane's 8-vector
//   B[k][j]: lane = j + 16*(k/8), element k%8
//   D[i][j]: lane = j + 16*(i/4), element i%4 of the lane's 4-vector
// The element operator uses M = 32 output rows (two 16-row tiles) and
// K = 32 input columns, both in "node-pad" order: index 4*node + comp with
// comp 3 a zero pad, so each lane's 8 B-operand values are two whole nodes
// and each lane's 4 D values are one whole node. tests/test_gpu_mfma_layout.py
// checks this layout on the device.
#pragma once

namespace tg {

__host__ __device__ constexpr int mfma_a_row(int tile, int lane) { return 16 * tile + (lane & 15); }
__host__ __device__ constexpr int mfma_a_col(int lane, int v) { return 8 * (lane >> 4) + v; }
__host__ __device__ constexpr int mfma_b_col(int lane) { return lane & 15; }          // N index
__host__ __device__ constexpr int mfma_b_row(int lane, int v) { return 8 * (lane >> 4) + v; }
__host__ __device__ constexpr int mfma_d_row(int tile, int lane, int v) { return 16 * tile + 4 * (lane >> 4) + v; }
__host__ __device__ constexpr int mfma_d_col(int lane) { return lane & 15; }

}  // namespace tg
