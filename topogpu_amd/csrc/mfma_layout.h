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

// Output-row order of the element operator in the MFMA kernel. D tile t holds
// the z-face t of the element (t = corner z offset); within a tile, row group
// g = (row % 16) / 4 holds one corner with x offset g & 1. The y offset of the
// group alternates between two layouts so that the y-line shared by
// neighbouring pencils py (layout py & 1) and py + 1 lands in the same lane
// half of both:
//   layout 0: groups 0,1 -> y offset 1, groups 2,3 -> y offset 0
//   layout 1: groups 0,1 -> y offset 0, groups 2,3 -> y offset 1
__host__ __device__ constexpr int mfma_out_cx(int g) { return g & 1; }
__host__ __device__ constexpr int mfma_out_cy(int layout, int g) { return (g >> 1) ^ (1 - layout); }

}  // namespace tg
