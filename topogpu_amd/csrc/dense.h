// Dense FP64 helpers shared by the coarsest-level solvers (gmg.hip, bcr.hip).
#pragma once

#include <rocblas/rocblas.h>

#include "common.h"

namespace tg {

// Largest matrix the blocked Gauss-Jordan inverse takes.
constexpr int kBlockedMaxN = 2048;

// In-place inverses of `batch` SPD n x n matrices at stride sa (doubles),
// n <= kBlockedMaxN, left exactly symmetric. A non-positive pivot sets the
// device status word workspace().flags[0], which the caller clears first and
// reads after. Asynchronous on st.
int spd_inverse_batched(double* A, int n, int batch, long long sa, hipStream_t st);

// The device's rocBLAS handle, bound to stream s.
int get_handle(hipStream_t s, rocblas_handle* out);

}  // namespace tg
