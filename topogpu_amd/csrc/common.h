// Shared definitions for the gfx950 hex8 SIMP kernels.
//
// Grid conventions (identical to the reference's pub_simp_solver.py:196-207 and
// bc_generator.py:104-105):
//   elements nx*ny*nz, element (i,j,k) -> (i*ny + j)*nz + k   (z fastest)
//   nodes (nx+1)(ny+1)(nz+1), node (i,j,k) -> (i*NY + j)*NZ + k
//   DOF = 3*node + component, component 0/1/2 = x/y/z
//   hex8 local node order (dx,dy,dz): (0,0,0) (1,0,0) (1,1,0) (0,1,0)
//                                     (0,0,1) (1,0,1) (1,1,1) (0,1,1)
// Vectors are full length (3*nodes). A per-node byte mask marks free DOFs
// (bit c set = component c free); every operator treats fixed DOFs of its
// input as zero and writes zero to fixed DOFs of its output, which makes it
// M K M with M = diag(mask), the embedding of the reference's K_ff.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "topogpu_hip.h"

namespace tg {

// Corner offsets of the local hex8 nodes.
__host__ __device__ constexpr int corner_x(int c) { return (c == 1 || c == 2 || c == 5 || c == 6) ? 1 : 0; }
__host__ __device__ constexpr int corner_y(int c) { return (c == 2 || c == 3 || c == 6 || c == 7) ? 1 : 0; }
__host__ __device__ constexpr int corner_z(int c) { return c >= 4 ? 1 : 0; }
// Local index of the corner with offset (x,y,z).
__host__ __device__ constexpr int corner_of(int x, int y, int z) {
  return z * 4 + (y ? (x ? 2 : 3) : (x ? 1 : 0));
}

struct Grid {
  int nx, ny, nz;        // elements
  __host__ __device__ int NX() const { return nx + 1; }
  __host__ __device__ int NY() const { return ny + 1; }
  __host__ __device__ int NZ() const { return nz + 1; }
  __host__ __device__ long long nodes() const { return (long long)(nx + 1) * (ny + 1) * (nz + 1); }
  __host__ __device__ long long elems() const { return (long long)nx * ny * nz; }
  __host__ __device__ long long node(int i, int j, int k) const { return ((long long)i * (ny + 1) + j) * (nz + 1) + k; }
  __host__ __device__ long long elem(int i, int j, int k) const { return ((long long)i * ny + j) * nz + k; }
};

inline Grid to_grid(const tg_grid* g) { return Grid{g->nelx, g->nely, g->nelz}; }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2). Remapping the linear block id so that each XCD receives a contiguous
// range of tiles keeps neighbouring tiles, which share halo nodes, in the
// same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  constexpr int NXCD = 8;
  const int per = nblocks / NXCD;          // full tiles per XCD
  const int full = per * NXCD;
  if (bid >= full) return bid;              // tail keeps its id
  return (bid % NXCD) * per + bid / NXCD;
}

// Set the thread-local error message and return the status code.
int set_error(int code, const char* fmt, ...);

#define TG_HIP_CHECK(expr)                                                         \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::tg::set_error(TG_ERR_HIP, "%s failed: %s (%s:%d)", #expr,           \
                             hipGetErrorString(_e), __FILE__, __LINE__);           \
  } while (0)

#define TG_LAUNCH_CHECK()                                                          \
  do {                                                                             \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return ::tg::set_error(TG_ERR_HIP, "kernel launch failed: %s (%s:%d)",       \
                             hipGetErrorString(_e), __FILE__, __LINE__);           \
  } while (0)

#define TG_REQUIRE(cond, ...)                                                      \
  do {                                                                             \
    if (!(cond)) return ::tg::set_error(TG_ERR_ARG, __VA_ARGS__);                  \
  } while (0)

int check_grid(const tg_grid* g);

// Unit element stiffness (24x24, row-major, E=1) in device constant memory,
// set by tg_init. Q8 holds the eight Galerkin child matrices P_s^T KE P_s.
extern __constant__ double c_ke64[576];
extern __constant__ float c_ke32[576];
extern __constant__ float c_q8_32[8 * 576];
bool ke_ready();
const double* ke_host();   // host copy of the unit stiffness given to tg_init

}  // namespace tg
