/* topogpu_hip.h -- C ABI of the MI355X (gfx950) SIMP state-solve library.
 *
 * Plain C: device pointers, sizes and an opaque stream handle (a hipStream_t
 * passed as void*; NULL = the default stream). Every entry returns 0 on
 * success or a TG_ERR_* code; tg_last_error() then describes the failure.
 * All work is enqueued on the given stream; nothing is host-synchronous
 * unless documented. Device buffers are owned by the caller; the library
 * never frees them. Reductions write their results to caller-provided device
 * scalars so that iterative solvers keep their scalars on the GPU.
 *
 * The library replaces the device side of the reference's SolverV4 path
 * (nbbllxx0/TopoGPU). Each entry cites the reference interface it replaces.
 * Reductions use one library-owned workspace per device: calls that reduce
 * (a non-NULL dot/norm output) must not run concurrently on one device.
 *
 * Vector layout: full length 3*(nelx+1)*(nely+1)*(nelz+1), node (i,j,k) at
 * (i*(nely+1)+j)*(nelz+1)+k, DOF = 3*node + component (reference
 * bc_generator.py:104-105). A per-node byte mask marks free DOFs (bit c set =
 * component c free); operators ignore fixed input DOFs and write zero to
 * fixed output DOFs, i.e. they apply M K M, the embedding of K_ff.
 */
#ifndef TOPOGPU_HIP_H
#define TOPOGPU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TG_OK = 0,
  TG_ERR_ARG = 1,     /* invalid argument or unsupported shape            */
  TG_ERR_HIP = 2,     /* HIP runtime / kernel launch failure              */
  TG_ERR_STATE = 3,   /* tg_init not called / library state missing       */
  TG_ERR_NOT_SPD = 4  /* dense SPD inverse: non-positive pivot          */
};

typedef struct {
  int32_t nelx, nely, nelz; /* elements per axis (>= 1) */
} tg_grid;

/* ---- library state --------------------------------------------------- */

/* Upload the 24x24 unit element stiffness (host, row-major, FP64) to device
 * constant memory of the current device, derive its FP32/BF16 forms and the
 * eight Galerkin child matrices, and allocate the reduction workspace.
 * Replaces the per-solver uploads of GPUFEMSolver._upload_cupy
 * (src/gpu_fem/fem_gpu.py:361-368) and FusedMatvec.__init__ KE packing
 * (src/gpu_fem/cuda_fused_matvec.py:279-292, 316-330). */
int tg_init(const double* ke_unit_host);
/* *dev = the device address of pinned host memory (hipHostMalloc) at host,
 * so that a kernel writes a result the host polls for straight into it, with
 * no copy launch in between. Any reduced scalar output of the entry points
 * below (the Krylov loop's r.r, the FP64 residual norms, dot products) may
 * be such an address, and so may a scalar input (tg_normalize64's ss).
 * TG_ERR_ARG when host is not such memory. No counterpart in the reference, whose host reads go
 * through CuPy's device-to-host copies (.get() / float()). */
int tg_mapped_pointer(const void* host, void** dev);
const char* tg_last_error(void);
int tg_version(void);

/* ---- operators --------------------------------------------------------- */

enum {
  /* level-0 hex8 operator, weights = E per element (FP32), vector FMA      */
  TG_OP_HEX_F32 = 1,
  /* level-0 hex8 operator, element GEMM on BF16 MFMA, FP32 accumulate       */
  TG_OP_HEX_BF16 = 2,
  /* as TG_OP_HEX_BF16 with hi/lo split operands (3 MFMA passes)            */
  TG_OP_HEX_BF16X3 = 3,
  /* Galerkin level-1 operator of a fine hex grid with even dims: element
   * matrix sum_s E_child(s) * Q_s, weights = fine E (FP32)                 */
  TG_OP_Q8_F32 = 4,
  /* stored element matrices (576 FP32 per element, row-major)               */
  TG_OP_STORED_F32 = 5,
  /* assembled 27-point block stencil from tg_assemble_stencil (243 FP32 per
   * node); input entries at fixed DOFs must be finite (their coefficients
   * are zero)                                                              */
  TG_OP_STENCIL_F32 = 6,
  /* as TG_OP_Q8_F32 with the eight child products on BF16 MFMA (FP32
   * accumulate), single pass / hi-lo split operands                        */
  TG_OP_Q8_BF16 = 7,
  TG_OP_Q8_BF16X3 = 8,
  /* as TG_OP_STENCIL_F32 with the coefficients stored as BF16 (from
   * tg_assemble_stencil_bf16: 243 x 2 bytes per node, half the bytes the
   * bandwidth-bound coarse levels read); products and sums in FP32        */
  TG_OP_STENCIL_BF16 = 9
};

typedef struct {
  int32_t kind;               /* TG_OP_*                                     */
  tg_grid grid;               /* grid the operator acts on                   */
  const uint8_t* mask;        /* per-node free mask of that grid             */
  const float* weights;       /* HEX*: E[elem]; Q8: fine E[fine elem]        */
  const float* elem_mats;     /* STORED: 576 per element                     */
  const void* stencil;        /* STENCIL_F32: float, STENCIL_BF16: BF16 bits;
                                 see tg_assemble_stencil                     */
} tg_op32;

enum {
  TG_EP_STORE = 0,      /* y = A u                                           */
  TG_EP_DOT = 1,        /* y = A u ;            *dot = sum(u*y)             */
  TG_EP_RESID = 2,      /* y = rhs - A u                                     */
  TG_EP_RESID_NORM = 3, /* y = rhs - A u ;      *dot = sum(y*y)             */
  TG_EP_CHEB = 4,       /* t = dinv*(rhs - A u); d = a1*t + a2*d;
                           x = u + d   (x is output only; x != u)           */
  TG_EP_CHEB_R = 5,     /* residual-update Chebyshev step on a correction u:
                           r -= A u; d = a1*dinv*r + a2*u; x += a3*u + d
                           (r, x in/out; d output, d != u)                  */
  TG_EP_RESID_UPDATE = 6, /* r -= A u (r in/out)                            */
  /* The first smoothing step fused with the Chebyshev start. The operator
   * is applied to d0 = (a3*dinv)*u, formed from the input u as it is read
   * (rounded as tg_cheb_init32 rounds it); with u == rhs this is the
   * start x = d0, r = rhs of smoothing A x = rhs from zero, and the
   * outputs equal tg_cheb_init32 followed by CHEB_R (a3 = 0) bit for bit.
   * Outputs must be distinct and differ from u and rhs.                  */
  TG_EP_CHEB_FIRST = 7, /* r = rhs - A d0; d = a1*dinv*r + a2*d0;
                           x = d0 + d                                       */
  TG_EP_RESID_FIRST = 8, /* degree-1 smoothing: r = rhs - A d0; x = d0
                           (x is zero at fixed DOFs)                        */
  TG_EP_CHEB_LAST = 9   /* last step of a smoothing: x as CHEB_R computes
                           it; r and d are not written                      */
};

typedef struct {
  int32_t mode;         /* TG_EP_*                                           */
  const float* rhs;
  const float* dinv;
  float* d;             /* CHEB: in/out (not read when a2 == 0); CHEB_R: out */
  float* x;             /* CHEB: output iterate; CHEB_R: in/out iterate     */
  float* r;             /* CHEB_R / RESID_UPDATE: residual in/out           */
  float a1, a2, a3;     /* CHEB / CHEB_R coefficients                       */
  double* dot;          /* device scalar for DOT / RESID_NORM               */
} tg_ep32;

/* Coarse-level operator M K M as a 27-point block stencil: for node n, row
 * component c, neighbour q = (dx+1)*9+(dy+1)*3+(dz+1) and column component
 * d, coef[(c*81 + 3*q + d) * nodes + n] = sum over the elements containing
 * n and its neighbour of the element matrix entry, zero for fixed rows,
 * fixed columns and neighbours outside the grid. mats: 576 FP64 per element
 * (row-major, reference local node order). Replaces the coarse CSR matrices
 * K_l = R K_{l-1} P of multigrid_v4.py:747-752 on the V-cycle path. */
int tg_assemble_stencil(const tg_grid* g, const uint8_t* mask, const double* mats, float* coef,
                        void* stream);
/* The same coefficients rounded to BF16 (round to nearest even of the FP32
 * value tg_assemble_stencil stores), as raw 16-bit patterns in the same
 * layout, for TG_OP_STENCIL_BF16. */
int tg_assemble_stencil_bf16(const tg_grid* g, const uint8_t* mask, const double* mats, uint16_t* coef,
                             void* stream);

/* y (or the epilogue outputs) = A u for a single-precision operator.
 * Replaces FusedMatvec.matvec / _launch (cuda_fused_matvec.py:336-409),
 * MatrixFreeKff.matvec in FP32 (solver_v2.py:1183-1223), the coarse CSR SpMV
 * (multigrid_v4.py:948-1001) and the elementwise work fused around them in
 * _chebyshev_steps (multigrid_v4.py:860-889). */
int tg_apply32(const tg_op32* op, const float* u, float* y, const tg_ep32* ep, void* stream);

/* Diagnostic: one v_mfma_f32_16x16x32_bf16 on one wavefront. a, b: 64 lanes
 * x 8 BF16 bit patterns (device), d: 64 lanes x 4 floats. Used by the tests
 * to confirm the operand layout the element kernel assumes. No reference
 * counterpart (the reference's WMMA fragments are opaque). */
int tg_mfma_probe(const uint16_t* a, const uint16_t* b, float* d, void* stream);

/* FP64 level-0 operator: y = M K(E) M u, or with rhs != NULL the residual
 * y = rhs - A u; if y32 != NULL also writes float(y); if sumsq != NULL writes
 * sum(y*y) (FP64). Replaces MatrixFreeKff.matvec in FP64
 * (solver_v2.py:1183-1223) as used for the outer operator and true residual
 * (solver_v4.py:177-178, multigrid_v4.py:994). */
int tg_apply64(const tg_grid* g, const uint8_t* mask, const double* E, const double* u,
               const double* rhs, double* y, float* y32, double* sumsq, void* stream);

/* tg_apply64 without the library's allocated workspace: the reduction
 * partials of sumsq live in a static device array of the library (one per
 * GPU, 65,536 FP64 partials, held until the process ends), so the call
 * itself allocates nothing and needs no workspace argument. No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc); tg_init, once at set-up, resolves the array's address. Grids
 * of more than 65,536 node tiles of 4 x 4 x 16 nodes (about 16M nodes) are
 * rejected with TG_ERR_ARG. Calls that overlap on one GPU share the array
 * and race; calls ordered on one stream do not. Results are bit-identical
 * to tg_apply64's. */
int tg_apply64_noalloc(const tg_grid* g, const uint8_t* mask, const double* E, const double* u,
                       const double* rhs, double* y, float* y32, double* sumsq, void* stream);

/* Inverse diagonal of M K(E) M (fixed DOFs -> 1). Replaces
 * MatrixFreeKff.extract_diagonal (solver_v2.py:1229-1252) and the D^-1 setup
 * (multigrid_v4.py:684-686). dinv64 or dinv32 may be NULL. */
int tg_diag_inv_hex(const tg_grid* g, const uint8_t* mask, const double* E, double* dinv64,
                    float* dinv32, void* stream);
/* Same for stored element matrices (FP64) of a coarse level
 * (assemble_level1_galerkin diagonal, multigrid_v4.py:268-269, 751-752). */
int tg_diag_inv_stored(const tg_grid* g, const uint8_t* mask, const double* mats64, float* dinv32,
                       void* stream);
/* Same for the implicit level-1 operator (weights = fine E, FP64). */
int tg_diag_inv_q8(const tg_grid* coarse, const uint8_t* mask, const double* fine_E,
                   float* dinv32, void* stream);

/* ---- multigrid ---------------------------------------------------------- */

/* rc = M_c R rf with R = P^T, P the row-normalised trilinear prolongation
 * of _build_scalar_prolongation (solver_v2.py:62-149) kron I3, coarse grid
 * max(1, n//2) per axis (multigrid_v4.py:525-529). Replaces R_gpu @ residual
 * (multigrid_v4.py:999). */
int tg_restrict32(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_c,
                  const float* rf, float* rc, void* stream);
/* xf = M_f P ec (accumulate = 0) or xf += M_f P ec (accumulate = 1).
 * Replaces P_gpu @ e_c (multigrid_v4.py:1001). */
int tg_prolong32(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_f,
                 const float* ec, float* xf, int accumulate, void* stream);

/* Galerkin coarse element matrices (FP64, 576 per coarse element):
 *   K_C = sum_{children f} (M_f P_f)^T K_f (M_f P_f)
 * with P_f the local 24x24 restriction of the prolongation above to child f
 * and M_f its free mask (use_mask = 0 skips the mask). K_f is E[f]*KE when
 * fine_mats == NULL (fine_E used), else fine_mats[f]. coarse_mats32 (may be
 * NULL) receives an FP32 copy. Replaces assemble_level1_galerkin and the
 * SpGEMM R@(K@P) (multigrid_v4.py:248-270, 747-752). */
int tg_galerkin_elems(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_f,
                      int use_mask, const double* fine_E, const double* fine_mats,
                      double* coarse_mats64, float* coarse_mats32, void* stream);

/* Galerkin element matrices `levels_down` (1 or 2) levels below a fine hex
 * grid when the Dirichlet masks are aligned with the coarse grids (no
 * per-child masking needed) and the fine dims are divisible by
 * 2^levels_down: K_C = sum_j E[desc_j(C)] * T_j with the 8 constant child
 * matrices P_t^T KE P_t (one level) or the 64 grandchild matrices
 * P_s^T P_t^T KE P_t P_s (two levels). One gather of the descendant moduli
 * (work: n_coarse_elem * 8^levels_down doubles) and one DGEMM. Same result
 * as tg_galerkin_elems applied once or twice with use_mask = 0; replaces
 * assemble_level1_galerkin + SpGEMM (multigrid_v4.py:248-270, 747-752). */
int tg_galerkin_from_fine(const tg_grid* fine, int levels_down, const double* fine_E, double* work,
                          double* coarse_mats64, float* coarse_mats32, void* stream);

/* Dense assembly of the free-DOF matrix of a level from element matrices:
 * A[i][j] (n_free x n_free, row-major FP64), free_index[dof] = compact index
 * or -1. Replaces K_coarsest.toarray() (multigrid_v4.py:781). */
int tg_assemble_dense(const tg_grid* g, const double* mats64, const int32_t* free_index,
                      int32_t n_free, double* A, void* stream);

/* In-place explicit inverse A^-1 (full symmetric) of a dense FP64 SPD
 * matrix A (n x n, row-major): Gauss-Jordan elimination without pivoting in
 * one workgroup with the matrix in registers (n <= 64), in 16-wide pivot
 * blocks in one workgroup with the matrix in LDS (n <= 140), in 16-wide
 * pivot blocks with tiled updates over the whole GPU (n <= 2048), rocSOLVER
 * Cholesky factorisation and inverse above.
 * Host-synchronous. Replaces cp.linalg.cholesky + solve_triangular
 * (multigrid_v4.py:781-801, 977-978): the V-cycle then applies the coarsest
 * solve as one matrix-vector product. Returns TG_ERR_NOT_SPD when a pivot is
 * not positive. */
int tg_dense_spd_inverse(double* A, int32_t n, void* stream);

/* tg_dense_spd_inverse for n <= 140 (the one-workgroup inverses), queued on
 * the stream without a host wait: *info (device memory) is set to 0, or to
 * the first non-positive pivot + 1, when the kernels run, and A then holds
 * A^-1 only if *info is 0. Lets the caller overlap the inverse, which
 * occupies one compute unit, with other work and read *info later. Same
 * reference lines as tg_dense_spd_inverse (multigrid_v4.py:781-801). */
int tg_dense_spd_inverse_async(double* A, int32_t n, int32_t* info, void* stream);

/* ec = scatter(Ainv * gather(rc)) over the free DOFs of the coarsest level;
 * fixed DOFs of ec are written as 0. free_list[i] = full DOF of compact i,
 * in ascending order.
 * Replaces _solve_coarsest (multigrid_v4.py:969-984). */
int tg_coarsest_apply(const double* Ainv, int32_t n_free, const int32_t* free_list,
                      int64_t n_full, const float* rc, float* ec, void* stream);

/* Coarsest solve by block cyclic reduction over the x-planes of the level
 * (each plane's 3*(nely+1)*(nelz+1) DOFs one block, at most 2048): an exact
 * FP64 block Cholesky factorisation of the block-tridiagonal operator in
 * odd-even plane order, for coarsest levels too large for the dense inverse.
 * Replaces cp.linalg.cholesky + solve_triangular (multigrid_v4.py:781-801,
 * 977-978) the way tg_dense_spd_inverse + tg_coarsest_apply do, with a
 * setup of batched plane-block inverses and DGEMMs per reduction level and
 * a solve of one GEMV kernel per level and direction.
 * bcr_size: doubles of the factor buffer (work) for grid g.
 * bcr_setup: assembles the blocks from the level's element matrices (mask
 *   as for the other level operators) and factors them into work.
 *   Host-synchronous; TG_ERR_NOT_SPD when a plane block is not SPD.
 * bcr_apply_b: e_coarse = A^-1 r_coarse for nrhs full-length slices (fixed
 *   DOFs come out 0); vec: 2 * nrhs * 3 * nodes FP64 scratch. */
int tg_bcr_size(const tg_grid* g, int64_t* n_work);
int tg_bcr_setup(const tg_grid* g, const uint8_t* mask, const double* mats64, double* work, void* stream);
int tg_bcr_apply_b(const tg_grid* g, const double* work, int32_t nrhs, const float* r_coarse,
                   float* e_coarse, double* vec, void* stream);

/* ---- Krylov vector kernels (FP32 vectors, FP64 reductions) -------------- */

/* *out = sum(a*b) (and *out2 = sum(a*c) when c != NULL), FP64 accumulate.
 * Replaces cp.dot host-synced reductions (solver_v2.py:588-619). */
int tg_dot32(int64_t n, const float* a, const float* b, const float* c, double* out,
             double* out2, void* stream);
int tg_dot64(int64_t n, const double* a, const double* b, double* out, void* stream);
/* PCG step: alpha = *rz / *pq; x += alpha p; r -= alpha q; *rr = sum(r*r).
 * Replaces the vector updates of _cupy_pcg (solver_v2.py:606-611). rr may be
 * the tg_mapped_pointer address of pinned host memory. */
int tg_pcg_xr(int64_t n, float* x, float* r, const float* p, const float* q, const double* rz,
              const double* pq, double* rr, void* stream);
/* Flexible (Polak-Ribiere) direction: beta = -(*zq) / (*pq); p = z + beta p.
 * beta = z_{k+1}.(r_{k+1}-r_k)/(z_k.r_k) rewritten with r_{k+1}-r_k = -alpha q. */
int tg_pcg_dir(int64_t n, float* p, const float* z, const double* zq, const double* pq,
               void* stream);
/* x64 += double(d32): iterative-refinement update of the FP64 iterate. */
int tg_axpy_f32_to_f64(int64_t n, double* x64, const float* d32, void* stream);
/* y = (accumulate ? y : 0) + a * x with a = scale * (alpha ? *alpha : 1) and
 * x = x64, or double(x32) when x64 is NULL; alpha is a device scalar. The
 * FGMRES vector updates of _cupy_fgmres (multigrid_v4.py:387 Gram-Schmidt
 * w -= H[i,j] V[i] with scale -1; :438 x += y_j z_j; z_j = M_op(V[j]) of
 * :380 taken to FP64 with accumulate = 0). */
int tg_axpy64(int64_t n, double* y, const double* x64, const float* x32, const double* alpha,
              double scale, int32_t accumulate, void* stream);
/* y = x / sqrt(*ss) and y32 = float(y) (either may be NULL; 0 when
 * *ss <= 0): the new Krylov vector w / ||w|| of _cupy_fgmres
 * (multigrid_v4.py:365, 388-392) with *ss = w.w from tg_dot64. */
int tg_normalize64(int64_t n, const double* x, const double* ss, double* y, float* y32, void* stream);
/* First Chebyshev step from a zero iterate: d = x = a * dinv * rhs and, when
 * r != NULL, r = rhs (_chebyshev_steps, multigrid_v4.py:874-879, A x0 = 0). */
int tg_cheb_init32(int64_t n, const float* rhs, const float* dinv, float a, float* d, float* x,
                   float* r, void* stream);
/* Power-iteration step for lambda_max(D^-1 K): *vdv = sum(v*v/dinv), then
 * v <- dinv * kv (kv = K v). The Rayleigh quotient (v.Kv)/(v.Dv) is
 * scale-free, so v is not renormalised (_estimate_fine_lambda_max,
 * multigrid_v4.py:273-295). vdv may be NULL: then only v is updated, with
 * no reduction (the steps before the last, whose quotient is not read). */
int tg_power_step32(int64_t n, float* v, const float* kv, const float* dinv, double* vdv,
                    void* stream);
/* Fill helpers (graph-safe replacements for memset). */
int tg_fill32(int64_t n, float* x, float value, void* stream);
int tg_fill64(int64_t n, double* x, double value, void* stream);
/* dst = src (n floats): p = z of _cupy_pcg (solver_v2.py:603) as a kernel. */
int tg_copy32(int64_t n, float* dst, const float* src, void* stream);

/* ---- replayable launch sequences (HIP graphs) ----------------------------
 * No counterpart in the reference, whose PCG loop (_cupy_pcg,
 * src/gpu_fem/solver_v2.py:552, called from solver_v4.py:193) issues every
 * kernel from the host on each iteration. Here one Krylov iteration (V-cycle
 * plus the PCG updates, ~50 launches) is captured once and replayed with a
 * single call, which removes the host's per-launch cost from the loop.
 * tg_graph_begin starts capturing the launches issued on `stream` (a created
 * stream, not the null stream); nothing runs while capturing. tg_graph_end
 * ends the capture and instantiates it into *exec. tg_graph_launch replays
 * it on any stream. Captured launches keep the arguments they had at capture
 * (pointers and scalars): a caller re-captures when one of them changes.
 * Captured sequences hold kernel launches only (no memset nodes).          */
int tg_graph_begin(void* stream);
int tg_graph_end(void* stream, void** exec);
int tg_graph_launch(void* exec, void* stream);
int tg_graph_destroy(void* exec);

/* ---- SIMP density / sensitivity / OC (FP64) ----------------------------- */

/* out[i] = (sum_k w_k * s(in)[i+off_k]) * (out_scale ? out_scale[i] : 1),
 * s(in)[j] = in[j] * (in_scale ? in_scale[j] : 1); 1 <= m <= 256 offsets
 * (dx,dy,dz), those outside the grid skipped. With in_scale = NULL and out_scale = 1/Hs
 * this is the density filter H (pub_simp_solver.py:215-243, H_mat.dot); with
 * in_scale = 1/Hs and out_scale = NULL it is H^T (:368, :371). */
int tg_stencil64(const tg_grid* g, int32_t m, const int32_t* offs_host, const double* w_host,
                 const double* in, const double* in_scale, const double* out_scale,
                 double* out, void* stream);

/* Passive elements: rho[e] = rho_min where mask[e] == 1 (void), 1 where
 * mask[e] == 2 (solid), unchanged where 0 (apply_passive_mask,
 * bc_generator.py:454-470, applied after each OC update when asked). */
int tg_passive64(int64_t n, const int32_t* mask, double* rho, double rho_min, void* stream);

/* rho_phys = heaviside(rho_f, beta, eta) (identity unless beta > 1e-6),
 * dh = d rho_phys / d rho_f, E = Emin + (E0-Emin) rho_phys^penal (FP64 and
 * FP32). pub_simp_solver.py:245-253, 347-350. Any output may be NULL. */
int tg_density_chain(int64_t n, const double* rho_f, double beta, double eta, double penal,
                     double emin, double e0, double* rho_phys, double* dh, double* E64,
                     float* E32, void* stream);

/* Element sensitivity: ce = u_e^T KE u_e from the FP64 full displacement,
 * dcf = -penal (E0-Emin) rho_phys^(penal-1) ce * dh (dh may be NULL = 1).
 * pub_simp_solver.py:361-367, solver_v4.py:246-251. ce may be NULL. */
int tg_sensitivity64(const tg_grid* g, const double* U, const double* rho_phys,
                     const double* dh, double penal, double e_range, double* ce, double* dcf,
                     void* stream);

/* F.u and u^T K u (K = sum_e E_e KE over elements) in double-double
 * precision: out = {Fu_hi, Fu_lo, uKu_hi, uKu_lo}. u must be zero on fixed
 * DOFs. 2 F.u - u^T K u equals the compliance F.u* to second order in the
 * solve error; with the double-double sums its finite differences resolve
 * per-element gradients (compliance, pub_simp_solver.py:356-360). */
int tg_compliance_dd(const tg_grid* g, const double* E, const double* u, const double* F, double* out,
                     void* stream);

/* OC bisection of pub_simp_solver.py:319-340 on the device.
 * oc_prepare: q = rho*sqrt(-min(dc,-1e-12)/max(dv,1e-12)),
 *   lo = clip(rho-move, 1e-3, 1), hi = clip(rho+move, 1e-3, 1) (clipping
 *   to [rho-move, rho+move] and then to [1e-3, 1], as the reference does,
 *   equals one clip to [lo, hi]); state <- {l1=0, l2=1e9} and the first
 *   batch of trial multipliers.
 * oc_bisect_step: one evaluation pass: the filtered+projected volume
 *   sum(heaviside(H clip(q/sqrt(lambda), lo, hi))) at 15 trial
 *   multipliers, then the bisection (lm=(l1+l2)/2, l1=lm if volume/n >
 *   volfrac else l2=lm, stop when l2-l1 < 1e-9 or after 100 steps) advances
 *   through every step these volumes decide -- at least two per pass, since
 *   the next two levels of midpoints are always among the trial multipliers;
 *   the others bracket the volume crossing, which decides whole runs of
 *   later midpoints by monotonicity. The decisions are those of the
 *   sequential loop. No-op once converged. Filter as in tg_stencil64.
 * oc_finalize: rho_new = clip(q/sqrt(lm_last)); *change = max|rho_new-rho|.
 * state: TG_OC_STATE_LEN device doubles {l1, l2, lm_last, done, steps,
 *   passes, ...}; entry 27 keeps the last bisection's result across
 *   oc_prepare calls (zero in a fresh state), around which the next first
 *   pass evaluates its multipliers. */
#define TG_OC_STATE_LEN 32
int tg_oc_prepare(int64_t n, const double* rho, const double* dc, const double* dv, double move,
                  double* q, double* lo, double* hi, double* state, void* stream);
int tg_oc_bisect_step(const tg_grid* g, int32_t m, const int32_t* offs_host,
                      const double* w_host, const double* hs_inv, const double* q,
                      const double* lo, const double* hi, double beta, double eta,
                      double volfrac, int use_heaviside, double* state, void* stream);
int tg_oc_finalize(int64_t n, const double* rho, const double* q, const double* lo,
                   const double* hi, const double* state, double* rho_new, double* change,
                   void* stream);

/* Mean-volume OC of the reference GPU loop (src/topogpu/solver.py:94-117,
 * `_oc_update`): no filter in the constraint, dc clamped to <= -1e-12, moves
 * clipped to [max(rho - move, rho_min), min(rho + move, 1)], hi doubled from
 * 1 while the mean stays above volfrac (hi < 1e40), then at most 80
 * bisections with the early stop |mean - volfrac| < 1e-8.
 * oc_mean_prepare: initialises state (TG_OC_MEAN_STATE_LEN doubles).
 * oc_mean_step: one evaluation pass (64 multipliers: the next 64 doublings,
 *   or the midpoints of the next 6 bisection levels) and the planner that
 *   walks the loop through them; state[7] != 0 once the loop has stopped.
 *   Calls after that return at once.
 * oc_mean_finalize: rho_new = cand(last midpoint), *change = max|rho_new-rho|.
 * The decisions and the last midpoint are those of the sequential loop. */
#define TG_OC_MEAN_STATE_LEN 80
int tg_oc_mean_prepare(double* state, void* stream);
int tg_oc_mean_step(int64_t n, const double* rho, const double* dc, double volfrac, double move,
                    double rho_min, double* state, void* stream);
int tg_oc_mean_finalize(int64_t n, const double* rho, const double* dc, double move, double rho_min,
                        const double* state, double* rho_new, double* change, void* stream);

/* ---- batches of right-hand sides (several load cases at once) ----------
 * The multi-load case runs its load cases as one block through every solver
 * kernel: each call below is the single-RHS entry point of the same name
 * without the "_b" suffix applied to nrhs (1..64) right-hand sides in one
 * launch (grid slice per RHS). A batched vector on a grid is nrhs contiguous
 * slices of 3 * nodes (Krylov helpers: of n); per-RHS scalars (dot outputs,
 * rz, pq, zq, rr, sumsq) are arrays of nrhs doubles. Operators, masks,
 * moduli, dinv and the coarsest inverse are shared by the batch. The
 * reference has no multi-RHS path; these replace running SolverV4 once per
 * load case. */
int tg_apply32_b(const tg_op32* op, int32_t nrhs, const float* u, float* y, const tg_ep32* ep,
                 void* stream);
int tg_apply64_b(const tg_grid* g, const uint8_t* mask, const double* E, int32_t nrhs, const double* u,
                 const double* rhs, double* y, float* y32, double* sumsq, void* stream);
int tg_restrict32_b(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_c, int32_t nrhs,
                    const float* r_fine, float* r_coarse, void* stream);
int tg_prolong32_b(const tg_grid* fine, const tg_grid* coarse, const uint8_t* mask_f, int32_t nrhs,
                   const float* e_coarse, float* x_fine, int accumulate, void* stream);
int tg_coarsest_apply_b(const double* Ainv, int32_t n_free, const int32_t* free_list, int64_t n_full,
                        int32_t nrhs, const float* r_coarse, float* e_coarse, void* stream);
int tg_dot32_b(int64_t n, int32_t nrhs, const float* a, const float* b, const float* c, double* out,
               double* out2, void* stream);
int tg_dot64_b(int64_t n, int32_t nrhs, const double* a, const double* b, double* out, void* stream);
int tg_pcg_xr_b(int64_t n, int32_t nrhs, float* x, float* r, const float* p, const float* q,
                const double* rz, const double* pq, double* rr, void* stream);
int tg_pcg_dir_b(int64_t n, int32_t nrhs, float* p, const float* z, const double* zq, const double* pq,
                 void* stream);
int tg_cheb_init32_b(int64_t n, int32_t nrhs, const float* rhs, const float* dinv, float a, float* d,
                     float* x, float* r, void* stream);
/* Weighted multi-load sensitivity: ce = sum_r w_r u_r,e^T KE u_r,e (weights
 * NULL = all 1), dcf as tg_sensitivity64 with that ce (the weighted
 * compliance's gradient). */
int tg_sensitivity64_b(const tg_grid* g, int32_t nrhs, const double* U, const double* weights,
                       const double* rho_phys, const double* dh, double penal, double e_range,
                       double* ce, double* dcf, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TOPOGPU_HIP_H */
