"""SolverV4's state-solve contract on the MI355X.

``HipSolverV4`` takes the constructor keywords of the reference's
``SolverV4`` (src/gpu_fem/solver_v4.py:38-52, with those of ``SolverV2``,
solver_v2.py:1519-1541, and ``GPUFEMSolver``, fem_gpu.py:174-191) and offers
its per-iteration interface:

* ``solve(rho_phys, penal) -> (compliance, dc_phys)``: host float64 in and out
  (fem_gpu.py:417; the body of ``_solve_cupy_matfree``, solver_v4.py:147-269):
  moduli E = Emin + (E0 - Emin) rho^penal, the GMG hierarchy, the state solve
  from the previous solution (dropped when the penalty moves by more than
  0.5), compliance F.u and dc = -penal (E0 - Emin) rho^(penal-1) u_e^T KE u_e;
* ``last_cg_iters``, ``last_outer_solver``, ``last_multigrid_quality`` (the
  V-cycle probe the reference runs after every solve) and, with
  ``enable_profiling``, ``last_timing`` with the keys of solver_v4.py:255-267;
* ``invalidate()`` drops the warm start (solver_v2.py:2400-2404,
  solver_v4.py:271-274).

So the reference GPU loop (src/topogpu/solver.py:199-240) runs on this class
unchanged: construct it where ``SolverV4(...)`` is constructed.

Keyword mapping: ``matfree_gmg_levels`` -> levels, ``gmg_fine_degree`` ->
level-0 Chebyshev degree, ``gmg_smooth_iters`` -> coarse-level degree,
``gmg_smoother_type``, ``gmg_cycle_type``, ``gmg_cheb_lower_frac``,
``gmg_level_precisions``, ``gmg_outer_solver`` ("auto": FGMRES when the fine
smoother is BF16, else PCG, solver_v4.py:187-190), ``gmg_restart``,
``cg_tol``, ``cg_maxiter``, ``enable_warm_start``. ``gmg_fine_smoother``
picks the level-0 operator precision: "bf16" single-pass BF16 MFMA in the
V-cycle with the split-BF16 Krylov product, "fp32" and "fp64" FP32 operators
(no level smooths in FP64 here). The connectivity is implicit in the
structured grid: ``edof`` is checked against it, the COO indices are not used
(the reference skips them in matrix-free mode, fem_gpu.py:229, 255-262).
Options of the assembled-CSR, Torch and CPU paths have no meaning here.
"""

from __future__ import annotations

import time

import numpy as np

from .. import fem

_FINE = {"bf16": "bf16", "fp32": "fp32", "fp64": "fp32"}


def structured_edof(dims) -> np.ndarray:
    """Element -> 24 DOF table of the structured hex grid (element (i,j,k) =
    (i ny + j) nz + k, local node order of pub_simp_solver.py:196-207)."""
    nx, ny, nz = dims
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    nodes = base[:, None, :] + fem.CORNERS[None, :, :]
    nid = (nodes[..., 0] * (ny + 1) + nodes[..., 1]) * (nz + 1) + nodes[..., 2]
    return (3 * nid[..., None] + np.arange(3)).reshape(-1, 24)


class HipSolverV4:
    def __init__(self, edof=None, row_idx=None, col_idx=None, KE_UNIT=None, free=None, F=None, ndof=None,
                 E0: float = 1.0, Emin: float = 1e-9, backend: str = "hip", cg_tol: float = 1e-5,
                 cg_maxiter: int = 1000, *, grid_dims=None, enable_warm_start: bool = True,
                 enable_matrix_free: bool = True, enable_fused_cuda: bool = True, enable_profiling: bool = False,
                 enable_matfree_gmg: bool = True, matfree_gmg_levels: int = 4, gmg_fine_smoother: str = "fp32",
                 gmg_fine_degree: int = 2, gmg_outer_solver: str = "auto", gmg_restart: int = 32,
                 gmg_smoother_type: str = "chebyshev", gmg_cheb_lower_frac: float = 1.0 / 30.0,
                 gmg_level_precisions=None, gmg_cycle_type: str = "v", gmg_smooth_iters: int = 2,
                 precision: str | None = None, **unused):
        # the validation of solver_v4.py:66-82, same messages
        if not enable_matfree_gmg:
            raise ValueError("HipSolverV4 runs the matrix-free GMG path only (enable_matfree_gmg=True).")
        if not enable_matrix_free:
            raise ValueError("SolverV4 requires enable_matrix_free=True when enable_matfree_gmg=True.")
        if grid_dims is None:
            raise ValueError("SolverV4 requires grid_dims=(nelx, nely, nelz) for the GMG path.")
        if gmg_fine_smoother not in {"fp64", "fp32", "bf16"}:
            raise ValueError(
                f"gmg_fine_smoother must be one of 'fp64', 'fp32', or 'bf16'; got {gmg_fine_smoother!r}.")
        if gmg_outer_solver not in {"auto", "pcg", "fgmres"}:
            raise ValueError(f"gmg_outer_solver must be 'auto', 'pcg', or 'fgmres'; got {gmg_outer_solver!r}.")
        if backend not in ("hip", "rocm", "auto", "cupy"):
            raise ValueError(f"HipSolverV4 runs on the hip backend; got backend={backend!r}")
        dims = tuple(int(d) for d in grid_dims)
        n_elem = int(np.prod(dims))
        n_dof = 3 * int(np.prod([d + 1 for d in dims]))
        if ndof is not None and int(ndof) != n_dof:
            raise ValueError(f"ndof={ndof} does not match grid_dims {dims} ({n_dof} DOFs)")
        if edof is not None:
            edof = np.asarray(edof)
            if edof.shape != (n_elem, 24) or not np.array_equal(edof, structured_edof(dims)):
                raise ValueError("edof must be the structured hex8 table of grid_dims (_edof_table_3d)")
        if KE_UNIT is not None and not np.allclose(np.asarray(KE_UNIT, dtype=np.float64), fem.unit_stiffness(),
                                                   rtol=0, atol=1e-12):
            raise ValueError("KE_UNIT must be the unit hex8 stiffness (E = 1, nu = 0.3) of KE_UNIT_3D")
        if free is None or F is None:
            raise ValueError("free and F are required")
        F = np.asarray(F, dtype=np.float64)
        if F.shape != (n_dof,):
            raise ValueError(f"F must have shape ({n_dof},)")
        self.ndof, self.E0, self.Emin = n_dof, float(E0), float(Emin)
        self.cg_tol, self.cg_maxiter = float(cg_tol), int(cg_maxiter)
        self.grid_dims = dims
        self._free = np.asarray(free, dtype=np.int64)
        self._enable_profiling = bool(enable_profiling)
        prec = precision or _FINE[gmg_fine_smoother]
        if gmg_level_precisions is not None:
            gmg_level_precisions = [prec] + [("fp32" if p == "fp64" else p) for p in list(gmg_level_precisions)[1:]]
        outer = gmg_outer_solver
        if outer == "auto":
            outer = "fgmres" if gmg_fine_smoother == "bf16" else "pcg"

        import torch

        from . import lib as L
        from .state import HipStateSolver

        self._L, self._torch = L, torch
        self.fea = HipStateSolver(dims, self._free, F, precision=prec, n_levels=int(matfree_gmg_levels),
                                  fine_degree=int(gmg_fine_degree), coarse_degree=int(gmg_smooth_iters),
                                  cheb_lower_frac=float(gmg_cheb_lower_frac), tol=self.cg_tol,
                                  max_krylov=self.cg_maxiter, warm_start=bool(enable_warm_start),
                                  smoother=gmg_smoother_type, cycle=gmg_cycle_type, outer_solver=outer,
                                  restart=int(gmg_restart), probe_quality=True,
                                  level_precisions=gmg_level_precisions)
        dev = self.fea.dev
        f64 = dict(dtype=torch.float64, device=dev)
        self._rho = torch.zeros(n_elem, **f64)
        self._rho_phys = torch.zeros(n_elem, **f64)
        self._E64 = torch.zeros(n_elem, **f64)
        self._E32 = torch.zeros(n_elem, dtype=torch.float32, device=dev)
        self._dc = torch.zeros(n_elem, **f64)
        self._ones = torch.ones(1, **f64)
        self._prev_penal: float | None = None
        self.last_cg_iters = 0
        self.last_outer_solver = ""
        self.last_multigrid_quality: dict = {}
        self.last_timing: dict = {}
        self.last_stats = None

    def _sync_time(self) -> float:
        if self._enable_profiling:
            self._torch.cuda.synchronize(self.fea.dev)
        return time.perf_counter()

    def solve(self, rho_phys: np.ndarray, penal: float) -> tuple[float, np.ndarray]:
        L, fea = self._L, self.fea
        rho = np.ascontiguousarray(rho_phys, dtype=np.float64).reshape(-1)
        if rho.size != self._rho.numel():
            raise ValueError(f"rho_phys must have {self._rho.numel()} entries")
        penal = float(penal)
        if self._prev_penal is not None and abs(penal - self._prev_penal) > 0.5:
            fea.reset_warm_start()
        self._prev_penal = penal
        st = fea.stream
        t0 = self._sync_time()
        self._rho.copy_(self._torch.from_numpy(rho))
        # E = Emin + (E0 - Emin) rho^penal (no projection: beta = 0); the
        # fine diagonal is built by the hierarchy's setup, so "diag_ms" holds
        # the moduli and "gmg_setup_ms" the diagonal with the rest
        L.call("tg_density_chain", rho.size, L.ptr(self._rho), 0.0, 0.5, penal, self.Emin, self.E0,
               L.ptr(self._rho_phys), None, L.ptr(self._E64), L.ptr(self._E32), st)
        emax = float(self._E64.max().item())
        t1 = self._sync_time()
        fea.set_modulus(self._E64, self._E32, emax)             # diagonal, Galerkin levels, coarsest, lambda
        t2 = self._sync_time()
        stats = fea.solve()
        t3 = self._sync_time()
        self.last_stats = stats
        self.last_cg_iters = int(stats.iters)
        self.last_outer_solver = fea.outer_solver
        self.last_multigrid_quality = dict(fea.last_multigrid_quality or {})
        if not stats.converged:
            print(f"[HipSolverV4] WARNING: {fea.outer_solver} did not converge in {self.cg_maxiter} iters "
                  f"(n_free={self._free.size:,}; rel. residual {stats.rel_residual:.2e}).")
        compliance = float(fea.compliance())
        L.call("tg_sensitivity64_b", fea.grid, 1, L.ptr(fea.u), L.ptr(self._ones), L.ptr(self._rho_phys), None,
               penal, self.E0 - self.Emin, None, L.ptr(self._dc), st)
        dc = self._dc.cpu().numpy()
        t4 = self._sync_time()
        if self._enable_profiling:
            self.last_timing = {
                "assemble_ms": 0.0, "diag_ms": (t1 - t0) * 1e3, "gmg_setup_ms": (t2 - t1) * 1e3,
                "cg_ms": (t3 - t2) * 1e3, "sensitivity_ms": (t4 - t3) * 1e3, "total_ms": (t4 - t0) * 1e3,
                "used_gmg": True, "matrix_free": True, "outer_solver": self.last_outer_solver,
                "cg_iters": self.last_cg_iters,
            }
        return compliance, dc

    def invalidate(self) -> None:
        """Forget the warm start; the hierarchy is rebuilt by every solve."""
        self.fea.reset_warm_start()
        self._prev_penal = None
