"""HipSolverV4: the reference's SolverV4 contract (solver_v4.py:38-274) on
the device, driven by the reference GPU loop (src/topogpu/solver.py:199-240)
with the oracle's mean-volume OC on the host."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from topogpu_amd.hip.solver_v4 import HipSolverV4, structured_edof

from gpu_util import cantilever_bc

pytestmark = pytest.mark.gpu

TIMING_KEYS = {"assemble_ms", "diag_ms", "gmg_setup_ms", "cg_ms", "sensitivity_ms", "total_ms", "used_gmg",
               "matrix_free", "outer_solver", "cg_iters"}
QUALITY_KEYS = {"finite", "pd", "z_norm", "jacobi_norm", "z_over_jacobi"}


def _construct(dims, free, F, ndof, **kw):
    """The construction of src/topogpu/solver.py:199-219 with the class
    swapped (the COO indices are not needed by the matrix-free path)."""
    edof = oracle.edof_table_3d(*dims)
    args = dict(edof=edof, row_idx=None, col_idx=None, KE_UNIT=oracle.KE_UNIT_3D, free=free.astype(np.int32), F=F,
                ndof=ndof, backend="cupy", enable_warm_start=True, enable_matrix_free=True, enable_fused_cuda=True,
                enable_matfree_gmg=True, matfree_gmg_levels=4, gmg_fine_smoother="fp32",
                gmg_smoother_type="chebyshev", grid_dims=dims, cg_tol=1e-9, cg_maxiter=800, enable_profiling=True)
    args.update(kw)
    return HipSolverV4(**args)


@pytest.mark.parametrize("fine,outer", [("fp32", "pcg"), ("bf16", "fgmres")])
def test_reference_gpu_loop_through_solver_v4_contract(gpu, fine, outer):
    dims = (24, 12, 6)
    free, F, ndof = cantilever_bc(dims)
    solver = _construct(dims, free, F, ndof, gmg_fine_smoother=fine)
    edof = oracle.edof_table_3d(*dims)
    rho = np.full(int(np.prod(dims)), 0.30)
    rho_ref = rho.copy()
    for it in range(4):
        compliance, dc = solver.solve(rho, penal=3.0)
        assert isinstance(compliance, float) and dc.dtype == np.float64 and dc.shape == rho.shape
        # the same iteration from the oracle: direct solve, SolverV4's dc
        E = oracle.EMIN + (oracle.E0 - oracle.EMIN) * rho_ref ** 3
        U = oracle.solve_state(edof, E, free, F, ndof)
        c_ref = float(F @ U)
        dc_ref = -3.0 * (oracle.E0 - oracle.EMIN) * rho_ref ** 2 * oracle.element_compliance(U, edof)
        assert compliance == pytest.approx(c_ref, rel=1e-7)
        assert np.abs(dc - dc_ref).max() <= 1e-6 * np.abs(dc_ref).max()
        assert solver.last_outer_solver == outer and solver.last_cg_iters > 0
        assert set(solver.last_timing) == TIMING_KEYS and solver.last_timing["cg_iters"] == solver.last_cg_iters
        assert set(solver.last_multigrid_quality) == QUALITY_KEYS and solver.last_multigrid_quality["pd"]
        rho = oracle.oc_update_mean(rho, dc, 0.30, 0.2, 1e-3)
        rho_ref = oracle.oc_update_mean(rho_ref, dc_ref, 0.30, 0.2, 1e-3)
        assert np.abs(rho - rho_ref).max() < 1e-6
    # a warm-started repeat returns at once; after invalidate() it solves from zero
    solver.solve(rho, penal=3.0)
    c1, _ = solver.solve(rho, penal=3.0)
    assert solver.last_cg_iters <= 1
    solver.invalidate()
    c2, _ = solver.solve(rho, penal=3.0)
    assert solver.last_cg_iters > 3 and c2 == pytest.approx(c1, rel=1e-8)


def test_solver_v4_penalty_jump_and_options(gpu):
    dims = (16, 8, 4)
    free, F, ndof = cantilever_bc(dims)
    s = _construct(dims, free, F, ndof, matfree_gmg_levels=2, gmg_fine_degree=3, gmg_smooth_iters=3,
                   gmg_smoother_type="jacobi", gmg_cycle_type="w", enable_profiling=False)
    assert s.fea.gmg.fine_degree == 3 and s.fea.gmg.coarse_degrees == [3]
    assert s.fea.gmg.smoother == "jacobi" and s.fea.gmg.cycle == "w" and len(s.fea.gmg.levels) == 2
    rho = np.full(int(np.prod(dims)), 0.4)
    s.solve(rho, 3.0)
    assert s.last_timing == {}                      # profiling off: no timing, as the reference
    s.solve(rho, 3.0)
    warm = s.last_cg_iters
    s.solve(rho, 3.6)                               # penalty jump > 0.5: from zero
    assert s.last_cg_iters > warm
    assert s.last_stats.converged
