"""State solve and full SIMP loop on the GPU against the FP64 oracle and the
reference's golden run_simp trajectories."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
import topogpu_amd as tg
from topogpu_amd.hip import lib as L
from topogpu_amd.hip.state import HipStateSolver

from gpu_util import DEV, cantilever_bc, h, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
@pytest.mark.parametrize("dims,levels", [((16, 8, 8), 3), ((24, 12, 6), 4), ((13, 7, 9), 3)])
def test_state_solve_matches_direct(gpu, precision, dims, levels):
    free, F, ndof = cantilever_bc(dims, point_support=True)
    rng = np.random.default_rng(21)
    E = 1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    U_ref = oracle.solve_state(oracle.edof_table_3d(*dims), E, free, F, ndof)
    s = HipStateSolver(dims, free, F, precision=precision, n_levels=levels, tol=1e-9, max_krylov=400)
    E64 = torch.from_numpy(E).to(DEV)
    s.set_modulus(E64, E64.float(), float(E.max()))
    st = s.solve()
    assert st.converged, st
    assert st.rel_residual <= 1e-9
    U = h(s.u)
    assert rel_l2(U, U_ref) < 1e-6
    assert s.compliance() == pytest.approx(float(F @ U_ref), rel=1e-8)
    # warm start from the converged solution returns without a Krylov step
    st2 = s.solve()
    assert st2.iters == 0 and st2.converged


def test_coarsest_not_spd_is_reported_by_the_solve(gpu):
    """The coarsest inverse runs asynchronously on a side stream (108 free
    DOFs here); a non-positive pivot still ends the solve with the library's
    not-SPD error, read with the first scalar the solve waits for."""
    dims = (16, 8, 8)
    free, F, ndof = cantilever_bc(dims)
    s = HipStateSolver(dims, free, F, n_levels=3)
    assert s.gmg.async_inverse and 64 < s.gmg.n_coarse <= 140
    E64 = torch.full((int(np.prod(dims)),), -1.0, dtype=torch.float64, device=DEV)   # K negative definite
    s.set_modulus(E64, E64.float(), -1.0)
    with pytest.raises(L.TopoHipError, match="not SPD \\(pivot 1\\)"):
        s.solve()


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_simp_c0_matches_reference_golden(gpu, golden, precision):
    """24x12x6 cantilever, 3 OC iterations: compliance history and physical
    density against the reference CPU run_simp (spsolve) trajectory."""
    g = golden("c0_cantilever_24x12x6")
    prob = tg.unit_cantilever((24, 12, 6), volfrac=0.30, rmin=1.5)
    res = tg.SIMPSolver(backend="hip", tol=1e-8, max_krylov=800, max_iter=3, move=0.2,
                        precision=precision).solve(prob)
    comp = np.array([r["compliance"] for r in res.history])
    assert np.allclose(comp, g["compliance_history"], rtol=1e-6, atol=0), comp / g["compliance_history"] - 1
    assert np.abs(res.rho_final - g["rho_final"]).max() < 1e-5
    assert res.summary["final_compliance"] == pytest.approx(float(g["final_compliance"]), rel=1e-6)
    assert res.summary["n_iter"] == int(g["n_iter"])


def test_simp_c0_at_stated_tolerance(gpu, golden):
    """Same case at the configured PCG tolerance 1e-5 (relative residual)."""
    g = golden("c0_cantilever_24x12x6")
    prob = tg.unit_cantilever((24, 12, 6), volfrac=0.30, rmin=1.5)
    res = tg.SIMPSolver(backend="hip", tol=1e-5, max_iter=3, move=0.2).solve(prob)
    comp = np.array([r["compliance"] for r in res.history])
    assert np.allclose(comp, g["compliance_history"], rtol=1e-5, atol=0), comp / g["compliance_history"] - 1
    assert np.abs(res.rho_final - g["rho_final"]).max() < 1e-3


def test_simp_odd_block_matches_reference_golden(gpu, golden):
    """10x5x3 block, side load, rmin 2, 6 iterations (odd coarse levels)."""
    g = golden("c0b_block_10x5x3_rmin2")
    spec = tg.ProblemSpec(Lx=10.0, Ly=5.0, Lz=3.0, nelx=10, nely=5, nelz=3, volfrac=0.4,
                          supports=[tg.EdgeSupport("left", "fixed")],
                          loads=[tg.PointLoad(x=10.0, y=0.0, z=1.0, fy=-1.0)])
    prob = tg.TopologyProblem("block", spec, filter_radius=2.0)
    res = tg.SIMPSolver(backend="hip", tol=1e-8, max_iter=6, move=0.2, n_levels=2).solve(prob)
    comp = np.array([r["compliance"] for r in res.history])
    assert np.allclose(comp, g["compliance_history"], rtol=1e-6, atol=0)
    assert np.abs(res.rho_final - g["rho_final"]).max() < 1e-5


def test_solverv4_formulation_matches_oracle_loop(gpu):
    """formulation="solverv4": the reference GPU loop (solver.py:186-249) --
    no filter, SolverV4's dc, mean-volume OC with rho_min -- on the 24x12x6
    cantilever for 4 iterations, against the same loop composed from the
    oracle (direct solve, oc_update_mean)."""
    dims = (24, 12, 6)
    prob = tg.unit_cantilever(dims, volfrac=0.30, rmin=1.5)
    res = tg.SIMPSolver(backend="hip", tol=1e-9, max_iter=4, move=0.2, rho_min=1e-3,
                        formulation="solverv4").solve(prob)
    free, F, ndof = cantilever_bc(dims)
    edof = oracle.edof_table_3d(*dims)
    rho = np.full(int(np.prod(dims)), 0.30)
    comp = []
    for _ in range(4):
        E = oracle.EMIN + (oracle.E0 - oracle.EMIN) * rho ** 3
        U = oracle.solve_state(edof, E, free, F, ndof)
        comp.append(float(F @ U))
        dc = -3.0 * (oracle.E0 - oracle.EMIN) * rho ** 2 * oracle.element_compliance(U, edof)
        rho = oracle.oc_update_mean(rho, dc, 0.30, 0.2, 1e-3)
    got = np.array([r["compliance"] for r in res.history])
    assert np.allclose(got, comp, rtol=1e-7, atol=0), got / np.array(comp) - 1
    assert np.abs(res.rho_final - rho).max() < 1e-6
    assert res.summary["n_iter"] == 4 and len(res.history) == 4
    assert res.history[-1]["rho_mean"] == pytest.approx(0.30, abs=1e-8)


@pytest.mark.parametrize("nrhs", [1, 2])
def test_graph_replay_equals_direct_launches(gpu, nrhs):
    """The Krylov iterations replayed from captured graphs (the V-cycle's
    first part, queued speculatively while the host reads r.r, and the rest)
    give the bits of the same iterations launched one by one: over three solves whose
    modulus changes (the second by more than 10 %, which re-estimates the
    smoother's lambda and forces a new capture), each replayed many times
    with host reads of the residual between replays."""
    dims = (16, 8, 8)
    free, F, ndof = cantilever_bc(dims, point_support=True)
    rng = np.random.default_rng(4)
    Fs = np.stack([F] + [rng.standard_normal(ndof) * (np.abs(F).max()) for _ in range(nrhs - 1)])
    Es = [1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3 for _ in range(3)]
    Es[1] = 1.5 * Es[1]
    out = {}
    for graphs, speculate in ((False, True), (True, True), (True, False)):
        s = HipStateSolver(dims, free, Fs, precision="bf16", n_levels=3, tol=1e-8, max_krylov=400, graphs=graphs,
                           speculate=speculate)
        res = []
        for E in Es:
            E64 = torch.from_numpy(E).to(DEV)
            s.set_modulus(E64, E64.float(), float(E.max()))
            st = s.solve()
            assert st.converged, st
            res.append((st.iters, st.restarts, h(s.u).copy()))
        out[graphs, speculate] = res
        if graphs:
            assert sorted(s._graphs) == ["begin", "finish", "finish_first"], "both parts of an iteration are captured"
    # speculation (queueing the next V-cycle's first part before reading
    # r.r) changes timing, never bits
    for key in ((True, True), (True, False)):
        for (i0, r0, u0), (i1, r1, u1) in zip(out[False, True], out[key]):
            assert (i0, r0) == (i1, r1)
            assert np.array_equal(u0, u1)


def _host_chebyshev(K, rhs, x, Dinv, lam_max, lower_frac, degree):
    """_chebyshev_steps (multigrid_v4.py:860-889) as the reference writes it:
    Saad's Algorithm 12.1 with M^-1 = D^-1, updating the iterate x with a
    full residual rhs - K x at every step."""
    lam_min = lower_frac * lam_max
    sigma, delta = 0.5 * (lam_max + lam_min), 0.5 * (lam_max - lam_min)
    r = rhs - K @ x
    alpha = 2.0 / (lam_max + lam_min)
    d = Dinv * r
    x = x + alpha * d
    for _ in range(1, degree):
        r = rhs - K @ x
        z = Dinv * r
        a_new = 1.0 / (sigma - delta ** 2 * alpha * 0.25)
        d = a_new * (z + delta ** 2 * alpha * 0.25 * d)
        x = x + d
        alpha = a_new
    return x


def _host_cycle(levels, lv, b, degrees, smooth, cycle):
    """The multigrid cycle of GalerkinMatFreeGMG (_vcycle / _wcycle,
    multigrid_v4.py:986-1042) in FP64 on the oracle's Galerkin hierarchy
    [(free, K, P)]: pre-smoothing from zero, one (V) or two (W) coarse
    corrections, post-smoothing from the corrected iterate, exact coarsest
    solve. smooth(lv, K, rhs, x, Dinv, degree) is the smoother."""
    _, K, _ = levels[lv]
    if lv == len(levels) - 1:
        return np.linalg.solve(K.toarray(), b)
    Dinv = 1.0 / K.diagonal()
    x = smooth(lv, K, b, np.zeros_like(b), Dinv, degrees[lv])
    P = levels[lv + 1][2]
    for _ in range(2 if cycle == "w" else 1):
        x = x + P @ _host_cycle(levels, lv + 1, P.T @ (b - K @ x), degrees, smooth, cycle)
    return smooth(lv, K, b, x, Dinv, degrees[lv])


def _jacobi(omega):
    """Damped Jacobi (multigrid_v4.py:918-932, 963-967)."""
    def smooth(lv, K, b, x, Dinv, degree):
        for _ in range(degree):
            x = x + omega * Dinv * (b - K @ x)
        return x
    return smooth


@pytest.mark.parametrize("smoother,cycle,degrees", [("jacobi", "v", (2, 3)), ("jacobi", "w", (2, 3)),
                                                    ("chebyshev", "v", (2, 3)), ("chebyshev", "v", (3, 2)),
                                                    ("chebyshev", "w", (2, 3))])
def test_smoother_and_cycles_match_host_restatement(gpu, smoother, cycle, degrees):
    """One preconditioner application on the device against the same cycle
    in FP64 on the oracle's Galerkin operators (FP32 products on the device:
    2e-5 relative):

    * smoother="jacobi": omega = min(0.5, 0.9 * 2 / lambda_max of level 0);
    * smoother="chebyshev": the reference's recurrence on
      [lambda_max / 30, lambda_max] with each level's lambda_max the device
      power-iteration estimate times the safety factor 1.1 -- the device
      runs it in residual-update form, the host as the reference writes it,
      so a wrong coefficient, bound or scaling on either side shows here."""
    dims = (16, 8, 8)
    free, F, ndof = cantilever_bc(dims, point_support=True)
    rng = np.random.default_rng(12)
    E = 1e-3 + rng.uniform(0.0, 1.0, int(np.prod(dims))) ** 3
    fine_deg, coarse_deg = degrees
    s = HipStateSolver(dims, free, F, precision="fp32", n_levels=3, fine_degree=fine_deg, coarse_degree=coarse_deg,
                       smoother=smoother, cycle=cycle)
    E64 = torch.from_numpy(E).to(DEV)
    s.set_modulus(E64, E64.float(), float(E.max()))
    g = s.gmg
    if smoother == "jacobi":
        smooth = _jacobi(min(0.5, 0.9 * 2.0 / g.levels[0].lam))
    else:
        assert g.safety == 1.1 and g.frac == pytest.approx(1.0 / 30.0)
        lams = [g.levels[lv].lam * g.safety for lv in range(2)]

        def smooth(lv, K, b, x, Dinv, degree):
            return _host_chebyshev(K, b, x, Dinv, lams[lv], 1.0 / 30.0, degree)
    K = oracle.assemble_k(oracle.edof_table_3d(*dims), E, ndof)[free, :][:, free]
    from oracle.gmg import galerkin_coarse_operators, level_dims
    levels = galerkin_coarse_operators(K, level_dims(*dims, 3), free)
    b = rng.standard_normal(free.size)
    r = np.zeros(ndof)
    r[free] = b
    st = torch.cuda.current_stream().cuda_stream
    z = h(g.apply(torch.from_numpy(r.astype(np.float32)).to(DEV), st))
    want = _host_cycle(levels, 0, b, [fine_deg, coarse_deg, coarse_deg], smooth, cycle)
    assert rel_l2(z[free], want) < 2e-5
    # and as a preconditioner: the solve converges to the direct solution
    st_ = s.solve()
    assert st_.converged
    U_ref = oracle.solve_state(oracle.edof_table_3d(*dims), E, free, F, ndof)
    assert rel_l2(h(s.u), U_ref) < 1e-4


@pytest.mark.parametrize("controller", [None, "A"])
def test_checkpoint_resume_equals_uninterrupted_run(gpu, tmp_path, controller):
    """checkpoint_dir / checkpoint_every (run_simp, pub_simp_solver.py:
    495-516, 597-603): a run stopped after a checkpoint, then a run that
    resumes from it, give the history and design of one uninterrupted run --
    also under a controller, whose parameters and best design the checkpoint
    carries."""
    from golden.controllers import CONTROLLERS

    ctl = (lambda: CONTROLLERS[controller]()) if controller else (lambda: None)
    dims, n, every = ((24, 12, 6), 6, 3) if controller is None else ((16, 8, 4), 12, 4)
    prob = tg.unit_cantilever(dims, volfrac=0.30, rmin=1.5)
    kw = dict(backend="hip", tol=1e-8, max_krylov=800, move=0.2, max_gray_for_best=0.9)
    full = tg.SIMPSolver(max_iter=n, callback=ctl(), **kw).solve(prob)
    d = str(tmp_path / "ckpt")
    first = tg.SIMPSolver(max_iter=every, checkpoint_dir=d, checkpoint_every=every, callback=ctl(), **kw).solve(prob)
    assert first.history[every - 1]["iteration"] == every
    assert (tmp_path / "ckpt" / f"ckpt_{dims[0]}x{dims[1]}x{dims[2]}_{prob.name}.npz").exists()
    resumed = tg.SIMPSolver(max_iter=n, checkpoint_dir=d, checkpoint_every=every, callback=ctl(), **kw).solve(prob)
    assert [r["iteration"] for r in resumed.history] == [r["iteration"] for r in full.history]
    c_full = np.array([r["compliance"] for r in full.history])
    c_res = np.array([r["compliance"] for r in resumed.history])
    assert np.array_equal(c_full, c_res), c_res / c_full - 1
    assert np.array_equal(full.rho_final, resumed.rho_final)
    assert resumed.summary["n_iter"] == full.summary["n_iter"]
    assert resumed.summary["best_iteration"] == full.summary["best_iteration"]
    assert np.array_equal(full.best_rho, resumed.best_rho)


@pytest.mark.parametrize("case", ["A", "B"])
def test_controlled_run_matches_reference_golden(gpu, golden, case, tmp_path):
    """run_simp with a controller (pub_simp_solver.py:412-657) on the GPU:
    actions, best tracking, restarts, stop and tail against the reference's
    own run with the same controller (16x8x4 cantilever)."""
    from golden.controllers import CONTROLLERS
    from test_oracle import CONTROLLED, check_controlled_run

    g = golden(f"c0c_controlled_16x8x4_{case}")
    prob = tg.unit_cantilever((16, 8, 4), volfrac=0.30, rmin=1.5)
    kw = dict(CONTROLLED[case])
    res = tg.SIMPSolver(backend="hip", tol=1e-8, max_krylov=800, move=0.2, max_iter=kw.pop("max_iter"),
                        callback=CONTROLLERS[case](), **kw).solve(prob)
    s = res.summary
    comp = [r["compliance"] for r in res.history]
    assert [r["iteration"] for r in res.history] == list(range(1, s["n_iter"] + 1))
    check_controlled_run(g, comp, res.rho_final, s["final_compliance"], s["n_iter"], dict(s, best_rho=res.best_rho),
                         s["params_log"], rtol=1e-6, rho_tol=1e-5)
    res.save(tmp_path / "evidence")              # the summary with its controller fields is JSON


@pytest.mark.parametrize("restart", [32, 5])
@pytest.mark.parametrize("dims,levels", [((16, 8, 8), 3), ((13, 7, 9), 3)])
def test_fgmres_state_solve_matches_direct(gpu, dims, levels, restart):
    """outer_solver="fgmres" (_cupy_fgmres, multigrid_v4.py:325-443): FP64
    right-preconditioned FGMRES around the BF16 V-cycle, with and without
    restarts, against the direct solve."""
    free, F, ndof = cantilever_bc(dims, point_support=True)
    rng = np.random.default_rng(22)
    E = 1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    U_ref = oracle.solve_state(oracle.edof_table_3d(*dims), E, free, F, ndof)
    s = HipStateSolver(dims, free, F, precision="bf16", n_levels=levels, tol=1e-9, max_krylov=400,
                       outer_solver="fgmres", restart=restart)
    E64 = torch.from_numpy(E).to(DEV)
    s.set_modulus(E64, E64.float(), float(E.max()))
    st = s.solve()
    assert st.converged, st
    if restart == 5:
        assert st.restarts > 2            # restart cycles were taken
    assert st.rel_residual <= 1e-8       # the Givens estimate tracks the true FP64 residual
    U = h(s.u)
    assert rel_l2(U, U_ref) < 1e-6
    assert s.compliance() == pytest.approx(float(F @ U_ref), rel=1e-8)
    st2 = s.solve()
    assert st2.converged and st2.iters <= 2
    with pytest.raises(ValueError):
        HipStateSolver(dims, free, np.stack([F, F]), outer_solver="fgmres")


def test_simp_c0_with_fgmres_matches_reference_golden(gpu, golden):
    """The golden 24x12x6 trajectory with the FGMRES outer solver."""
    g = golden("c0_cantilever_24x12x6")
    prob = tg.unit_cantilever((24, 12, 6), volfrac=0.30, rmin=1.5)
    res = tg.SIMPSolver(backend="hip", tol=1e-8, max_krylov=800, max_iter=3, move=0.2,
                        outer_solver="fgmres").solve(prob)
    comp = np.array([r["compliance"] for r in res.history])
    assert all(r["outer_solver"] == "fgmres" for r in res.history)
    assert np.allclose(comp, g["compliance_history"], rtol=1e-6, atol=0), comp / g["compliance_history"] - 1
    assert np.abs(res.rho_final - g["rho_final"]).max() < 1e-5


def test_probe_quality_matches_host(gpu):
    """probe_quality (multigrid_v4.py:821-832): the diagnostics after a solve
    equal the same quantities computed on the host from the V-cycle's output
    and the level-0 inverse diagonal."""
    dims = (16, 8, 8)
    free, F, ndof = cantilever_bc(dims, point_support=True)
    rng = np.random.default_rng(23)
    E = 1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    s = HipStateSolver(dims, free, F, n_levels=3, tol=1e-8, probe_quality=True)
    E64 = torch.from_numpy(E).to(DEV)
    s.set_modulus(E64, E64.float(), float(E.max()))
    assert s.solve().converged
    q = s.last_multigrid_quality
    F32 = h(s.F).astype(np.float32)
    s.r.copy_(torch.from_numpy(F32).to(DEV))
    z = h(s.gmg.apply(s.r, s.stream)).astype(np.float64)
    jac = (h(s.gmg.levels[0].dinv).astype(np.float32) * F32).astype(np.float64)
    assert q["finite"] and q["pd"] == bool(F32.astype(np.float64) @ z > 0.0) and q["pd"]
    assert q["z_norm"] == pytest.approx(np.linalg.norm(z), rel=1e-10)
    assert q["jacobi_norm"] == pytest.approx(np.linalg.norm(jac), rel=1e-10)
    assert q["z_over_jacobi"] == pytest.approx(np.linalg.norm(z) / np.linalg.norm(jac), rel=1e-10)


def test_warm_start_policy(gpu, monkeypatch):
    """solver_v4.py:162-164, 192-209: a penalty jump above 0.5 drops the warm
    start; a warm-started solve that does not converge is retried from zero."""
    from topogpu_amd.bc import generate_bc
    from topogpu_amd.hip.simp import HipSIMP

    prob = tg.unit_cantilever((16, 8, 4), volfrac=0.30, rmin=1.5)
    bc = generate_bc(prob.spec)
    eng = HipSIMP(prob.nel, bc.free_dofs, bc.F, volfrac=0.3, rmin=1.5, move=0.2, tol=1e-8, max_krylov=400)
    resets, runs = [], []
    reset, run = eng.fea.reset_warm_start, eng.fea._pcg_run
    monkeypatch.setattr(eng.fea, "reset_warm_start", lambda: (resets.append(1), reset())[1])
    monkeypatch.setattr(eng.fea, "_pcg_run", lambda zero: (runs.append(zero), run(zero))[1])
    eng.step()
    eng.penal = 3.5
    eng.step()
    assert resets == [] and runs == [False, False]
    eng.penal = 4.1
    eng.step()
    assert resets == [1] and eng.fea.last.converged
    eng.fea.max_krylov = 1                    # too small a budget: the warm start fails, then zero does
    runs.clear()
    eng.step()
    assert runs == [False, True] and not eng.fea.last.converged


@pytest.mark.parametrize("schedule", [["fp32", "bf16", "bf16"], ["bf16", "fp64"]])
def test_level_precisions_schedule(gpu, schedule):
    """level_precisions (multigrid_v4.py:673-682): level 0 takes the first
    entry, the coarse levels the rest (missing ones the most accurate kind);
    every schedule still solves to the direct solution."""
    from topogpu_amd.hip import lib as L

    dims = (24, 12, 6)
    # the face clamp keeps level 1 a q8 level; the point supports make it a
    # stencil level (its free mask is not the injection of the fine one)
    free, F, ndof = cantilever_bc(dims, point_support=schedule[0] == "bf16")
    rng = np.random.default_rng(24)
    E = 1e-9 + rng.uniform(1e-3, 1.0, int(np.prod(dims))) ** 3
    U_ref = oracle.solve_state(oracle.edof_table_3d(*dims), E, free, F, ndof)
    s = HipStateSolver(dims, free, F, n_levels=4, tol=1e-9, max_krylov=400, level_precisions=schedule)
    assert s.precision == schedule[0]
    bf16 = [len(schedule) > k and schedule[k] == "bf16" for k in range(1, 3)]
    if s.gmg.l1_q8:
        assert s.gmg.levels[1].kind == (L.TG_OP_Q8_BF16 if bf16[0] else L.TG_OP_Q8_F32)
    else:
        assert s.gmg.levels[1].kind == (L.TG_OP_STENCIL_BF16 if bf16[0] else L.TG_OP_STENCIL_F32)
    assert s.gmg.levels[2].kind == (L.TG_OP_STENCIL_BF16 if bf16[1] else L.TG_OP_STENCIL_F32)
    E64 = torch.from_numpy(E).to(DEV)
    s.set_modulus(E64, E64.float(), float(E.max()))
    st = s.solve()
    assert st.converged and st.rel_residual <= 1e-9
    assert rel_l2(h(s.u), U_ref) < 1e-6
    with pytest.raises(ValueError):
        HipStateSolver(dims, free, F, n_levels=4, level_precisions=["bf16", "fp16"])


def test_passive_regions_enforced_when_asked(gpu):
    """apply_passive=True enforces the passive mask after each OC update
    (apply_passive_mask, bc_generator.py:454-470); by default the mask is
    computed but, as in the reference loops, not applied."""
    from topogpu_amd.problem_spec import CircularRegion, RectangularRegion

    spec = tg.ProblemSpec(Lx=16.0, Ly=8.0, Lz=4.0, nelx=16, nely=8, nelz=4, volfrac=0.3,
                          supports=[tg.EdgeSupport("left", "fixed")],
                          loads=[tg.PointLoad(x=16.0, y=4.0, z=0.0, fz=-1.0)],
                          passive_regions=[CircularRegion(8.0, 4.0, 2.0, "void"),
                                           RectangularRegion(12.0, 0.0, 16.0, 2.0, "solid")])
    from topogpu_amd.bc import generate_bc
    from topogpu_amd.hip.simp import HipSIMP

    prob = tg.TopologyProblem("passive", spec, filter_radius=1.5)
    b = generate_bc(spec)
    mask = b.passive_mask
    assert (mask == 1).any() and (mask == 2).any()
    kw = dict(backend="hip", tol=1e-8, max_iter=3, move=0.2)
    res = tg.SIMPSolver(apply_passive=True, **kw).solve(prob)
    assert res.summary["n_iter"] == 3
    off = tg.SIMPSolver(**kw).solve(prob)
    eng = HipSIMP(prob.nel, b.free_dofs, b.F, volfrac=0.3, rmin=1.5, move=0.2, tol=1e-8, passive=b.passive_mask)
    rho0 = eng.rho.cpu().numpy()
    assert np.all(rho0[mask == 1] == 1e-3) and np.all(rho0[mask == 2] == 1.0)   # the start is clamped too
    step = eng.step()
    rho = eng.rho.cpu().numpy()
    assert np.all(rho[mask == 1] == 1e-3) and np.all(rho[mask == 2] == 1.0)
    assert step.change == np.abs(rho - rho0).max()          # the change of the clamped design
    assert not np.array_equal(res.rho_final, off.rho_final)
