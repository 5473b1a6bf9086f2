"""Public SIMP facade with the ``hip`` backend.

Mirrors ``SIMPSolver`` / ``OptimizationResult`` / ``solve`` of the reference
(src/topogpu/solver.py:39-253): same constructor signature, same validation
errors, same result object. ``backend="hip"`` runs the whole SIMP iteration
on an MI355X: the reference's CPU algorithm (``run_simp``, density filter +
Heaviside projection + filtered-volume OC, src/gpu_fem/pub_simp_solver.py:
412-657) with the state solve replaced by the GMG-PCG of ``topogpu_amd.hip``.
BCs and loads come from the problem itself (the reference's ``run_simp``
``bc_override`` path). With ``formulation="solverv4"`` it runs the loop of
the reference GPU backend instead (``_solve_solverv4``, src/topogpu/
solver.py:186-249): no filter or projection, mean-volume OC with
``rho_min``, exactly ``max_iter`` iterations, the design density as
``rho_final``. Other backend names are not served here: the reference keeps
serving them.
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

from .bc import generate_bc
from .evidence import EvidenceBundle
from .problem import TopologyProblem

HIP_BACKENDS = {"hip", "rocm"}


@dataclass(slots=True)
class SIMPModel:
    penal: float = 3.0
    rho_min: float = 1.0e-3


@dataclass(slots=True)
class OCOptimizer:
    move: float = 0.10


@dataclass(slots=True)
class OptimizationResult:
    problem: TopologyProblem
    rho_final: np.ndarray
    history: list[dict[str, Any]]
    summary: dict[str, Any]
    best_rho: np.ndarray | None = None   # best valid physical density (run_simp's best_rho)

    def save(self, path) -> EvidenceBundle:
        """Evidence folder (src/topogpu/solver.py:46-61)."""
        bundle = EvidenceBundle(Path(path))
        bundle.write_history(self.history)
        bundle.write_density(self.rho_final)
        bundle.write_render_metadata({"density_threshold": 0.5, "scalar": "density",
                                      "scalar_normalization": "per_case_0_1", "camera": "orthographic_default",
                                      "source": "OptimizationResult.save"})
        bundle.write_summary(self.summary)
        bundle.write_manifest()
        return bundle


def save_checkpoint(path: str, state: dict[str, Any]) -> None:
    """Arrays to <path>.npz, everything else to <path>.json, each written to
    a temporary name and renamed (_save_checkpoint, pub_simp_solver.py:
    397-403)."""
    arrays = {k: v for k, v in state.items() if isinstance(v, np.ndarray)}
    scalars = {k: v for k, v in state.items() if not isinstance(v, np.ndarray)}
    np.savez_compressed(path + ".tmp", **arrays)
    os.replace(path + ".tmp.npz", path + ".npz")
    with open(path + ".json.tmp", "w") as f:
        json.dump(scalars, f)
    os.replace(path + ".json.tmp", path + ".json")


def load_checkpoint(path: str) -> dict[str, Any] | None:
    """The state save_checkpoint wrote, or None (_load_checkpoint,
    pub_simp_solver.py:405-410). Arrays load without pickles."""
    if not (os.path.exists(path + ".npz") and os.path.exists(path + ".json")):
        return None
    with np.load(path + ".npz", allow_pickle=False) as z:
        data: dict[str, Any] = {k: z[k] for k in z.files}
    with open(path + ".json") as f:
        data.update(json.load(f))
    return data


def grayness(rho: np.ndarray) -> float:
    return float(4.0 * np.mean(rho * (1.0 - rho)))


class SIMPSolver:
    """Bounded SIMP loop; ``backend="hip"`` runs it on the GPU.

    Extra keyword options of the hip backend (all optional):
    precision ("bf16" | "bf16x3" | "fp32"), n_levels, fine_degree,
    coarse_degree, smoother ("chebyshev" | "jacobi"), omega, cycle ("v" |
    "w"), inner_rtol, use_heaviside, beta, min_iter, conv_tol,
    formulation ("filtered": the reference CPU algorithm, default;
    "solverv4": the reference GPU loop), weights (multi-load),
    checkpoint_dir and checkpoint_every (filtered formulation: every
    checkpoint_every iterations the design, history and iteration count go
    to <checkpoint_dir>/ckpt_<nx>x<ny>x<nz>_<problem name>.npz/.json, and a
    run that finds them there resumes from them, as run_simp does,
    pub_simp_solver.py:495-516, 597-603), callback (run_simp's controller,
    topogpu_amd.control), min_penal_for_best, max_gray_for_best,
    tail_default_iters, beta_max, seed, outer_solver ("pcg" | "fgmres": the
    state solve's outer Krylov method, SolverV4's gmg_outer_solver) and
    restart (FGMRES restart length, gmg_restart), apply_passive (enforce the
    spec's passive regions after each OC update, apply_passive_mask,
    bc_generator.py:454-470; off by default, as the reference loops leave
    them unapplied).
    """

    def __init__(self, backend: str = "hip", linear_solver: str = "auto", optimizer: str = "oc",
                 tol: float = 1.0e-5, max_krylov: int = 800, max_iter: int = 12, move: float = 0.10,
                 rho_min: float = 1.0e-3, **hip_options: Any) -> None:
        if optimizer != "oc":
            raise ValueError("Only optimizer='oc' is implemented in v0.1.0.")
        self.backend = backend
        self.linear_solver = linear_solver
        self.tol = tol
        self.max_krylov = max_krylov
        self.max_iter = max_iter
        self.move = move
        self.rho_min = rho_min
        self.hip_options = dict(hip_options)

    def solve(self, problem: TopologyProblem) -> OptimizationResult:
        problem.validate()
        if self.backend in HIP_BACKENDS:
            return self._solve_hip(problem)
        raise ValueError(f"Unknown backend {self.backend!r}.")

    def _solve_hip(self, problem: TopologyProblem) -> OptimizationResult:
        """run_simp (pub_simp_solver.py:412-657) with the step on the GPU:
        the main loop with best-design tracking, the optional controller,
        checkpoints and convergence test, then the optional tail phase and
        the final re-evaluation."""
        import torch

        from . import control as C
        from .hip.simp import HipSIMP

        if self.linear_solver not in {"auto", "pcg_gmg"}:
            raise ValueError(f"backend='hip' supports linear_solver 'auto' or 'pcg_gmg'; got {self.linear_solver!r}")
        spec = problem.spec
        if not problem.is_3d:
            raise ValueError("The hip backend currently requires a 3D problem.")
        opts = dict(self.hip_options)
        if opts.get("formulation", "filtered") == "solverv4":
            return self._solve_hip_solverv4(problem, opts)
        min_iter = int(opts.pop("min_iter", min(2, self.max_iter)))
        conv_tol = float(opts.pop("conv_tol", 1.0e-3))
        compliance_tol = float(opts.pop("compliance_tol", 5.0e-4))
        window = int(opts.pop("compliance_window", 5))
        ckpt_dir = opts.pop("checkpoint_dir", None)
        ckpt_every = int(opts.pop("checkpoint_every", 10))
        callback = opts.pop("callback", None)
        seed = opts.pop("seed", None)
        min_penal_best = float(opts.pop("min_penal_for_best", 3.0))
        max_gray_best = float(opts.pop("max_gray_for_best", 0.25))
        tail_default_iters = int(opts.pop("tail_default_iters", 18))
        beta_max = float(opts.pop("beta_max", 32.0))
        if seed is not None:
            np.random.seed(seed)
        bc = generate_bc(spec)
        force = bc.F if problem.force_override is None else np.asarray(problem.force_override, dtype=np.float64)
        rmin = problem.filter_radius or spec.rmin or 1.5
        passive = bc.passive_mask if opts.pop("apply_passive", False) else None
        engine = HipSIMP(problem.nel, bc.free_dofs, force, volfrac=spec.volfrac, rmin=rmin, move=self.move,
                         tol=self.tol, max_krylov=self.max_krylov, passive=passive, **opts)
        nx, ny, nz = problem.nel
        params = C.RunParams(nelx=nx, nely=ny, nelz=nz, volfrac=spec.volfrac, penal=engine.penal, rmin=rmin,
                             move=self.move, max_iter=self.max_iter, tol=conv_tol, compliance_tol=compliance_tol,
                             compliance_window=window, min_iter=min_iter, tail_default_iters=tail_default_iters,
                             use_heaviside=engine.use_heaviside, beta_init=engine.beta, beta_max=beta_max,
                             eta=engine.eta, min_penal_for_best=min_penal_best, max_gray_for_best=max_gray_best,
                             seed=seed, checkpoint_dir=ckpt_dir, checkpoint_every=ckpt_every)
        ctl = C.Controls(engine.penal, rmin, self.move, engine.beta)
        rho0 = engine.rho.cpu().numpy()
        if callback is not None and hasattr(callback, "initial_action"):
            init = callback.initial_action(params)
            if init:
                C.apply_action(init, ctl, rho0, rho0, False)
        best = C.BestTracker(rho0, grayness(rho0), min_penal_best, max_gray_best)

        def to_device(rho: np.ndarray) -> None:
            engine.rho.copy_(torch.from_numpy(np.ascontiguousarray(rho, dtype=np.float64)))

        def run_step(it: int, phase: str) -> tuple[Any, np.ndarray, float]:
            engine.penal, engine.move, engine.beta = ctl.penal, ctl.move, ctl.beta
            t0 = time.perf_counter()
            step = engine.step()
            wall = time.perf_counter() - t0
            comp_hist.append(step.compliance)
            rho_phys = engine.rho_phys.cpu().numpy()
            gray = grayness(rho_phys)
            history.append({
                "iteration": it, "compliance": float(step.compliance), "outer_iters": int(step.iters),
                "rho_mean": float(rho_phys.mean()), "rho_min": float(rho_phys.min()),
                "rho_max": float(rho_phys.max()), "grayness": gray, "wall_s": wall,
                "outer_solver": engine.fea.outer_solver, "change": float(step.change),
                "ir_restarts": int(step.restarts),
                "rel_residual": float(step.rel_residual), "converged": bool(step.converged),
            })
            params_log.append({"iter": it, "phase": phase, "penal": ctl.penal, "rmin": ctl.rmin, "move": ctl.move,
                               "beta": ctl.beta, "compliance": float(step.compliance),
                               "best_compliance": float(best.compliance), "grayness": gray,
                               "best_is_valid": best.is_valid,
                               "volume_fraction": float(rho_phys.sum() / rho_phys.size)})
            return step, rho_phys, gray

        history: list[dict[str, Any]] = []
        comp_hist: list[float] = []
        params_log: list[dict[str, Any]] = []
        t_total = time.perf_counter()
        start = 1
        ckpt = None
        if ckpt_dir is not None:
            os.makedirs(ckpt_dir, exist_ok=True)
            ckpt = os.path.join(ckpt_dir, f"ckpt_{nx}x{ny}x{nz}_{problem.name}")
            saved = load_checkpoint(ckpt)
            if saved is not None:
                # resume (pub_simp_solver.py:498-514): the design, the best
                # design, the controls, the history and, when saved, the
                # displacements the next solve starts from
                to_device(saved["rho"])
                if "u" in saved and saved["u"].size == engine.fea.u.numel():
                    engine.fea.u.copy_(torch.from_numpy(np.asarray(saved["u"], dtype=np.float64)))
                best.restore(saved)
                ctl.penal, ctl.rmin = float(saved["penal"]), float(saved["rmin"])
                ctl.move, ctl.beta = float(saved["move"]), float(saved["beta"])
                comp_hist = [float(c) for c in saved.get("compliance_hist", [])]
                history = list(saved.get("history", []))
                start = int(saved["iteration"]) + 1
        iteration = start - 1
        for iteration in range(start, self.max_iter + 1):
            if abs(ctl.rmin - engine.rmin) > 1e-6:
                engine.set_rmin(ctl.rmin)
            step, rho_phys, gray = run_step(iteration, "main")
            # the log row holds the best after this iteration's update
            # (pub_simp_solver.py:525-546, 576-581)
            best.update(iteration, step.compliance, rho_phys, gray, ctl.penal)
            params_log[-1]["best_compliance"] = float(best.compliance)
            params_log[-1]["best_is_valid"] = best.is_valid
            win = comp_hist[-window:]
            conv_c = len(win) == window and abs(win[-1] - win[0]) / max(abs(win[0]), 1e-10) < compliance_tol
            conv = step.change < conv_tol and conv_c
            if callback is not None:
                rho_new = engine.rho.cpu().numpy()
                state = C.step_state(iteration, comp_hist, window, conv, rho_phys, gray, (nx, ny, nz), best, ctl)
                action = callback(state, rho_new)
                if action:
                    restart = bool(action.get("restart", False)) and best.is_valid
                    rho_new = C.apply_action(action, ctl, rho_new, best.rho, best.is_valid)
                    if restart:
                        # the displacements belong to another design: the
                        # next solve starts from zero (pub_simp_solver.py:
                        # 588-589 invalidates the solver here)
                        engine.fea.reset_warm_start()
                # the next design is the controller's (it may also have
                # changed the array in place)
                to_device(rho_new)
                if action and action.get("stop", False) and iteration >= min_iter:
                    break
            if ckpt is not None and iteration % ckpt_every == 0:
                state = {"rho": engine.rho.cpu().numpy(), "u": engine.fea.u.cpu().numpy(),
                         "compliance_hist": np.array(comp_hist), "penal": ctl.penal, "rmin": ctl.rmin,
                         "move": ctl.move, "beta": ctl.beta, "iteration": iteration, "history": history}
                state.update(best.checkpoint())
                save_checkpoint(ckpt, state)
            if conv and iteration >= min_iter:
                break

        tail = C.tail_defaults(callback, params)
        tail_run = 0
        if tail.get("enabled", False) and tail.get("tail_iters", 0) > 0:
            # pub_simp_solver.py:612-633: restart from the best valid design
            # (clipped to [1e-3, 1]) or, with none, from the uniform start
            if tail.get("restart_from_best", True):
                to_device(np.clip(best.rho, 1e-3, 1.0) if best.is_valid else np.full(engine.n_elem, spec.volfrac))
            ctl.penal, ctl.rmin = float(tail.get("penal", ctl.penal)), float(tail.get("rmin", ctl.rmin))
            ctl.move, ctl.beta = float(tail.get("move", ctl.move)), float(tail.get("beta", ctl.beta))
            engine.set_rmin(ctl.rmin)
            engine.fea.reset_warm_start()             # pub_simp_solver.py:622
            for k in range(1, int(tail["tail_iters"]) + 1):
                run_step(iteration + k, "tail")
                tail_run = k
        # final re-evaluation with the last controls (the filter is the one
        # last built, pub_simp_solver.py:635-636)
        engine.penal, engine.beta = ctl.penal, ctl.beta
        final_c, rho_final, _ = engine.evaluate()
        summary = {
            "problem": problem.to_dict(), "backend": "hip",
            "linear_solver": "pcg_gmg" if self.linear_solver == "auto" else self.linear_solver,
            "precision": engine.fea.precision, "n_levels": len(engine.fea.gmg.levels),
            "n_iter": iteration + tail_run, "total_wall_s": time.perf_counter() - t_total,
            "final_compliance": float(final_c), "final_grayness": grayness(rho_final),
            "final": history[-1] if history else {},
        }
        summary.update(best.summary())
        if callback is not None:
            # plain Python values, so that the summary is JSON (a controller
            # may return numpy scalars)
            tail_cfg = {k: (v.item() if isinstance(v, np.generic) else v) for k, v in tail.items()}
            summary.update({"n_tail_iters": tail_run, "tail_config": tail_cfg, "params_log": params_log})
        return OptimizationResult(problem, rho_final, history, summary, best_rho=best.rho)


    def _solve_hip_solverv4(self, problem: TopologyProblem, opts: dict[str, Any]) -> OptimizationResult:
        """The reference GPU loop (solver.py:186-249) on the device: solve,
        mean-volume OC, one history row per iteration, max_iter iterations."""
        from .hip.simp import HipSIMP

        spec = problem.spec
        bc = generate_bc(spec)
        force = bc.F if problem.force_override is None else np.asarray(problem.force_override, dtype=np.float64)
        rmin = problem.filter_radius or spec.rmin or 1.5   # unused by the formulation; kept for the engine
        passive = bc.passive_mask if opts.pop("apply_passive", False) else None
        engine = HipSIMP(problem.nel, bc.free_dofs, force, volfrac=spec.volfrac, rmin=rmin, move=self.move,
                         rho_min=self.rho_min, tol=self.tol, max_krylov=self.max_krylov, passive=passive, **opts)
        history: list[dict[str, Any]] = []
        t_total = time.perf_counter()
        for iteration in range(1, self.max_iter + 1):
            t0 = time.perf_counter()
            step = engine.step()
            wall = time.perf_counter() - t0
            rho = engine.rho.cpu().numpy()
            history.append({
                "iteration": iteration, "compliance": float(step.compliance), "outer_iters": int(step.iters),
                "rho_mean": float(rho.mean()), "rho_min": float(rho.min()), "rho_max": float(rho.max()),
                "grayness": grayness(rho), "wall_s": wall, "outer_solver": engine.fea.outer_solver,
                "change": float(step.change), "ir_restarts": int(step.restarts),
                "rel_residual": float(step.rel_residual), "converged": bool(step.converged),
            })
        rho_final = engine.rho.cpu().numpy()
        summary = {
            "problem": problem.to_dict(), "backend": "hip",
            "linear_solver": "pcg_gmg" if self.linear_solver == "auto" else self.linear_solver,
            "formulation": "solverv4", "precision": engine.fea.precision, "n_levels": len(engine.fea.gmg.levels),
            "n_iter": self.max_iter, "total_wall_s": time.perf_counter() - t_total,
            "final": history[-1] if history else {},
        }
        return OptimizationResult(problem, rho_final, history, summary)


def solve(problem: TopologyProblem, **kwargs: Any) -> OptimizationResult:
    return SIMPSolver(**kwargs).solve(problem)
