"""Continuation control of the SIMP loop: the per-iteration state a
controller sees, the actions it may return, best-design tracking and the
tail phase, as ``run_simp`` has them (src/gpu_fem/pub_simp_solver.py:
115-135, 263-270, 379-395, 470-633).

A controller (``callback``) is any callable ``callback(state, rho_new)``
returning ``None`` or an action dict with optional keys ``penal``, ``rmin``,
``move``, ``beta`` (each clipped to its range), ``restart`` (replace the next
design by the best valid physical density) and ``stop`` (end the main loop,
honoured from ``min_iter`` on). It may also define ``initial_action(params)``
(an action applied before the first iteration) and ``finalize_tail(params)``
(overrides of the tail settings, ``tail_defaults``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

# a best design is only recorded from this iteration on while no valid one
# exists (early compliance spikes), pub_simp_solver.py:480
MIN_ITER_FOR_FALLBACK = 10


@dataclass
class RunParams:
    """The run parameters a controller's initial_action / finalize_tail
    receive: the fields of SIMPParams (pub_simp_solver.py:89-113)."""
    nelx: int
    nely: int
    nelz: int
    volfrac: float
    penal: float = 3.0
    rmin: float = 1.5
    move: float = 0.2
    max_iter: int = 100
    tol: float = 0.01
    compliance_tol: float = 5e-4
    compliance_window: int = 5
    min_iter: int = 50
    tail_default_iters: int = 18
    use_heaviside: bool = True
    beta_init: float = 1.0
    beta_max: float = 32.0
    eta: float = 0.5
    min_penal_for_best: float = 3.0
    max_gray_for_best: float = 0.25
    seed: int | None = None
    checkpoint_dir: str | None = None
    checkpoint_every: int = 10


@dataclass
class StepState:
    """What a controller sees after each main-loop iteration
    (pub_simp_solver.py:115-135)."""
    iteration: int
    compliance: float
    best_compliance: float
    best_iteration: int
    volume_fraction: float
    grayness: float
    best_grayness: float
    checkerboard: float
    obj_slope: float
    rel_change_1: float
    rel_change_5: float
    stagnation_counter: int
    penal: float
    rmin: float
    move: float
    beta: float
    converged: bool
    best_is_valid: bool
    compliance_history: list[float] = field(default_factory=list)


@dataclass
class Controls:
    """The four continuation parameters an action may change."""
    penal: float
    rmin: float
    move: float
    beta: float


def apply_action(action: dict[str, Any], ctl: Controls, rho_new: np.ndarray, best_rho: np.ndarray,
                 best_is_valid: bool) -> np.ndarray:
    """Apply an action to ctl in place; returns the next design
    (_apply_action, pub_simp_solver.py:379-386): a restart takes a copy of
    the best valid design, and each parameter is clipped to its range."""
    if action.get("restart", False) and best_is_valid:
        rho_new = best_rho.copy()
    if "penal" in action:
        ctl.penal = float(np.clip(action["penal"], 1.0, 5.0))
    if "rmin" in action:
        ctl.rmin = float(np.clip(action["rmin"], 1.1, 4.0))
    if "move" in action:
        ctl.move = float(np.clip(action["move"], 0.03, 0.4))
    if "beta" in action:
        ctl.beta = float(np.clip(action["beta"], 1.0, 64.0))
    return rho_new


def tail_defaults(callback, params: RunParams) -> dict[str, Any]:
    """Tail settings (_tail_defaults, pub_simp_solver.py:388-395): off unless
    the controller's finalize_tail turns it on."""
    tail = {"enabled": False, "tail_iters": params.tail_default_iters, "restart_from_best": True,
            "penal": 4.5, "rmin": 1.2, "move": 0.05, "beta": min(params.beta_max, 32.0)}
    if callback is not None and hasattr(callback, "finalize_tail"):
        custom = callback.finalize_tail(params)
        if custom:
            tail.update(custom)
    return tail


def checkerboard_3d(rho: np.ndarray, nelx: int, nely: int, nelz: int) -> float:
    """Standard deviation of the 2x2x2 block means over the even-aligned
    blocks (_checkerboard_3d, pub_simp_solver.py:263-270); 0 when an axis
    has fewer than two elements."""
    if nelx < 2 or nely < 2 or nelz < 2:
        return 0.0
    r = rho.reshape(nelx, nely, nelz)
    bx, by, bz = nelx // 2, nely // 2, nelz // 2
    blocks = r[:2 * bx, :2 * by, :2 * bz].reshape(bx, 2, by, 2, bz, 2)
    return float(np.std(blocks.mean(axis=(1, 3, 5))))


class BestTracker:
    """Best valid design and the pre-valid fallback (pub_simp_solver.py:
    470-480, 525-546): a design counts as valid once penal reaches
    min_penal_for_best and its grayness is below max_gray_for_best."""

    def __init__(self, rho: np.ndarray, grayness: float, min_penal: float, max_gray: float):
        self.min_penal, self.max_gray = float(min_penal), float(max_gray)
        self.rho, self.compliance, self.iteration = rho.copy(), np.inf, 0
        self.is_valid, self.grayness = False, float(grayness)
        self.pre_rho, self.pre_compliance, self.pre_iteration = rho.copy(), np.inf, 0
        self.pre_grayness = float(grayness)
        self.stagnation = 0

    def update(self, iteration: int, compliance: float, rho_phys: np.ndarray, gray: float, penal: float) -> None:
        gate_ok = penal >= self.min_penal and gray < self.max_gray
        if compliance < self.compliance and gate_ok:
            self.compliance, self.rho, self.iteration = compliance, rho_phys.copy(), iteration
            self.grayness, self.is_valid, self.stagnation = gray, True, 0
            return
        if compliance < self.pre_compliance and not self.is_valid and iteration >= MIN_ITER_FOR_FALLBACK:
            self.pre_compliance, self.pre_rho = compliance, rho_phys.copy()
            self.pre_iteration, self.pre_grayness = iteration, gray
        if compliance >= self.compliance or not gate_ok:
            self.stagnation += 1

    def restore(self, ckpt: dict[str, Any]) -> None:
        """State from a checkpoint (pub_simp_solver.py:499-509): the
        pre-valid fallback restarts from the saved best."""
        self.rho = np.asarray(ckpt["best_rho"], dtype=np.float64)
        self.compliance = float(ckpt["best_compliance"])
        self.iteration = int(ckpt["best_iteration"])
        self.is_valid = bool(int(ckpt["best_is_valid"]))
        self.pre_compliance, self.pre_iteration = self.compliance, self.iteration
        self.pre_grayness, self.pre_rho = self.grayness, self.rho.copy()
        self.grayness = float(ckpt["best_grayness"])

    def checkpoint(self) -> dict[str, Any]:
        return {"best_rho": self.rho, "best_compliance": self.compliance, "best_iteration": self.iteration,
                "best_is_valid": int(self.is_valid), "best_grayness": self.grayness}

    def summary(self) -> dict[str, Any]:
        return {"best_compliance": float(self.compliance), "best_iteration": int(self.iteration),
                "best_grayness": float(self.grayness), "best_is_valid": bool(self.is_valid),
                "pre_valid_best_compliance": float(self.pre_compliance),
                "pre_valid_best_iteration": int(self.pre_iteration),
                "pre_valid_best_grayness": float(self.pre_grayness)}


def step_state(iteration: int, compliance_hist: list[float], window: int, converged: bool, rho_phys: np.ndarray,
               gray: float, dims, best: BestTracker, ctl: Controls) -> StepState:
    """The StepState of one iteration (pub_simp_solver.py:554-575)."""
    win = compliance_hist[-window:]
    slope = (win[-1] - win[0]) / max(len(win) - 1, 1)
    h = compliance_hist
    rel1 = 0.0 if len(h) < 2 else (h[-1] - h[-2]) / max(abs(h[-2]), 1e-10)
    rel5 = 0.0 if len(h) < 6 else (h[-1] - h[-6]) / max(abs(h[-6]), 1e-10)
    return StepState(iteration=iteration, compliance=h[-1], best_compliance=best.compliance,
                     best_iteration=best.iteration, volume_fraction=float(rho_phys.sum() / rho_phys.size),
                     grayness=gray, best_grayness=best.grayness, checkerboard=checkerboard_3d(rho_phys, *dims),
                     obj_slope=slope, rel_change_1=rel1, rel_change_5=rel5, stagnation_counter=best.stagnation,
                     penal=ctl.penal, rmin=ctl.rmin, move=ctl.move, beta=ctl.beta, converged=converged,
                     best_is_valid=best.is_valid, compliance_history=list(h))
