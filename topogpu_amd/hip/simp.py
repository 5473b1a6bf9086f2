"""One SIMP design step on the device: filter -> projection -> modulus ->
state solve -> compliance sensitivity -> filter transpose -> OC bisection.

``formulation="filtered"`` (default) follows ``_fea_and_sensitivity`` and
``_oc_update`` of the reference CPU path (src/gpu_fem/pub_simp_solver.py:
319-377): density filter, Heaviside projection, OC on the filtered volume.
``formulation="solverv4"`` follows the reference GPU loop
(src/topogpu/solver.py:94-117, 186-249): the design density is the physical
density (no filter, no projection), dc is SolverV4's element sensitivity
(solver_v4.py:246-251) and the OC bisects on the plain mean volume with the
``rho_min`` bound. Everything except the state solve's convergence test and a
few scalars stays on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import fem
from . import lib as L
from .readback import Readback
from .state import HipStateSolver

E0, EMIN = 1.0, 1.0e-9


@dataclass
class StepResult:
    compliance: float
    change: float
    iters: int
    restarts: int
    rel_residual: float
    converged: bool
    oc_steps: int


class HipSIMP:
    def __init__(self, dims, free, F, *, volfrac: float, rmin: float = 1.5, penal: float = 3.0,
                 move: float = 0.2, use_heaviside: bool = True, beta: float = 1.0, eta: float = 0.5,
                 rho0: np.ndarray | None = None, weights=None, formulation: str = "filtered",
                 rho_min: float = 1.0e-3, passive: np.ndarray | None = None, **solver_kw):
        if formulation not in ("filtered", "solverv4"):
            raise ValueError(f"formulation must be 'filtered' or 'solverv4'; got {formulation!r}")
        self.formulation, self.rho_min = formulation, float(rho_min)
        self.dims = tuple(int(d) for d in dims)
        self.n_elem = int(np.prod(self.dims))
        self.volfrac, self.rmin, self.penal, self.move = float(volfrac), float(rmin), float(penal), float(move)
        self.use_heaviside, self.beta, self.eta = bool(use_heaviside), float(beta), float(eta)
        self._prev_penal: float | None = None
        # passive elements (0 free, 1 void, 2 solid) enforced after each OC
        # update; None: every element is designable
        pm = None
        if passive is not None:
            pm = np.ascontiguousarray(passive, dtype=np.int32).ravel()
            if pm.size != self.n_elem or not np.isin(pm, (0, 1, 2)).all():
                raise ValueError("passive must hold 0, 1 or 2 per element")
        self.fea = HipStateSolver(self.dims, free, F, **solver_kw)
        self.dev = self.fea.dev
        self.passive = None if pm is None else torch.from_numpy(pm).to(self.dev)
        # multi-load: the objective is sum_k w_k F_k.u_k (equal weights of 1
        # by default, so it is the sum of the single-load compliances)
        w = np.ones(self.fea.nrhs) if weights is None else np.asarray(weights, dtype=np.float64).ravel()
        if w.size != self.fea.nrhs:
            raise ValueError(f"weights must have one entry per load case ({self.fea.nrhs})")
        self.weights_host = w.copy()
        self.weights = torch.from_numpy(w.copy()).to(self.dev)
        self.grid = self.fea.grid
        self.set_rmin(self.rmin)
        f64 = dict(dtype=torch.float64, device=self.dev)
        rho = np.full(self.n_elem, self.volfrac) if rho0 is None else np.asarray(rho0, dtype=np.float64)
        rho = rho.copy()
        if pm is not None:
            # the passive elements start at their bounds, as after every update
            rho[pm == 1] = self._passive_lo()
            rho[pm == 2] = 1.0
        self.rho = torch.from_numpy(rho).to(self.dev)
        self.rho_new = torch.zeros(self.n_elem, **f64)
        self.rho_f = torch.zeros(self.n_elem, **f64)
        self.rho_phys = torch.zeros(self.n_elem, **f64)
        self.dh = torch.zeros(self.n_elem, **f64)
        self.E64 = torch.zeros(self.n_elem, **f64)
        self.E32 = torch.zeros(self.n_elem, dtype=torch.float32, device=self.dev)
        self.dcf = torch.zeros(self.n_elem, **f64)
        self.dc = torch.zeros(self.n_elem, **f64)
        self.dv = torch.zeros(self.n_elem, **f64)
        self.q = torch.zeros(self.n_elem, **f64)
        self.lo = torch.zeros(self.n_elem, **f64)
        self.hi = torch.zeros(self.n_elem, **f64)
        self.oc_state = torch.zeros(max(L.OC_STATE_LEN, L.OC_MEAN_STATE_LEN), **f64)
        self.scal = torch.zeros(4, **f64)
        self._rb = Readback(4)            # OC decisions, step count and change
        self._rb_emax = Readback(1)

    def set_rmin(self, rmin: float) -> None:
        """(Re)build the density filter for radius rmin: its offsets and
        weights, and the inverse row sums 1 / Hs."""
        self.rmin = float(rmin)
        offs, w = fem.filter_stencil(self.rmin)
        self.offs = np.ascontiguousarray(offs.ravel(), dtype=np.int32)
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.m = int(w.size)
        self.hs_inv = torch.from_numpy(1.0 / self._row_sums()).to(self.dev)

    def _row_sums(self) -> np.ndarray:
        """Hs = W 1 on the host (geometry only)."""
        nx, ny, nz = self.dims
        hs = np.zeros((nx, ny, nz))
        for (dx, dy, dz), wk in zip(self.offs.reshape(-1, 3), self.w):
            cnt = np.zeros((nx, ny, nz))
            sx = slice(max(0, -dx), nx - max(0, dx))
            sy = slice(max(0, -dy), ny - max(0, dy))
            sz = slice(max(0, -dz), nz - max(0, dz))
            cnt[sx, sy, sz] = 1.0
            hs += wk * cnt
        return hs.ravel()

    @property
    def stream(self):
        return self.fea.stream

    # ------------------------------------------------------------------
    def filter(self, x, out, transpose: bool = False) -> None:
        """out = H x (or H^T x)."""
        ins = self.hs_inv if transpose else None
        outs = None if transpose else self.hs_inv
        L.call("tg_stencil64", self.grid, self.m, self.offs.ctypes.data, self.w.ctypes.data, L.ptr(x),
               L.ptr(ins), L.ptr(outs), L.ptr(out), self.stream)

    def physical(self, rho=None) -> None:
        """rho_f = H rho; rho_phys, dh, E (pub_simp_solver.py:347-350); with
        the solverv4 formulation rho_phys = rho and E = Emin + (E0 - Emin)
        rho^p (solver_v4.py:167-168)."""
        rho = self.rho if rho is None else rho
        if self.formulation == "solverv4":
            L.call("tg_density_chain", self.n_elem, L.ptr(rho), 0.0, self.eta, self.penal, EMIN, E0,
                   L.ptr(self.rho_phys), L.ptr(self.dh), L.ptr(self.E64), L.ptr(self.E32), self.stream)
            return
        self.filter(rho, self.rho_f)
        beta = self.beta if self.use_heaviside else 0.0
        L.call("tg_density_chain", self.n_elem, L.ptr(self.rho_f), beta, self.eta, self.penal, EMIN, E0,
               L.ptr(self.rho_phys), L.ptr(self.dh), L.ptr(self.E64), L.ptr(self.E32), self.stream)

    def state(self):
        """Set moduli, build the hierarchy and solve; returns (compliance,
        stats). A penalty change above 0.5 since the last solve drops the
        warm start (solver_v4.py:162-164)."""
        stats = self._solve()
        return float(self.weights_host @ self.fea.compliances()), stats

    def _solve(self):
        """Moduli to the hierarchy, then the state solve. max E reaches the
        host through a copy queued before the hierarchy's setup, which
        needs it only at its end (whether to re-estimate lambda_max), so the
        host does not wait on the GPU for it."""
        if self._prev_penal is not None and abs(self.penal - self._prev_penal) > 0.5:
            self.fea.reset_warm_start()
        self._prev_penal = self.penal
        emax = self._rb_emax.start(self.E64.max().reshape(1))
        self.fea.set_modulus(self.E64, self.E32, lambda: float(emax.wait()[0]))
        return self.fea.solve()

    def sensitivities(self) -> None:
        """dc = H^T(dc_phys * dh), dv = H^T dh (pub_simp_solver.py:361-371);
        with several load cases dc_phys is the weighted sum over the cases."""
        if self.formulation == "solverv4":
            # dc = -p (E0 - Emin) rho^(p-1) u_e^T KE u_e (solver_v4.py:246-251)
            L.call("tg_sensitivity64_b", self.grid, self.fea.nrhs, L.ptr(self.fea.u), L.ptr(self.weights),
                   L.ptr(self.rho_phys), None, self.penal, E0 - EMIN, None, L.ptr(self.dc), self.stream)
            return
        L.call("tg_sensitivity64_b", self.grid, self.fea.nrhs, L.ptr(self.fea.u), L.ptr(self.weights),
               L.ptr(self.rho_phys), L.ptr(self.dh), self.penal, E0 - EMIN, None, L.ptr(self.dcf), self.stream)
        self.filter(self.dcf, self.dc, transpose=True)
        self.filter(self.dh, self.dv, transpose=True)

    def oc_mean(self) -> tuple[float, int]:
        """Mean-volume OC of the reference GPU loop (solver.py:94-117)."""
        st = self.stream
        L.call("tg_oc_mean_prepare", L.ptr(self.oc_state), st)
        # one doubling pass, then six bisection steps per pass (80 at most);
        # calls after the stop return at once, so the next four are queued
        # before the host reads whether the last four stopped it
        def queue(k):
            for _ in range(k):
                L.call("tg_oc_mean_step", self.n_elem, L.ptr(self.rho), L.ptr(self.dc), self.volfrac, self.move,
                       self.rho_min, L.ptr(self.oc_state), st)
        queue(4)
        passes = 4
        while True:
            self._rb.start(self.oc_state[7:8])
            ahead = 4 if passes < 40 else 0
            queue(ahead)
            if self._rb.wait()[0] or passes >= 40:
                break
            passes += ahead
        L.call("tg_oc_mean_finalize", self.n_elem, L.ptr(self.rho), L.ptr(self.dc), self.move, self.rho_min,
               L.ptr(self.oc_state), L.ptr(self.rho_new), L.ptr(self.scal[0:1]), st)
        n_steps, change = self._rb.read(self.oc_state[3:4], self._apply_passive(self.rho_new))
        n_steps, change = int(n_steps), float(change)
        self.rho, self.rho_new = self.rho_new, self.rho
        return change, n_steps

    def oc(self, max_steps: int = 100) -> tuple[float, int]:
        """Filtered-volume OC bisection (pub_simp_solver.py:319-340); the
        mean-volume update with the solverv4 formulation."""
        if self.formulation == "solverv4":
            return self.oc_mean()
        st = self.stream
        L.call("tg_oc_prepare", self.n_elem, L.ptr(self.rho), L.ptr(self.dc), L.ptr(self.dv), self.move,
               L.ptr(self.q), L.ptr(self.lo), L.ptr(self.hi), L.ptr(self.oc_state), st)
        use_h = int(self.use_heaviside and self.beta > 1e-6)
        # each call evaluates one batch of multipliers and advances the
        # bisection by at least two steps (usually far more; from the second
        # iteration on, two or three calls finish it); calls after
        # convergence return at once, so two are queued first and then one
        # more before each read of whether those before them converged
        def queue(k):
            for _ in range(k):
                L.call("tg_oc_bisect_step", self.grid, self.m, self.offs.ctypes.data, self.w.ctypes.data,
                       L.ptr(self.hs_inv), L.ptr(self.q), L.ptr(self.lo), L.ptr(self.hi), self.beta, self.eta,
                       self.volfrac, use_h, L.ptr(self.oc_state), st)
        limit = (max_steps + 1) // 2
        queue(2)
        passes = 2
        while True:
            self._rb.start(self.oc_state[3:4])
            ahead = 1 if passes < limit else 0
            queue(ahead)
            if self._rb.wait()[0] or passes >= limit:
                break
            passes += ahead
        L.call("tg_oc_finalize", self.n_elem, L.ptr(self.rho), L.ptr(self.q), L.ptr(self.lo), L.ptr(self.hi),
               L.ptr(self.oc_state), L.ptr(self.rho_new), L.ptr(self.scal[0:1]), st)
        n_steps, change = self._rb.read(self.oc_state[4:5], self._apply_passive(self.rho_new))
        n_steps, change = int(n_steps), float(change)
        self.rho, self.rho_new = self.rho_new, self.rho
        return change, n_steps

    def _passive_lo(self) -> float:
        return self.rho_min if self.formulation == "solverv4" else 1.0e-3

    def _apply_passive(self, rho_new) -> torch.Tensor:
        """Void elements to the lower density bound, solid ones to 1
        (apply_passive_mask, bc_generator.py:454-470; rho_min = 1e-3 there,
        the filtered OC's bound). Returns, on the device (one element),
        max |rho_new - rho| of the design the next iteration uses: the
        finalize kernel's value, or, with passive elements, recomputed after
        the clamp."""
        if self.passive is None:
            return self.scal[0:1]
        L.call("tg_passive64", self.n_elem, L.ptr(self.passive), L.ptr(rho_new), self._passive_lo(), self.stream)
        return (rho_new - self.rho).abs().max().reshape(1)

    def step(self) -> StepResult:
        self.physical()
        stats = self._solve()
        # the compliance is read once the OC has needed the host anyway
        comp = self.fea.compliances_start()
        self.sensitivities()
        change, n_oc = self.oc()
        c = float(self.weights_host @ comp.wait())
        return StepResult(c, change, stats.iters, stats.restarts, stats.rel_residual, stats.converged, n_oc)

    def evaluate(self):
        """Final re-evaluation (pub_simp_solver.py:635-641): compliance and the
        physical density of the current design."""
        self.physical()
        c, stats = self.state()
        return c, self.rho_phys.cpu().numpy(), stats
