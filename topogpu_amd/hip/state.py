"""State solve K_ff u = F on the device: GMG-preconditioned flexible PCG in
FP32 inside FP64 iterative refinement.

Replaces ``SolverV4._solve_cupy_matfree`` (src/gpu_fem/solver_v4.py:147-269)
with ``_cupy_pcg`` (src/gpu_fem/solver_v2.py:552-624) as the Krylov method:

* outer loop (FP64): r = F - K u with the FP64 operator; stop when
  ||r|| <= tol ||F|| (the reference's criterion, checked on the true FP64
  residual instead of the recursive one);
* inner loop (FP32 vectors, FP64 dot products): PCG on K d = r with the
  level-0 operator in the chosen precision (see PRECISIONS) and the GMG
  V-cycle as preconditioner. BF16 rounding inside the operators makes the
  preconditioner slightly nonlinear, so the direction update uses the
  flexible (Polak-Ribiere) beta = z_new.(r_new - r_old)/(z.r);
* u += d, and the next restart recomputes the FP64 residual.

All Krylov scalars stay on the device; the host reads one scalar per
iteration for the convergence test. Vectors are full length with fixed DOFs
held at zero.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .. import fem
from . import graph
from . import lib as L
from .gmg import GMG, launch_op
from .readback import Mapped, Readback

# precision -> (Krylov operator, level-0 smoother operator, level-0
#               post-correction operator, level-1 post-correction operator,
#               level-1 smoother operator, coarse-level stencil storage)
#   bf16       : split BF16 MFMA (hi/lo operands, 3 passes) for the PCG
#                product, single-pass BF16 MFMA in the whole V-cycle (split
#                post-corrections gave the same iteration counts on the
#                160x80x80 workload at 2% more time per iteration), BF16
#                coarse-level stencil coefficients
#   bf16-plain : single-pass BF16 MFMA everywhere (the inner solve then
#                stalls near the BF16 rounding level; kept for study)
#   bf16x3     : split BF16 MFMA everywhere
#   fp32       : FP32 vector FMA everywhere (the plain FP32 path)
# The level-1 entries apply when level 1 is the implicit Galerkin (Q8) form.
# The post-correction product (r -= A c after the prolonged coarse
# correction c) and the power iteration for lambda_max use the
# post-correction operator of their level.
PRECISIONS = {
    "bf16": (L.TG_OP_HEX_BF16X3, L.TG_OP_HEX_BF16, L.TG_OP_HEX_BF16, L.TG_OP_Q8_BF16, L.TG_OP_Q8_BF16,
             L.TG_OP_STENCIL_BF16),
    "bf16-plain": (L.TG_OP_HEX_BF16, L.TG_OP_HEX_BF16, L.TG_OP_HEX_BF16, L.TG_OP_Q8_BF16, L.TG_OP_Q8_BF16,
                   L.TG_OP_STENCIL_BF16),
    "bf16x3": (L.TG_OP_HEX_BF16X3, L.TG_OP_HEX_BF16X3, L.TG_OP_HEX_BF16X3, L.TG_OP_Q8_BF16X3, L.TG_OP_Q8_BF16X3,
               L.TG_OP_STENCIL_F32),
    "fp32": (L.TG_OP_HEX_F32, L.TG_OP_HEX_F32, L.TG_OP_HEX_F32, L.TG_OP_Q8_F32, L.TG_OP_Q8_F32,
             L.TG_OP_STENCIL_F32),
}


@dataclass
class SolveStats:
    iters: int            # preconditioner applications (the reference's count)
    restarts: int         # FP64 residual evaluations
    converged: bool       # every right-hand side met its tolerance
    rel_residual: float   # max over right-hand sides of ||F - K u|| / ||F|| in FP64


class HipStateSolver:
    """Device state solver for one structured grid and one or more load cases.

    ``F`` is one load vector (n,) or a batch (nrhs, n). A batch is solved as
    nrhs independent PCG recurrences that share every launch: one operator
    product, one V-cycle and one set of vector kernels per iteration serve all
    load cases (the operator's moduli and the hierarchy are shared). Each load
    case keeps its own Krylov scalars and tolerance; the block iterates until
    every case has converged.
    """

    def __init__(self, dims, free: np.ndarray, F: np.ndarray, *, precision: str = "bf16", n_levels="auto",
                 fine_degree: int = 1, coarse_degree=(1, 2, 10, 4), cheb_lower_frac: float = 1.0 / 30.0,
                 cheb_safety: float = 1.1,
                 tol: float = 1e-5, max_krylov: int = 800, inner_rtol: float = 1e-2, max_restarts: int = 50,
                 warm_start: bool = True, level1: str = "q8", graphs: bool = True, speculate: bool = True,
                 smoother: str = "chebyshev", omega: float = 0.5, cycle: str = "v", outer_solver: str = "pcg",
                 restart: int = 32, probe_quality: bool = False, level_precisions=None, coarse_solver: str = "auto",
                 device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("backend='hip' needs a ROCm GPU (torch.cuda.is_available() is False)")
        if level_precisions is not None:
            # level 0's entry is the fine-level precision; the rest go to the
            # coarse levels (GalerkinMatFreeGMG level_precisions)
            level_precisions = list(level_precisions)
            if not level_precisions:
                raise ValueError("level_precisions must name at least level 0")
            precision = level_precisions[0]
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}; got {precision!r}")
        self.dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.dims = tuple(int(d) for d in dims)
        self.ke = fem.unit_stiffness()
        with torch.cuda.device(self.dev):
            L.init_device(self.ke, self.dev.index)
        self.grid = L.grid(self.dims)
        self.n_nodes = int(np.prod([d + 1 for d in self.dims]))
        self.n = 3 * self.n_nodes
        self.n_elem = int(np.prod(self.dims))
        mask = fem.node_mask_from_free(free, self.n_nodes)
        self.mask = torch.from_numpy(mask).to(self.dev)
        if n_levels == "auto":
            n_levels = fem.auto_levels(self.dims, mask)
        F = np.array(F, dtype=np.float64, ndmin=2)
        if F.ndim != 2 or F.shape[1] != self.n or not 1 <= F.shape[0] <= 64:
            raise ValueError(f"F must have shape ({self.n},) or (nrhs, {self.n}) with 1 <= nrhs <= 64")
        self.nrhs = int(F.shape[0])
        if outer_solver not in ("pcg", "fgmres"):
            raise ValueError(f"outer_solver must be 'pcg' or 'fgmres'; got {outer_solver!r}")
        if outer_solver == "fgmres" and self.nrhs != 1:
            raise ValueError("outer_solver='fgmres' solves one load case at a time (nrhs = 1)")
        if int(restart) < 1:
            raise ValueError("restart must be >= 1")
        self.outer_solver, self.restart = outer_solver, int(restart)
        self._krylov = None
        # one extra V-cycle after each solve for its diagnostics
        self.probe = bool(probe_quality)
        self.last_multigrid_quality: dict | None = None
        F[:, np.setdiff1d(np.arange(self.n), free)] = 0.0
        self.F = torch.from_numpy(F.ravel()).to(self.dev)
        self.bnorms = np.linalg.norm(F, axis=1)
        self.bnorm = float(self.bnorms.max())
        self.precision = precision
        self.tol, self.max_krylov = float(tol), int(max_krylov)
        self.inner_rtol, self.max_restarts = float(inner_rtol), int(max_restarts)
        self.warm_start = warm_start
        ops = PRECISIONS[precision]
        self.gmg = GMG(self.dims, mask, n_levels=n_levels, fine_op=ops[1], fine_accurate_op=ops[2],
                       l1_q8_op=ops[4], l1_q8_accurate_op=ops[3], fine_degree=fine_degree,
                       coarse_degree=coarse_degree, cheb_lower_frac=cheb_lower_frac, cheb_safety=cheb_safety,
                       level1=level1, nrhs=self.nrhs, coarse_stencil=ops[5], smoother=smoother, omega=omega,
                       cycle=cycle, coarse_precisions=None if level_precisions is None else level_precisions[1:],
                       coarse_solver=coarse_solver, device=self.dev)
        nb = self.nrhs * self.n
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.u = torch.zeros(nb, dtype=torch.float64, device=self.dev)
        self.r64 = torch.zeros(nb, dtype=torch.float64, device=self.dev)
        self.r = torch.zeros(nb, **f32)
        self.d = torch.zeros(nb, **f32)
        self.p = torch.zeros(nb, **f32)
        self.q = torch.zeros(nb, **f32)
        # device scalars, one row of nrhs per quantity:
        # 0 rz, 1 pq, 3 zq (rows 2, 4 and 5 unused: r.r, the FP64
        # residual norms and the compliances go to mapped host memory)
        self.s = torch.zeros(8, self.nrhs, dtype=torch.float64, device=self.dev)
        # host reads of device scalars through pinned memory: r.r in the
        # Krylov loop, the other residual norms and sums, the compliances
        # r.r of the Krylov loop: written by the update kernel straight into
        # pinned host memory (row 2 of s stays unused)
        self._rb_rr = Mapped(self.nrhs)
        self._rb = Readback(max(8 * self.nrhs, int(restart) + 2))
        self._rb_res = Mapped(self.nrhs)
        self._rb_c = Mapped(self.nrhs)
        self.E64 = None
        self.E32 = None
        self.last: SolveStats | None = None
        # one Krylov iteration as two captured launch sequences: the V-cycle's
        # first part ("begin") and the rest with the PCG updates ("finish",
        # "finish_first")
        self.graphs = bool(graphs)
        self._graphs: dict = {}
        # queue the next V-cycle's first part while the host waits for r.r
        self.speculate = bool(speculate)

    @property
    def stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _slice(self, t: torch.Tensor, k: int) -> torch.Tensor:
        return t[k * self.n:(k + 1) * self.n]

    # ------------------------------------------------------------------
    def fine_op(self) -> L.TgOp32:
        return L.TgOp32(PRECISIONS[self.precision][0], self.grid, L.ptr(self.mask), L.ptr(self.E32), None)

    def set_modulus(self, E64: torch.Tensor, E32: torch.Tensor, emax: float) -> None:
        """Element moduli for the next solve (FP64 for the true residual,
        FP32 for the low-precision operators); builds the GMG hierarchy."""
        self.E64, self.E32 = E64, E32
        self.gmg.setup(E64, E32, emax, self.stream)

    def residual64(self, speculate: bool = False) -> np.ndarray:
        """r64 = F - K u (FP64), r = float(r64) for every load case; returns
        the nrhs norms ||r64||. With speculate, the V-cycle's first part on
        the new r is queued before the host waits for the norms (the next
        call of _iteration may then pass begun=True)."""
        L.call("tg_apply64_b", self.grid, L.ptr(self.mask), L.ptr(self.E64), self.nrhs, L.ptr(self.u),
               L.ptr(self.F), L.ptr(self.r64), L.ptr(self.r), self._rb_res.device_ptr, self.stream)
        self._rb_res.start()
        if speculate:
            self._begin(self.stream)
        rr = self._rb_res.wait()
        self.gmg.check(block=False)
        return np.sqrt(np.maximum(rr, 0.0))

    def solve(self) -> SolveStats:
        stats = self._solve_fgmres() if self.outer_solver == "fgmres" else self._solve_pcg()
        self.gmg.join(self.stream)
        self.gmg.check()
        if self.probe:
            self.last_multigrid_quality = self.probe_quality()
        return stats

    def probe_quality(self) -> dict:
        """Diagnostics of the V-cycle as a preconditioner on F (probe_quality,
        multigrid_v4.py:821-832, called after every solve by solver_v4.py:
        232): z = M F finite, F.z > 0, ||z||, ||D^-1 F|| and their ratio.
        F enters the V-cycle in FP32 as the PCG residual does; the sums are
        FP64. With several load cases every value is a list, one per case."""
        st, n, nr = self.stream, self.n, self.nrhs
        if not hasattr(self, "_probe_s"):
            self._probe_s = torch.zeros(3, nr, dtype=torch.float64, device=self.dev)
            self._one = torch.ones(1, dtype=torch.float64, device=self.dev)
        ps = self._probe_s
        L.call("tg_normalize64", nr * n, L.ptr(self.F), L.ptr(self._one), None, L.ptr(self.r), st)   # r = float(F)
        z = self.gmg.apply(self.r, st)
        L.call("tg_dot32_b", n, nr, L.ptr(z), L.ptr(self.r), L.ptr(z), L.ptr(ps[0]), L.ptr(ps[1]), st)
        L.call("tg_cheb_init32_b", n, nr, L.ptr(self.r), L.ptr(self.gmg.levels[0].dinv), 1.0, L.ptr(self.p),
               L.ptr(self.q), None, st)                                                      # p = D^-1 F
        L.call("tg_dot32_b", n, nr, L.ptr(self.p), L.ptr(self.p), None, L.ptr(ps[2]), None, st)
        fz, zz, jj = self._rb.read(ps).reshape(3, nr)
        zn, jn = np.sqrt(np.maximum(zz, 0.0)), np.sqrt(np.maximum(jj, 0.0))
        q = {"finite": np.isfinite(zz), "pd": fz > 0.0, "z_norm": zn, "jacobi_norm": jn,
             "z_over_jacobi": zn / np.maximum(jn, 1e-30)}
        if nr == 1:
            return {k: (bool(v[0]) if v.dtype == bool else float(v[0])) for k, v in q.items()}
        return {k: [bool(x) if v.dtype == bool else float(x) for x in v] for k, v in q.items()}

    def reset_warm_start(self) -> None:
        """Start the next solve from zero (solver_v4.py:162-164 drops the
        previous solution when the penalty jumps)."""
        L.call("tg_fill64", self.nrhs * self.n, L.ptr(self.u), 0.0, self.stream)

    def _solve_pcg(self) -> SolveStats:
        """Flexible PCG in FP32 inside FP64 residual restarts; a warm start
        that does not converge within the budget is retried once from zero
        (solver_v4.py:192-209)."""
        stats = self._pcg_run(zero=not self.warm_start)
        if not stats.converged and self.warm_start:
            retry = self._pcg_run(zero=True)
            self.last = SolveStats(retry.iters, stats.restarts + retry.restarts, retry.converged,
                                   retry.rel_residual)
        return self.last

    def _pcg_run(self, zero: bool) -> SolveStats:
        st, n, nr = self.stream, self.n, self.nrhs
        if zero:
            L.call("tg_fill64", nr * n, L.ptr(self.u), 0.0, st)
        target = self.tol * self.bnorms
        its = restarts = 0
        # each FP64 residual queues the next restart's first V-cycle part
        # before the host reads it (wasted, harmlessly, after the last one)
        spec = self.speculate
        rnorm = self.residual64(spec)
        restarts += 1
        if not np.all(np.isfinite(rnorm)):
            raise FloatingPointError("non-finite residual at the start of the state solve")
        far = rnorm > 10.0 * self.bnorms
        if not zero and far.any():
            # a warm start whose residual exceeds ten times ||F|| is worse
            # than zero: drop it before spending iterations on it
            for k in np.flatnonzero(far):
                L.call("tg_fill64", n, L.ptr(self._slice(self.u, k)), 0.0, st)
            rnorm = self.residual64(spec)
            restarts += 1
        s = self.s
        while np.any(rnorm > target) and its < self.max_krylov and restarts <= self.max_restarts:
            inner_target = np.maximum(self.inner_rtol * rnorm, 0.5 * target)
            self._iteration(True, st, begun=spec)
            its += 1
            while True:
                # the next V-cycle's first part runs while the host waits for
                # r.r (wasted, harmlessly, when this was the last iteration)
                speculate = self.speculate
                rr = self._read_rr(st, speculate)
                if not np.all(np.isfinite(rr)):
                    raise FloatingPointError("non-finite inner residual")
                if np.all(np.sqrt(np.maximum(rr, 0.0)) <= inner_target) or its >= self.max_krylov:
                    break
                self._iteration(False, st, begun=speculate)
                its += 1
            L.call("tg_axpy_f32_to_f64", nr * n, L.ptr(self.u), L.ptr(self.d), st)
            rnorm = self.residual64(spec)
            restarts += 1
        rel = rnorm / np.maximum(self.bnorms, 1e-300)
        self.last = SolveStats(its, restarts, bool(np.all(rnorm <= target)), float(rel.max()))
        return self.last

    # ------------------------------------------------------------------
    def _solve_fgmres(self) -> SolveStats:
        """The reference's other outer solver: right-preconditioned restarted
        FGMRES in FP64 with the GMG V-cycle as the (variable) preconditioner
        (_cupy_fgmres, multigrid_v4.py:325-443, chosen by solver_v4.py:
        187-190, 214-233): from the warm start, and once more from zero when
        that does not converge. The Krylov vectors stay on the device; the
        modified Gram-Schmidt coefficients are device scalars consumed by the
        next update, and the host reads one Hessenberg column per step for
        the Givens rotations and the convergence test."""
        st, n = self.stream, self.n
        if self.bnorm <= 1e-30:
            L.call("tg_fill64", n, L.ptr(self.u), 0.0, st)
            self.last = SolveStats(0, 0, True, 0.0)
            return self.last
        if not self.warm_start:
            L.call("tg_fill64", n, L.ptr(self.u), 0.0, st)
        its, cycles, conv = self._fgmres_run()
        if not conv and self.warm_start:
            L.call("tg_fill64", n, L.ptr(self.u), 0.0, st)
            its, c2, conv = self._fgmres_run()
            cycles += c2
        rnorm = self.residual64()
        rel = rnorm / np.maximum(self.bnorms, 1e-300)
        self.last = SolveStats(its, cycles + 1, bool(conv), float(rel.max()))
        return self.last

    def _fgmres_run(self) -> tuple[int, int, bool]:
        st, n, m, tol = self.stream, self.n, self.restart, self.tol
        if self._krylov is None or self._krylov[0].shape[0] != m + 1:
            f64 = dict(dtype=torch.float64, device=self.dev)
            self._krylov = (torch.zeros(m + 1, n, **f64), torch.zeros(m, n, **f64), torch.zeros(n, **f64),
                            torch.zeros(m + 2, **f64), torch.zeros(m, **f64))
        V, Z, w, h, coef = self._krylov
        b_norm = self.bnorm
        total = cycles = 0
        converged = False
        while total < self.max_krylov:
            beta = float(self.residual64()[0])            # r64 = F - K u and its norm
            cycles += 1
            if not np.isfinite(beta):
                raise FloatingPointError("non-finite residual in FGMRES")
            if beta / b_norm <= tol:
                converged = True
                break
            L.call("tg_normalize64", n, L.ptr(self.r64), self._rb_res.device_ptr, L.ptr(V[0]), L.ptr(self.r), st)
            H = np.zeros((m + 1, m))
            cs, sn, g = np.zeros(m), np.zeros(m), np.zeros(m + 1)
            g[0] = beta
            j_final = 0
            for j in range(min(m, self.max_krylov - total)):
                z = self.gmg.apply(self.r, st)             # M V[j] (FP32 input, the copy of V[j])
                L.call("tg_axpy64", n, L.ptr(Z[j]), None, L.ptr(z), None, 1.0, 0, st)
                L.call("tg_apply64_b", self.grid, L.ptr(self.mask), L.ptr(self.E64), 1, L.ptr(Z[j]), None,
                       L.ptr(w), None, None, st)
                for i in range(j + 1):                     # modified Gram-Schmidt
                    L.call("tg_dot64_b", n, 1, L.ptr(V[i]), L.ptr(w), L.ptr(h[i:i + 1]), st)
                    L.call("tg_axpy64", n, L.ptr(w), L.ptr(V[i]), None, L.ptr(h[i:i + 1]), -1.0, 1, st)
                L.call("tg_dot64_b", n, 1, L.ptr(w), L.ptr(w), L.ptr(h[j + 1:j + 2]), st)
                L.call("tg_normalize64", n, L.ptr(w), L.ptr(h[j + 1:j + 2]), L.ptr(V[j + 1]), L.ptr(self.r), st)
                col = self._rb.read(h[:j + 2])
                H[:j + 1, j] = col[:j + 1]
                h_norm = float(np.sqrt(max(col[j + 1], 0.0)))
                H[j + 1, j] = h_norm
                for i in range(j):                         # previous Givens rotations
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                den = np.sqrt(H[j, j] ** 2 + H[j + 1, j] ** 2)
                cs[j], sn[j] = (1.0, 0.0) if den < 1e-300 else (H[j, j] / den, H[j + 1, j] / den)
                H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
                H[j + 1, j] = 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                total += 1
                j_final = j + 1
                if not np.isfinite(g[j + 1]):
                    raise FloatingPointError("non-finite FGMRES residual estimate")
                if abs(g[j + 1]) / b_norm <= tol or h_norm <= 1e-14:
                    converged = True
                    break
            if j_final:
                R = np.triu(H[:j_final, :j_final])
                try:
                    y = np.linalg.solve(R, g[:j_final])
                except np.linalg.LinAlgError:
                    y = np.linalg.lstsq(R, g[:j_final], rcond=None)[0]
                coef[:j_final].copy_(torch.from_numpy(np.ascontiguousarray(y, dtype=np.float64)))
                for k in range(j_final):                   # u += sum_k y_k Z[k]
                    L.call("tg_axpy64", n, L.ptr(self.u), L.ptr(Z[k]), None, L.ptr(coef[k:k + 1]), 1.0, 1, st)
            if converged:
                break
        return total, cycles, converged

    def _read_rr(self, st, speculate: bool) -> np.ndarray:
        """r.r of the last iteration on the host. With speculate, the next
        V-cycle's first part (gmg.begin, which touches only the hierarchy's
        buffers) is queued behind the copy before the host waits, so the GPU
        is not idle while the host reads the value and queues the rest."""
        self._rb_rr.start()
        if speculate:
            self._begin(st)
        rr = self._rb_rr.wait()
        self.gmg.check(block=False)
        return rr

    def _begin(self, st) -> None:
        """The V-cycle's first part on self.r (gmg.begin, which writes only
        the hierarchy's buffers), replayed or launched."""
        if self._replaying():
            self._replay("begin", st)
        else:
            self.gmg.begin(self.r, st)

    def _issue_iteration(self, first: bool, st, begun: bool = False) -> None:
        """One flexible-PCG iteration on stream st: z = M r, then
        first: d = 0, p = z;  later: p = z + beta p (beta = -(z.q)/(p.q));
        then q = A p and x, r updated with alpha = (r.z)/(p.q), rr = r.r.
        Scalars stay on the device (rows of self.s)."""
        n, nr, s = self.n, self.nrhs, self.s
        if first:
            L.call("tg_fill32", nr * n, L.ptr(self.d), 0.0, st)
        z = self.gmg.finish(st) if begun else self.gmg.apply(self.r, st)
        if first:
            L.call("tg_dot32_b", n, nr, L.ptr(z), L.ptr(self.r), None, L.ptr(s[0]), None, st)
            L.call("tg_copy32", nr * n, L.ptr(self.p), L.ptr(z), st)
        else:
            L.call("tg_dot32_b", n, nr, L.ptr(z), L.ptr(self.r), L.ptr(self.q), L.ptr(s[0]), L.ptr(s[3]), st)
            L.call("tg_pcg_dir_b", n, nr, L.ptr(self.p), L.ptr(z), L.ptr(s[3]), L.ptr(s[1]), st)
        ep = L.ep(L.TG_EP_DOT, dot=s[1])
        launch_op(self.gmg.timer, 0, self.fine_op(), L.ptr(self.p), L.ptr(self.q), ep, st, n, self.n_elem, nr)
        L.call("tg_pcg_xr_b", n, nr, L.ptr(self.d), L.ptr(self.r), L.ptr(self.p), L.ptr(self.q),
               L.ptr(s[0]), L.ptr(s[1]), self._rb_rr.device_ptr, st)

    def _replaying(self) -> bool:
        """Graph replay is on and no timer wants per-launch events."""
        timer = self.gmg.timer
        return self.graphs and not (timer is not None and timer.events)

    def _replay(self, name: str, st) -> None:
        """Replay the captured sequence `name` ("begin": the V-cycle's first
        part; "finish" / "finish_first": the rest of a later / first
        iteration), capturing it again when what it baked in changed (the
        hierarchy's coefficients and buffers, the modulus buffer)."""
        if name != "begin":
            self.gmg.join(st)     # outside the capture: the wait is not recorded
        key = (self.gmg.graph_key(), L.ptr(self.E32))
        g = self._graphs.get(name)
        if g is None or g.key != key:
            if g is not None:
                g.close()
            issue = {"begin": lambda cs: self.gmg.begin(self.r, cs),
                     "finish": lambda cs: self._issue_iteration(False, cs, begun=True),
                     "finish_first": lambda cs: self._issue_iteration(True, cs, begun=True)}[name]
            g = self._graphs[name] = graph.capture(key, issue, self.gmg, self.dev)
        g.launch(st)
        graph.account(self.gmg.timer, g)

    def _iteration(self, first: bool, st, begun: bool = False) -> None:
        """Issue one iteration (whose V-cycle's first part is already queued
        when begun): two graph replays, or direct launches when graphs are
        off or a timer records per-launch events."""
        if not self._replaying():
            self._issue_iteration(first, st, begun)
            return
        if not begun:
            self._replay("begin", st)
        self._replay("finish_first" if first else "finish", st)

    def compliances_start(self) -> Mapped:
        """Queue F_k . u_k in FP64 for every load case (solver_v4.py:244),
        written into mapped host memory; the returned reader's wait() gives
        the values, so the host may queue further work before it blocks."""
        L.call("tg_dot64_b", self.n, self.nrhs, L.ptr(self.F), L.ptr(self.u), self._rb_c.device_ptr, self.stream)
        return self._rb_c.start()

    def compliances(self) -> np.ndarray:
        """F_k . u_k in FP64 for every load case (solver_v4.py:244)."""
        return self.compliances_start().wait()

    def compliance(self) -> float:
        """F . u in FP64 (solver_v4.py:244); the sum over load cases of a batch."""
        return float(self.compliances().sum())
