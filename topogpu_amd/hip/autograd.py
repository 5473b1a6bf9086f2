"""Compliance of a design density as a differentiable torch function.

``HipCompliance(dims, free, F, rmin=...)(rho)`` returns the compliance
C(rho) = F . u with K(E(heaviside(H rho))) u = F, as a 0-dim FP64 tensor that
autograd can differentiate with respect to the design density ``rho``.

Forward runs the device path of one SIMP evaluation (filter, projection,
moduli, GMG-preconditioned state solve, compliance; pub_simp_solver.py:342-360).
Backward is the hand-written adjoint: compliance is self-adjoint (the adjoint
state equals u), so

    dC/drho = H^T ( -p (E0 - Emin) rho_phys^(p-1) (u_e^T KE u_e) * heaviside'(H rho) )

(pub_simp_solver.py:361-371), computed by the element sensitivity kernel and
the transposed filter stencil. The state u is not kept between forward and
backward: backward recomputes it from a zero start (``recompute=True``), so a
chain of many evaluations holds one state vector, not one per evaluation.

``HipCompliance.variational(rho)`` evaluates 2 F.u - u.K u, which equals
F.u* to second order in the solve error, with double-double sums; finite
differences of compliance are taken on it (``central_difference``).
"""

from __future__ import annotations

import torch

from . import lib as L
from .simp import HipSIMP


class HipCompliance:
    def __init__(self, dims, free, F, *, rmin: float = 1.5, penal: float = 3.0, use_heaviside: bool = True,
                 beta: float = 1.0, eta: float = 0.5, recompute: bool = True, **solver_kw):
        solver_kw.setdefault("tol", 1e-10)
        self.model = HipSIMP(dims, free, F, volfrac=0.5, rmin=rmin, penal=penal, use_heaviside=use_heaviside,
                             beta=beta, eta=eta, **solver_kw)
        self.recompute = recompute
        self.last_stats = None

    @property
    def device(self):
        return self.model.dev

    def __call__(self, rho: torch.Tensor) -> torch.Tensor:
        return _ComplianceFn.apply(rho, self)

    # ------------------------------------------------------------------
    def _solve(self, rho: torch.Tensor, zero_start: bool) -> float:
        m = self.model
        if rho.dtype != torch.float64 or rho.device != m.dev or rho.numel() != m.n_elem:
            raise ValueError("rho must be an FP64 tensor of n_elem entries on the solver's device")
        m.rho.copy_(rho.detach().reshape(-1))
        m.physical()
        if zero_start:
            L.call("tg_fill64", m.fea.nrhs * m.fea.n, L.ptr(m.fea.u), 0.0, m.stream)
        c, stats = m.state()
        self.last_stats = stats
        if not stats.converged:
            raise RuntimeError(f"state solve did not converge (rel. residual {stats.rel_residual:.2e})")
        return c

    def forward_value(self, rho: torch.Tensor) -> float:
        return self._solve(rho, zero_start=False)

    def gradient(self, rho: torch.Tensor) -> torch.Tensor:
        """dC/drho for the design rho (recomputes the state)."""
        self._solve(rho, zero_start=self.recompute)
        self.model.sensitivities()
        return self.model.dc.clone()

    def variational_parts(self, rho: torch.Tensor) -> tuple[float, float, float, float]:
        """After a solve at rho: (F.u, u^T K u) as double-double pairs
        (Fu_hi, Fu_lo, uKu_hi, uKu_lo); with several load cases, the weighted
        sums over the cases (exact for unit weights)."""
        self._solve(rho, zero_start=False)
        m = self.model
        fea = m.fea
        out = torch.zeros(fea.nrhs, 4, dtype=torch.float64, device=m.dev)
        for k in range(fea.nrhs):
            L.call("tg_compliance_dd", fea.grid, L.ptr(fea.E64), L.ptr(fea._slice(fea.u, k)),
                   L.ptr(fea._slice(fea.F, k)), L.ptr(out[k]), m.stream)
        parts = out.cpu().numpy() * m.weights_host[:, None]
        fu, uku = _dd_sum(parts[:, 0], parts[:, 1]), _dd_sum(parts[:, 2], parts[:, 3])
        return fu + uku

    def variational(self, rho: torch.Tensor) -> float:
        """2 F.u - u^T K u after a solve at rho."""
        fh, fl, kh, kl = self.variational_parts(rho)
        return (2.0 * fh - kh) + (2.0 * fl - kl)

    @staticmethod
    def central_difference(parts_plus, parts_minus, h: float) -> float:
        """(C(rho + h) - C(rho - h)) / 2h from two variational_parts results,
        differencing the double-double components separately so that no
        rounding of the O(1) compliance enters the O(h) difference."""
        fh_p, fl_p, kh_p, kl_p = parts_plus
        fh_m, fl_m, kh_m, kl_m = parts_minus
        d = 2.0 * (fh_p - fh_m) - (kh_p - kh_m) + (2.0 * (fl_p - fl_m) - (kl_p - kl_m))
        return d / (2.0 * h)


def _dd_sum(hi, lo) -> tuple[float, float]:
    """Sum of double-double numbers (hi_k, lo_k) as one double-double pair
    (Knuth's TwoSum on the high parts, low parts added to the error term)."""
    s, e = 0.0, 0.0
    for h, l in zip(hi.tolist(), lo.tolist()):
        t = s + h
        bv = t - s
        e += (s - (t - bv)) + (h - bv) + l
        s = t
    t = s + e
    return t, e - (t - s)


class _ComplianceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, comp):
        c = comp.forward_value(rho)
        ctx.comp = comp
        ctx.save_for_backward(rho)
        return torch.tensor(c, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, grad_c):
        (rho,) = ctx.saved_tensors
        g = ctx.comp.gradient(rho)
        return grad_c * g.reshape(rho.shape), None
