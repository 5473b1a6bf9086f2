"""Galerkin geometric multigrid on the device (V-cycle preconditioner).

Restates the hierarchy of ``GalerkinMatFreeGMG`` (src/gpu_fem/multigrid_v4.py:
446-1042) on gfx950 kernels:

* level 0 is the matrix-free hex8 operator (BF16 MFMA or FP32);
* level 1 is the exact Galerkin operator P^T K P; when the fine dims are even
  and the Dirichlet masks are aligned with the coarse grid its element
  matrices are sum_s E_child(s) * Q_s and are applied implicitly from the
  fine E (no stored matrices on the V-cycle path), otherwise stored;
* levels >= 2 are Galerkin element matrices built on the device from the
  level above (replacing the SpGEMM R @ (K @ P), multigrid_v4.py:747-752),
  assembled into 27-point block stencils for the V-cycle (FP32
  coefficients, or BF16 ones: these levels are bound by reading them, and
  BF16 coefficients leave the PCG iteration count within 1 % on the
  160x80x80 workload);
* the coarsest level is solved with an explicit FP64 dense inverse computed
  once per setup (Gauss-Jordan elimination; it replaces the dense Cholesky
  factorisation of multigrid_v4.py:774-811).

Smoothing is Chebyshev-Jacobi on [lambda_max * lower_frac, lambda_max] with
the coefficient recurrence of ``_chebyshev_steps`` (multigrid_v4.py:860-889),
written in residual-update form: the residual is carried along
(r <- r - A d) and each operator application acts on the smoother step d,
never on the full iterate. A low-precision product then errs relative to
the step rather than to the whole (smooth, large) iterate, and the
pre-smoother hands its residual to the restriction without another product.
Each step is one operator launch with the update fused into its epilogue.
The one smooth product of the cycle, A * (P e_c) after the coarse
correction, uses the accurate operator of the level. Coarse levels compute
in FP32.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import fem
from . import lib as L
from .readback import Readback


class _Level:
    pass


# Coarsest levels above BCR_MIN_DOFS free DOFs go to block cyclic reduction:
# on 160x80x80, 5 levels (1,080 coarsest DOFs) run at 9.60 ms per SIMP
# iteration with it against 10.29 ms with the dense inverse (whose setup is
# ~70 pivot-step launches); 6 levels (135 DOFs) favour the dense inverse,
# 9.62 against 9.78 ms. A plane block goes through the blocked Gauss-Jordan
# inverse, hence its size limit.
BCR_MIN_DOFS = 1024
BCR_MAX_PLANE = 2048
# largest coarsest level inverted by tg_dense_spd_inverse_async (one workgroup)
ASYNC_INVERSE_MAX = 140

KERNEL_NAMES = {L.TG_OP_HEX_F32: "hex_f32", L.TG_OP_HEX_BF16: "hex_bf16_mfma",
                L.TG_OP_HEX_BF16X3: "hex_bf16x3_mfma", L.TG_OP_Q8_F32: "galerkin_q8_f32",
                L.TG_OP_STORED_F32: "stored_f32", L.TG_OP_STENCIL_F32: "stencil_f32", L.TG_OP_STENCIL_BF16: "stencil_bf16",
                L.TG_OP_Q8_BF16: "galerkin_q8_bf16_mfma", L.TG_OP_Q8_BF16X3: "galerkin_q8_bf16x3_mfma"}

_EP_VEC_TRAFFIC = {L.TG_EP_STORE: 4, L.TG_EP_DOT: 4, L.TG_EP_RESID: 8, L.TG_EP_RESID_NORM: 8, L.TG_EP_CHEB: 20,
                   L.TG_EP_CHEB_R: 24, L.TG_EP_RESID_UPDATE: 8, L.TG_EP_CHEB_FIRST: 16, L.TG_EP_RESID_FIRST: 12,
                   L.TG_EP_CHEB_LAST: 16}
_COEF_BYTES = {L.TG_OP_STENCIL_F32: 972, L.TG_OP_STENCIL_BF16: 486}
EP_NAMES = {0: "STORE", 1: "DOT", 2: "RESID", 3: "RESID_NORM", 4: "CHEB", 5: "CHEB_R", 6: "RESID_UPDATE",
            7: "CHEB_FIRST", 8: "RESID_FIRST", 9: "CHEB_LAST", -1: "-"}


def hex_bytes(n_dof: int, n_elem: int, mode: int) -> float:
    """Algorithmic HBM bytes of one level-0 operator launch: FP32 input, E,
    the node mask, plus the epilogue's own vector reads and writes (4 bytes
    per DOF per vector)."""
    return 4.0 * n_dof + 4.0 * n_elem + n_dof / 3.0 + _EP_VEC_TRAFFIC[mode] * n_dof


def op_bytes(kind: int, n_dof: int, n_elem: int, mode: int) -> float:
    """Algorithmic bytes of one operator launch of any level: the level-0
    count, with E read on the fine grid (8 fine elements per coarse one) for
    the implicit level-1 operator, and the 27-point coefficients (243 per
    node) instead of E for a stencil level."""
    if kind in _COEF_BYTES:
        return hex_bytes(n_dof, 0, mode) + _COEF_BYTES[kind] * (n_dof / 3.0)
    if kind in (L.TG_OP_Q8_F32, L.TG_OP_Q8_BF16, L.TG_OP_Q8_BF16X3):
        return hex_bytes(n_dof, 8 * n_elem, mode)
    return hex_bytes(n_dof, n_elem, mode)


def kernel_key(lv: int, kind: int) -> str:
    return f"L{lv} {KERNEL_NAMES[kind]}"


def launch_op(timer, lv: int, op, u_ptr, y_ptr, ep, stream, n_dof: int, n_elem: int, nrhs: int = 1) -> None:
    """tg_apply32 over nrhs right-hand sides, accounted (level, operator,
    epilogue) when a timer is attached."""
    if timer is None:
        L.call("tg_apply32_b", op, nrhs, u_ptr, y_ptr, ep, stream)
        return
    timer.launch(kernel_key(lv, op.kind), nrhs * op_bytes(op.kind, n_dof, n_elem, ep.mode),
                 lambda st=stream: L.call("tg_apply32_b", op, nrhs, u_ptr, y_ptr, ep, st), ep.mode)


def launch_transfer(timer, name: str, lv: int, nbytes: float, *args) -> None:
    """tg_restrict32_b / tg_prolong32_b between levels lv and lv + 1,
    accounted when a timer is attached (args end with the stream)."""
    if timer is None:
        L.call(name, *args)
        return
    head, stream = args[:-1], args[-1]
    timer.launch(f"L{lv} {name[3:-2]}", nbytes, lambda st=stream: L.call(name, *head, st))


class GMG:
    def __init__(self, dims, mask0: np.ndarray, *, n_levels: int = 4, fine_op: int = L.TG_OP_HEX_BF16,
                 fine_accurate_op: int | None = None, l1_q8_op: int = L.TG_OP_Q8_F32,
                 l1_q8_accurate_op: int | None = None, fine_degree: int = 2, coarse_degree=2,
                 cheb_lower_frac: float = 1.0 / 30.0, cheb_safety: float = 1.1, device: torch.device,
                 power_iters: int = 20, coarse_power_iters: int = 10, seed: int = 7, level1: str = "q8",
                 nrhs: int = 1, coarse_stencil: int = L.TG_OP_STENCIL_F32, coarse_precisions=None,
                 smoother: str = "chebyshev",
                 omega: float = 0.5, cycle: str = "v", coarse_solver: str = "auto"):
        if n_levels < 2:
            raise ValueError("the GMG preconditioner needs n_levels >= 2")
        if coarse_stencil not in (L.TG_OP_STENCIL_F32, L.TG_OP_STENCIL_BF16):
            raise ValueError("coarse_stencil must be TG_OP_STENCIL_F32 or TG_OP_STENCIL_BF16")
        self.dev = device
        self.nrhs = int(nrhs)
        if not 1 <= self.nrhs <= 64:
            raise ValueError("nrhs must be 1..64")
        self.fine_degree = max(1, int(fine_degree))
        # an int, or one degree per level from level 1 down (the last entry
        # repeats for deeper levels)
        cd = [coarse_degree] if np.isscalar(coarse_degree) else list(coarse_degree)
        if not cd:
            raise ValueError("coarse_degree must not be empty")
        self.coarse_degrees = [max(1, int(d)) for d in cd]
        self.frac = float(cheb_lower_frac)
        self.safety = float(cheb_safety)
        # smoother_type / omega / cycle_type of GalerkinMatFreeGMG
        # (multigrid_v4.py:473-511): Chebyshev-Jacobi or damped Jacobi with
        # omega = min(omega, 0.9 * 2 / lambda_max of level 0) on every level
        # (:705, :918-932, :963-967); V- or W-cycle (:1010-1042)
        if smoother not in ("chebyshev", "jacobi"):
            raise ValueError(f"smoother must be 'chebyshev' or 'jacobi'; got {smoother!r}")
        if cycle not in ("v", "w"):
            raise ValueError(f"cycle must be 'v' or 'w'; got {cycle!r}")
        if coarse_solver not in ("auto", "dense", "bcr"):
            raise ValueError(f"coarse_solver must be 'auto', 'dense' or 'bcr'; got {coarse_solver!r}")
        self.smoother, self.omega, self.cycle = smoother, float(omega), cycle
        self.power_iters = power_iters
        self.coarse_power_iters = coarse_power_iters
        self.dims = fem.level_dims(dims, n_levels)
        self.levels: list[_Level] = []
        rng = np.random.default_rng(seed)
        mask = mask0
        for lv, d in enumerate(self.dims):
            if lv > 0:
                mask = fem.injection_mask(self.dims[lv - 1], d, self.levels[-1].mask_host)
            Lv = _Level()
            Lv.dims = d
            Lv.grid = L.grid(d)
            Lv.mask_host = mask
            Lv.mask = torch.from_numpy(mask).to(device)
            Lv.nodes = int(np.prod([n + 1 for n in d]))
            Lv.n = 3 * Lv.nodes
            Lv.n_elem = int(np.prod(d))
            f32 = dict(dtype=torch.float32, device=device)
            nb = self.nrhs * Lv.n                 # one slice of Lv.n per right-hand side
            Lv.x = torch.zeros(nb, **f32)         # iterate
            Lv.r = torch.zeros(nb, **f32)         # residual carried by the smoother
            Lv.da = torch.zeros(nb, **f32)        # smoother steps (ping-pong)
            Lv.db = torch.zeros(nb, **f32)
            Lv.c = torch.zeros(nb, **f32)         # prolonged coarse correction
            Lv.rhs = torch.zeros(nb, **f32)
            Lv.dinv = torch.ones(Lv.n, **f32)     # shared by the batch
            Lv.v0 = torch.from_numpy(rng.standard_normal(Lv.n).astype(np.float32)).to(device)
            Lv.lam = float("nan")
            Lv.mats64 = None
            Lv.stencil = None
            self.levels.append(Lv)
        self.levels[0].kind = fine_op
        self.levels[0].kind_accurate = fine_accurate_op if fine_accurate_op is not None else fine_op
        for lv in range(1, n_levels):
            f, c = self.levels[lv - 1], self.levels[lv]
            c.consistent = fem.masks_consistent(f.dims, c.dims, f.mask_host, c.mask_host)
            c.kind = coarse_stencil
        l1 = self.levels[1]
        if level1 not in ("q8", "stencil"):
            raise ValueError(f"level1 must be 'q8' or 'stencil'; got {level1!r}")
        for lv in range(1, n_levels):
            self.levels[lv].kind_accurate = self.levels[lv].kind
        if level1 == "q8" and n_levels > 2 and all(n % 2 == 0 for n in self.dims[0]) and l1.consistent:
            l1.kind = l1_q8_op
            l1.kind_accurate = l1_q8_accurate_op if l1_q8_accurate_op is not None else l1_q8_op
        self.l1_q8 = l1.kind in (L.TG_OP_Q8_F32, L.TG_OP_Q8_BF16, L.TG_OP_Q8_BF16X3)
        if coarse_precisions is not None:
            # a per-level schedule for levels 1.. (level_precisions[1:],
            # multigrid_v4.py:673-682; missing entries are the most accurate
            # kind, as the reference pads with "fp64"): "bf16" selects BF16
            # operands, "bf16x3" the split BF16 level-1 product, "fp32" and
            # "fp64" FP32 operands (no level smooths in FP64 here)
            for lv in range(1, n_levels):
                pr = coarse_precisions[lv - 1] if lv - 1 < len(coarse_precisions) else "fp64"
                if pr not in ("bf16", "bf16x3", "fp32", "fp64"):
                    raise ValueError(f"level precision must be bf16, bf16x3, fp32 or fp64; got {pr!r}")
                Lv = self.levels[lv]
                if lv == 1 and self.l1_q8:
                    Lv.kind = {"bf16": L.TG_OP_Q8_BF16, "bf16x3": L.TG_OP_Q8_BF16X3}.get(pr, L.TG_OP_Q8_F32)
                else:
                    Lv.kind = L.TG_OP_STENCIL_BF16 if pr == "bf16" else L.TG_OP_STENCIL_F32
                Lv.kind_accurate = Lv.kind
        # Galerkin build plan: aligned levels come straight from the fine E by
        # one gather + DGEMM (levels 1 and 2); the rest from the level above.
        even = lambda d: all(n % 2 == 0 for n in d)   # noqa: E731
        self.l1_from_fine = even(self.dims[0]) and l1.consistent
        self.l2_from_fine = (n_levels > 2 and self.l1_from_fine and even(self.dims[1])
                             and self.levels[2].consistent)
        self.need_l1_mats = not self.l1_q8 or n_levels == 2 or not self.l2_from_fine
        work = 0
        if self.l1_from_fine and self.need_l1_mats:
            work = max(work, 8 * self.levels[1].n_elem)
        if self.l2_from_fine:
            work = max(work, 64 * self.levels[2].n_elem)
        self.gwork = torch.empty(max(work, 1), dtype=torch.float64, device=device)
        last = self.levels[-1]
        free = fem.free_from_node_mask(last.mask_host)
        self.n_coarse = int(free.size)
        if self.n_coarse == 0:
            raise ValueError("coarsest level has no free DOFs")
        # Coarsest solve (multigrid_v4.py:781-801, 977-978): the explicit dense
        # inverse, one GEMV per cycle, up to BCR_MIN_DOFS free DOFs; above,
        # block cyclic reduction over the level's x-planes (an exact block
        # Cholesky whose setup is batched plane-block work instead of an n^3
        # dense factorisation and inverse), when a plane fits a block.
        plane = 3 * (last.dims[1] + 1) * (last.dims[2] + 1)
        bcr_ok = plane <= BCR_MAX_PLANE and last.dims[0] >= 1
        if coarse_solver == "bcr" and not bcr_ok:
            raise ValueError(f"coarse_solver='bcr' needs x-planes of at most {BCR_MAX_PLANE} DOFs; got {plane}")
        self.coarse_bcr = coarse_solver == "bcr" or (coarse_solver == "auto" and bcr_ok
                                                     and self.n_coarse > BCR_MIN_DOFS)
        self.coarse_A = self.bcr_work = self.bcr_vec = None
        if self.coarse_bcr:
            nw = C.c_int64(0)
            L.call("tg_bcr_size", last.grid, C.byref(nw))
            self.bcr_work = torch.zeros(int(nw.value), dtype=torch.float64, device=device)
            self.bcr_vec = torch.zeros(2 * self.nrhs * last.n, dtype=torch.float64, device=device)
        else:
            fidx = -np.ones(last.n, dtype=np.int32)
            fidx[free] = np.arange(free.size, dtype=np.int32)
            self.coarse_fidx = torch.from_numpy(fidx).to(device)
            self.coarse_list = torch.from_numpy(free.astype(np.int32)).to(device)
            self.coarse_A = torch.zeros(self.n_coarse * self.n_coarse, dtype=torch.float64, device=device)
        self.scal = torch.zeros(8, dtype=torch.float64, device=device)
        self._rb = Readback(2)
        # (v.Kv, v.Dv) of every level's power iteration, read at once
        self._lam_scal = torch.zeros(2 * len(self.levels), dtype=torch.float64, device=device)
        self._rb_lam = Readback(2 * len(self.levels))
        # The one-workgroup coarsest inverse occupies one compute unit for
        # about 160 us at 135 DOFs: it runs, with the dense assembly before
        # it, on a side stream, beside the levels' stencil assembly and the
        # start of the solve, and the first use of the coarsest level waits
        # for it (join); its status reaches the host with the next read
        # (check).
        self.async_inverse = not self.coarse_bcr and self.n_coarse <= ASYNC_INVERSE_MAX
        if self.async_inverse:
            self._side = torch.cuda.Stream(device)
            self._ev_dense = torch.cuda.Event()
            self._ev_inv = torch.cuda.Event()
            self._inv_info = torch.zeros(1, dtype=torch.int32, device=device)
            self._rb_info = Readback(1, dtype=torch.int32)
        self._join_pending = self._check_pending = False
        self._emax_at_estimate = float("nan")
        self.is_setup = False
        self.timer = None          # optional KernelTimer (operator and transfer launches)
        self.E64 = self.E32 = None

    # ------------------------------------------------------------------ ops
    def op(self, lv: int, accurate: bool = False) -> L.TgOp32:
        Lv = self.levels[lv]
        kind = Lv.kind_accurate if accurate else Lv.kind
        if kind in (L.TG_OP_STENCIL_F32, L.TG_OP_STENCIL_BF16):
            return L.TgOp32(kind, Lv.grid, L.ptr(Lv.mask), None, None, L.ptr(Lv.stencil))
        return L.TgOp32(kind, Lv.grid, L.ptr(Lv.mask), L.ptr(self.E32), None, None)

    def _launch(self, lv, op, u, y, ep, stream, nrhs=None):
        Lv = self.levels[lv]
        launch_op(self.timer, lv, op, L.ptr(u), L.ptr(y), ep, stream, Lv.n, Lv.n_elem,
                  self.nrhs if nrhs is None else nrhs)

    # ---------------------------------------------------------------- setup
    def setup(self, E64: torch.Tensor, E32: torch.Tensor, emax, stream) -> None:
        """Per-SIMP-iteration setup (GalerkinMatFreeGMG.setup, multigrid_v4.py:657-813).
        emax (max E, which decides whether lambda_max is estimated again) is
        a number or a function returning it; the function is called only
        once the hierarchy's kernels are queued."""
        self.E64, self.E32 = E64, E32
        self.join(stream)         # the previous inverse still owns coarse_A
        self.check()
        lv0 = self.levels[0]
        L.call("tg_diag_inv_hex", lv0.grid, L.ptr(lv0.mask), L.ptr(E64), None, L.ptr(lv0.dinv), stream)
        last_lv = len(self.levels) - 1
        # the Galerkin chain first, then the coarsest factorisation (which
        # needs only the last level's matrices), then each level's stencil
        # and diagonal, which may run beside an asynchronous inverse
        for lv in range(1, len(self.levels)):
            f, c = self.levels[lv - 1], self.levels[lv]
            if lv == 1 and not self.need_l1_mats:
                continue                                # applied implicitly (Q8)
            if c.mats64 is None:
                c.mats64 = torch.empty(576 * c.n_elem, dtype=torch.float64, device=self.dev)
            m32 = None
            if lv == 1 and self.l1_from_fine:
                L.call("tg_galerkin_from_fine", lv0.grid, 1, L.ptr(E64), L.ptr(self.gwork), L.ptr(c.mats64),
                       m32, stream)
            elif lv == 2 and self.l2_from_fine:
                L.call("tg_galerkin_from_fine", lv0.grid, 2, L.ptr(E64), L.ptr(self.gwork), L.ptr(c.mats64),
                       m32, stream)
            else:
                L.call("tg_galerkin_elems", f.grid, c.grid, L.ptr(f.mask), int(not c.consistent),
                       L.ptr(E64) if lv == 1 else None, None if lv == 1 else L.ptr(f.mats64),
                       L.ptr(c.mats64), m32, stream)
        last = self.levels[-1]
        if self.coarse_bcr:
            L.call("tg_bcr_setup", last.grid, L.ptr(last.mask), L.ptr(last.mats64), L.ptr(self.bcr_work), stream)
        elif self.async_inverse:
            self._queue_inverse(stream)
        else:
            L.call("tg_assemble_dense", last.grid, L.ptr(last.mats64), L.ptr(self.coarse_fidx), self.n_coarse,
                   L.ptr(self.coarse_A), stream)
            L.call("tg_dense_spd_inverse", L.ptr(self.coarse_A), self.n_coarse, stream)
        for lv in range(1, last_lv):
            c = self.levels[lv]
            if lv == 1 and self.l1_q8:
                L.call("tg_diag_inv_q8", c.grid, L.ptr(c.mask), L.ptr(E64), L.ptr(c.dinv), stream)
                continue
            bf16 = c.kind == L.TG_OP_STENCIL_BF16
            if c.stencil is None:
                c.stencil = torch.empty(243 * c.nodes, dtype=torch.int16 if bf16 else torch.float32,
                                        device=self.dev)
            L.call("tg_assemble_stencil_bf16" if bf16 else "tg_assemble_stencil", c.grid, L.ptr(c.mask),
                   L.ptr(c.mats64), L.ptr(c.stencil), stream)
            L.call("tg_diag_inv_stored", c.grid, L.ptr(c.mask), L.ptr(c.mats64), L.ptr(c.dinv), stream)
        emax = float(emax() if callable(emax) else emax)
        prev = self._emax_at_estimate
        reestimate = np.isnan(prev) or abs(emax - prev) / max(abs(prev), 1e-12) > 0.10
        if reestimate:
            # every level's power iteration is queued before the one read
            for lv in range(last_lv):
                iters = self.power_iters if lv == 0 else self.coarse_power_iters
                self._queue_lambda(lv, iters, stream, self._lam_scal[2 * lv:2 * lv + 2])
            q = self._rb_lam.read(self._lam_scal[:2 * last_lv])
            for lv in range(last_lv):
                self.levels[lv].lam = self._quotient(q[2 * lv], q[2 * lv + 1])
            self._emax_at_estimate = emax
        self.is_setup = True

    def _queue_inverse(self, stream) -> None:
        """The coarsest matrix's dense assembly and inverse on the side
        stream, once the last level's Galerkin matrices are queued on
        `stream`; the inverse's status is copied to the host behind it."""
        self._ev_dense.record(self._torch_stream(stream))
        self._side.wait_event(self._ev_dense)
        last, side = self.levels[-1], self._side.cuda_stream
        L.call("tg_assemble_dense", last.grid, L.ptr(last.mats64), L.ptr(self.coarse_fidx), self.n_coarse,
               L.ptr(self.coarse_A), side)
        L.call("tg_dense_spd_inverse_async", L.ptr(self.coarse_A), self.n_coarse, L.ptr(self._inv_info), side)
        with torch.cuda.stream(self._side):
            self._rb_info.start(self._inv_info)
        self._ev_inv.record(self._side)
        self._join_pending = self._check_pending = True

    def _torch_stream(self, handle) -> torch.cuda.Stream:
        cur = torch.cuda.current_stream(self.dev)
        return cur if cur.cuda_stream == (handle or 0) else torch.cuda.ExternalStream(handle, device=self.dev)

    def join(self, stream) -> None:
        """Make `stream` wait for the asynchronous coarsest inverse (once per
        setup). Called before anything on `stream` reads or rewrites
        coarse_A, outside graph capture."""
        if self._join_pending:
            self._torch_stream(stream).wait_event(self._ev_inv)
            self._join_pending = False

    def check(self, block: bool = True) -> None:
        """Raise if the last asynchronous coarsest inverse met a pivot <= 0
        (as tg_dense_spd_inverse returns TG_ERR_NOT_SPD). Waits for the
        inverse; with block=False it returns at once while the inverse is
        still running."""
        if self._check_pending and (block or self._rb_info.event.query()):
            self._check_pending = False
            info = int(self._rb_info.wait()[0])
            if info != 0:
                raise L.TopoHipError(f"dense inverse: coarsest matrix not SPD (pivot {info})")

    def estimate_lambda(self, lv: int, iters: int, stream) -> float:
        """lambda_max(D^-1 A) by power iteration in the D inner product
        (_estimate_fine_lambda_max, multigrid_v4.py:273-295). The start
        vector is a fixed random vector rather than all ones: on a uniform
        density the all-ones vector is almost a rigid translation, has almost
        no high-frequency content and 20 steps then underestimate lambda_max
        badly on large grids."""
        self._queue_lambda(lv, iters, stream, self.scal[0:2])
        return self._quotient(*self._rb.read(self.scal[0:2]))

    def _queue_lambda(self, lv: int, iters: int, stream, s) -> None:
        """Queue estimate_lambda's power iteration on level lv; the last
        step's v.Kv and v.Dv land in s[0] and s[1]. It uses the level's x
        and r as v and Kv."""
        Lv = self.levels[lv]
        v, kv = Lv.x[:Lv.n], Lv.r[:Lv.n]
        v.copy_(Lv.v0)
        store, ep = L.ep(L.TG_EP_STORE), L.ep(L.TG_EP_DOT, dot=s[0:1])
        # the Rayleigh quotient is scale-free and v grows by about lambda_max
        # per step, so the steps run back to back and only the last step
        # reduces: its quotient is the one read
        for k in range(iters):
            last = k == iters - 1
            self._launch(lv, self.op(lv, accurate=True), v, kv, ep if last else store, stream, nrhs=1)
            L.call("tg_power_step32", Lv.n, L.ptr(v), L.ptr(kv), L.ptr(Lv.dinv), L.ptr(s[1:2]) if last else None,
                   stream)

    @staticmethod
    def _quotient(vkv: float, vdv: float) -> float:
        lam = vkv / vdv if vdv > 1e-300 and np.isfinite(vkv / vdv) else 1.0
        return max(lam, 1e-6)

    # --------------------------------------------------------------- cycle
    def _cheb_coeffs(self, lv: int, degree: int):
        """Per-step (a1, a2) of the steps s_k = a1 D^-1 r + a2 s_{k-1} of the
        recurrence of multigrid_v4.py:870-887: alpha_0 = 1/sigma,
        alpha_k = 1/(sigma - delta^2 alpha_{k-1}/4), and the reference's
        vector d, which is z_0 = D^-1 r_0 after the first step (x moves by
        alpha_0 z_0) and the step itself afterwards, d_k = alpha_k (z_k +
        delta^2 alpha_{k-1}/4 d_{k-1}). So in terms of steps the momentum
        coefficient is alpha_1 delta^2 / 4 at k = 1 (s_0 = alpha_0 z_0) and
        alpha_k delta^2 alpha_{k-1} / 4 for k >= 2.

        Damped Jacobi is the same residual-update sequence with every step
        d_k = omega D^-1 r_{k-1}: (omega, 0) for each of the degree steps."""
        if self.smoother == "jacobi":
            lam0 = self.levels[0].lam if np.isfinite(self.levels[0].lam) else 1.0
            w = min(self.omega, 0.9 * 2.0 / max(lam0, 1e-6))
            return [(w, 0.0)] * degree
        Lv = self.levels[lv]
        lmax = (Lv.lam if np.isfinite(Lv.lam) else 1.0) * self.safety
        lmin = self.frac * lmax
        sigma, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        alpha = 1.0 / sigma
        out = [(alpha, 0.0)]
        for k in range(1, degree):
            a_new = 1.0 / (sigma - delta * delta * alpha * 0.25)
            out.append((a_new, a_new * delta * delta * 0.25 * (1.0 if k == 1 else alpha)))
            alpha = a_new
        return out

    def _presmooth(self, lv: int, rhs, degree: int, stream):
        """x from zero; leaves r = rhs - A x in Lv.r.

        The Chebyshev start d0 = alpha0 Dinv rhs (x = d0) is not written
        out: the first operator application forms it from rhs as it reads
        its input (CHEB_FIRST, or RESID_FIRST for degree 1), with the same
        rounding as a separate start, so the results are identical."""
        Lv = self.levels[lv]
        co = self._cheb_coeffs(lv, degree)
        op = self.op(lv)
        if degree == 1:
            self._launch(lv, op, rhs, None,
                         L.ep(L.TG_EP_RESID_FIRST, rhs=rhs, dinv=Lv.dinv, x=Lv.x, r=Lv.r, a3=co[0][0]), stream)
            return
        a1, a2 = co[1]
        self._launch(lv, op, rhs, None, L.ep(L.TG_EP_CHEB_FIRST, rhs=rhs, dinv=Lv.dinv, d=Lv.da, x=Lv.x, r=Lv.r,
                                             a1=a1, a2=a2, a3=co[0][0]), stream)
        d, dn = Lv.da, Lv.db
        for a1, a2 in co[2:]:
            self._launch(lv, op, d, None, L.ep(L.TG_EP_CHEB_R, dinv=Lv.dinv, d=dn, x=Lv.x, r=Lv.r, a1=a1, a2=a2), stream)
            d, dn = dn, d
        self._launch(lv, op, d, None, L.ep(L.TG_EP_RESID_UPDATE, r=Lv.r), stream)

    def _postsmooth(self, lv: int, degree: int, stream):
        """x += c (the prolonged correction, in Lv.c) then Chebyshev from there.
        The last step writes only x (CHEB_LAST): r and d are dead after it."""
        Lv = self.levels[lv]
        co = self._cheb_coeffs(lv, degree)
        last = lambda i: L.TG_EP_CHEB_LAST if i == degree - 1 else L.TG_EP_CHEB_R
        # r -= A c ; d1 = alpha0 Dinv r ; x += c + d1   (accurate operator: c is smooth)
        self._launch(lv, self.op(lv, accurate=True), Lv.c, None,
                     L.ep(last(0), dinv=Lv.dinv, d=Lv.da, x=Lv.x, r=Lv.r, a1=co[0][0], a2=0.0, a3=1.0), stream)
        d, dn = Lv.da, Lv.db
        op = self.op(lv)
        for i, (a1, a2) in enumerate(co[1:], start=1):
            self._launch(lv, op, d, None, L.ep(last(i), dinv=Lv.dinv, d=dn, x=Lv.x, r=Lv.r, a1=a1, a2=a2), stream)
            d, dn = dn, d

    def _degree(self, lv: int) -> int:
        return self.fine_degree if lv == 0 else self.coarse_degrees[min(lv, len(self.coarse_degrees)) - 1]

    def _vcycle_down(self, lv: int, rhs, stream) -> None:
        """First half of a level's cycle: pre-smoothing, then the restriction
        of its residual to the next level's right-hand side."""
        Lv, C = self.levels[lv], self.levels[lv + 1]
        self._presmooth(lv, rhs, self._degree(lv), stream)
        self._restrict(lv, stream)

    def _vcycle_up(self, lv: int, stream):
        """Second half: the coarser cycle, prolongation, post-smoothing. A
        W-cycle corrects twice from the coarser level (multigrid_v4.py:
        1010-1042): x += P e_c and r -= A P e_c (CHEB_R with a1 = a2 = 0,
        a3 = 1: no smoothing step), restrict, cycle again."""
        Lv, C = self.levels[lv], self.levels[lv + 1]
        ec = self._vcycle(lv + 1, C.rhs, stream)
        if self.cycle == "w":
            self._prolong(lv, ec, stream)
            self._launch(lv, self.op(lv, accurate=True), Lv.c, None,
                         L.ep(L.TG_EP_CHEB_R, dinv=Lv.dinv, d=Lv.da, x=Lv.x, r=Lv.r, a1=0.0, a2=0.0, a3=1.0), stream)
            self._restrict(lv, stream)
            ec = self._vcycle(lv + 1, C.rhs, stream)
        self._prolong(lv, ec, stream)
        self._postsmooth(lv, self._degree(lv), stream)
        return Lv.x

    def _restrict(self, lv: int, stream) -> None:
        """C.rhs = M_c P^T r of level lv (reads r, the coarse mask; writes the
        coarse right-hand side)."""
        Lv, C = self.levels[lv], self.levels[lv + 1]
        launch_transfer(self.timer, "tg_restrict32_b", lv, self.nrhs * (4.0 * Lv.n + 4.0 * C.n + C.nodes),
                        Lv.grid, C.grid, L.ptr(C.mask), self.nrhs, L.ptr(Lv.r), L.ptr(C.rhs), stream)

    def _prolong(self, lv: int, ec, stream) -> None:
        """Lv.c = M P ec (reads ec, the fine mask; writes the correction)."""
        Lv, C = self.levels[lv], self.levels[lv + 1]
        launch_transfer(self.timer, "tg_prolong32_b", lv, self.nrhs * (4.0 * C.n + 4.0 * Lv.n + Lv.nodes),
                        Lv.grid, C.grid, L.ptr(Lv.mask), self.nrhs, L.ptr(ec), L.ptr(Lv.c), 0, stream)

    def _vcycle(self, lv: int, rhs, stream):
        Lv = self.levels[lv]
        if lv == len(self.levels) - 1:
            if self.coarse_bcr:
                L.call("tg_bcr_apply_b", Lv.grid, L.ptr(self.bcr_work), self.nrhs, L.ptr(rhs), L.ptr(Lv.x),
                       L.ptr(self.bcr_vec), stream)
            else:
                L.call("tg_coarsest_apply_b", L.ptr(self.coarse_A), self.n_coarse, L.ptr(self.coarse_list), Lv.n,
                       self.nrhs, L.ptr(rhs), L.ptr(Lv.x), stream)
            return Lv.x
        self._vcycle_down(lv, rhs, stream)
        return self._vcycle_up(lv, stream)

    def graph_key(self) -> tuple:
        """What a captured V-cycle bakes in besides the level buffers made at
        construction: the smoother coefficients (from each level's lambda),
        the modulus buffers and the setup-time buffers."""
        return (tuple(float(Lv.lam) for Lv in self.levels), self.fine_degree, tuple(self.coarse_degrees),
                self.safety, self.frac, L.ptr(self.E32), L.ptr(self.E64),
                tuple(L.ptr(Lv.stencil) for Lv in self.levels), L.ptr(self.coarse_A), L.ptr(self.bcr_work),
                L.ptr(self.bcr_vec), self.nrhs, self.smoother,
                self.omega, self.cycle)

    def apply(self, r, stream):
        """z = M r for every right-hand side of the batch (r: nrhs slices of
        the level-0 length); returns a level-0 buffer valid until the next
        apply."""
        if not self.is_setup:
            raise RuntimeError("GMG.setup() must be called before apply()")
        self.join(stream)
        return self._vcycle(0, r, stream)

    # apply() in two parts, so a caller can queue the first while it waits
    # on the host for something else: begin() writes only the hierarchy's own
    # buffers (level-0 pre-smoothing and the restriction to level 1), so it
    # may be issued speculatively and abandoned. begin() does not touch the
    # coarsest level, so it may run beside the asynchronous inverse.
    def begin(self, r, stream) -> None:
        if not self.is_setup:
            raise RuntimeError("GMG.setup() must be called before apply()")
        self._vcycle_down(0, r, stream)

    def finish(self, stream):
        """The rest of the cycle begin() started; returns z as apply() does."""
        self.join(stream)
        return self._vcycle_up(0, stream)
