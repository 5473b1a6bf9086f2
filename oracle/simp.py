"""FP64 restatement of the reference CPU SIMP path (test infrastructure only).

Reference: ``src/gpu_fem/pub_simp_solver.py`` of nbbllxx0/TopoGPU. Only the 3-D
structured-hex8 path is restated: ``run_simp`` without and with a controller
(best tracking, actions, tail phase); its 2-D presets and checkpoint files
are not.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

E0 = 1.0
EMIN = 1.0e-9

# Local node order of a hex8 element as (dx, dy, dz) corner offsets.
# pub_simp_solver.py:202-205 (_edof_table_3d) and :159-160 (node sign table).
HEX_CORNERS = np.array(
    [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
     (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
    dtype=np.int64,
)


def ke_unit_3d(nu: float = 0.3, E: float = 1.0, n_gauss: int = 2) -> np.ndarray:
    """24x24 trilinear-hex stiffness on the [-1,1]^3 cube (edge length 2).

    Follows ``_build_ke_unit_3d`` (pub_simp_solver.py:157-183): isotropic D with
    Lame parameters, engineering shear strains ordered (xy, yz, zx), DOF order
    node-major with (x, y, z) per node, Gauss quadrature. The shape-function
    gradients are written here from N_a = prod(1 + s_a * xi)/8 rather than the
    reference's expanded table. ``n_gauss`` > 2 is used by the tests: the
    integrand is polynomial of degree <= 2 per coordinate, so 2 and 3 points
    must agree.
    """
    signs = 2.0 * HEX_CORNERS.astype(np.float64) - 1.0          # (8, 3) node coords
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2.0 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    pts, wts = np.polynomial.legendre.leggauss(n_gauss)
    K = np.zeros((24, 24))
    for a, wa in zip(pts, wts):
        for b, wb in zip(pts, wts):
            for c, wc in zip(pts, wts):
                xi = np.array([a, b, c])
                fac = 1.0 + signs * xi[None, :]                 # (8, 3)
                grad = np.empty((8, 3))
                for d in range(3):
                    others = [o for o in range(3) if o != d]
                    grad[:, d] = 0.125 * signs[:, d] * fac[:, others[0]] * fac[:, others[1]]
                # Jacobian d(x)/d(xi) = sum_a x_a grad_a^T = I on this cube.
                J = signs.T @ grad
                gx = grad @ np.linalg.inv(J)                    # physical gradients
                B = np.zeros((6, 24))
                B[0, 0::3] = gx[:, 0]
                B[1, 1::3] = gx[:, 1]
                B[2, 2::3] = gx[:, 2]
                B[3, 0::3] = gx[:, 1]
                B[3, 1::3] = gx[:, 0]
                B[4, 1::3] = gx[:, 2]
                B[4, 2::3] = gx[:, 1]
                B[5, 0::3] = gx[:, 2]
                B[5, 2::3] = gx[:, 0]
                K += (B.T @ D @ B) * abs(np.linalg.det(J)) * wa * wb * wc
    return K


KE_UNIT_3D = ke_unit_3d()


def node_id(i, j, k, nely: int, nelz: int):
    """Global node index, x slowest and z fastest (bc_generator.py:104-105)."""
    return (i * (nely + 1) + j) * (nelz + 1) + k


def edof_table_3d(nelx: int, nely: int, nelz: int) -> np.ndarray:
    """Element -> 24 global DOFs, element index i*nely*nelz + j*nelz + k.

    pub_simp_solver.py:196-207. DOF = 3*node + component.
    """
    ex, ey, ez = np.meshgrid(np.arange(nelx), np.arange(nely), np.arange(nelz), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    nodes = np.stack(
        [node_id(ex + c[0], ey + c[1], ez + c[2], nely, nelz) for c in HEX_CORNERS], axis=1
    )
    return (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24).astype(np.int64)


def filter_offsets(rmin: float):
    """Neighbour offsets and raw weights of the cone filter.

    Pairs at centre distance d <= rmin - 1e-10 get weight rmin - d; the element
    itself gets rmin (pub_simp_solver.py:226-234). Returned as (offsets (m,3),
    weights (m,)), self first.
    """
    r = int(np.floor(rmin))
    offs, wts = [(0, 0, 0)], [float(rmin)]
    for di in range(-r, r + 1):
        for dj in range(-r, r + 1):
            for dk in range(-r, r + 1):
                if di == dj == dk == 0:
                    continue
                d = float(np.sqrt(di * di + dj * dj + dk * dk))
                if d <= rmin - 1e-10:
                    offs.append((di, dj, dk))
                    wts.append(rmin - d)
    return np.array(offs, dtype=np.int64), np.array(wts, dtype=np.float64)


def build_density_filter(nelx: int, nely: int, nelz: int, rmin: float) -> sp.csr_matrix:
    """Row-normalised cone filter H (pub_simp_solver.py:215-243).

    The reference finds pairs with a KD-tree; on the structured grid the same
    pair set is the fixed offset stencil clipped at the domain boundary.
    """
    n = nelx * nely * nelz
    offs, wts = filter_offsets(rmin)
    ex, ey, ez = np.meshgrid(np.arange(nelx), np.arange(nely), np.arange(nelz), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    rows, cols, vals = [], [], []
    for (di, dj, dk), w in zip(offs, wts):
        jx, jy, jz = ex + di, ey + dj, ez + dk
        ok = (jx >= 0) & (jx < nelx) & (jy >= 0) & (jy < nely) & (jz >= 0) & (jz < nelz)
        rows.append(np.nonzero(ok)[0])
        cols.append((jx[ok] * nely + jy[ok]) * nelz + jz[ok])
        vals.append(np.full(int(ok.sum()), w))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    hs = np.bincount(rows, weights=vals, minlength=n)
    return sp.csr_matrix((vals / hs[rows], (rows, cols)), shape=(n, n))


def heaviside(rho_f: np.ndarray, beta: float, eta: float = 0.5) -> np.ndarray:
    """Smooth Heaviside projection (pub_simp_solver.py:245-248)."""
    if beta < 1e-6:
        return rho_f.copy()
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    return (np.tanh(beta * eta) + np.tanh(beta * (rho_f - eta))) / den


def heaviside_deriv(rho_f: np.ndarray, beta: float, eta: float = 0.5) -> np.ndarray:
    """d heaviside / d rho_f (pub_simp_solver.py:250-253)."""
    if beta < 1e-6:
        return np.ones_like(rho_f)
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    return beta * (1.0 - np.tanh(beta * (rho_f - eta)) ** 2) / den


def assemble_k(edof: np.ndarray, E_e: np.ndarray, ndof: int, KE: np.ndarray = KE_UNIT_3D) -> sp.csc_matrix:
    """K = sum_e E_e * P_e^T KE P_e (pub_simp_solver.py:350-352)."""
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    vals = (E_e[:, None] * KE.ravel()[None, :]).ravel()
    K = sp.csc_matrix((vals, (rows, cols)), shape=(ndof, ndof))
    K.sum_duplicates()
    return K


def solve_state(edof, E_e, free, F, ndof) -> np.ndarray:
    """Direct solve K_ff U_f = F_f, fixed DOFs zero (pub_simp_solver.py:40-49, 356-358)."""
    K = assemble_k(edof, E_e, ndof)
    U = np.zeros(ndof)
    U[free] = spla.spsolve(K[free][:, free], F[free])
    return U


def element_compliance(U: np.ndarray, edof: np.ndarray, KE: np.ndarray = KE_UNIT_3D) -> np.ndarray:
    """ce_e = u_e^T KE u_e (pub_simp_solver.py:361-363)."""
    Ue = U[edof]
    return np.einsum("ei,ei->e", Ue @ KE, Ue)


def matfree_matvec(u: np.ndarray, E_e: np.ndarray, edof: np.ndarray, free_mask: np.ndarray,
                   KE: np.ndarray = KE_UNIT_3D) -> np.ndarray:
    """y = M K M u on full-length vectors, M = diag(free_mask).

    Same operator as ``MatrixFreeKff.matvec`` (solver_v2.py:1183-1223) embedded
    in full-length vectors: fixed DOFs of the input are ignored and fixed DOFs
    of the output are zero.
    """
    uf = np.where(free_mask, u, 0.0)
    fe = (uf[edof] @ KE) * E_e[:, None]          # KE symmetric
    y = np.bincount(edof.ravel(), weights=fe.ravel(), minlength=u.shape[0])
    return np.where(free_mask, y, 0.0)


def matfree_diag(E_e: np.ndarray, edof: np.ndarray, free_mask: np.ndarray, ndof: int,
                 KE: np.ndarray = KE_UNIT_3D) -> np.ndarray:
    """diag(K_ff) embedded, fixed DOFs set to 1 (solver_v2.py:1229-1252,
    multigrid_v4.py:685 replaces |d| <= 1e-12 by 1)."""
    contrib = E_e[:, None] * np.diag(KE)[None, :]
    d = np.bincount(edof.ravel(), weights=contrib.ravel(), minlength=ndof)
    return np.where(free_mask, d, 1.0)


def oc_update_filtered(rho, dc, dv, volfrac, move, H, beta, eta, use_heaviside, return_state=False):
    """OC with filtered+projected volume constraint (pub_simp_solver.py:319-340).

    Bisection on lambda in [0, 1e9], at most 100 steps, stop when the bracket
    is narrower than 1e-9; returns the candidate of the last midpoint (and,
    with ``return_state``, also the last midpoint and the number of steps).
    """
    n = rho.shape[0]
    dc_s = np.minimum(dc, -1e-12)
    dv_s = np.maximum(dv, 1e-12)
    l1, l2 = 0.0, 1e9
    rn = rho
    lm, steps = 0.5 * (l1 + l2), 0
    for _ in range(100):
        steps += 1
        lm = 0.5 * (l1 + l2)
        rn = rho * np.sqrt(-dc_s / (lm * dv_s))
        rn = np.clip(rn, rho - move, rho + move)
        rn = np.clip(rn, 1e-3, 1.0)
        rf = H @ rn
        rp = heaviside(rf, beta, eta) if (use_heaviside and beta > 1e-6) else rf
        if rp.sum() / n > volfrac:
            l1 = lm
        else:
            l2 = lm
        if l2 - l1 < 1e-9:
            break
    return (rn, lm, steps) if return_state else rn


def oc_update_mean(rho, dc, volfrac, move, rho_min):
    """OC with a plain mean-volume constraint (src/topogpu/solver.py:94-117)."""
    dc_safe = np.minimum(dc, -1e-12)
    lo, hi = 0.0, 1.0

    def cand(lam):
        return np.clip(rho * np.sqrt(np.maximum(-dc_safe / max(lam, 1e-40), 0.0)),
                       np.maximum(rho - move, rho_min), np.minimum(rho + move, 1.0))

    while cand(hi).mean() > volfrac and hi < 1e40:
        lo, hi = hi, hi * 2.0
    rn = rho
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        rn = cand(mid)
        if rn.mean() > volfrac:
            lo = mid
        else:
            hi = mid
        if abs(float(rn.mean()) - volfrac) < 1e-8:
            break
    return rn


def fea_and_sensitivity(rho_design, penal, move, beta, eta, use_heaviside, H, *, volfrac,
                        edof, free, F, ndof, solver=None):
    """One SIMP step (pub_simp_solver.py:342-377).

    Returns (compliance, rho_new, change, rho_phys, extras) where extras holds
    U, dc_phys, dc_design and dv_design for the tests.
    """
    n = rho_design.shape[0]
    rho_f = H @ rho_design
    proj = use_heaviside and beta > 1e-6
    rho_phys = heaviside(rho_f, beta, eta) if proj else rho_f
    E_e = EMIN + (E0 - EMIN) * rho_phys ** penal
    U = (solver or solve_state)(edof, E_e, free, F, ndof)
    compliance = float(F @ U)
    ce = element_compliance(U, edof)
    dc_phys = -penal * (E0 - EMIN) * rho_phys ** (penal - 1.0) * ce
    dh = heaviside_deriv(rho_f, beta, eta) if proj else np.ones_like(rho_f)
    dc_design = H.T @ (dc_phys * dh)
    dv_design = H.T @ dh
    rho_new = oc_update_filtered(rho_design, dc_design, dv_design, volfrac, move, H,
                                 beta, eta, use_heaviside)
    change = float(np.max(np.abs(rho_new - rho_design)))
    extras = dict(U=U, E_e=E_e, ce=ce, dc_phys=dc_phys, dc_design=dc_design,
                  dv_design=dv_design, rho_f=rho_f)
    return compliance, rho_new, change, rho_phys, extras


def run_simp(nelx, nely, nelz, volfrac, free, F, *, rmin=1.5, penal=3.0, move=0.2,
             max_iter=100, min_iter=50, tol=0.01, compliance_tol=5e-4, compliance_window=5,
             use_heaviside=True, beta=1.0, eta=0.5, solver=None):
    """Main loop of ``run_simp`` without callbacks/tail/checkpoints.

    pub_simp_solver.py:412-657 with ``bc_override`` supplying (free, F): start
    from rho = volfrac, one FEA+sensitivity+OC step per iteration, stop when
    the max density change is below ``tol`` AND the compliance window has
    settled (:548-552) once ``iteration >= min_iter`` (:605-606); finally
    re-evaluate to return the physical density (:635-641).
    """
    n = nelx * nely * nelz
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    edof = edof_table_3d(nelx, nely, nelz)
    H = build_density_filter(nelx, nely, nelz, rmin)
    kw = dict(volfrac=volfrac, edof=edof, free=free, F=F, ndof=ndof, solver=solver)
    rho = volfrac * np.ones(n)
    hist, changes, rho_hist = [], [], []
    iteration = 0
    for iteration in range(1, max_iter + 1):
        c, rho_new, change, rho_phys, _ = fea_and_sensitivity(
            rho, penal, move, beta, eta, use_heaviside, H, **kw)
        hist.append(c)
        changes.append(change)
        win = hist[-compliance_window:]
        conv_c = (len(win) == compliance_window
                  and abs(win[-1] - win[0]) / max(abs(win[0]), 1e-10) < compliance_tol)
        conv = change < tol and conv_c
        rho = rho_new
        rho_hist.append(rho.copy())
        if conv and iteration >= min_iter:
            break
    final_c, _, _, rho_final, _ = fea_and_sensitivity(
        rho, penal, move, beta, eta, use_heaviside, H, **kw)
    return dict(compliance_history=hist, rho_final=rho_final, final_compliance=final_c,
                final_grayness=float(4.0 * np.mean(rho_final * (1.0 - rho_final))),
                n_iter=iteration, changes=changes, rho_design_history=rho_hist)


def _grayness(rho):
    return float(4.0 * np.mean(rho * (1.0 - rho)))


def run_simp_controlled(nelx, nely, nelz, volfrac, free, F, *, callback, rmin=1.5, penal=3.0, move=0.2,
                        max_iter=100, min_iter=50, tol=0.01, compliance_tol=5e-4, compliance_window=5,
                        use_heaviside=True, beta=1.0, eta=0.5, min_penal_for_best=3.0,
                        max_gray_for_best=0.25, tail_default_iters=18, beta_max=32.0, solver=None):
    """``run_simp`` with a controller (pub_simp_solver.py:412-657, without
    checkpoints): initial action, best / pre-valid tracking behind the
    penal/grayness gate, per-iteration actions (clipped parameters, restart
    from the best valid design, stop), the filter rebuilt at the top of an
    iteration after rmin changed, the tail phase and the final evaluation
    with the last parameters and the last built filter.

    The controller sees ``state.iteration`` and the other StepState fields
    the tests need; ``params`` carries ``tail_default_iters`` and
    ``beta_max``."""
    from types import SimpleNamespace

    n = nelx * nely * nelz
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    edof = edof_table_3d(nelx, nely, nelz)
    kw = dict(volfrac=volfrac, edof=edof, free=free, F=F, ndof=ndof, solver=solver)
    params = SimpleNamespace(tail_default_iters=tail_default_iters, beta_max=beta_max, min_iter=min_iter)

    def clip_action(act, p, r, m, b, rho_next, best_rho, valid):
        # _apply_action, pub_simp_solver.py:379-386
        if act.get("restart", False) and valid:
            rho_next = best_rho.copy()
        p = float(np.clip(act["penal"], 1.0, 5.0)) if "penal" in act else p
        r = float(np.clip(act["rmin"], 1.1, 4.0)) if "rmin" in act else r
        m = float(np.clip(act["move"], 0.03, 0.4)) if "move" in act else m
        b = float(np.clip(act["beta"], 1.0, 64.0)) if "beta" in act else b
        return p, r, m, b, rho_next

    rho = volfrac * np.ones(n)
    H = build_density_filter(nelx, nely, nelz, rmin)
    h_rmin = rmin
    if hasattr(callback, "initial_action"):
        act = callback.initial_action(params)
        if act:
            penal, rmin, move, beta, rho = clip_action(act, penal, rmin, move, beta, rho, rho, False)
    best = dict(rho=rho.copy(), c=np.inf, it=0, valid=False, gray=_grayness(rho))
    pre = dict(rho=rho.copy(), c=np.inf, it=0, gray=_grayness(rho))
    hist, log = [], []
    iteration = 0
    for iteration in range(1, max_iter + 1):
        if abs(rmin - h_rmin) > 1e-6:
            H, h_rmin = build_density_filter(nelx, nely, nelz, rmin), rmin
        c, rho_new, change, rho_phys, _ = fea_and_sensitivity(rho, penal, move, beta, eta, use_heaviside, H, **kw)
        hist.append(c)
        gray = _grayness(rho_phys)
        if c < best["c"] and penal >= min_penal_for_best and gray < max_gray_for_best:
            best.update(rho=rho_phys.copy(), c=c, it=iteration, valid=True, gray=gray)
        elif c < pre["c"] and not best["valid"] and iteration >= 10:
            pre.update(rho=rho_phys.copy(), c=c, it=iteration, gray=gray)
        win = hist[-compliance_window:]
        conv = (change < tol and len(win) == compliance_window
                and abs(win[-1] - win[0]) / max(abs(win[0]), 1e-10) < compliance_tol)
        log.append(dict(phase="main", penal=penal, rmin=rmin, move=move, beta=beta, best_compliance=best["c"],
                        best_is_valid=best["valid"]))
        act = callback(SimpleNamespace(iteration=iteration, compliance=c, converged=conv,
                                       best_is_valid=best["valid"]), rho_new)
        if act:
            penal, rmin, move, beta, rho_new = clip_action(act, penal, rmin, move, beta, rho_new, best["rho"],
                                                           best["valid"])
            if act.get("stop", False) and iteration >= min_iter:
                rho = rho_new
                break
        rho = rho_new
        if conv and iteration >= min_iter:
            break
    tail = {"enabled": False, "tail_iters": tail_default_iters, "restart_from_best": True,
            "penal": 4.5, "rmin": 1.2, "move": 0.05, "beta": min(beta_max, 32.0)}
    if hasattr(callback, "finalize_tail"):
        tail.update(callback.finalize_tail(params) or {})
    n_tail = 0
    if tail["enabled"] and tail["tail_iters"] > 0:
        if tail["restart_from_best"]:
            rho = np.clip(best["rho"], 1e-3, 1.0) if best["valid"] else volfrac * np.ones(n)
        penal, rmin = float(tail["penal"]), float(tail["rmin"])
        move, beta = float(tail["move"]), float(tail["beta"])
        H = build_density_filter(nelx, nely, nelz, rmin)
        for n_tail in range(1, int(tail["tail_iters"]) + 1):
            c, rho, _, _, _ = fea_and_sensitivity(rho, penal, move, beta, eta, use_heaviside, H, **kw)
            hist.append(c)
            log.append(dict(phase="tail", penal=penal, rmin=rmin, move=move, beta=beta,
                            best_compliance=best["c"], best_is_valid=best["valid"]))
    final_c, _, _, rho_final, _ = fea_and_sensitivity(rho, penal, move, beta, eta, use_heaviside, H, **kw)
    return dict(compliance_history=hist, rho_final=rho_final, final_compliance=final_c,
                final_grayness=_grayness(rho_final), n_iter=iteration + n_tail, best_rho=best["rho"],
                best_compliance=best["c"], best_iteration=best["it"], best_is_valid=best["valid"],
                best_grayness=best["gray"], pre_valid_best_compliance=pre["c"],
                pre_valid_best_iteration=pre["it"], params_log=log)
