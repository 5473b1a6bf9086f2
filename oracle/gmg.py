"""Multigrid hierarchy definitions (test infrastructure only).

Restates the grid transfer and Galerkin definitions the reference GMG uses:
level sizes ``max(1, n//2)`` (multigrid_v4.py:525-529), row-normalised
trilinear prolongation (solver_v2.py:62-149), injection free sets
(solver_v2.py:717-764) and coarse operators K_c = P_f^T K_ff P_f restricted to
free DOFs (multigrid_v4.py:549-552, 747-752). Used on small meshes only.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def level_dims(nelx: int, nely: int, nelz: int, n_levels: int):
    dims = [(nelx, nely, nelz)]
    for _ in range(1, n_levels):
        x, y, z = dims[-1]
        dims.append((max(1, x // 2), max(1, y // 2), max(1, z // 2)))
    return dims


def _axis_weights(n_f: int, n_c: int) -> np.ndarray:
    """1-D interpolation matrix (n_f+1, n_c+1): fine node i takes weight 1 from
    coarse node i/2 when i is even, 1/2 from floor(i/2) and floor(i/2)+1 when
    odd; parents beyond the coarse grid are dropped (solver_v2.py:100-129)."""
    W = np.zeros((n_f + 1, n_c + 1))
    for i in range(n_f + 1):
        if i % 2 == 0:
            if i // 2 <= n_c:
                W[i, i // 2] = 1.0
        else:
            for I in (i // 2, i // 2 + 1):
                if I <= n_c:
                    W[i, I] = 0.5
    return W


def scalar_prolongation(dims_f, dims_c) -> sp.csr_matrix:
    """Trilinear node prolongation with every row scaled to sum 1
    (solver_v2.py:62-149, row normalisation :146-148)."""
    Wx, Wy, Wz = (_axis_weights(f, c) for f, c in zip(dims_f, dims_c))
    P = sp.kron(sp.kron(sp.csr_matrix(Wx), sp.csr_matrix(Wy)), sp.csr_matrix(Wz)).tocsr()
    rs = np.asarray(P.sum(axis=1)).ravel()
    rs = np.where(rs > 1e-12, rs, 1.0)
    return (sp.diags(1.0 / rs) @ P).tocsr()


def coarse_free_injection(dims_f, dims_c, free_f: np.ndarray) -> np.ndarray:
    """Coarse DOF free iff the DOF of fine node (2I, 2J, 2K) is free, using the
    reference's flat index (including its wrap-around when 2K > nelz_f, which
    only happens for 1-element-thick levels) (solver_v2.py:738-764)."""
    nxf, nyf, nzf = dims_f
    nxc, nyc, nzc = dims_c
    ndof_f = 3 * (nxf + 1) * (nyf + 1) * (nzf + 1)
    I, J, K = np.meshgrid(np.arange(nxc + 1), np.arange(nyc + 1), np.arange(nzc + 1), indexing="ij")
    fine_node = (2 * I * (nyf + 1) * (nzf + 1) + 2 * J * (nzf + 1) + 2 * K).ravel()
    mask_f = np.zeros(ndof_f, dtype=bool)
    mask_f[free_f] = True
    ndof_c = 3 * (nxc + 1) * (nyc + 1) * (nzc + 1)
    is_free = np.zeros(ndof_c, dtype=bool)
    for d in range(3):
        fd = 3 * fine_node + d
        cd = 3 * np.arange(fine_node.size) + d
        ok = fd < ndof_f
        is_free[cd[ok]] = mask_f[fd[ok]]
    return np.nonzero(is_free)[0]


def vector_prolongation_free(dims_f, dims_c, free_f, free_c) -> sp.csr_matrix:
    """P restricted to free x free (multigrid_v4.py:548-552)."""
    P = sp.kron(scalar_prolongation(dims_f, dims_c), sp.eye(3), format="csr")
    return P[free_f, :][:, free_c].tocsr()


def galerkin_coarse_operators(K_ff: sp.spmatrix, dims, free_fine: np.ndarray):
    """[(free_l, K_l, P_l)] for every level: K_0 = K_ff, K_{l+1} = P^T K_l P."""
    out = [(np.asarray(free_fine), sp.csr_matrix(K_ff), None)]
    for lv in range(len(dims) - 1):
        free_f = out[-1][0]
        free_c = coarse_free_injection(dims[lv], dims[lv + 1], free_f)
        P = vector_prolongation_free(dims[lv], dims[lv + 1], free_f, free_c)
        Kc = (P.T @ out[-1][1] @ P).tocsr()
        out.append((free_c, Kc, P))
    return out
