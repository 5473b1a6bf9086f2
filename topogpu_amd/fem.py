"""Host-side geometry and constants for the structured hex8 SIMP path.

Mesh conventions follow the reference (``src/gpu_fem/pub_simp_solver.py:196-207``,
``src/gpu_fem/bc_generator.py:104-105``): element (i,j,k) -> (i*ny + j)*nz + k,
node (i,j,k) -> (i*(ny+1) + j)*(nz+1) + k, DOF = 3*node + component.
"""

from __future__ import annotations

import numpy as np

# (dx, dy, dz) of the eight local nodes, reference order (pub_simp_solver.py:202-205).
CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                    (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.int64)


def unit_stiffness(nu: float = 0.3, E: float = 1.0) -> np.ndarray:
    """24x24 trilinear hex stiffness on [-1,1]^3 (pub_simp_solver.py:157-183).

    B = strain-displacement matrix with engineering shears (xy, yz, zx), D =
    isotropic elasticity, integrated by 2x2x2 Gauss (exact here because the
    Jacobian is the identity and the integrand is at most quadratic per axis).
    Built with einsum over all Gauss points at once.
    """
    s = 2.0 * CORNERS - 1.0                                    # node signs (8, 3)
    g = np.array(np.meshgrid(*[[-1.0, 1.0]] * 3, indexing="ij")).reshape(3, -1).T / np.sqrt(3.0)
    one = 1.0 + g[:, None, :] * s[None, :, :]                  # (8 gp, 8 nodes, 3)
    dN = np.empty_like(one)
    dN[..., 0] = s[None, :, 0] * one[..., 1] * one[..., 2] / 8.0
    dN[..., 1] = s[None, :, 1] * one[..., 0] * one[..., 2] / 8.0
    dN[..., 2] = s[None, :, 2] * one[..., 0] * one[..., 1] / 8.0
    B = np.zeros((8, 6, 24))
    for a in range(3):
        B[:, a, a::3] = dN[..., a]
    for row, (p, q) in zip((3, 4, 5), ((0, 1), (1, 2), (2, 0))):
        B[:, row, p::3] = dN[..., q]
        B[:, row, q::3] = dN[..., p]
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    D = np.diag([lam + 2 * mu] * 3 + [mu] * 3)
    D[:3, :3] += lam * (1.0 - np.eye(3))
    return np.einsum("gai,ab,gbj->ij", B, D, B)


def node_mask_from_free(free: np.ndarray, n_nodes: int) -> np.ndarray:
    """Per-node byte mask, bit c set when DOF 3*node+c is free."""
    m = np.zeros(3 * n_nodes, dtype=np.uint8)
    m[np.asarray(free, dtype=np.int64)] = 1
    m = m.reshape(n_nodes, 3)
    return (m[:, 0] | (m[:, 1] << 1) | (m[:, 2] << 2)).astype(np.uint8)


def free_from_node_mask(mask: np.ndarray) -> np.ndarray:
    bits = (mask[:, None] >> np.arange(3)[None, :]) & 1
    return np.nonzero(bits.ravel())[0]


def coarse_dims(dims):
    return tuple(max(1, n // 2) for n in dims)


def level_dims(dims, n_levels: int):
    out = [tuple(dims)]
    for _ in range(1, n_levels):
        out.append(coarse_dims(out[-1]))
    return out


def auto_levels(dims, mask0: np.ndarray, max_coarse: int = 192, max_levels: int = 12) -> int:
    """Fewest levels (>= 2) whose coarsest level has at most ``max_coarse``
    free DOFs. The coarsest level is inverted densely once per design
    iteration; up to 192 unknowns this is one workgroup holding the matrix in
    registers (Gauss-Jordan, n^3 FMA), while a larger coarsest level needs a
    library factorisation whose launch sequence costs milliseconds, more than
    the smoothing of the extra levels."""
    dims_l, mask = tuple(dims), mask0
    for n_levels in range(2, max_levels + 1):
        nxt = coarse_dims(dims_l)
        mask = injection_mask(dims_l, nxt, mask)
        dims_l = nxt
        n_free = int(((mask[:, None] >> np.arange(3)[None, :]) & 1).sum())
        if n_free <= max_coarse or dims_l == (1, 1, 1):
            return n_levels
    return max_levels


def injection_mask(dims_f, dims_c, mask_f: np.ndarray) -> np.ndarray:
    """Coarse node mask by injection: coarse node (I,J,K) copies the mask of
    fine node (2I, 2J, 2K) (solver_v2.py:717-764). On one-element-thick axes
    (n_f = n_c = 1) the reference's flat index wraps into a neighbouring node
    line; here the fine index is clamped to the grid instead."""
    nf = [n + 1 for n in dims_f]
    nc = [n + 1 for n in dims_c]
    idx = [np.minimum(2 * np.arange(c), f - 1) for c, f in zip(nc, nf)]
    I, J, K = np.meshgrid(*idx, indexing="ij")
    return mask_f.reshape(nf)[I, J, K].ravel().copy()


def _axis_parents(nf: int, nc: int):
    """Per fine node: (lo, hi, wlo, whi) of the row-normalised trilinear P
    along one axis (solver_v2.py:100-148)."""
    i = np.arange(nf + 1)
    lo = i // 2
    odd = (i & 1) == 1
    two = odd & (lo + 1 <= nc)
    hi = np.where(two, lo + 1, lo)
    wlo = np.where(two, 0.5, 1.0)
    whi = np.where(two, 0.5, 0.0)
    return lo, hi, wlo, whi


def masks_consistent(dims_f, dims_c, mask_f: np.ndarray, mask_c: np.ndarray) -> bool:
    """True when every fixed fine DOF interpolates only from fixed coarse DOFs.

    Then M_f P M_c = P M_c, so P_free^T K_ff P_free equals the unmasked
    Galerkin product restricted to free coarse DOFs and the coarse element
    matrices need no per-child masks.
    """
    nfx, nfy, nfz = (n + 1 for n in dims_f)
    ncs = [n + 1 for n in dims_c]
    px, py, pz = (_axis_parents(f, c) for f, c in zip(dims_f, dims_c))
    mf = mask_f.reshape(nfx, nfy, nfz)
    mc = mask_c.reshape(ncs)
    for a in range(8):
        sel = [(a >> 2) & 1, (a >> 1) & 1, a & 1]
        idx, w = [], []
        for p, s in zip((px, py, pz), sel):
            idx.append(p[1] if s else p[0])
            w.append(p[3] if s else p[2])
        I, J, K = np.meshgrid(*idx, indexing="ij")
        W = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
        parent = mc[I, J, K]
        # fixed fine bit with a free parent bit of the same component
        bad = (~mf & parent & 7) != 0
        if np.any(bad & (W > 0)):
            return False
    return True


def filter_stencil(rmin: float):
    """Offsets (m, 3) int32 and raw weights (m,) of the cone filter.

    Pairs at centre distance d <= rmin - 1e-10 weigh rmin - d, the element
    itself rmin (pub_simp_solver.py:226-234). The offset set is symmetric, so
    the raw weight matrix W is symmetric.
    """
    r = int(np.floor(rmin))
    ax = np.arange(-r, r + 1)
    d = np.array(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, -1).T
    dist = np.sqrt((d.astype(np.float64) ** 2).sum(axis=1))
    keep = (dist <= rmin - 1e-10) & (dist > 0)
    offs = np.vstack([np.zeros((1, 3), dtype=np.int64), d[keep]])
    w = np.concatenate([[float(rmin)], rmin - dist[keep]])
    return offs.astype(np.int32), w.astype(np.float64)
