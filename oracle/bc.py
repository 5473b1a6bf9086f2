"""Boundary-condition restatement (test infrastructure only).

Follows ``generate_bc`` of ``src/gpu_fem/bc_generator.py`` for 3-D specs:
face supports (EdgeSupport), nearest-node point supports and point loads,
trapezoidal face pressure (DistributedLoad). The spec is duck-typed: any
object with the ProblemSpec field names works, supports/loads are recognised
by their attributes (``edge`` -> face, otherwise point).
"""

from __future__ import annotations

import numpy as np

_MASKS = {
    "fixed": (0, 1, 2), "pin_x": (0,), "pin_y": (1,), "pin_z": (2,),
    "roller_x": (1, 2), "roller_y": (0, 2), "roller_z": (0, 1),
}


def node_coords(spec) -> np.ndarray:
    """Node positions i*Lx/nelx etc., x slowest (bc_generator.py:63-71)."""
    ix, iy, iz = np.meshgrid(np.arange(spec.nelx + 1), np.arange(spec.nely + 1),
                             np.arange(spec.nelz + 1), indexing="ij")
    return np.column_stack([ix.ravel() * (spec.Lx / spec.nelx), iy.ravel() * (spec.Ly / spec.nely),
                            iz.ravel() * (spec.Lz / spec.nelz)])


def face_nodes(face: str, nelx: int, nely: int, nelz: int) -> np.ndarray:
    """Node ids on a named face in the reference's loop order (bc_generator.py:136-166)."""
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sel = {"left": ii == 0, "right": ii == nelx, "bottom": jj == 0, "top": jj == nely,
           "front": kk == 0, "back": kk == nelz}.get(face)
    if sel is None:
        raise ValueError(f"Unknown 3D face: {face}")
    ids = (ii * ny + jj) * nz + kk
    return ids[sel].ravel()            # C order == the reference's nested loops


def nearest_node(coords: np.ndarray, p) -> int:
    """argmin of Euclidean distance, first index on ties (bc_generator.py:108-111)."""
    return int(np.argmin(np.linalg.norm(coords - np.asarray(p, dtype=float)[None, :], axis=1)))


def tributary_weights(pts: np.ndarray, spec, face: str) -> np.ndarray:
    """Trapezoidal tributary areas on a face (bc_generator.py:290-311)."""
    if face in ("left", "right"):
        h1, h2, c1, c2 = spec.Ly / spec.nely, spec.Lz / spec.nelz, pts[:, 1], pts[:, 2]
    elif face in ("bottom", "top"):
        h1, h2, c1, c2 = spec.Lx / spec.nelx, spec.Lz / spec.nelz, pts[:, 0], pts[:, 2]
    else:
        h1, h2, c1, c2 = spec.Lx / spec.nelx, spec.Ly / spec.nely, pts[:, 0], pts[:, 1]
    w1 = np.full(len(pts), h1)
    w1[np.abs(c1 - c1.min()) < 0.01 * h1] = h1 / 2
    w1[np.abs(c1 - c1.max()) < 0.01 * h1] = h1 / 2
    w2 = np.full(len(pts), h2)
    w2[np.abs(c2 - c2.min()) < 0.01 * h2] = h2 / 2
    w2[np.abs(c2 - c2.max()) < 0.01 * h2] = h2 / 2
    return w1 * w2


def generate_bc(spec):
    """(fixed, free, F, ndof) for a 3-D spec (bc_generator.py:381-447)."""
    nelx, nely, nelz = spec.nelx, spec.nely, spec.nelz
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    coords = node_coords(spec)
    fixed = set()
    for s in spec.supports:
        offs = _MASKS.get(s.constraint, (0, 1, 2))
        if hasattr(s, "edge"):
            nodes = face_nodes(s.edge, nelx, nely, nelz)
        else:
            nodes = [nearest_node(coords, (s.x, s.y, s.z))]
        for n in nodes:
            for o in offs:
                fixed.add(3 * int(n) + o)
    fixed = np.array(sorted(fixed), dtype=np.int64)
    free = np.setdiff1d(np.arange(ndof), fixed)
    F = np.zeros(ndof)
    for ld in spec.loads:
        if hasattr(ld, "edge"):
            nodes = face_nodes(ld.edge, nelx, nely, nelz)
            w = tributary_weights(coords[nodes], spec, ld.edge)
            comp = {"left": 0, "right": 0, "bottom": 1, "top": 1, "front": 2, "back": 2}[ld.edge]
            sign = 1.0 if ld.edge in ("right", "top", "back") else -1.0
            np.add.at(F, 3 * nodes + comp, ld.magnitude * sign * w)
        else:
            n = nearest_node(coords, (ld.x, ld.y, ld.z))
            F[3 * n] += ld.fx
            F[3 * n + 1] += ld.fy
            F[3 * n + 2] += ld.fz
    F[fixed] = 0.0
    if np.max(np.abs(F[free])) < 1e-30:
        raise ValueError("Force vector is zero on all free DOFs.")
    return fixed, free, F, ndof
