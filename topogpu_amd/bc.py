"""ProblemSpec -> (fixed, free, F, ndof) for 3-D structured meshes.

Semantics of ``generate_bc`` (src/gpu_fem/bc_generator.py:381-447): face
supports fix every node of the face, point supports and point loads snap to
the nearest node (lowest node index on ties), face pressure uses trapezoidal
tributary areas along the face's two in-plane axes, loads on fixed DOFs are
dropped, and an all-zero free load is an error.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .problem_spec import (CircularRegion, DistributedLoad, EdgeSupport, PointLoad, PointSupport,
                           RectangularRegion)

# face -> (normal axis, index on that axis is 0 or n, outward sign)
_FACES = {"left": (0, 0, -1.0), "right": (0, 1, 1.0), "bottom": (1, 0, -1.0), "top": (1, 1, 1.0),
          "front": (2, 0, -1.0), "back": (2, 1, 1.0)}


@dataclass
class BoundaryConditions:
    fixed_dofs: np.ndarray
    free_dofs: np.ndarray
    F: np.ndarray
    ndof: int
    passive_mask: np.ndarray | None = None   # per element: 0 free, 1 void, 2 solid


def passive_mask(spec) -> np.ndarray | None:
    """Passive elements by centroid (build_passive_mask, bc_generator.py:
    84-93, 318-358): a circular region is a cylinder along z (distance in x-y
    at most the radius), a rectangular one an x-y box prism, boundaries
    inclusive; later regions overwrite earlier ones. 1 = void, 2 = solid;
    None when no element is passive."""
    if not spec.passive_regions:
        return None
    (nx, ny, nz), (lx, ly, _) = _grid_axes(spec)
    hx, hy = lx / nx, ly / ny
    cx = np.repeat(np.arange(nx) * hx + 0.5 * hx, ny * nz)          # element e = (i*ny + j)*nz + k
    cy = np.tile(np.repeat(np.arange(ny) * hy + 0.5 * hy, nz), nx)
    mask = np.zeros(nx * ny * nz, dtype=np.int32)
    for r in spec.passive_regions:
        if isinstance(r, CircularRegion):
            dx, dy = cx - r.cx, cy - r.cy
            inside = np.sqrt(dx * dx + dy * dy) <= r.radius
        elif isinstance(r, RectangularRegion):
            inside = (cx >= r.x0) & (cx <= r.x1) & (cy >= r.y0) & (cy <= r.y1)
        else:
            raise TypeError(f"unsupported passive region {r!r}")
        mask[inside] = 1 if r.kind == "void" else 2
    return mask if np.any(mask) else None


def _grid_axes(spec):
    n = (spec.nelx, spec.nely, spec.nelz)
    L = (spec.Lx, spec.Ly, spec.Lz)
    return n, L


def _face_node_ids(face: str, n) -> np.ndarray:
    if face not in _FACES:
        raise ValueError(f"Unknown 3D face: {face}")
    axis, end, _ = _FACES[face]
    ranges = [np.arange(k + 1) for k in n]
    ranges[axis] = np.array([n[axis] if end else 0])
    I, J, K = np.meshgrid(*ranges, indexing="ij")
    return ((I * (n[1] + 1) + J) * (n[2] + 1) + K).ravel()


def _nearest_node(n, L, p) -> int:
    axes = [np.arange(k + 1) * (l / k) for k, l in zip(n, L)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    d = np.sqrt((X.ravel() - p[0]) ** 2 + (Y.ravel() - p[1]) ** 2 + (Z.ravel() - p[2]) ** 2)
    return int(np.argmin(d))


def _trapezoid(k: int, h: float) -> np.ndarray:
    w = np.full(k + 1, h)
    w[0] = w[-1] = h / 2.0
    return w


def generate_bc(spec) -> BoundaryConditions:
    errors = spec.validate()
    if errors:
        raise ValueError("ProblemSpec validation failed:\n" + "\n".join(f"  - {e}" for e in errors))
    if not spec.is_3d:
        raise ValueError("the hip backend requires a 3D problem (Lz > 0 and nelz > 0)")
    n, L = _grid_axes(spec)
    ndof = 3 * (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
    fixed = np.zeros(ndof, dtype=bool)
    for s in spec.supports:
        if isinstance(s, EdgeSupport):
            nodes = _face_node_ids(s.edge, n)
        elif isinstance(s, PointSupport):
            nodes = np.array([_nearest_node(n, L, (s.x, s.y, s.z))])
        else:
            raise TypeError(f"unsupported support {s!r}")
        for c in s.dof_mask(3):
            fixed[3 * nodes + c] = True
    F = np.zeros(ndof)
    for ld in spec.loads:
        if isinstance(ld, PointLoad):
            node = _nearest_node(n, L, (ld.x, ld.y, ld.z))
            F[3 * node: 3 * node + 3] += (ld.fx, ld.fy, ld.fz)
        elif isinstance(ld, DistributedLoad):
            axis, _, sign = _FACES[ld.edge]
            inplane = [a for a in range(3) if a != axis]
            h = [L[a] / n[a] for a in range(3)]
            w1 = _trapezoid(n[inplane[0]], h[inplane[0]])
            w2 = _trapezoid(n[inplane[1]], h[inplane[1]])
            shape = [1, 1, 1]
            shape[inplane[0]] = n[inplane[0]] + 1
            shape[inplane[1]] = n[inplane[1]] + 1
            wface = (w1[:, None] * w2[None, :]).reshape(shape).ravel()
            nodes = _face_node_ids(ld.edge, n)
            np.add.at(F, 3 * nodes + axis, ld.magnitude * sign * wface)
        else:
            raise TypeError(f"unsupported load {ld!r}")
    F[fixed] = 0.0
    free = np.nonzero(~fixed)[0]
    if free.size == 0 or np.max(np.abs(F[free])) < 1e-30:
        raise ValueError("Force vector is zero on all free DOFs. "
                         "Check that loads are not placed on fully-fixed supports.")
    return BoundaryConditions(np.nonzero(fixed)[0], free, F, ndof, passive_mask(spec))
