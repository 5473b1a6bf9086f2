"""TopologyProblem (src/topogpu/problem.py:13-58) and case builders."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .problem_spec import EdgeSupport, PointLoad, ProblemSpec


@dataclass(slots=True)
class TopologyProblem:
    """Structured-grid SIMP problem: a ProblemSpec plus filter radius and an
    optional full-length force vector that replaces the spec's loads."""

    name: str
    spec: ProblemSpec
    filter_radius: float | None = None
    force_override: np.ndarray | None = None
    role: str = "user_case"
    metadata: dict[str, Any] = field(default_factory=dict)

    @property
    def nel(self) -> tuple[int, int, int]:
        return (self.spec.nelx, self.spec.nely, self.spec.nelz)

    @property
    def volfrac(self) -> float:
        return float(self.spec.volfrac)

    @property
    def is_3d(self) -> bool:
        return self.spec.is_3d

    def validate(self) -> None:
        errors = self.spec.validate()
        if errors:
            raise ValueError(f"Invalid TopoGPU problem {self.name!r}: {'; '.join(errors)}")

    def to_dict(self) -> dict[str, Any]:
        data = self.spec.to_dict()
        data.update({"name": self.name, "filter_radius": self.filter_radius, "role": self.role,
                     "metadata": dict(self.metadata), "has_force_override": self.force_override is not None})
        return data


def cantilever_3d(nel=(24, 12, 6), volfrac: float = 0.30, filter_radius: float = 1.5, support: str = "xmin",
                  load: str = "tip_patch_z") -> TopologyProblem:
    """Gallery cantilever (src/topogpu/gallery.py:10-50): x=0 face clamped,
    unit tip load at (Lx, Ly/2, Lz/2) with Lx=2, Ly=1, Lz=0.5."""
    if support != "xmin":
        raise ValueError("Only support='xmin' is currently implemented.")
    if load not in {"tip_patch_z", "tip_point_y"}:
        raise ValueError("Supported loads are 'tip_patch_z' and 'tip_point_y'.")
    force = {"tip_patch_z": {"fz": -1.0}, "tip_point_y": {"fy": -1.0}}[load]
    nelx, nely, nelz = nel
    spec = ProblemSpec(Lx=2.0, Ly=1.0, Lz=0.5, nelx=nelx, nely=nely, nelz=nelz, volfrac=volfrac,
                       supports=[EdgeSupport(edge="left", constraint="fixed")],
                       loads=[PointLoad(x=2.0, y=0.5, z=0.25, **force)], rmin=filter_radius)
    return TopologyProblem("cantilever_3d", spec, filter_radius=filter_radius, role="example",
                           metadata={"support": support, "load": load})


def unit_cantilever(nel, volfrac: float = 0.30, rmin: float = 1.5, load: str = "point") -> TopologyProblem:
    """Unit-element cantilever used by the benchmarks: x=0 face clamped.

    load="point": unit -z force at node (nelx, nely/2, 0).
    load="line":  total -1 in z along the x=nelx, z=0 edge, trapezoidal split
                  (end nodes half weight), given as ``force_override``.
    """
    nelx, nely, nelz = nel
    spec = ProblemSpec(Lx=float(nelx), Ly=float(nely), Lz=float(nelz), nelx=nelx, nely=nely, nelz=nelz,
                       volfrac=volfrac, supports=[EdgeSupport("left", "fixed")],
                       loads=[PointLoad(x=float(nelx), y=nely / 2.0, z=0.0, fz=-1.0)], rmin=rmin)
    force = None
    if load == "line":
        ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
        force = np.zeros(ndof)
        w = np.full(nely + 1, 1.0 / nely)
        w[0] = w[-1] = 0.5 / nely
        j = np.arange(nely + 1)
        nodes = (nelx * (nely + 1) + j) * (nelz + 1) + 0
        force[3 * nodes + 2] = -w
    elif load != "point":
        raise ValueError("load must be 'point' or 'line'")
    return TopologyProblem(f"cantilever_{nelx}x{nely}x{nelz}_{load}", spec, filter_radius=rmin,
                           force_override=force, role="benchmark", metadata={"load": load})


def multi_load_bracket(nel=(192, 64, 64), n_loads: int = 8, seed: int = 3, volfrac: float = 0.30,
                       rmin: float = 1.5) -> TopologyProblem:
    """Multi-load bracket: y=0 face clamped, ``n_loads`` load cases, each a
    unit point load at a node of the y=nely face.

    From ``np.random.default_rng(seed)``: the x indices of the nodes
    (0..nelx), then their z indices (0..nelz), then one Gaussian triple per
    case, normalised to unit length (a direction uniform on the sphere).
    ``force_override`` holds the cases as rows, shape (n_loads, ndof); the
    objective is the sum of their compliances. The reference solves one load
    case at a time; this case runs them as one block of right-hand sides.
    """
    nelx, nely, nelz = nel
    rng = np.random.default_rng(seed)
    ix = rng.integers(0, nelx + 1, n_loads)
    iz = rng.integers(0, nelz + 1, n_loads)
    dirs = rng.standard_normal((n_loads, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    force = np.zeros((n_loads, ndof))
    nodes = (ix * (nely + 1) + nely) * (nelz + 1) + iz
    for k in range(n_loads):
        force[k, 3 * nodes[k]: 3 * nodes[k] + 3] = dirs[k]
    # the spec's own load (first case) keeps validate() and generate_bc() whole
    spec = ProblemSpec(Lx=float(nelx), Ly=float(nely), Lz=float(nelz), nelx=nelx, nely=nely, nelz=nelz,
                       volfrac=volfrac, supports=[EdgeSupport("bottom", "fixed")],
                       loads=[PointLoad(x=float(ix[0]), y=float(nely), z=float(iz[0]), fx=dirs[0, 0],
                                        fy=dirs[0, 1], fz=dirs[0, 2])], rmin=rmin)
    return TopologyProblem(f"bracket_{nelx}x{nely}x{nelz}_{n_loads}loads", spec, filter_radius=rmin,
                           force_override=force, role="benchmark",
                           metadata={"n_loads": n_loads, "seed": seed, "nodes": nodes.tolist()})
