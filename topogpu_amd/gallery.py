"""Case gallery (src/topogpu/gallery.py:10-93) and the tool-paper cases it
bridges to (experiments/tool_paper/fast_new_topology_probe.py:31-235).

``cantilever_3d`` and ``side_load_cantilever`` build the gallery's
cantilever; ``tool_case(name, dims)`` builds one of the ten tool-paper cases
of ``TOOL_CASES``: a patch-loaded cantilever or bracket (x=0 face clamped;
each patch spreads its load vector evenly over the x=Lx face nodes within
``patch_radius_rel * min(Ly, Lz)`` of its centre in y-z, or over the 12
nearest face nodes when none is that close) or a portal bridge (x=0 face
clamped, x=Lx face on x-rollers, -y point load at (0.52 Lx, Ly, Lz/2)). The
patch load is returned as ``force_override``; the spec keeps a 1e-12 point
load at the first patch centre so that it validates, as the reference's.
"""

from __future__ import annotations

import numpy as np

from .bc import _face_node_ids, generate_bc
from .problem import TopologyProblem, cantilever_3d
from .problem_spec import EdgeSupport, PointLoad, ProblemSpec

__all__ = ["cantilever_3d", "side_load_cantilever", "tool_case", "TOOL_CASES", "CaseGallery"]


def side_load_cantilever(nel=(56, 36, 28), volfrac: float = 0.24, filter_radius: float = 2.3) -> TopologyProblem:
    """The gallery cantilever with its -y tip point load (gallery.py:53-64)."""
    return cantilever_3d(nel=nel, volfrac=volfrac, filter_radius=filter_radius, load="tip_point_y")


def _patch(load_rel, radius_rel, vec):
    return {"load_rel": load_rel, "patch_radius_rel": radius_rel, "load_vector": vec}


# (kind, dims, volfrac, rmin, patches or None, role); roles default to
# "tool_paper_case"
TOOL_CASES = {
    "tool_long_cantilever_vf16": ("patch_cantilever", (72, 36, 36), 0.16, 2.4,
                                  [_patch((1.0, 0.50, 0.18), 0.18, (0.0, 0.0, -1.0))], None),
    "tool_portal_bridge_vf18": ("portal_bridge", (80, 40, 32), 0.18, 2.8, None, None),
    "tool_asymmetric_bracket_vf14": ("asym_bracket", (72, 48, 32), 0.14, 2.4,
                                     [_patch((1.0, 0.72, 0.28), 0.16, (0.0, -0.7, -0.7))], None),
    "tool_short_cantilever_vf25": ("patch_cantilever", (48, 32, 24), 0.25, 2.2,
                                   [_patch((1.0, 0.50, 0.50), 0.20, (0.0, 0.0, -1.0))],
                                   "production_timing_candidate"),
    "tool_deep_cantilever_vf20": ("patch_cantilever", (64, 48, 32), 0.20, 2.4,
                                  [_patch((1.0, 0.50, 0.35), 0.18, (0.0, 0.0, -1.0))],
                                  "production_timing_candidate"),
    "tool_oblique_cantilever_vf22": ("patch_cantilever", (64, 40, 32), 0.22, 2.4,
                                     [_patch((1.0, 0.55, 0.35), 0.18, (0.0, -0.6, -0.8))],
                                     "production_timing_candidate"),
    "tool_side_load_cantilever_vf24": ("patch_cantilever", (56, 36, 28), 0.24, 2.3,
                                       [_patch((1.0, 0.62, 0.50), 0.20, (0.0, -1.0, 0.0))],
                                       "production_timing_candidate"),
    "tool_dual_load_cantilever_vf26": ("patch_cantilever", (56, 36, 28), 0.26, 2.3,
                                       [_patch((1.0, 0.35, 0.65), 0.16, (0.0, -0.45, -0.55)),
                                        _patch((1.0, 0.72, 0.30), 0.14, (0.0, 0.35, -0.65))],
                                       "production_timing_candidate"),
    "tool_high_volume_bracket_vf28": ("asym_bracket", (72, 48, 32), 0.28, 2.4,
                                      [_patch((1.0, 0.72, 0.28), 0.18, (0.0, -0.55, -0.45))],
                                      "non_cantilever_candidate"),
    "tool_portal_bridge_vf30": ("portal_bridge", (80, 40, 32), 0.30, 2.8, None, "non_cantilever_candidate"),
}


def _parse_dims(dims):
    d = tuple(int(x) for x in dims.lower().split("x"))
    if len(d) != 3:
        raise ValueError("--dims must have form nelx x nely x nelz, e.g. 72x36x36")
    return d


def patch_force(nel, lengths, patches) -> np.ndarray:
    """Full-length load vector of the patches on the x = Lx face."""
    (nx, ny, nz), (_, ly, lz) = nel, lengths
    nodes = _face_node_ids("right", nel)
    # y-z coordinates of the face nodes (node = (nx*(ny+1) + j)*(nz+1) + k)
    jk = nodes - nx * (ny + 1) * (nz + 1)
    yz = np.stack([(jk // (nz + 1)) * (ly / ny), (jk % (nz + 1)) * (lz / nz)], axis=1)
    F = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    for p in patches:
        centre = np.array([float(p["load_rel"][1]) * ly, float(p["load_rel"][2]) * lz])
        radius = float(p["patch_radius_rel"]) * min(ly, lz)
        dist = np.linalg.norm(yz - centre[None, :], axis=1)
        sel = nodes[dist <= radius]
        if sel.size == 0:
            sel = nodes[np.argsort(dist)[:12]]
        vec = np.asarray(p["load_vector"], dtype=float) / float(sel.size)
        for node in sel:
            F[3 * int(node): 3 * int(node) + 3] += vec
    return F


def tool_case(name: str, dims: str | None = None) -> TopologyProblem:
    """One of the tool-paper cases (gallery.py:67-86 -> build_problem,
    fast_new_topology_probe.py:224-235), optionally on other grid dims
    ("nelx x nely x nelz")."""
    if name not in TOOL_CASES:
        raise KeyError(f"Unknown case {name}. Available: {', '.join(TOOL_CASES)}")
    kind, nel, volfrac, rmin, patches, role = TOOL_CASES[name]
    if dims:
        nel = _parse_dims(dims)
    nx, ny, nz = nel
    if kind == "portal_bridge":
        lx, ly, lz = float(nx), float(ny), float(nz)
        spec = ProblemSpec(Lx=lx, Ly=ly, Lz=lz, nelx=nx, nely=ny, nelz=nz, volfrac=volfrac,
                           supports=[EdgeSupport("left", "fixed"), EdgeSupport("right", "roller_x")],
                           loads=[PointLoad(x=lx * 0.52, y=ly, z=lz * 0.5, fy=-1.0)], rmin=rmin)
        force = generate_bc(spec).F
    else:
        lx, ly, lz = (2.0, 1.25, 0.75) if kind == "asym_bracket" else (2.4, 1.0, 1.0)
        first = patches[0]["load_rel"]
        spec = ProblemSpec(Lx=lx, Ly=ly, Lz=lz, nelx=nx, nely=ny, nelz=nz, volfrac=volfrac,
                           supports=[EdgeSupport("left", "fixed")],
                           loads=[PointLoad(x=lx, y=float(first[1]) * ly, z=float(first[2]) * lz, fz=-1e-12)],
                           rmin=rmin)
        force = patch_force(nel, (lx, ly, lz), patches)
    return TopologyProblem(name=name, spec=spec, filter_radius=float(rmin), force_override=force,
                           role=role or "tool_paper_case",
                           metadata={"source": "experiments.tool_paper.fast_new_topology_probe"})


class CaseGallery:
    """Namespace for the gallery constructors (gallery.py:89-93)."""

    cantilever_3d = staticmethod(cantilever_3d)
    side_load_cantilever = staticmethod(side_load_cantilever)
    tool_case = staticmethod(tool_case)
