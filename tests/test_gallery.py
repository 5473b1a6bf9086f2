"""Gallery and tool-paper cases (src/topogpu/gallery.py:10-93,
experiments/tool_paper/fast_new_topology_probe.py:31-235).

The tool-paper module itself does not import in this container (it needs
pyvista and CuPy at import time), so its patch loads are checked against
what they must be: the nodes of the x = Lx face within the patch radius,
found here by brute force over the oracle's node coordinates, or the 12
nearest when none is that close; each patch's load vector split evenly; the
total force equal to the sum of the load vectors. The portal bridge goes
through generate_bc and is checked against the reference's own generate_bc
(tests/golden/tp_portal_bridge_24x12x10, written from the same spec).
"""

from __future__ import annotations

from types import SimpleNamespace as NS

import numpy as np
import pytest

import oracle
import topogpu_amd as tg
from oracle.bc import face_nodes, node_coords
from topogpu_amd.bc import generate_bc
from topogpu_amd.gallery import TOOL_CASES


def _expected_patch_force(prob, patches):
    sp = prob.spec
    coords = node_coords(sp)
    nodes = face_nodes("right", sp.nelx, sp.nely, sp.nelz)
    F = np.zeros(3 * coords.shape[0])
    for p in patches:
        c = np.array([p["load_rel"][1] * sp.Ly, p["load_rel"][2] * sp.Lz])
        d = np.sqrt(((coords[nodes, 1:] - c) ** 2).sum(axis=1))
        sel = nodes[d <= p["patch_radius_rel"] * min(sp.Ly, sp.Lz)]
        if sel.size == 0:
            sel = nodes[np.argsort(d)[:12]]
        for n in sel:
            F[3 * n:3 * n + 3] += np.asarray(p["load_vector"]) / sel.size
    return F


@pytest.mark.parametrize("name", [n for n, c in TOOL_CASES.items() if c[0] != "portal_bridge"])
def test_patch_cases(name):
    kind, nel, volfrac, rmin, patches, role = TOOL_CASES[name]
    prob = tg.tool_case(name)
    assert prob.nel == nel and prob.volfrac == volfrac and prob.filter_radius == rmin == prob.spec.rmin
    assert prob.role == (role or "tool_paper_case")
    assert (prob.spec.Lx, prob.spec.Ly, prob.spec.Lz) == ((2.0, 1.25, 0.75) if kind == "asym_bracket" else (2.4, 1.0, 1.0))
    F = prob.force_override
    assert np.allclose(F, _expected_patch_force(prob, patches), rtol=0, atol=1e-15)
    total = np.sum([p["load_vector"] for p in patches], axis=0)
    assert np.allclose(F.reshape(-1, 3).sum(axis=0), total, rtol=0, atol=1e-12)
    loaded = np.nonzero(F.reshape(-1, 3).any(axis=1))[0]
    assert np.all(loaded // ((nel[1] + 1) * (nel[2] + 1)) == nel[0])         # all on the x = Lx face
    b = generate_bc(prob.spec)                                                 # the placeholder load validates
    assert np.all(F[b.fixed_dofs] == 0.0)


def test_patch_fallback_to_twelve_nearest():
    """No face node within the radius on a coarse grid: the 12 nearest take
    the load (fast_new_topology_probe.py:182-183)."""
    prob = tg.tool_case("tool_long_cantilever_vf16", dims="4x3x3")
    F = prob.force_override
    assert np.count_nonzero(F.reshape(-1, 3).any(axis=1)) == 12
    assert np.allclose(F, _expected_patch_force(prob, TOOL_CASES["tool_long_cantilever_vf16"][4]), atol=1e-15)
    with pytest.raises(ValueError, match="nelx x nely x nelz"):
        tg.tool_case("tool_long_cantilever_vf16", dims="4x3")
    with pytest.raises(KeyError, match="Unknown case"):
        tg.tool_case("nope")


def test_portal_bridge_matches_reference_generate_bc(golden):
    g = golden("tp_portal_bridge_24x12x10")
    prob = tg.tool_case("tool_portal_bridge_vf30", dims="24x12x10")
    assert prob.role == "non_cantilever_candidate" and prob.filter_radius == 2.8
    b = generate_bc(prob.spec)
    assert np.array_equal(b.free_dofs, g["free"])
    assert np.allclose(prob.force_override, g["F"], rtol=0, atol=1e-16)


def test_gallery_cantilevers():
    s = tg.side_load_cantilever()
    assert s.nel == (56, 36, 28) and s.volfrac == 0.24 and s.filter_radius == 2.3
    assert s.spec.loads[0].fy == -1.0 and s.spec.loads[0].fz == 0.0 and s.metadata["load"] == "tip_point_y"
    c = tg.CaseGallery.cantilever_3d()
    assert c.nel == (24, 12, 6) and c.spec.loads[0].fz == -1.0
    assert tg.CaseGallery.tool_case is tg.tool_case
    with pytest.raises(ValueError):
        tg.cantilever_3d(support="xmax")
