"""CPU tests of the C ABI surface and the host-side logic of topogpu_amd."""

from __future__ import annotations

import csv
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
import topogpu_amd as tg
from oracle import gmg as ogmg
from topogpu_amd import bc, fem
from topogpu_amd.hip import lib

ROOT = Path(__file__).resolve().parent.parent


def _header_symbols():
    text = (ROOT / "include" / "topogpu_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tg_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    names = _header_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert sorted(lib.EXPORTED) == names


def test_library_reports_errors_without_gpu():
    L = lib.load()
    rc = L.tg_fill32(10, None, 0.0, None)
    assert rc == 1
    assert b"fill32" in L.tg_last_error()


def test_product_ke_equals_oracle():
    assert np.abs(fem.unit_stiffness() - oracle.KE_UNIT_3D).max() < 1e-15


def _specs():
    yield tg.ProblemSpec(Lx=24, Ly=12, Lz=6, nelx=24, nely=12, nelz=6, volfrac=0.3,
                         supports=[tg.EdgeSupport("left")], loads=[tg.PointLoad(24, 6, 0, fz=-1)])
    yield tg.ProblemSpec(Lx=2, Ly=1, Lz=0.5, nelx=8, nely=4, nelz=3, volfrac=0.3,
                         supports=[tg.EdgeSupport("left"), tg.PointSupport(1.0, 0.0, 0.25, "pin_y")],
                         loads=[tg.PointLoad(2.0, 0.5, 0.25, fy=-1.0), tg.DistributedLoad("top", -0.5)])
    yield tg.ProblemSpec(Lx=6, Ly=4, Lz=4, nelx=6, nely=4, nelz=4, volfrac=0.4,
                         supports=[tg.EdgeSupport("bottom", "roller_y"), tg.EdgeSupport("front", "pin_z"),
                                   tg.EdgeSupport("left", "pin_x")],
                         loads=[tg.DistributedLoad("back", -1.0), tg.PointLoad(3.3, 2.7, 4.0, fx=0.4)])


def test_bc_generation_matches_oracle():
    for spec in _specs():
        mine = bc.generate_bc(spec)
        fixed, free, F, ndof = oracle.generate_bc(spec)
        assert mine.ndof == ndof
        assert np.array_equal(mine.free_dofs, free)
        assert np.array_equal(mine.fixed_dofs, fixed)
        assert np.abs(mine.F - F).max() < 1e-15


def test_bc_errors():
    spec = tg.ProblemSpec(Lx=4, Ly=2, Lz=2, nelx=4, nely=2, nelz=2, supports=[tg.EdgeSupport("left")],
                          loads=[tg.PointLoad(0.0, 1.0, 1.0, fz=-1.0)])
    with pytest.raises(ValueError, match="zero on all free DOFs"):
        bc.generate_bc(spec)
    with pytest.raises(ValueError, match="3D"):
        bc.generate_bc(tg.ProblemSpec(supports=[tg.EdgeSupport("left")], loads=[tg.PointLoad(2, 1, fy=-1)]))


@pytest.mark.parametrize("rmin", [1.5, 2.0, 2.6, 3.0])
def test_filter_stencil_matches_oracle(rmin):
    o_offs, o_w = oracle.simp.filter_offsets(rmin)
    offs, w = fem.filter_stencil(rmin)
    a = sorted(zip(map(tuple, o_offs), np.round(o_w, 14)))
    b = sorted(zip(map(tuple, offs.astype(np.int64)), np.round(w, 14)))
    assert a == b
    assert tuple(offs[0]) == (0, 0, 0) and w[0] == rmin


def test_masks_consistency_rule():
    dims = (8, 4, 4)
    nn = 9 * 5 * 5
    face = np.full(nn, 7, np.uint8)
    face.reshape(9, 5, 5)[0] = 0
    dc = fem.coarse_dims(dims)
    assert fem.masks_consistent(dims, dc, face, fem.injection_mask(dims, dc, face))
    point = face.copy()
    point.reshape(9, 5, 5)[3, 2, 2] = 0          # odd x index: parents are free
    assert not fem.masks_consistent(dims, dc, point, fem.injection_mask(dims, dc, point))


def test_level_dims_match_oracle():
    for dims in [(24, 12, 6), (160, 80, 80), (7, 3, 1)]:
        assert fem.level_dims(dims, 5) == ogmg.level_dims(*dims, 5)


def test_solver_facade_errors():
    with pytest.raises(ValueError, match="optimizer"):
        tg.SIMPSolver(optimizer="mma")
    prob = tg.cantilever_3d((4, 3, 2))
    with pytest.raises(ValueError, match="Unknown backend"):
        tg.SIMPSolver(backend="cuda").solve(prob)
    bad = tg.TopologyProblem("bad", tg.ProblemSpec(nelx=1, supports=[], loads=[]))
    with pytest.raises(ValueError, match="Invalid TopoGPU problem"):
        tg.SIMPSolver(backend="hip").solve(bad)


def test_problem_roundtrip_and_gallery():
    p = tg.cantilever_3d(nel=(4, 3, 2), volfrac=0.25)
    p.validate()
    assert p.nel == (4, 3, 2) and p.volfrac == 0.25 and p.metadata["support"] == "xmin"
    d = p.spec.to_dict()
    assert tg.ProblemSpec.from_dict(d).to_dict() == d
    u = tg.unit_cantilever((8, 4, 4), load="line")
    assert u.force_override.sum() == pytest.approx(-1.0)
    assert np.count_nonzero(u.force_override) == 5


def test_evidence_bundle(tmp_path):
    res = tg.OptimizationResult(tg.cantilever_3d((4, 3, 2)), np.ones(24), [{"iteration": 1, "compliance": 2.0}],
                                {"status": "ok"})
    b = res.save(tmp_path / "run")
    with open(tmp_path / "run" / "ARTIFACT_MANIFEST.csv", newline="") as fh:
        rows = list(csv.DictReader(fh))
    assert {r["path"] for r in rows} == {"history.csv", "summary.json", "rho_final.npy", "render_metadata.json"}
    assert all(len(r["sha256"]) == 64 for r in rows)
    assert b.root == tmp_path / "run"


def test_checkpoint_files_roundtrip(tmp_path):
    """Arrays in <path>.npz, scalars in <path>.json (pub_simp_solver.py:
    397-410); nothing is loaded unless both files exist."""
    from topogpu_amd.solver import load_checkpoint, save_checkpoint

    path = str(tmp_path / "ckpt_4x3x2_case")
    assert load_checkpoint(path) is None
    rho = np.linspace(0.0, 1.0, 24)
    save_checkpoint(path, {"rho": rho, "compliance_hist": np.array([3.0, 2.5]), "iteration": 2,
                           "history": [{"iteration": 1}, {"iteration": 2}], "move": 0.2})
    assert sorted(p.name for p in tmp_path.iterdir()) == ["ckpt_4x3x2_case.json", "ckpt_4x3x2_case.npz"]
    got = load_checkpoint(path)
    assert np.array_equal(got["rho"], rho) and got["iteration"] == 2 and got["move"] == 0.2
    assert list(got["compliance_hist"]) == [3.0, 2.5] and got["history"][1] == {"iteration": 2}
    (tmp_path / "ckpt_4x3x2_case.json").unlink()
    assert load_checkpoint(path) is None


def test_passive_mask_matches_bruteforce():
    """build_passive_mask (bc_generator.py:84-93, 318-358) restated as a
    loop over elements: centroids (i + 1/2) h, a cylinder along z and an x-y
    box with inclusive boundaries, later regions overwriting earlier ones."""
    from topogpu_amd.problem_spec import CircularRegion, RectangularRegion

    spec = tg.ProblemSpec(Lx=8.0, Ly=4.0, Lz=2.0, nelx=8, nely=4, nelz=2, volfrac=0.4,
                          supports=[tg.EdgeSupport("left", "fixed")],
                          loads=[tg.PointLoad(x=8.0, y=0.0, z=1.0, fy=-1.0)],
                          passive_regions=[RectangularRegion(0.5, 0.5, 2.5, 1.5, "solid"),
                                           CircularRegion(2.5, 1.5, 1.0, "void")])
    got = bc.generate_bc(spec).passive_mask
    want = np.zeros(64, dtype=np.int32)
    for i in range(8):
        for j in range(4):
            for k in range(2):
                x, y = i + 0.5, j + 0.5
                v = 0
                if 0.5 <= x <= 2.5 and 0.5 <= y <= 1.5:
                    v = 2
                if ((x - 2.5) ** 2 + (y - 1.5) ** 2) ** 0.5 <= 1.0:
                    v = 1
                want[(i * 4 + j) * 2 + k] = v
    assert np.array_equal(got, want) and (got == 1).sum() == 10 and (got == 2).sum() == 6
    spec.passive_regions = [CircularRegion(100.0, 100.0, 1.0, "void")]
    assert bc.generate_bc(spec).passive_mask is None


def test_solver_v4_contract_validation():
    """HipSolverV4 rejects what SolverV4 rejects (solver_v4.py:66-82), with
    the same messages, before touching the GPU."""
    from topogpu_amd.hip.solver_v4 import HipSolverV4, structured_edof

    import oracle

    dims = (4, 3, 2)
    assert np.array_equal(structured_edof(dims), oracle.edof_table_3d(*dims))
    ndof = 3 * 5 * 4 * 3
    base = dict(free=np.arange(ndof), F=np.ones(ndof), ndof=ndof, enable_matrix_free=True, grid_dims=dims)
    with pytest.raises(ValueError, match="requires grid_dims"):
        HipSolverV4(**{**base, "grid_dims": None})
    with pytest.raises(ValueError, match="requires enable_matrix_free=True"):
        HipSolverV4(**{**base, "enable_matrix_free": False})
    with pytest.raises(ValueError, match="gmg_fine_smoother must be one of"):
        HipSolverV4(**base, gmg_fine_smoother="fp16")
    with pytest.raises(ValueError, match="gmg_outer_solver must be"):
        HipSolverV4(**base, gmg_outer_solver="gmres")
    with pytest.raises(ValueError, match="edof must be the structured"):
        HipSolverV4(**base, edof=oracle.edof_table_3d(*dims)[::-1])
    with pytest.raises(ValueError, match="KE_UNIT"):
        HipSolverV4(**base, KE_UNIT=2.0 * oracle.KE_UNIT_3D)
    with pytest.raises(ValueError, match="ndof"):
        HipSolverV4(**{**base, "ndof": ndof + 3})
