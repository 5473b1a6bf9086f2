"""The 1M-element cantilever workload (x=0 face clamped, total -1 in z on the
x=nelx, z=0 edge, volfrac 0.3, rmin 1.5, move 0.2, Galerkin GMG-PCG,
tol 1e-5) and the other reference golden cases on the GPU.

* 32x16x16 with the workload's solver settings (4 levels, Chebyshev degree 3)
  against the reference's own run_simp trajectory (tests/golden/
  c2r_line_32x16x16, written with spsolve): compliance history within
  1e-5 relative at tol 1e-5, 1e-6 at tol 1e-8;
* the full 160x80x80 size, first two SIMP iterations: every state solve's
  true residual recomputed on the host by the oracle's FP64 matrix-free
  operator, and the compliances of the bf16, bf16x3 and fp32 paths (and the
  4-level / degree-3 solver) within 1e-5 of a tol 1e-10 solve of the same
  design, whose residual the oracle certifies too;
* face pressure (DistributedLoad) and the tool-paper portal bridge against
  reference goldens; the mean-volume OC against the reference function's
  own outputs.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

import oracle
import topogpu_amd as tg
from topogpu_amd.bc import generate_bc
from topogpu_amd.hip import lib as L
from topogpu_amd.hip.simp import HipSIMP
from topogpu_amd.problem_spec import ProblemSpec

from conftest import GOLDEN
from gpu_util import DEV, h, stream, t

pytestmark = pytest.mark.gpu


def _golden_problem(name: str):
    case = json.loads((GOLDEN / f"{name}.json").read_text())["case"]
    spec = ProblemSpec.from_dict(case["spec"])
    return tg.TopologyProblem(name, spec, filter_radius=case["params"]["rmin"]), case["params"]


@pytest.mark.parametrize("precision,tol,rtol", [("bf16", 1e-5, 1e-5), ("fp32", 1e-5, 1e-5), ("bf16", 1e-8, 1e-6)])
def test_c2_line_load_matches_reference_golden(gpu, golden, precision, tol, rtol):
    """The workload's solver (4 levels, degree 3) on the 32x16x16 line-load
    cantilever, 5 OC iterations, against the reference run_simp."""
    g = golden("c2r_line_32x16x16")
    prob = tg.unit_cantilever((32, 16, 16), volfrac=0.30, rmin=1.5, load="line")
    assert np.array_equal(prob.force_override, g["F"])
    res = tg.SIMPSolver(backend="hip", tol=tol, max_krylov=800, max_iter=5, move=0.2, precision=precision,
                        n_levels=4, fine_degree=3, coarse_degree=3).solve(prob)
    assert res.summary["n_levels"] == 4
    comp = np.array([r["compliance"] for r in res.history])
    rel = np.abs(comp / g["compliance_history"] - 1)
    rho_err = float(np.abs(res.rho_final - g["rho_final"]).max())
    print(f"c2r {precision} tol {tol}: compliance rel err {rel.max():.2e}, rho_final max abs err {rho_err:.2e}")
    assert rel.max() < rtol, rel
    assert all(r["converged"] and r["rel_residual"] <= tol for r in res.history)
    assert res.summary["final_compliance"] == pytest.approx(float(g["final_compliance"]), rel=rtol)
    assert rho_err < (1e-3 if tol == 1e-5 else 1e-5)


@pytest.mark.parametrize("name", ["c0d_pressure_12x6x6", "tp_portal_bridge_24x12x10"])
def test_golden_cases_match_reference(gpu, golden, name):
    """Face pressure with trapezoidal tributary areas on a non-unit mesh, and
    the tool-paper portal bridge (face rollers, rmin 2.8), against the
    reference run_simp."""
    g = golden(name)
    prob, p = _golden_problem(name)
    b = generate_bc(prob.spec)
    assert np.array_equal(b.free_dofs, g["free"])
    assert np.allclose(b.F, g["F"], rtol=1e-15, atol=1e-16)
    res = tg.SIMPSolver(backend="hip", tol=1e-8, max_krylov=800, max_iter=p["max_iter"], move=p["move"]).solve(prob)
    comp = np.array([r["compliance"] for r in res.history])
    assert np.allclose(comp, g["compliance_history"], rtol=1e-6, atol=0), comp / g["compliance_history"] - 1
    assert np.abs(res.rho_final - g["rho_final"]).max() < 1e-5
    assert res.summary["n_iter"] == int(g["n_iter"])


def test_oc_mean_matches_reference_fixture_on_device(gpu, golden):
    """The device mean-volume OC on the inputs of the reference function's
    own fixture: the updated densities bit for bit."""
    g = golden("oc_mean_cases")
    cases = json.loads((GOLDEN / "oc_mean_cases.json").read_text())["cases"]
    for i, c in enumerate(cases):
        rho, dc, want = g[f"rho_{i}"], g[f"dc_{i}"], g[f"rho_new_{i}"]
        n = rho.size
        state = torch.zeros(L.OC_MEAN_STATE_LEN, dtype=torch.float64, device=DEV)
        out = torch.zeros(n, dtype=torch.float64, device=DEV)
        chg = torch.zeros(1, dtype=torch.float64, device=DEV)
        dr, ddc = t(rho), t(dc)
        L.call("tg_oc_mean_prepare", L.ptr(state), stream())
        for _ in range(40):
            L.call("tg_oc_mean_step", n, L.ptr(dr), L.ptr(ddc), c["volfrac"], c["move"], c["rho_min"], L.ptr(state),
                   stream())
        L.call("tg_oc_mean_finalize", n, L.ptr(dr), L.ptr(ddc), c["move"], c["rho_min"], L.ptr(state), L.ptr(out),
               L.ptr(chg), stream())
        assert h(state)[7] == 1.0, c["name"]
        assert np.array_equal(h(out), want), c["name"]


# -------------------------------------------------------------- full size
DIMS = (160, 80, 80)


def _host_residual(E, U, F, edof, free_mask):
    """||F - K(E) U|| / ||F|| with the oracle's FP64 matrix-free operator."""
    r = F - oracle.matfree_matvec(U, E, edof, free_mask)
    r[~free_mask] = 0.0
    return float(np.linalg.norm(r) / np.linalg.norm(F[free_mask]))


@pytest.fixture(scope="module")
def c2_full(gpu):
    """Reference designs and compliances of the first two SIMP iterations at
    160x80x80 from a tol 1e-10 solve; each residual certified on the host."""
    prob = tg.unit_cantilever(DIMS, volfrac=0.30, rmin=1.5, load="line")
    bc = generate_bc(prob.spec)
    F = prob.force_override.copy()
    F[bc.fixed_dofs] = 0.0
    free_mask = np.zeros(F.size, bool)
    free_mask[bc.free_dofs] = True
    edof = oracle.edof_table_3d(*DIMS)
    eng = HipSIMP(DIMS, bc.free_dofs, F, volfrac=0.3, rmin=1.5, move=0.2, precision="bf16x3", tol=1e-10,
                  max_krylov=2000)
    designs, comps = [], []
    for it in range(2):
        designs.append(eng.rho.cpu().numpy().copy())
        eng.physical()
        c, st = eng.state()
        assert st.converged
        res = _host_residual(eng.E64.cpu().numpy(), eng.fea.u.cpu().numpy(), F, edof, free_mask)
        print(f"c2 full reference solve {it}: compliance {c:.12g}, host FP64 residual {res:.2e}")
        assert res <= 1e-9
        comps.append(c)
        eng.sensitivities()
        eng.oc()
    del eng
    torch.cuda.empty_cache()
    return dict(bc=bc, F=F, free_mask=free_mask, edof=edof, designs=designs, comps=comps)


@pytest.mark.parametrize("precision,levels,degree", [("bf16", "auto", None), ("bf16x3", "auto", None),
                                                     ("fp32", "auto", None), ("bf16", 4, 3)])
def test_c2_full_size_residual_and_compliance(c2_full, precision, levels, degree):
    d = c2_full
    kw = dict(n_levels=levels)
    if degree is not None:
        kw.update(fine_degree=degree, coarse_degree=degree)
    eng = HipSIMP(DIMS, d["bc"].free_dofs, d["F"], volfrac=0.3, rmin=1.5, move=0.2, precision=precision, tol=1e-5,
                  max_krylov=800, **kw)
    for it in range(2):
        eng.rho.copy_(torch.from_numpy(d["designs"][it]))
        eng.physical()
        c, st = eng.state()
        res = _host_residual(eng.E64.cpu().numpy(), eng.fea.u.cpu().numpy(), d["F"], d["edof"], d["free_mask"])
        rel = abs(c / d["comps"][it] - 1)
        print(f"c2 full {precision} levels={levels} degree={degree} iteration {it}: its {st.iters}, "
              f"device residual {st.rel_residual:.2e}, host FP64 residual {res:.2e}, compliance rel err {rel:.2e}")
        assert st.converged and res <= 1e-5
        assert abs(res - st.rel_residual) <= 1e-3 * res + 1e-12
        assert rel <= 1e-5
    del eng
    torch.cuda.empty_cache()


@pytest.mark.parametrize("name,dims,iters", [("tool_short_cantilever_vf25", "24x16x12", 3),
                                             ("tool_dual_load_cantilever_vf26", "20x12x10", 3),
                                             ("tool_asymmetric_bracket_vf14", "24x16x10", 2)])
def test_tool_cases_match_oracle(gpu, name, dims, iters):
    """Tool-paper patch-load cases (reduced dims, their own rmin) on the GPU
    against the oracle's run_simp on the same load vector."""
    prob = tg.tool_case(name, dims=dims)
    b = generate_bc(prob.spec)
    res = tg.SIMPSolver(backend="hip", tol=1e-8, max_krylov=800, max_iter=iters, move=0.2).solve(prob)
    F = prob.force_override.copy()
    F[b.fixed_dofs] = 0.0
    ref = oracle.run_simp(*prob.nel, prob.volfrac, b.free_dofs, F, rmin=prob.filter_radius, move=0.2,
                          max_iter=iters, min_iter=min(2, iters), tol=1e-3)
    comp = np.array([r["compliance"] for r in res.history])
    assert np.allclose(comp, ref["compliance_history"], rtol=1e-6, atol=0), comp / ref["compliance_history"] - 1
    assert np.abs(res.rho_final - ref["rho_final"]).max() < 1e-5
