"""Write golden run_simp trajectories from the reference CPU path.

Run in a container where the reference sources are mounted read-only at
/root/reference (never on the GPU box):

    python tests/golden/make_golden.py [case ...]

It imports the reference's pure NumPy/SciPy CPU modules
(``gpu_fem.pub_simp_solver``, ``gpu_fem.bc_generator``,
``gpu_fem.problem_spec``), runs ``run_simp`` with ``bc_override`` built by the
reference's own ``generate_bc`` for each case, and stores inputs and outputs
as small ``.npz`` fixtures next to this script. Only data is stored.

Why an import: the composition of filter -> Heaviside -> solve -> sensitivity
chain -> filtered-volume OC bisection over several iterations is not fixed by
any closed form; the individual pieces are pinned by the oracle tests. The
controlled cases drive run_simp's continuation logic (actions, best
tracking, tail) with the deterministic controllers of controllers.py.
pyamg is absent here, so the reference solves with ``spsolve``
(pub_simp_solver.py:44-49): the direct solve is the stricter oracle.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

REF_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent

# Each case: spec fields for the reference ProblemSpec + SIMPParams overrides.
CASES = {
    # 24x12x6 cantilever: x=0 face clamped, unit -z load at node (24, 6, 0);
    # volfrac 0.30, rmin 1.5, 3 OC iterations, move 0.2.
    "c0_cantilever_24x12x6": dict(
        spec=dict(Lx=24.0, Ly=12.0, Lz=6.0, nelx=24, nely=12, nelz=6, volfrac=0.30,
                  supports=[dict(type="edge", edge="left", constraint="fixed")],
                  loads=[dict(type="point", x=24.0, y=6.0, z=0.0, fz=-1.0)]),
        params=dict(rmin=1.5, max_iter=3, min_iter=2, move=0.2, tol=1e-3),
    ),
    # Odd-sized block with a side (-y) tip load, a second filter radius and
    # more iterations, so convergence bookkeeping and odd dims are exercised.
    "c0b_block_10x5x3_rmin2": dict(
        spec=dict(Lx=10.0, Ly=5.0, Lz=3.0, nelx=10, nely=5, nelz=3, volfrac=0.4,
                  supports=[dict(type="edge", edge="left", constraint="fixed")],
                  loads=[dict(type="point", x=10.0, y=0.0, z=1.0, fy=-1.0)]),
        params=dict(rmin=2.0, max_iter=6, min_iter=2, move=0.2, tol=1e-3),
    ),
    # run_simp with a controller (controllers.py): actions, best tracking
    # and the tail phase on a 16x8x4 cantilever.
    "c0c_controlled_16x8x4_A": dict(
        spec=dict(Lx=16.0, Ly=8.0, Lz=4.0, nelx=16, nely=8, nelz=4, volfrac=0.30,
                  supports=[dict(type="edge", edge="left", constraint="fixed")],
                  loads=[dict(type="point", x=16.0, y=4.0, z=0.0, fz=-1.0)]),
        params=dict(rmin=1.5, max_iter=12, min_iter=2, move=0.2, tol=1e-3, max_gray_for_best=0.9),
        controller="A",
    ),
    "c0c_controlled_16x8x4_B": dict(
        spec=dict(Lx=16.0, Ly=8.0, Lz=4.0, nelx=16, nely=8, nelz=4, volfrac=0.30,
                  supports=[dict(type="edge", edge="left", constraint="fixed")],
                  loads=[dict(type="point", x=16.0, y=4.0, z=0.0, fz=-1.0)]),
        params=dict(rmin=1.5, max_iter=10, min_iter=2, move=0.2, tol=1e-3, beta_max=4.0),
        controller="B",
    ),
    # The 1M-element workload's load and BCs at 32x16x16: x=0 face clamped,
    # total -1 in z spread over the x=nelx, z=0 edge (trapezoidal: end nodes
    # half weight; force="line" below), volfrac 0.3, rmin 1.5, move 0.2,
    # 5 OC iterations.
    "c2r_line_32x16x16": dict(
        spec=dict(Lx=32.0, Ly=16.0, Lz=16.0, nelx=32, nely=16, nelz=16, volfrac=0.30,
                  supports=[dict(type="edge", edge="left", constraint="fixed")],
                  loads=[dict(type="point", x=32.0, y=8.0, z=0.0, fz=-1.0)]),
        params=dict(rmin=1.5, max_iter=5, min_iter=2, move=0.2, tol=1e-3),
        force="line",
    ),
    # Face pressure (DistributedLoad, trapezoidal tributary areas on a
    # non-unit mesh: h = 1/6 in every direction) on the top face of a
    # 12x6x6 beam clamped at x=0.
    "c0d_pressure_12x6x6": dict(
        spec=dict(Lx=2.0, Ly=1.0, Lz=1.0, nelx=12, nely=6, nelz=6, volfrac=0.30,
                  supports=[dict(type="edge", edge="left", constraint="fixed")],
                  loads=[dict(type="distributed", edge="top", magnitude=-1.0)]),
        params=dict(rmin=1.5, max_iter=4, min_iter=2, move=0.2, tol=1e-3),
    ),
    # The tool-paper portal bridge (fast_new_topology_probe.py:200-219:
    # x=0 face clamped, x=Lx face on x-rollers, -y point load at
    # (0.52 Lx, Ly, 0.5 Lz), rmin 2.8) at 24x12x10.
    "tp_portal_bridge_24x12x10": dict(
        spec=dict(Lx=24.0, Ly=12.0, Lz=10.0, nelx=24, nely=12, nelz=10, volfrac=0.30,
                  supports=[dict(type="edge", edge="left", constraint="fixed"),
                            dict(type="edge", edge="right", constraint="roller_x")],
                  loads=[dict(type="point", x=24.0 * 0.52, y=12.0, z=5.0, fy=-1.0)]),
        params=dict(rmin=2.8, max_iter=4, min_iter=2, move=0.2, tol=1e-3),
    ),
}


def line_load(nelx: int, nely: int, nelz: int) -> np.ndarray:
    """Total -1 in z over the nodes (nelx, j, 0), j = 0..nely, end nodes at
    half weight."""
    F = np.zeros(3 * (nelx + 1) * (nely + 1) * (nelz + 1))
    w = np.full(nely + 1, 1.0 / nely)
    w[0] = w[-1] = 0.5 / nely
    for j in range(nely + 1):
        F[3 * ((nelx * (nely + 1) + j) * (nelz + 1)) + 2] = -w[j]
    return F


# generate_bc inputs: every face under pressure, point supports of every
# constraint kind, face rollers, a non-unit and anisotropic mesh, and load
# points exactly between nodes (nearest-node ties).
BC_SPECS = [
    dict(Lx=2.4, Ly=1.0, Lz=0.75, nelx=8, nely=4, nelz=3, volfrac=0.3,
         supports=[dict(type="edge", edge="right" if face == "left" else "left", constraint="fixed")],
         loads=[dict(type="distributed", edge=face, magnitude=m)])
    for face, m in (("right", -1.0), ("top", -0.5), ("bottom", 2.0), ("front", -1.5), ("back", 0.7),
                    ("left", -1.0))
] + [
    dict(Lx=6.0, Ly=3.0, Lz=2.0, nelx=6, nely=3, nelz=2, volfrac=0.3,
         supports=[dict(type="edge", edge="bottom", constraint=c),
                   dict(type="point", x=3.0, y=3.0, z=1.0, constraint=c2)],
         loads=[dict(type="point", x=6.0, y=1.5, z=1.0, fx=0.3, fy=-1.0, fz=0.2),
                dict(type="point", x=2.5, y=3.0, z=0.5, fz=-0.4),
                dict(type="distributed", edge="back", magnitude=-0.25)])
    for c, c2 in (("fixed", "pin_x"), ("roller_x", "pin_y"), ("roller_y", "pin_z"), ("roller_z", "roller_x"),
                  ("pin_x", "roller_y"), ("pin_y", "roller_z"), ("pin_z", "fixed"))
] + [
    dict(Lx=80.0, Ly=40.0, Lz=32.0, nelx=20, nely=10, nelz=8, volfrac=0.18,
         supports=[dict(type="edge", edge="left", constraint="fixed"),
                   dict(type="edge", edge="right", constraint="roller_x")],
         loads=[dict(type="point", x=80.0 * 0.52, y=40.0, z=16.0, fy=-1.0)]),
]


def bc_fixture(generate_bc, ProblemSpec) -> None:
    out = {}
    for i, d in enumerate(BC_SPECS):
        bc = generate_bc(ProblemSpec.from_dict(d))
        out[f"fixed_{i}"] = np.asarray(bc.fixed_dofs, dtype=np.int64)
        out[f"F_{i}"] = np.asarray(bc.F, dtype=np.float64)
    np.savez_compressed(OUT / "bc_cases.npz", **out)
    (OUT / "bc_cases.json").write_text(json.dumps(dict(specs=BC_SPECS, generator="tests/golden/make_golden.py",
                                                       reference="src/gpu_fem/bc_generator.py generate_bc"),
                                                  indent=1) + "\n")
    print("bc_cases", len(BC_SPECS))


def oc_mean_fixture(oc_update) -> None:
    """Inputs and outputs of the GPU loop's mean-volume OC (src/topogpu/
    solver.py:94-117) on cases that reach each of its exits: the early stop
    inside the bisection, the bracket doubled up to its 1e40 cap (2^133),
    all 80 bisections with the mean above or below the target throughout,
    dc >= 0 entries (clamped to -1e-12) and densities at the bounds."""
    rng = np.random.default_rng(11)
    cases = []
    n = 1000
    rho = rng.uniform(0.05, 0.95, n)
    cases.append(("typical", rho, -rng.uniform(1e-2, 10.0, n), 0.5, 0.2, 1e-3))
    cases.append(("huge_dc_doubling", rho, -1e60 * rng.uniform(0.5, 2.0, n), 0.5, 0.2, 1e-3))
    cases.append(("target_below_reach", np.full(n, 0.5), -rng.uniform(0.1, 1.0, n), 0.05, 0.2, 1e-3))
    cases.append(("target_above_reach", np.full(n, 0.5), -rng.uniform(0.1, 1.0, n), 0.95, 0.2, 1e-3))
    dc = -rng.uniform(1e-3, 5.0, n)
    dc[::7] = rng.uniform(0.0, 1.0, dc[::7].size)
    rho_b = rng.uniform(1e-3, 1.0, n)
    rho_b[::5] = 1e-3
    rho_b[1::5] = 1.0
    cases.append(("positive_dc_and_bounds", rho_b, dc, 0.4, 0.1, 1e-3))
    cases.append(("rho_min_0.05", rng.uniform(0.0, 0.4, 1728), -rng.uniform(1e-4, 1.0, 1728) ** 3, 0.3, 0.2, 0.05))
    cases.append(("small_n_plateaus", np.array([0.3, 0.3, 0.3]), np.array([-1.0, -1.0, -1.0]), 0.3, 0.2, 1e-3))
    out, meta = {}, []
    for i, (name, r, d, vf, mv, rmn) in enumerate(cases):
        out[f"rho_{i}"], out[f"dc_{i}"] = r, d
        out[f"rho_new_{i}"] = np.asarray(oc_update(r.copy(), d.copy(), vf, mv, rmn), dtype=np.float64)
        meta.append(dict(name=name, volfrac=vf, move=mv, rho_min=rmn))
    np.savez_compressed(OUT / "oc_mean_cases.npz", **out)
    (OUT / "oc_mean_cases.json").write_text(json.dumps(dict(cases=meta, generator="tests/golden/make_golden.py",
                                                            reference="src/topogpu/solver.py _oc_update"),
                                                       indent=1) + "\n")
    print("oc_mean_cases", [m["name"] for m in meta])


INJECTION_DIMS = [(8, 4, 4), (7, 5, 3), (6, 3, 1), (3, 1, 1), (2, 1, 2), (1, 1, 1), (5, 2, 1), (12, 6, 6)]


def injection_fixture(injection) -> None:
    """Coarse free DOFs by injection (_coarse_free_dofs_injection,
    solver_v2.py:717-764) from fine grids with seeded random free sets
    (default_rng(13), 70 % free), coarse dims max(1, n // 2) -- including
    axes one element thick, where the reference's flat node index runs past
    the end of the line."""
    rng = np.random.default_rng(13)
    out, meta = {}, []
    for i, (nx, ny, nz) in enumerate(INJECTION_DIMS):
        ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        free = np.nonzero(rng.random(ndof) < 0.7)[0].astype(np.int64)
        c = tuple(max(1, n // 2) for n in (nx, ny, nz))
        out[f"free_{i}"] = free
        out[f"coarse_free_{i}"] = np.asarray(injection(nx, ny, nz, *c, free), dtype=np.int64)
        meta.append(dict(fine=[nx, ny, nz], coarse=list(c)))
    np.savez_compressed(OUT / "injection_cases.npz", **out)
    (OUT / "injection_cases.json").write_text(json.dumps(dict(cases=meta, generator="tests/golden/make_golden.py",
                                                              reference="src/gpu_fem/solver_v2.py "
                                                                        "_coarse_free_dofs_injection"),
                                                         indent=1) + "\n")
    print("injection_cases", len(meta))


def prolongation_fixture(prolongation) -> None:
    """The scalar trilinear prolongation (_build_scalar_prolongation,
    solver_v2.py:62-149) as dense matrices for the fine dims of
    INJECTION_DIMS, coarse dims max(1, n // 2)."""
    out, meta = {}, []
    for i, d in enumerate(INJECTION_DIMS):
        c = tuple(max(1, n // 2) for n in d)
        out[f"P_{i}"] = np.asarray(prolongation(*d, *c).toarray(), dtype=np.float64)
        meta.append(dict(fine=list(d), coarse=list(c)))
    np.savez_compressed(OUT / "prolongation_cases.npz", **out)
    (OUT / "prolongation_cases.json").write_text(json.dumps(dict(cases=meta, generator="tests/golden/make_golden.py",
                                                                 reference="src/gpu_fem/solver_v2.py "
                                                                           "_build_scalar_prolongation"),
                                                            indent=1) + "\n")
    print("prolongation_cases", len(meta))


def evidence_fixture(OptimizationResult, TopologyProblem, ProblemSpec) -> None:
    """A folder written by the reference's OptimizationResult.save (src/
    topogpu/solver.py:46-61) from the fixed inputs of evidence_inputs.py."""
    import shutil

    root = OUT / "evidence_ref"
    if root.exists():
        shutil.rmtree(root)
    spec = ProblemSpec.from_dict(CASES["c0_cantilever_24x12x6"]["spec"])
    prob = TopologyProblem(name="c0", spec=spec, filter_radius=1.5)
    from evidence_inputs import evidence_inputs

    rho, history, summary = evidence_inputs()
    OptimizationResult(problem=prob, rho_final=rho, history=history, summary=summary).save(root)
    print("evidence_ref", sorted(p.name for p in root.iterdir()))


def main() -> None:
    sys.path.insert(0, str(REF_SRC))
    from gpu_fem.bc_generator import generate_bc
    from gpu_fem.problem_spec import ProblemSpec
    from gpu_fem.pub_simp_solver import SIMPParams, run_simp

    sys.path.insert(0, str(OUT))
    from controllers import CONTROLLERS

    names = sys.argv[1:] or [*CASES, "bc_cases", "oc_mean_cases", "evidence_ref", "injection_cases",
                              "prolongation_cases"]
    for name in names:
        if name == "prolongation_cases":
            from gpu_fem.solver_v2 import _build_scalar_prolongation
            prolongation_fixture(_build_scalar_prolongation)
            continue
        if name == "injection_cases":
            from gpu_fem.solver_v2 import _coarse_free_dofs_injection
            injection_fixture(_coarse_free_dofs_injection)
            continue
        if name == "bc_cases":
            bc_fixture(generate_bc, ProblemSpec)
            continue
        if name == "oc_mean_cases":
            from topogpu.solver import _oc_update
            oc_mean_fixture(_oc_update)
            continue
        if name == "evidence_ref":
            from topogpu.problem import TopologyProblem
            from topogpu.solver import OptimizationResult
            evidence_fixture(OptimizationResult, TopologyProblem, ProblemSpec)
            continue
        case = CASES[name]
        spec = ProblemSpec.from_dict(case["spec"])
        bc = generate_bc(spec)
        p = case["params"]
        params = SIMPParams(nelx=spec.nelx, nely=spec.nely, nelz=spec.nelz,
                            volfrac=spec.volfrac, **p)
        ctrl = CONTROLLERS[case["controller"]]() if "controller" in case else None
        bc_override = bc.bc_override
        if case.get("force") == "line":
            F_line = line_load(spec.nelx, spec.nely, spec.nelz)
            F_line[bc.fixed_dofs] = 0.0
            bc.F = F_line
            bc_override = lambda nx, ny, nz: (bc.fixed_dofs, bc.free_dofs, F_line)  # noqa: E731
        res = run_simp(params, callback=ctrl, verbose=False, bc_override=bc_override)
        extra = {}
        if ctrl is not None:
            log = res["params_log"]
            extra = dict(
                best_rho=np.asarray(res["best_rho"], dtype=np.float64),
                best_compliance=np.float64(res["best_compliance"]),
                best_iteration=np.int64(res["best_iteration"]),
                best_is_valid=np.int64(res["best_is_valid"]),
                best_grayness=np.float64(res["best_grayness"]),
                pre_valid_best_compliance=np.float64(res["pre_valid_best_compliance"]),
                pre_valid_best_iteration=np.int64(res["pre_valid_best_iteration"]),
                log_phase=np.array([r["phase"] == "tail" for r in log], dtype=np.int64),
                log_controls=np.array([[r["penal"], r["rmin"], r["move"], r["beta"]] for r in log]),
                log_best_compliance=np.array([r["best_compliance"] for r in log]),
                log_best_is_valid=np.array([r["best_is_valid"] for r in log], dtype=np.int64),
            )
        np.savez_compressed(
            OUT / f"{name}.npz",
            free=bc.free_dofs.astype(np.int64),
            F=bc.F.astype(np.float64),
            compliance_history=np.asarray(res["compliance_history"], dtype=np.float64),
            rho_final=np.asarray(res["rho_final"], dtype=np.float64),
            final_compliance=np.float64(res["final_compliance"]),
            final_grayness=np.float64(res["final_grayness"]),
            n_iter=np.int64(res["n_iter"]),
            **extra,
        )
        meta = dict(case=case, generator="tests/golden/make_golden.py",
                    reference="nbbllxx0/TopoGPU v0.1.0 src/gpu_fem/pub_simp_solver.py run_simp",
                    solver="scipy.sparse.linalg.spsolve (pyamg absent)")
        (OUT / f"{name}.json").write_text(json.dumps(meta, indent=1) + "\n")
        print(name, res["compliance_history"], res["final_compliance"], res["n_iter"])


if __name__ == "__main__":
    main()
