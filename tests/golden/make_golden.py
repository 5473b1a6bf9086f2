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
}


def main() -> None:
    sys.path.insert(0, str(REF_SRC))
    from gpu_fem.bc_generator import generate_bc
    from gpu_fem.problem_spec import ProblemSpec
    from gpu_fem.pub_simp_solver import SIMPParams, run_simp

    sys.path.insert(0, str(OUT))
    from controllers import CONTROLLERS

    names = sys.argv[1:] or list(CASES)
    for name in names:
        case = CASES[name]
        spec = ProblemSpec.from_dict(case["spec"])
        bc = generate_bc(spec)
        p = case["params"]
        params = SIMPParams(nelx=spec.nelx, nely=spec.nely, nelz=spec.nelz,
                            volfrac=spec.volfrac, **p)
        ctrl = CONTROLLERS[case["controller"]]() if "controller" in case else None
        res = run_simp(params, callback=ctrl, verbose=False, bc_override=bc.bc_override)
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
