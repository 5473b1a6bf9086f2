"""Write golden run_simp trajectories from the reference CPU path.

Run in a container where the reference sources are mounted read-only at
/root/reference (never on the GPU box):

    python tests/golden/make_golden.py

It imports the reference's pure NumPy/SciPy CPU modules
(``gpu_fem.pub_simp_solver``, ``gpu_fem.bc_generator``,
``gpu_fem.problem_spec``), runs ``run_simp`` with ``bc_override`` built by the
reference's own ``generate_bc`` for each case, and stores inputs and outputs
as small ``.npz`` fixtures next to this script. Only data is stored.

Why an import: the composition of filter -> Heaviside -> solve -> sensitivity
chain -> filtered-volume OC bisection over several iterations is not fixed by
any closed form; the individual pieces are pinned by the oracle tests.
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
}


def main() -> None:
    sys.path.insert(0, str(REF_SRC))
    from gpu_fem.bc_generator import generate_bc
    from gpu_fem.problem_spec import ProblemSpec
    from gpu_fem.pub_simp_solver import SIMPParams, run_simp

    for name, case in CASES.items():
        spec = ProblemSpec.from_dict(case["spec"])
        bc = generate_bc(spec)
        p = case["params"]
        params = SIMPParams(nelx=spec.nelx, nely=spec.nely, nelz=spec.nelz,
                            volfrac=spec.volfrac, **p)
        res = run_simp(params, verbose=False, bc_override=bc.bc_override)
        np.savez_compressed(
            OUT / f"{name}.npz",
            free=bc.free_dofs.astype(np.int64),
            F=bc.F.astype(np.float64),
            compliance_history=np.asarray(res["compliance_history"], dtype=np.float64),
            rho_final=np.asarray(res["rho_final"], dtype=np.float64),
            final_compliance=np.float64(res["final_compliance"]),
            final_grayness=np.float64(res["final_grayness"]),
            n_iter=np.int64(res["n_iter"]),
        )
        meta = dict(case=case, generator="tests/golden/make_golden.py",
                    reference="nbbllxx0/TopoGPU v0.1.0 src/gpu_fem/pub_simp_solver.py run_simp",
                    solver="scipy.sparse.linalg.spsolve (pyamg absent)")
        (OUT / f"{name}.json").write_text(json.dumps(meta, indent=1) + "\n")
        print(name, res["compliance_history"], res["final_compliance"], res["n_iter"])


if __name__ == "__main__":
    main()
