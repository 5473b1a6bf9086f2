"""FP64 CPU oracle for the SIMP state-solve hot path.

TEST INFRASTRUCTURE ONLY. This package is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and by nothing else. The product (``topogpu_amd``) never imports
it and never falls back to it.

It restates, in plain NumPy/SciPy, the reference CPU path of nbbllxx0/TopoGPU
(``src/gpu_fem/pub_simp_solver.py``, ``src/gpu_fem/bc_generator.py``,
``src/topogpu/solver.py``) plus the multigrid transfer/Galerkin definitions of
``src/gpu_fem/solver_v2.py`` and ``src/gpu_fem/multigrid_v4.py``. Every function
cites the reference lines it follows. How each function is pinned (golden
vectors, closed forms, brute force, finite differences) is listed in
``oracle/PINNING.md``; a function without such a check would be marked
"parity unpinned" there and in its docstring.
"""

from .simp import (  # noqa: F401
    E0,
    EMIN,
    KE_UNIT_3D,
    build_density_filter,
    edof_table_3d,
    element_compliance,
    fea_and_sensitivity,
    heaviside,
    heaviside_deriv,
    ke_unit_3d,
    matfree_matvec,
    oc_update_filtered,
    oc_update_mean,
    run_simp,
    run_simp_controlled,
    assemble_k,
    solve_state,
)
from .bc import generate_bc  # noqa: F401
from .gmg import (  # noqa: F401
    coarse_free_injection,
    galerkin_coarse_operators,
    level_dims,
    scalar_prolongation,
    vector_prolongation_free,
)
