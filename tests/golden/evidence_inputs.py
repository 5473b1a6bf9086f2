"""Fixed inputs of the result-folder fixture ``evidence_ref/`` (data only:
shared by tests/golden/make_golden.py, which saved them with the reference's
OptimizationResult.save, and tests/test_evidence.py)."""

from __future__ import annotations

import numpy as np


def evidence_inputs():
    """Density, history rows and summary with what a run's records hold:
    ints, floats of every magnitude and sign, bools, None, strings that need
    CSV quoting, rows with different key sets, nested and non-ASCII JSON."""
    rho = np.linspace(1e-3, 1.0, 24).reshape(4, 3, 2).ravel()
    history = [
        {"iteration": 1, "compliance": 12.198912345678901, "outer_iters": 12, "rho_mean": 0.3, "grayness": 0.84,
         "wall_s": 0.012, "outer_solver": "pcg", "converged": True},
        {"iteration": 2, "compliance": 5.949318, "outer_iters": 13, "rho_mean": 0.30000000000000004,
         "grayness": 0.7, "wall_s": 1e-05, "outer_solver": "fgmres", "converged": False, "change": -2.5e-20},
        {"iteration": 3, "compliance": 3.9278, "outer_iters": 0, "outer_solver": 'a "quoted", comma\nline',
         "note": None},
    ]
    summary = {"problem": {"name": "c0", "nelx": 24, "loads": [{"type": "point", "fz": -1.0}]},
               "backend": "hip", "linear_solver": "pcg_gmg", "n_iter": 3, "total_wall_s": 0.5,
               "final": history[-1], "ratio": 1.5, "unicode": "ρ ≤ 1", "empty": {},
               "list": [1, 2.5, None, True]}
    return rho, history, summary
