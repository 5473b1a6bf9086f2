"""Deterministic controllers for the run_simp continuation fixtures: their
actions depend on the iteration number only, so the same schedule runs
wherever the trajectory's floating-point details differ."""

from __future__ import annotations


class ScheduleA:
    """Initial action, the validity gate passed from iteration 2, a filter
    radius change, a restart from the best design, a stop, and a tail that
    restarts from the best design with its own parameters."""

    def initial_action(self, params):
        return {"penal": 2.5, "beta": 1.5}

    def __call__(self, state, rho_new):
        return {2: {"penal": 3.0},
                4: {"rmin": 2.0, "move": 0.15, "beta": 2.0},
                6: {"restart": True, "penal": 3.2},
                8: {"stop": True}}.get(state.iteration)

    def finalize_tail(self, params):
        return {"enabled": True, "tail_iters": 3, "penal": 3.5, "rmin": 1.5, "move": 0.1, "beta": 4.0}


class ScheduleB:
    """No valid best (the default grayness gate is not met): a restart
    request is ignored, a stop with changed parameters, and a tail on its
    default parameters from the uniform design."""

    def __call__(self, state, rho_new):
        return {2: {"beta": 3.0, "rmin": 1.8, "restart": True},
                5: {"stop": True, "penal": 3.3, "rmin": 2.5, "move": 0.5}}.get(state.iteration)

    def finalize_tail(self, params):
        return {"enabled": True, "tail_iters": 2}


CONTROLLERS = {"A": ScheduleA, "B": ScheduleB}
