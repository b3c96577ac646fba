"""Reynolds-number-dependent step parameters used by the GUI step.

Mirrors ``get_re_parameters`` (solver/params.py:308-413): a table of
Smagorinsky constant, hyper-viscosity ratio, Brinkman eta, dt limit and CFL
target by Re band, with a coarse-grid adjustment below 512 cells in x.  The
GUI solver applies C_s, nu_h, brinkman_eta, dt_max and nu_hyper_ratio from
it (solver/simulation/base_solver.py:197-211).
"""
from __future__ import annotations

# (upper Re bound, C_s, nu_hyper_ratio, advection_scheme, nu_h, brinkman_eta, dt_max, cfl_target)
_BANDS = (
    (1e3, 0.08, 0.0, "rk3_simple", 0.0, 1e-4, 0.005, 0.3),
    (1e4, 0.10, 0.01, "rk3_simple", 0.0, 1e-4, 0.008, 0.25),
    (1e5, 0.12, 0.03, "rk3_simple", 0.0, 1e-6, 0.008, 0.35),
    (1e6, 0.15, 0.05, "rk3_high", 1e-5, 1e-7, 0.006, 0.3),
    (1e7, 0.18, 0.08, "rk3_high", 2e-5, 1e-7, 0.004, 0.25),
    (float("inf"), 0.20, 0.1, "rk3_high", 3e-5, 1e-7, 0.003, 0.2),
)


def get_re_parameters(Re: float, grid_nx: int = 512) -> dict:
    """Parameters for Reynolds number ``Re`` on a grid with ``grid_nx``
    cells in x (the band bounds are inclusive: Re <= 1000, <= 10000, ...)."""
    for bound, cs, hyper, scheme, nu_h, eta, dt_max, cfl in _BANDS:
        if Re <= bound:
            break
    params = {"C_s": cs, "nu_hyper_ratio": hyper, "advection_scheme": scheme, "nu_h": nu_h,
              "brinkman_eta": eta, "dt_max": dt_max, "cfl_target": cfl, "use_les": True}
    if grid_nx < 512:
        params["C_s"] *= 1.2
        params["nu_hyper_ratio"] *= 1.5
        params["brinkman_eta"] *= 1.2
        params["cfl_target"] *= 0.8
    return params
