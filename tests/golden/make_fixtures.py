"""Write the golden fixtures of this directory from the oracle (FP64).

    python tests/golden/make_fixtures.py

Small seeded cases, inputs and outputs stored together:
  step_pure_48x32.npz      two steps of the collocated von Karman step
                           (oracle.reference_step.step_pure: RK3 + Brinkman,
                           predictor BCs, rhs, reference multigrid with 3
                           V-cycles on 4 levels, projection) on a 48 x 32 grid
                           of a 3 x 2 channel with a NACA 4412 at 8 deg
  textbook_direct_48x24.npz  a seeded rhs and the exact solution of the
                           to-tolerance operator (oracle.textbook_mg.direct_solve)
  smagorinsky_48x24.npz    constant Smagorinsky nu_sgs (Cs 0.17) of seeded
                           velocity fields
The regression test (tests/test_golden.py) recomputes them with the oracle;
the GPU test compares the HIP path with the stored outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import reference_step as R  # noqa: E402
from oracle import textbook_mg as T  # noqa: E402

NX, NY, LX, LY = 48, 24, 4.0, 1.0
SX, SY, SLX, SLY = 48, 32, 3.0, 2.0  # the step's grid: 4 reference multigrid levels


def step_case():
    dx, dy, X, Y = R.grid_coords(SX, SY, SLX, SLY)
    m, p, t = R.parse_naca_4digit("4412")
    sdf = np.nan_to_num(R.naca_sdf(X, Y, 1.0, 8.0, 0.8, 1.0, m, p, t), nan=0.0)
    mask = R.mask_from_sdf(sdf, 0.05)
    u, v = R.von_karman_ic(SX, SY, 1.0, 1e-2, seed=7, dtype=np.float64)
    return dict(u0=u, v0=v, mask=mask, sdf=sdf, dx=dx, dy=dy, nu=1e-2, U_inf=1.0, dt=2e-3)


def step_outputs(c):
    out = {}
    u, v = c["u0"], c["v0"]
    for k in (1, 2):
        u, v, p, parts = R.step_pure(u, v, c["mask"], c["sdf"], c["dx"], c["dy"], c["nu"],
                                     c["U_inf"], c["dt"], v_cycles=3, levels=4,
                                     return_parts=True)
        out.update({f"u{k}": u, f"v{k}": v, f"p{k}": p})
        if k == 1:
            out.update(u_star1=parts["u_star"], v_star1=parts["v_star"], rhs1=parts["rhs"])
    return out


def textbook_case():
    dx, dy = LX / NX, LY / NY
    rng = np.random.default_rng(17)
    rhs = rng.normal(size=(NX, NY)) * 50.0
    return dict(rhs=rhs, dx=dx, dy=dy)


def textbook_outputs(c):
    return {"p": T.direct_solve(c["rhs"], c["dx"], c["dy"])}


def smagorinsky_case():
    dx, dy, X, Y = R.grid_coords(NX, NY, LX, LY)
    rng = np.random.default_rng(23)
    u = 1.0 + 0.3 * np.sin(2 * X) * np.cos(3 * Y) + 0.05 * rng.normal(size=X.shape)
    v = 0.2 * np.cos(X) * np.sin(5 * Y) + 0.05 * rng.normal(size=X.shape)
    return dict(u=u, v=v, dx=dx, dy=dy, delta=(dx * dy) ** 0.5, cs=0.17)


def smagorinsky_outputs(c):
    return {"nu_sgs": R.constant_smagorinsky(c["u"], c["v"], c["dx"], c["dy"], c["delta"],
                                             c["cs"])}


FIXTURES = {
    "step_pure_48x32.npz": (step_case, step_outputs),
    "textbook_direct_48x24.npz": (textbook_case, textbook_outputs),
    "smagorinsky_48x24.npz": (smagorinsky_case, smagorinsky_outputs),
}


def main():
    for name, (mk, run) in FIXTURES.items():
        c = mk()
        np.savez_compressed(os.path.join(HERE, name), **c, **run(c))
        print("wrote", name)


if __name__ == "__main__":
    main()
