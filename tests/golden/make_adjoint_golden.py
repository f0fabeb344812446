"""Converged oracle states of the adjoint test cases (``tests/adjoint_cases.py``): the 1088-cell
square cylinder at Re 40 (the drag case) and the same flow with the cylinder held at 350 K (the
heat case), each run by the oracle to scaled residuals of 1e-14. The CPU adjoint tests load
them instead of re-converging the oracle (minutes of NumPy); each test that uses one first
checks that it is a fixed point of the oracle's own outer iteration.

Run from the repository root: ``python tests/golden/make_adjoint_golden.py``."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle.simplec_oracle as O  # noqa: E402
from adjoint_cases import case, energy_case, small_mesh, u_ctrl0  # noqa: E402

NU = 1.0 / 40.0
T_WALL = 350.0


def converged_state(c):
    o = O.Simplec(small_mesh(), c)
    while not o.finished:
        o.step()
    assert o.converged
    out = {"U": o.U, "p": o.p, "F": o.F, "Fb": o.Fb, "gp": o._gp, "gU": np.stack(o._gU, axis=0),
           "iteration": np.int64(o.iteration), "res_continuity_norm": np.float64(o._res_max5["continuity"])}
    if o.energy:
        out.update(T=o.T, gT=o._gT)
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    st = converged_state(case(O, u_ctrl0(), NU)[0])
    np.savez_compressed(os.path.join(here, "adjoint_drag_converged.npz"), **st)
    st = converged_state(energy_case(O, u_ctrl0(), NU, T_WALL)[0])
    np.savez_compressed(os.path.join(here, "adjoint_heat_converged.npz"), **st)


if __name__ == "__main__":
    main()
