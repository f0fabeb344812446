"""Write the reference's converged FP64 solutions of the heated-duct workload at reduced sizes.

Run once, in a container where the reference checkout is mounted (its path is the first
argument, default ``/root/reference``), with the reference's NumPy path forced. For each size
the reference's SIMPLEC runs the heated duct (``make_golden.duct_solver``: parabolic inlet,
Re = 500, energy with flux walls, SA-AMG-PCG at tol 0.1) until its scaled residuals are below
1e-8, and ``duct<nx>x<ny>_converged.npz`` records U, p, T, the iteration count and the
residuals. These are the FP64 solutions that mixed precision is judged against at equal error.

``channel16x8x8_converged.npz``: the 3-D periodic channel workload (``make_golden.channel_mesh``,
``channel_case``: tanh-stretched hexes, constant pressure gradient, Re_bulk 2000) at 16 x 8 x 8
with a user pressure relaxation of 0.3, so that the stability guard is off and the iteration
converges (from rest at the automatic relaxation the reference trips its guard repeatedly);
run to scaled residuals below 1e-8.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import BE, SimpleSolver, channel_case, channel_mesh, duct_mesh, duct_solver  # noqa: E402


def converged(nx, ny, max_iter):
    s = duct_solver(duct_mesh(nx, ny), max_iter)
    s.case.tol = 1e-8
    t0 = time.time()
    while not s.finished:
        s.step()
        if s.iteration % 500 == 0:
            print(nx, ny, s.iteration, s.residuals()["scaled"], f"{time.time() - t0:.0f} s", flush=True)
    r = s.residuals()["scaled"]
    print(nx, ny, "done", s.iteration, s.converged, r, flush=True)
    np.savez_compressed(os.path.join(HERE, f"duct{nx}x{ny}_converged.npz"), U=s.U, p=s.p, T=s.T,
                        iterations=np.int64(s.iteration), converged=np.bool_(s.converged),
                        relax_p=np.float64(s.relax_p), guard_events=np.int64(len(s.guard_events)),
                        **{"res_" + k: np.float64(v) for k, v in r.items()})


def channel_converged():
    case = channel_case(20000)
    case.tol, case.relax_p = 1e-8, 0.3
    s = SimpleSolver(channel_mesh(16, 8, 8), case, BE)
    while not s.finished:
        s.step()
    r = s.residuals()["scaled"]
    print("channel done", s.iteration, s.converged, r, flush=True)
    np.savez_compressed(os.path.join(HERE, "channel16x8x8_converged.npz"), U=s.U, p=s.p,
                        iterations=np.int64(s.iteration), converged=np.bool_(s.converged),
                        **{"res_" + k: np.float64(v) for k, v in r.items()})


if __name__ == "__main__":
    which = sys.argv[2] if len(sys.argv) > 2 else "all"
    if which in ("all", "duct"):
        converged(40, 20, 5000)
        converged(250, 125, 20000)
    if which in ("all", "channel"):
        channel_converged()
