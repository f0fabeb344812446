"""SIMPLEC outer iterations per second on one MI355X.

Workload: the heated 4:1 duct, 1000 x 500 structured cells split into 1,000,000 triangles
with interior nodes jittered by +-20 % of the spacing (seed 0), parabolic inlet u = 4 y (1-y),
Re = 500, energy on (inlet T = 300, wall heat flux 5000, Pr = 0.7), FP64 SIMPLEC with
SA-AMG-PCG (pressure tolerance 0.1) and Jacobi-BiCGSTAB. A step is one outer iteration; the
AMG hierarchy is rebuilt every 10 iterations inside the timed region, as the solver does.

``value`` is the FP64 solver's rate over the start-up window (``--warmup`` then ``--steps``
iterations from rest, with the stability guard's rollbacks); ``steady`` times the same solver
after iteration ``--steady-from``. ``mixed`` times the mixed-precision solver (FP32 transport
assembly and BiCGSTAB corrections, FP32 AMG-PCG refined in FP64) over the same windows.
``equal_error`` runs both solvers to convergence on the 250 x 125 duct and compares them with
the reference's converged FP64 solution. ``batch`` times 256 cavities in one launch,
``adjoint`` the drag gradient on a square cylinder. ``long_run`` (``--long-run SECONDS``, off by
default) runs both precisions from rest towards convergence at full size for that long each.

Prints one JSON line (see ``--help``). ``roofline`` is measured live for the dominant kernel,
the BiCGSTAB p-update fused with the face-operator SpMV (its launches in three further outer
iterations bracketed by HIP events on the solver's stream); ``cpu_baseline`` times the
reference's CPU path (the oracle's host code on the reference's compiled C++/OpenMP kernels)
on a bounded sample of the same workload, and ``gpu_vs_oracle`` compares the GPU solver with
that oracle run: fields after the first iteration, iteration counts and rollbacks after that.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _workload(nx, ny, precision="fp64"):
    from gputest_amd import case as C
    from gputest_amd.mesh import duct_mesh

    def parabolic(x):
        import numpy as np

        return np.stack([4.0 * x[:, 1] * (1.0 - x[:, 1]), np.zeros(x.shape[0])], axis=1)

    mu = 1.0 / 500.0
    B = C.BoundarySpec
    case = C.UnstructuredCase(
        boundaries={"inlet": B("inlet", velocity=[1.0, 0.0], value=300.0, profile=parabolic),
                    "outlet": B("outlet", value=300.0),
                    "bottom": B("wall", thermal="flux", value=5000.0),
                    "top": B("wall", thermal="flux", value=5000.0)},
        viscosity=mu, energy=True, conductivity=mu / 0.7, cp=1.0, init_T=300.0, pressure_tol=0.1,
        max_iter=10 ** 9, tol=0.0, precision=precision)
    return duct_mesh(nx, ny), case, parabolic


# whether solvers run their energy step beside the next iteration's flow work (--overlap-energy)
OVERLAP_ENERGY = True


def _solver(mesh, case):
    from gputest_amd.simple import SimpleSolver

    s = SimpleSolver(mesh, case, "cuda")
    s.overlap_energy = OVERLAP_ENERGY
    return s


def bi_pv_bytes(n: int, nf: int) -> int:
    """Compulsory HBM bytes of one launch of the BiCGSTAB p-update + SpMV kernel on the face
    operator in its 64-row slice layout: reads r, p_prev, v_prev, inv, rhat, diag (6 n
    doubles), the slice bases (n / 64 int32) and row degrees (n bytes), neighbours (2 nf
    int32) and coefficients (2 nf doubles); writes p and v (2 n doubles). Neighbour values
    gathered again count once; the previous iteration's partial sums (read from L2 by every
    workgroup) are not counted."""
    return int(n * 8 * 6 + (n + 63) // 64 * 4 + n + 2 * nf * (4 + 8) + n * 8 * 2)


def measure_roofline(s, steps=3):
    """Average duration of the dominant kernel (bi_pv, BiCGSTAB p-update + LDU SpMV) over the
    launches of ``steps`` further outer iterations, from HIP events the library records around
    each launch on the solver's stream."""
    import torch

    from gputest_amd import _native as nat

    L = nat.lib()
    # the kernel alone: the energy step serial for these iterations (beside the flow its
    # kernels would share the GPU with the timed ones, and its solve is not timed)
    overlap, s.overlap_energy = s.overlap_energy, False
    torch.cuda.synchronize()
    L.ftmi_timer_enable(1)
    try:
        for _ in range(steps):
            s.step()
        torch.cuda.synchronize()
    finally:
        s.overlap_energy = overlap
    ms, cnt = ctypes.c_double(0.0), ctypes.c_int64(0)
    L.ftmi_timer_read(ctypes.byref(ms), ctypes.byref(cnt))
    L.ftmi_timer_enable(0)
    t = ms.value * 1e-3 / max(cnt.value, 1)
    nbytes = bi_pv_bytes(s.nc, s.fv.n_if)
    gbs = nbytes / t / 1e9
    traffic = None
    path = os.path.join(ROOT, "profiles", "pmc_traffic_duct1M.json")
    if os.path.exists(path):
        with open(path) as fh:
            traffic = json.load(fh).get("bi_pv_hbm_bytes_per_launch")
    return {"kernel": "bi_pv<LduSellOp<double>> (BiCGSTAB p-update fused with the face-operator SpMV, 64-row slice layout)",
            "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_launch": nbytes,
            "us_per_launch": round(t * 1e6, 2), "launches_timed": int(cnt.value)}


def pcg_iteration_bytes(s) -> int:
    """Compulsory HBM bytes of one AMG-PCG iteration of the pressure solve, from the solver's
    live hierarchy (FP64 values, int32 indices): pcg_ap (face operator + z, p_prev in, p, q out;
    in the 64-row slice layout the row extents are a base per slice and a degree byte per row),
    pcg_xr (p, q, x, r in; x, r out), with the projection pcg_z (r, z; without it r.z comes from
    the last smoothing sweep, which reads both anyway), and one V(2,2) cycle. Per level of n rows
    and nnz(A): two fused pre-smoothing sweeps and the residual read A, D^-1 / b / x and write
    one vector each; restriction reads R and r and writes the coarse b; prolongation reads P,
    the coarse x and x and writes x; the two post-smoothing sweeps read A, D^-1, b, x and
    write y. The dense coarse solve reads its inverse."""
    n, nf = s.nc, s.fv.n_if
    rows = (n + 63) // 64 * 4 + n if s.soffp is not None else 4 * (n + 1)
    b = 8 * n + rows + 24 * nf + 32 * n  # pcg_ap
    b += 48 * n  # pcg_xr
    if s.singular_p:
        b += 16 * n  # pcg_z (mean removal and r.z)
    amg = s._amg
    for A, P, R, _ in amg.levels:
        nl, nc = A.n, P.ncols
        a_bytes = 12 * A.nnz + 4 * (nl + 1)
        b += 2 * (a_bytes + 24 * nl)  # pre2, resid
        b += 12 * R.nnz + 4 * (nc + 1) + 8 * nl + 8 * nc  # restriction
        b += 12 * P.nnz + 4 * (nl + 1) + 8 * nc + 16 * nl  # prolongation
        b += 2 * (a_bytes + 32 * nl)  # post-smoothing
    m = amg.coarse_n
    b += 8 * m * m + 16 * m
    return int(b)


def pressure_solve_roofline(s, info, iters=100) -> dict:
    """Bandwidth of the AMG-PCG pressure solve against the algorithmic bytes of one PCG
    iteration (``pcg_iteration_bytes``). Measured on the workload's own system: the latest
    outer iteration's pressure system, with its hierarchy, solved again from zero alone on the
    GPU for ``iters`` iterations (a tolerance it cannot reach; the solve's start-up is
    amortised over them; the median of three solves), right after the timed window. Beside it,
    the timed window's own figure: its pressure stage without
    the AMG set-up per PCG iteration, which includes each solve's start-up and host polls and,
    with the energy step beside the flow, the energy kernels sharing the GPU."""
    import torch

    from gputest_amd.krylov import pcg_ldu

    torch.cuda.synchronize()
    x = torch.zeros_like(s.rhs_p)
    # the solver's own operator layout (its sliced copy holds the same system)
    kw = {"sliced_off": s.soffp}
    pcg_ldu(s.fv, s.diag_p, s.offp, s.rhs_p, x, 1e-30, 3, s._amg, s.singular_p, **kw)  # first use
    ts = []
    for _ in range(3):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pinfo = pcg_ldu(s.fv, s.diag_p, s.offp, s.rhs_p, x, 1e-30, iters, s._amg, s.singular_p, **kw)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) / max(pinfo.iterations, 1))
    t = sorted(ts)[1]  # the median of three solves
    nbytes = pcg_iteration_bytes(s)
    gbs = nbytes / t / 1e9
    its = info["pressure_iters_total"]
    tw = (info["stage_seconds"]["pressure"] - info["amg_setup_seconds"]) / max(its, 1)
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_pcg_iteration": nbytes,
            "us_per_pcg_iteration": round(t * 1e6, 2), "pcg_iterations": int(pinfo.iterations),
            "method": "the pressure system of the timed window's last iteration solved again from zero, alone "
                      "on the GPU; median of three solves",
            "window": {"us_per_pcg_iteration": round(tw * 1e6, 2), "pcg_iterations": its,
                       "frac": round(nbytes / tw / 1e9 / HBM_PEAK_GBS, 4),
                       "note": "pressure stage of the timed window without the AMG set-up, per PCG iteration: "
                               "each solve's start-up, the host polls and the overlapped energy kernels included"}}


def window_difference(a, b) -> dict:
    """Relative L2 difference of the fields of two solvers after the same outer iterations from
    the same start (pressure mean-free): how far the mixed-precision trajectory is from FP64's."""
    import torch

    def rel(x, y):
        return float(torch.linalg.vector_norm(x - y) / torch.clamp(torch.linalg.vector_norm(y), min=1e-300))

    out = {"U": rel(a.U, b.U), "p": rel(a.p - a.p.mean(), b.p - b.p.mean())}
    if a.energy:
        out["T"] = rel(a.T, b.T)
    return {k: float(f"{v:.3g}") for k, v in out.items()}


def copy_rate(nbytes=256 << 20, reps=10):
    """Rate of a plain device-to-device copy (the HIP runtime's copy kernel), read plus write
    bytes over the event time of `reps` copies: a reference rate beside the kernels' fractions
    of the 8 TB/s peak. It is not a ceiling: a sparse kernel whose gathered data stays in the
    Infinity Cache or L2 can move more fabric bytes per second than a streaming copy."""
    import torch

    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (t0.elapsed_time(t1) * 1e-3) / 1e9
    return {"gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_copy": 2 * nbytes}


def calibration_copy(nbytes=256 << 20):
    """One device-to-device copy of known size: the reference dispatch against which the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run are scaled."""
    import torch

    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    return nbytes


def batch_leg(count, steps, warmup, seed=0):
    """256 cases, lid-driven cavities and backward-facing steps 50/50 (cells log-uniform in
    [16k, 200k], quads/triangles 50/50, Re uniform in [100, 1000], seeded) in one concatenated
    mesh, mixed precision, outer iterations replayed from HIP graphs, the streaming kernels'
    workgroup cap and the V-cycle's threads per row autotuned."""
    import torch

    from gputest_amd.batch import BatchSolver, autotune, sample_cases

    specs = sample_cases(count, seed=seed)
    print(f"[bench:batch] building {count} cases", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = BatchSolver(specs, viscosity=0.01, precision="mixed", use_graph=True, max_iter=10 ** 9, tol=0.0)
    setup = time.perf_counter() - t0
    print(f"[bench:batch] {count} cases, {b.n_cells} cells, set-up {setup:.1f} s", file=sys.stderr, flush=True)
    tune = autotune(b.solver, cache=None)
    cap = tune["cap"]
    print(f"[bench:batch] tuned {tune}", file=sys.stderr, flush=True)
    for _ in range(warmup):
        b.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = []
    for _ in range(steps):
        t1 = time.perf_counter()
        b.step()  # the host reads the reduced scalars at the end of every step: a step is synchronous
        per.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"cases": count, "median_ms_per_step": round(float(sorted(per)[len(per) // 2]) * 1e3, 2), "step_cases": sum(not sp.closed for sp in specs), "cells": b.n_cells, "steps": steps,
            "it_per_s": round(steps / dt, 3),
            "ms_per_step": round(dt / steps * 1e3, 2), "cell_updates_per_s": round(b.n_cells * steps / dt, 1),
            "grid_cap": cap, "tpr_bias": tune["tpr_bias"], "tuning_ms": tune["ms_per_iteration"],
            "graph_recordings": b.solver.graph_captures, "graph_redos": b.solver.graph_redos,
            "graph_record_s": round(b.solver.graph_record_seconds, 3),
            "amg_rebuilds": b.solver.amg_rebuilds,
            "setup_s": round(setup, 1), "precision": "mixed", "residuals": b.solver.residuals()["scaled"]}


def adjoint_leg(n_body, tol, max_iter, gmres_tol):
    """Square cylinder D = 1 in a 20D x 10D channel (n_body cells per side, graded, jittered),
    Re = 40, uniform inlet given by 64 control values: FP64 forward to convergence, then the
    discrete-adjoint gradient of the drag with respect to the 64 values and nu."""
    import numpy as np
    import torch

    from gputest_amd import case as C
    from gputest_amd.adjoint import InletProfile, drag_gradient
    from gputest_amd.mesh import square_cylinder_mesh

    mesh = square_cylinder_mesh(n_body)
    prof = InletProfile(-5.0, 5.0, np.ones(64))
    B = C.BoundarySpec
    case = C.UnstructuredCase(boundaries={"inlet": B("inlet", profile=prof), "outlet": B("outlet"),
                                          "bottom": B("symmetry"), "top": B("symmetry"), "cylinder": B("wall")},
                              viscosity=1.0 / 40.0, max_iter=max_iter, tol=tol)
    s = _solver(mesh, case)
    t0 = time.perf_counter()
    while not s.finished:
        s.step()
        if s.iteration % 5000 == 0:
            print(f"[bench:adjoint] forward iteration {s.iteration}: {s.residuals()['scaled']}", file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    print(f"[bench:adjoint] {mesh.n_cells} cells, forward {s.iteration} iterations in {tf:.1f} s, "
          f"residuals {s.residuals()['scaled']}", file=sys.stderr, flush=True)
    gu, gnu, J, info = drag_gradient(s, prof, tol=gmres_tol, restart=300, maxiter=3000)
    return {"cells": mesh.n_cells, "grads_per_s": round(1.0 / info.seconds, 5), "adjoint_s": round(info.seconds, 2),
            "gmres_iterations": info.gmres_iterations, "gmres_residual": info.residual,
            "forward_iterations": s.iteration, "forward_s": round(tf, 2), "forward_converged": bool(s.converged),
            "Cd": round(2.0 * J, 6), "dCd_dnu": round(2.0 * float(gnu), 6), "parameters": 65}


def sst_leg(nx, ny, steps, warmup):
    """k-omega SST on the reference's turbulence validation case scaled to nx x ny cells
    (flowtrace/validation/turbulent.py:47-75): a half channel of length 4, wall at y = 0 and
    symmetry at y = 1, quads graded geometrically from a first cell of 0.065 / ny (ny <= 300:
    the grading's root bracket overflows beyond), periodic in x, driven by dp/dx = -1 at
    Re_tau = 392.24, started from U = 15; FP64."""
    import numpy as np

    from gputest_amd import case as C
    from gputest_amd.mesh import cartesian_mesh_2d, geometric, make_periodic

    ys = geometric(ny, 1.0, 0.065 / ny)
    mesh = cartesian_mesh_2d(nx, ny, xs=np.linspace(0.0, 4.0, nx + 1), ys=ys, names=("in", "out", "wall", "centre"),
                             types=("wall", "wall", "wall", "symmetry"))
    mesh = make_periodic(mesh, "in", "out")
    B = C.BoundarySpec
    case = C.UnstructuredCase(boundaries={"wall": B("wall"), "centre": B("symmetry")}, viscosity=1.0 / 392.24,
                              body_force=[1.0, 0.0, 0.0], turbulence="k-omega-sst", turb_reference_velocity=15.0,
                              init_velocity=[15.0, 0.0, 0.0], relax_u=0.9, pressure_tol=0.1, max_iter=10 ** 9,
                              tol=0.0)
    s, dt, info = run_timed(mesh, case, warmup, steps, "sst")
    y = s.turb.fields()["y-plus"]
    return {"workload": f"SST half channel Re_tau 392, {nx}x{ny} graded quads, periodic", "cells": s.nc,
            "steps": steps, "it_per_s": round(steps / dt, 3), "ms_per_step": round(dt / steps * 1e3, 2),
            "precision": "fp64", "max_y_plus": round(float(y.max()), 3),
            "residuals": s.residuals()["scaled"], "last_linear": s.last_linear, **info}


def channel_leg(nx, ny, nz, steps, levels=3, coarse_sweeps=4):
    """The 3-D channel workload: nx x ny x nz hexes on [0,4]x[0,1]x[0,1], y tanh-stretched
    (factor 2), periodic in x, driven by the constant pressure gradient of Re_bulk 2000, laminar,
    from rest at the automatic pressure relaxation (the stability guard on), ``steps`` outer
    iterations each, FP64 and mixed precision (FP64 continuity residual), both with the pressure
    preconditioner cut at ``levels`` levels, the coarsest solved by ``coarse_sweeps`` Jacobi
    sweeps. (The reference's rule, coarsening to 300 unknowns, stops here at a fourth level of
    8,251 unknowns with 1,390 entries per row, too dense for the device sparse products, and
    would invert it densely.)"""
    import torch

    from gputest_amd import case as C
    from gputest_amd.mesh import channel_mesh

    t0 = time.perf_counter()
    mesh = channel_mesh(nx, ny, nz)
    print(f"[bench:channel] mesh {mesh.n_cells} cells in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    mu = 1.0 / 2000.0
    out = {"workload": f"3-D periodic channel {nx}x{ny}x{nz} hexes ({mesh.n_cells} cells), Re_bulk 2000, laminar",
           "steps": steps}
    rates = {}
    for precision in ("fp64", "mixed"):
        kw = {"amg_levels": levels, "amg_coarse_sweeps": coarse_sweeps}
        if precision == "mixed":
            kw["precision"] = "mixed"
        case = C.UnstructuredCase(viscosity=mu, periodic_pressure_gradient=[-12.0 * mu, 0.0, 0.0],
                                  pressure_tol=0.1, max_iter=10 ** 9, tol=0.0, **kw)
        s = _solver(mesh, case)
        s.step()  # set-up of the hierarchy and first-use costs out of the timed window
        g0 = len(s.guard_events)
        pits = []
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps - 1):
            s.step()
            pits.append(s.last_pressure_iters)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        rates[precision] = (steps - 1) / dt
        out[precision] = {"it_per_s": round(rates[precision], 3), "ms_per_step": round(dt / (steps - 1) * 1e3, 3),
                          "guard_events": len(s.guard_events) - g0, "pressure_iters_total": int(sum(pits)),
                          "amg_levels": s._amg.n_levels if s._amg is not None else None,
                          "coarse_unknowns": s._amg.coarse_n if s._amg is not None else None,
                          "residuals": {k: float(f"{v:.3g}") for k, v in s.residuals()["scaled"].items()}}
        print(f"[bench:channel] {precision}: {out[precision]}", file=sys.stderr, flush=True)
        del s
    out["mixed"]["speedup_vs_fp64"] = round(rates["mixed"] / rates["fp64"], 3)
    out["note"] = ("from rest the flow's cross-stream mode makes the reference's own FP64 iteration roll back "
                   "repeatedly (guard events); the converged comparison of mixed against the reference's FP64 "
                   "solution is tests/test_gpu_mixed.py (16x8x8 with the reference's hierarchy, 32x16x8 with "
                   "the cut one)")
    return out


def _oracle_case(nx, ny):
    import oracle.simplec_oracle as O

    mesh, _, parabolic = _workload(nx, ny)
    mu = 1.0 / 500.0
    Bo = O.Boundary
    case = O.Case(boundaries={"inlet": Bo("inlet", velocity=[1.0, 0.0], value=300.0, profile=parabolic),
                              "outlet": Bo("outlet", value=300.0),
                              "bottom": Bo("wall", thermal="flux", value=5000.0),
                              "top": Bo("wall", thermal="flux", value=5000.0)},
                  viscosity=mu, energy=True, conductivity=mu / 0.7, init_T=300.0, pressure_tol=0.1,
                  max_iter=10 ** 9, tol=0.0)
    return mesh, case


def _first_fields(s):
    """U, p, T and the linear iteration counts of a solver (oracle or GPU) after its first
    outer iteration."""
    import numpy as np

    if isinstance(s.U, np.ndarray):
        U, p, T = s.U, s.p, s.T
    else:
        U, p, T = s.U.cpu().numpy().T, s.p.cpu().numpy(), s.T.cpu().numpy()
    return {"U": U.copy(), "p": p.copy(), "T": T.copy(),
            "momentum_iters": list(s.last_linear["momentum_iters"]),
            "pressure_iters": int(s.last_linear["pressure_iters"])}


def cpu_baseline(nx, ny, warmup, iters):
    """The reference's CPU path on the same workload: the oracle's restatement of the NumPy host
    code with the reference's own C++/OpenMP kernels (``oracle/native_cpu.py``, compiled from
    ``flowtrace/native/fvm_kernels.cpp`` into ``oracle/_ref``) for the gradients, face
    interpolation, assembly, incidence and matrix products and BiCGSTAB; AMG and PCG in
    NumPy/SciPy as in the reference without pyamg. Timed the reference's way
    (``hardware/planner.py:312-331``): set-up excluded, ``warmup`` iterations, then the median of
    ``iters`` synchronised outer iterations, AMG rebuilt at its cadence. Falls back to the plain
    NumPy oracle (one thread) when ``oracle/_ref`` is absent. Returns the baseline and what the
    GPU is compared with: the fields after the first iteration, the pressure iteration count of
    every iteration and the stability guard's rollbacks."""
    import numpy as np

    import oracle.simplec_oracle as O
    from oracle import native_cpu as N

    mesh, case = _oracle_case(nx, ny)
    s = O.Simplec(mesh, case)
    kernels = N.available()
    if kernels:
        N.with_reference_kernels(s)
    cores = N.threads() if kernels else 1
    ref = {"kernels": kernels, "pressure_iters": []}
    ts = []
    for k in range(warmup + iters):
        t0 = time.perf_counter()
        s.step()
        if k >= warmup:
            ts.append(time.perf_counter() - t0)
        ref["pressure_iters"].append(int(s.last_linear["pressure_iters"]))
        if k == 0:
            ref["first"] = _first_fields(s)
    ref["rollbacks"] = [int(e["detected_at"]) for e in s.guard_events]
    med = float(np.median(ts))
    how = (f"oracle host code + the reference's C++/OpenMP kernels (oracle/_ref) on {cores} threads"
           if kernels else "NumPy/SciPy oracle, single-threaded")
    out = {"value": round(1.0 / med, 4), "unit": "it/s", "cores": cores, "kind": "port",
           "sample": f"median of {iters} synchronised outer iterations of the same {mesh.n_cells}-cell duct after "
                     f"{warmup} warm-up iterations (AMG rebuilt every 10 iterations); {how}",
           "mean_it_per_s": round(iters / float(np.sum(ts)), 4)}
    return out, ref


def _rel(a, b):
    import numpy as np

    return float(f"{float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)):.3g}")


def _field_errors(x, ref):
    return {"U": _rel(x["U"], ref["U"]), "p": _rel(x["p"] - x["p"].mean(), ref["p"] - ref["p"].mean()),
            "T": _rel(x["T"] - 300.0, ref["T"] - 300.0)}


def gpu_vs_oracle(nx, ny, ref) -> dict:
    """The GPU FP64 solver against the CPU baseline's oracle run on the same workload from rest.
    ``iteration_1``: relative L2 of the fields after the first outer iteration (pressure
    mean-free, temperature as T - 300) and the linear iteration counts. ``reference_spread``: the
    same comparison between the plain NumPy oracle and the oracle on the reference's compiled
    kernels, two CPU implementations of the reference that differ only in round-off: the
    74-iteration AMG-PCG stopped at relative residual 0.1 amplifies a 1e-15 difference in its
    right-hand side to about 1e-6 in the pressure correction, so neither side can be closer to
    the other than this. ``trajectory``: the pressure iteration count of every outer iteration
    and the iterations at which the stability guard rolled back, for both."""

    mesh, case, _ = _workload(nx, ny)
    s = _solver(mesh, case)
    pits = []
    first = None
    for k in range(len(ref["pressure_iters"])):
        s.step()
        pits.append(int(s.last_linear["pressure_iters"]))
        if k == 0:
            first = _first_fields(s)
    rb = [int(e["detected_at"]) for e in s.guard_events]
    del s
    same = 0
    while same < len(pits) and pits[same] == ref["pressure_iters"][same]:
        same += 1
    out = {"iteration_1": {**_field_errors(first, ref["first"]),
                           "pressure_iters": [first["pressure_iters"], ref["first"]["pressure_iters"]],
                           "momentum_iters": [first["momentum_iters"], ref["first"]["momentum_iters"]]},
           "trajectory": {"iterations": len(pits), "pressure_iters_gpu": pits,
                          "pressure_iters_oracle": ref["pressure_iters"], "equal_counts_through": same,
                          "rollbacks_gpu": rb, "rollbacks_oracle": ref["rollbacks"]}}
    if ref["kernels"]:
        import oracle.simplec_oracle as O

        mesh_o, case_o = _oracle_case(nx, ny)
        plain = O.Simplec(mesh_o, case_o)
        plain.step()
        out["reference_spread"] = {**_field_errors(_first_fields(plain), ref["first"]),
                                   "pressure_iters": plain.last_linear["pressure_iters"],
                                   "between": "plain NumPy oracle and the oracle on the reference's compiled "
                                              "kernels, first iteration"}
    out["note"] = ("from rest this workload's flow turns unstable near iteration 11 in both solvers and the "
                   "stability guard rolls back to the start with smaller relaxation; from there the two "
                   "trajectories part, as two CPU runs of the reference that differ in round-off do, so fields "
                   "are compared after the first iteration and the trajectories by their iteration counts and "
                   "rollbacks; converged fields are compared in equal_error")
    return out


def equal_error_leg(nx=250, ny=125):
    """Equal L2 error: the FP64 and mixed solvers on the duct workload at nx x ny run to
    convergence (FP64 to scaled residuals of 1e-8; mixed to its test at 1e-6, then on to the
    FP64 solution's iteration count) and compared with the reference's own converged FP64
    solution (``tests/golden/duct<nx>x<ny>_converged.npz``); also the rates of both."""
    import numpy as np
    import torch

    g = np.load(os.path.join(ROOT, "tests", "golden", f"duct{nx}x{ny}_converged.npz"))
    until = int(g["iterations"])
    out = {"workload": f"heated duct {nx}x{ny}x2 jittered triangles", "reference_iterations": until}

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    for precision in ("fp64", "mixed"):
        mesh, case, _ = _workload(nx, ny, precision)
        case.tol, case.max_iter = (1e-8 if precision == "fp64" else 1e-6), 3 * until
        s = _solver(mesh, case)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while not s.finished:
            s.step()
        passed = s.iteration
        if precision == "mixed" and s.converged and until > s.iteration:
            s.case.tol, s.converged, s.case.max_iter = 0.0, False, until
            while not s.finished:
                s.step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        U, p, T = s.U.cpu().numpy().T, s.p.cpu().numpy(), s.T.cpu().numpy()
        out[precision] = {"converged_at": passed, "iterations": s.iteration, "it_per_s": round(s.iteration / dt, 2),
                          "rel_l2_vs_reference": {"U": float(f"{rel(U, g['U']):.3g}"), "p": float(f"{rel(p, g['p']):.3g}"),
                                                  "T-T_in": float(f"{rel(T - 300.0, g['T'] - 300.0):.3g}")}}
        print(f"[bench:equal-error] {precision}: {out[precision]}", file=sys.stderr, flush=True)
    return out


def long_run_leg(mesh, nx, ny, seconds):
    """How far the workload gets towards convergence at full size. The FP64 solver runs from
    rest for ``seconds`` (or to scaled residuals of 1e-8); the mixed solver then runs as many
    iterations. Reported per precision: the rate over the whole run, the largest scaled residual
    every 30 s, the relaxation the stability guard left, and how far u still is from the fully
    developed profile 4 y (1 - y) that the parabolic inlet makes the exact solution; and the
    relative L2 difference of the two end states, after the same iterations."""
    import numpy as np
    import torch

    out, ends, until = {"seconds_fp64": seconds}, {}, None
    for precision in ("fp64", "mixed"):
        _, case, _ = _workload(nx, ny, precision)
        case.tol, case.max_iter = 1e-8, (10 ** 9 if until is None else until)
        s = _solver(mesh, case)
        y = s.cell_centres()[:, 1]
        hist = []
        torch.cuda.synchronize()
        t0 = last = time.perf_counter()
        while not s.finished:
            s.step()
            now = time.perf_counter()
            if now - last >= 30.0 or (until is None and now - t0 >= seconds):
                r = max(s.residuals()["scaled"].values())
                hist.append([s.iteration, round(now - t0, 1), float(f"{r:.3g}")])
                print(f"[bench:long-run] {precision} {hist[-1]}", file=sys.stderr, flush=True)
                last = now
                if until is None and now - t0 >= seconds:
                    break
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        U = s.U.cpu().numpy()
        out[precision] = {"iterations": s.iteration, "seconds": round(dt, 1), "it_per_s": round(s.iteration / dt, 2),
                          "converged": bool(s.converged), "guard_events": len(s.guard_events),
                          "settings": {"algorithm": s.algo, "relax_u": round(float(s.relax_u), 4),
                                       "relax_p": round(float(s.relax_p), 4)},
                          "residuals": {k: float(f"{v:.3g}") for k, v in s.residuals()["scaled"].items()},
                          "u_vs_developed_profile": _rel(U[0], 4.0 * y * (1.0 - y)),
                          "history": hist}
        ends[precision] = (U, s.p.cpu().numpy(), s.T.cpu().numpy())
        until = s.iteration
        del s
    (U0, p0, T0), (U1, p1, T1) = ends["fp64"], ends["mixed"]
    out["mixed_vs_fp64_end"] = {"U": _rel(U1, U0), "p": _rel(p1 - np.mean(p1), p0 - np.mean(p0)),
                                "T-T_in": _rel(T1 - 300.0, T0 - 300.0)}
    out["speedup_vs_fp64"] = round(out["mixed"]["it_per_s"] / out["fp64"]["it_per_s"], 3)
    return out


def run_timed(mesh, case, warmup, steps, tag):
    """Warm-up, then `steps` timed outer iterations bracketed by device synchronisation."""
    import torch

    t0 = time.perf_counter()
    s = _solver(mesh, case)
    print(f"[bench:{tag}] set-up {time.perf_counter() - t0:.1f} s: {s.nc} cells, {s.fv.n_if} interior faces",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        s.step()
    print(f"[bench:{tag}] warm-up done: {s.last_linear}", file=sys.stderr, flush=True)
    for k in s.timers:
        s.timers[k] = 0.0
    s.amg_setup_seconds = 0.0
    s.amg_rebuilds = 0
    s.amg_reuses = 0
    s.amg_refreshes = 0
    s.amg_timing = {}
    s.amg_levels_kept = []
    guard0 = len(s.guard_events)
    pits, per_step = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        s.step()  # ends with the iteration's one host read, so the step's work is done here
        per_step.append(time.perf_counter() - ts)
        pits.append(s.last_pressure_iters)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    guards = len(s.guard_events) - guard0
    print(f"[bench:{tag}] timed {steps} steps in {dt:.3f} s; stages {dict(s.timers)}; amg set-up "
          f"{s.amg_setup_seconds:.3f} s in {s.amg_rebuilds} rebuilds + {s.amg_refreshes} refreshes ({s.amg_reuses} hierarchies reused); guard events {guards}; "
          f"pressure iterations {pits}; last {s.last_linear}", file=sys.stderr, flush=True)
    info = {"stage_seconds": {k: round(v, 4) for k, v in s.timers.items()},
            "amg_setup_seconds": round(s.amg_setup_seconds, 3), "amg_rebuilds": s.amg_rebuilds,
            "amg_refreshes": s.amg_refreshes,
            # due rebuilds whose matrix was bit-identical to a kept hierarchy's (used as it is;
            # counted in neither amg_rebuilds nor amg_refreshes)
            "amg_reuses": s.amg_reuses,
            "amg_setup_host_seconds": {k: round(v, 4) for k, v in s.amg_timing.items() if k != "levels_kept"},
            "amg_levels_kept": list(s.amg_levels_kept),
            "guard_events": guards, "pressure_iters_total": int(sum(p or 0 for p in pits)),
            # the reference's own timing method (median of synchronised outer iterations,
            # hardware/planner.py:312-331); `value` stays the whole-run rate
            "median_ms_per_step": round(float(sorted(per_step)[len(per_step) // 2]) * 1e3, 3)}
    return s, dt, info


def steady_window(s, until, steps):
    """Rate of ``steps`` outer iterations timed once the solver has run to iteration ``until``
    (past the start-up transient and the stability guard's rollbacks)."""
    import torch

    while s.iteration < until:
        s.step()
    g0, r0 = len(s.guard_events), s.amg_rebuilds
    pits = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
        pits.append(s.last_pressure_iters)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"from_iteration": until, "steps": steps, "it_per_s": round(steps / dt, 3),
            "ms_per_step": round(dt / steps * 1e3, 3), "guard_events": len(s.guard_events) - g0,
            "amg_rebuilds": s.amg_rebuilds - r0, "pressure_iters_total": int(sum(p or 0 for p in pits))}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=1000)
    ap.add_argument("--ny", type=int, default=500)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-warmup", type=int, default=3)
    ap.add_argument("--steady-from", type=int, default=80, help="iteration after which the steady window is timed")
    ap.add_argument("--overlap-energy", type=int, default=1,
                    help="1: each iteration's energy step runs beside the next iteration's flow work (default); "
                         "0: serial, as for contention-free kernel statistics")
    ap.add_argument("--equal-error", type=int, default=1, help="run the converged equal-error comparison (1) or not")
    ap.add_argument("--long-run", type=float, default=0.0, help="also run FP64 and mixed from rest for this many "
                    "seconds each towards convergence and compare their end states (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed", type=int, default=1, help="also time the mixed-precision solver (1) or not (0)")
    ap.add_argument("--batch", type=int, default=256, help="also time a batch of this many cavities (256: the "
                    "log-uniform 16k-200k-cell batch; 0: skip)")
    ap.add_argument("--adjoint", type=int, default=32, help="also time the square-cylinder adjoint with this many "
                    "cells per body side (32: 47,532 triangles; 100: about 0.47M; 0: skip)")
    ap.add_argument("--sst", type=int, default=0, help="also time k-omega SST on the duct mesh (1) or not (0)")
    ap.add_argument("--channel", type=int, default=1, help="also time the 3-D channel, 128x64x64 hexes, FP64 "
                    "against mixed precision with a three-level pressure cycle, 200 outer iterations (1) or not (0)")
    args = ap.parse_args()
    global OVERLAP_ENERGY
    OVERLAP_ENERGY = bool(args.overlap_energy)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the SIMPLEC path on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    mesh, case, _ = _workload(args.nx, args.ny)
    s, dt, info = run_timed(mesh, case, args.warmup, args.steps, "fp64")
    pres = pressure_solve_roofline(s, info)
    mixed = None
    if args.mixed:
        _, mcase, _ = _workload(args.nx, args.ny, "mixed")
        sm, dtm, infom = run_timed(mesh, mcase, args.warmup, args.steps, "mixed")
        mixed = {"value": round(args.steps / dtm, 3), "ms_per_step": round(dtm / args.steps * 1e3, 3),
                 "speedup_vs_fp64": round(dt / dtm, 3), "dtype": "f32 solves + f64 refinement",
                 "window_rel_l2_vs_fp64": window_difference(sm, s),
                 "window_note": "fields of the two solvers after the same start-up window: two trajectories "
                                "(rollbacks and AMG schedules differ), not the error of mixed precision; the "
                                "equal-error comparison at convergence is `equal_error`", **infom}
        mixed["steady"] = steady_window(sm, args.steady_from, args.steps)
        del sm
    steady = steady_window(s, args.steady_from, args.steps)
    if mixed is not None:
        mixed["steady"]["speedup_vs_fp64"] = round(mixed["steady"]["it_per_s"] / steady["it_per_s"], 3)
    print(f"[bench] steady windows: fp64 {steady}, mixed {mixed and mixed['steady']}", file=sys.stderr, flush=True)
    roof = measure_roofline(s)
    roof["copy_rate"] = copy_rate()
    calibration_copy()
    out = {"metric": "SIMPLEC outer it/s at 1M cells (FP64)", "value": round(args.steps / dt, 3), "unit": "it/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded jittered-triangle mesh generated on the box)",
           "config": {"workload": f"heated duct {args.nx}x{args.ny}x2 jittered triangles ({s.nc} cells), "
                                  f"FP64 SIMPLEC + energy, SA-AMG-PCG tol 0.1",
                      "cells": s.nc, "interior_faces": s.fv.n_if, "precision": "fp64",
                      **info},
           "steady": steady, "roofline": roof, "pressure_solve": pres}
    del s
    if mixed is not None:
        out["mixed"] = mixed
    if args.equal_error:
        out["equal_error"] = equal_error_leg()
    if args.batch:
        out["batch"] = batch_leg(args.batch, steps=5, warmup=3)
    if args.adjoint:
        out["adjoint"] = adjoint_leg(args.adjoint, tol=1e-10, max_iter=40000, gmres_tol=1e-8)
    if args.sst:
        out["sst"] = sst_leg(4 * args.nx, args.ny // 2, args.steps, args.warmup)
    if args.channel:
        out["channel"] = channel_leg(128, 64, 64, 200)
    if args.long_run > 0:
        out["long_run"] = long_run_leg(mesh, args.nx, args.ny, args.long_run)
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"], ref = cpu_baseline(args.nx, args.ny, args.cpu_warmup, args.cpu_iters)
        out["gpu_vs_oracle"] = gpu_vs_oracle(args.nx, args.ny, ref)
        print(f"[bench] cpu baseline {out['cpu_baseline']}; gpu vs oracle {out['gpu_vs_oracle']}", file=sys.stderr,
              flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
