"""Batched solves: many independent cases in one set of launches (no reference counterpart;
the reference batches at the process level, ``flowtrace/runtime/batch.py:34-151``).

The cases' meshes are concatenated into one disconnected mesh (node, cell and face arrays
appended with offsets, zones renamed per case), so every kernel of the outer iteration is one
launch over the cells and faces of all cases, and the linear operators are block diagonal.
``BatchSolver`` runs one ``SimpleSolver`` on that mesh, replaying each outer iteration from a
HIP graph (``SimpleSolver.use_graph``), with the streaming kernels' workgroup cap and the
V-cycle's threads per row picked by ``autotune``. The cases are lid-driven cavities (``CavitySpec``) and backward-facing
steps (``StepSpec``), sampled by ``sample_cases``. Each case keeps its own boundary conditions;
Reynolds numbers differ by lid or inlet speed at a shared viscosity. The Krylov reductions run over the whole batch, so the iterates
differ from separate solves, while each converged field is the case's own steady solution (the
batch is checked per case against the oracle). Residuals are normalised and convergence is
judged case by case (``SimpleSolver.set_case_partition``); the stability guard and the
relaxation factors are shared by a launch group. Closed cases fix
their pressure level at a reference cell each (``SimpleSolver.set_reference_cells``).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from .case import BoundarySpec, UnstructuredCase
from .mesh import UnstructuredMesh, Zone, cartesian_mesh_2d, extract_cells, non_orthogonality_deg, sort_faces_by_zone
from .simple import SimpleSolver, auto_relax_p


def concatenate_meshes(meshes: Sequence[UnstructuredMesh], prefixes: Sequence[str]) -> UnstructuredMesh:
    """One mesh holding ``meshes`` side by side (no shared faces); zone ``z`` of mesh ``k`` is
    renamed ``prefixes[k] + z``."""
    if any(m.periodic for m in meshes):
        raise NotImplementedError("periodic pairs are not supported in a batch")
    dim = meshes[0].dim
    nodes, fptr, fnodes, own, nei, fz, cz = [], [np.zeros(1, np.int64)], [], [], [], [], []
    zones: Dict[int, Zone] = {}
    node_off = cell_off = fn_off = 0
    zid = 1
    for m, pre in zip(meshes, prefixes):
        if m.dim != dim:
            raise ValueError("all meshes of a batch must have the same dimension")
        remap = {}
        for k in sorted(m.zones):
            z = m.zones[k]
            remap[k] = zid
            zones[zid] = Zone(zid, pre + z.name, z.type, z.kind)
            zid += 1
        lut = np.zeros(max(m.zones) + 1, np.int64)
        for k, v in remap.items():
            lut[k] = v
        nodes.append(m.nodes)
        fptr.append(m.face_ptr[1:] + fn_off)
        fnodes.append(m.face_nodes + node_off)
        own.append(m.owner + cell_off)
        nei.append(np.where(m.neighbour >= 0, m.neighbour + cell_off, -1))
        fz.append(lut[m.face_zone])
        cz.append(lut[m.cell_zone])
        node_off += m.n_nodes
        cell_off += m.n_cells
        fn_off += int(m.face_ptr[-1])
    czs = np.concatenate(cz)
    for k, z in zones.items():
        if z.kind == "cell":
            w = np.flatnonzero(czs == k)
            zones[k] = Zone(k, z.name, z.type, "cell", int(w[0]), int(w[-1])) if w.size else z
    out = UnstructuredMesh(dim=dim, nodes=np.concatenate(nodes), face_ptr=np.concatenate(fptr),
                           face_nodes=np.concatenate(fnodes), owner=np.concatenate(own),
                           neighbour=np.concatenate(nei), face_zone=np.concatenate(fz), n_cells=cell_off,
                           cell_zone=czs, zones=zones, name=f"batch{len(meshes)}")
    return sort_faces_by_zone(out)


@dataclass
class CavitySpec:
    """A lid-driven cavity of the batch: n x n cells on the unit square (quads, or 2 n^2
    triangles), lid speed giving Re = lid * 1 / nu."""
    n: int
    triangles: bool
    reynolds: float

    closed = True  # no inflow or outflow: its pressure level is fixed at a reference cell

    def mesh(self) -> UnstructuredMesh:
        return cartesian_mesh_2d(self.n, self.n, triangulate=self.triangles)

    @property
    def cells(self) -> int:
        return self.n * self.n * (2 if self.triangles else 1)

    def boundaries(self, prefix: str, viscosity: float) -> Dict[str, BoundarySpec]:
        lid = self.reynolds * viscosity  # unit side
        out = {prefix + "top": BoundarySpec("wall", velocity=[lid, 0.0])}
        for side in ("left", "right", "bottom"):
            out[prefix + side] = BoundarySpec("wall")
        return out


@dataclass
class StepSpec:
    """A backward-facing step of the batch (expansion ratio 2): the channel [0, length] x [0, 2]
    without the step block [0, 1] x [0, 1], so that the flow enters over the step through the
    inlet x = 0, y in [1, 2], and leaves through the outlet x = length; n cells per step height
    h = 1 (quads, or two triangles per quad). Parabolic inlet of mean velocity U with
    Re = U h / nu (the reference's step case, ``structured/case.py:215-231``: length 20 h,
    step length h)."""
    n: int
    triangles: bool
    reynolds: float
    length: float = 20.0

    closed = False

    def mesh(self) -> UnstructuredMesh:
        n, per = self.n, 2 if self.triangles else 1
        nx, ny = int(round(self.length * n)), 2 * n
        m = cartesian_mesh_2d(nx, ny, lx=self.length, ly=2.0, names=("inlet", "outlet", "bottom", "top"),
                              types=("velocity-inlet", "pressure-outlet", "wall", "wall"), triangulate=self.triangles)
        q = np.arange(m.n_cells) // per  # the quad of each cell: i ny + j
        keep = ~((q // ny < n) & (q % ny < n))
        return extract_cells(m, keep, name="step", type="wall")

    @property
    def cells(self) -> int:
        return self.n * self.n * (2 * int(round(self.length)) - 1) * (2 if self.triangles else 1)

    def boundaries(self, prefix: str, viscosity: float) -> Dict[str, BoundarySpec]:
        U = self.reynolds * viscosity

        def parabolic(x, U=U):  # mean U over y in [1, 2]
            return np.stack([6.0 * U * (x[:, 1] - 1.0) * (2.0 - x[:, 1]), np.zeros(x.shape[0])], axis=1)

        out = {prefix + "inlet": BoundarySpec("inlet", velocity=[U, 0.0], profile=parabolic),
               prefix + "outlet": BoundarySpec("outlet")}
        for side in ("bottom", "top", "step"):
            out[prefix + side] = BoundarySpec("wall")
        return out


def _log_uniform_n(rng, cells, per_n2: float) -> int:
    target = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
    return max(4, int(round(np.sqrt(target / per_n2))))


def sample_cavities(count: int = 256, cells=(16_000, 200_000), reynolds=(100.0, 1000.0),
                    seed: int = 0) -> List[CavitySpec]:
    """Cell counts log-uniform in ``cells``, triangles/quads 50/50, Re uniform in ``reynolds``
    (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        tri = bool(k % 2)
        n = _log_uniform_n(rng, cells, 2.0 if tri else 1.0)
        out.append(CavitySpec(n, tri, float(rng.uniform(*reynolds))))
    return out


def sample_cases(count: int = 256, cells=(16_000, 200_000), reynolds=(100.0, 1000.0), seed: int = 0,
                 step_length: float = 20.0) -> list:
    """Cavities and backward-facing steps, 50/50, each half 50/50 quads/triangles (case k:
    triangles when k is odd, a step when k // 2 is odd); cell counts log-uniform in ``cells``,
    Re uniform in ``reynolds`` (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        tri, step = bool(k % 2), bool((k // 2) % 2)
        per = (2.0 if tri else 1.0) * ((2 * round(step_length) - 1) if step else 1.0)
        n = _log_uniform_n(rng, cells, per)
        re = float(rng.uniform(*reynolds))
        out.append(StepSpec(n, tri, re, step_length) if step else CavitySpec(n, tri, re))
    return out


class BatchSolver:
    """Independent cavities and steps solved together on one concatenated mesh."""

    def __init__(self, specs: Sequence, viscosity: float = 0.01, precision: str = "mixed",
                 device="cuda", use_graph: bool = True, meshes: Optional[Sequence[UnstructuredMesh]] = None,
                 **case_kw):
        t0 = time.perf_counter()
        self.specs = list(specs)
        meshes = list(meshes) if meshes is not None else [s.mesh() for s in self.specs]
        prefixes = [f"c{k}/" for k in range(len(meshes))]
        self.mesh = concatenate_meshes(meshes, prefixes)
        self.cell_offsets = np.concatenate([[0], np.cumsum([m.n_cells for m in meshes])])
        bnd = {}
        for k, s in enumerate(self.specs):
            bnd.update(s.boundaries(prefixes[k], viscosity))
        self.case = UnstructuredCase(boundaries=bnd, viscosity=viscosity, precision=precision, **case_kw)
        self.solver = SimpleSolver(self.mesh, self.case, device)
        self.solver.use_graph = use_graph
        # each closed case (a cavity) fixes its pressure level at a reference cell
        closed = [int(self.cell_offsets[k]) for k, s in enumerate(self.specs) if s.closed]
        if closed:
            self.solver.set_reference_cells(closed)
        self.solver.set_case_partition(self.cell_offsets)  # each cavity converges on its own
        self.setup_seconds = time.perf_counter() - t0

    @property
    def n_cells(self) -> int:
        return self.mesh.n_cells

    def step(self) -> float:
        return self.solver.step()

    def run(self, steps: int):
        for _ in range(steps):
            if self.solver.finished:
                break
            self.solver.step()

    def fields(self, k: int) -> Dict[str, np.ndarray]:
        """Case k's velocity components and pressure (host arrays; mean-free for a closed
        case, whose pressure level is arbitrary)."""
        a, b = int(self.cell_offsets[k]), int(self.cell_offsets[k + 1])
        f = self.solver.fields()
        out = {name: v[a:b].copy() for name, v in f.items() if v.shape[0] == self.n_cells}
        if self.specs[k].closed:
            out["pressure"] = out["pressure"] - out["pressure"].mean()
        return out


# ---------------------------------------------------------------- job-list driver
@dataclass
class BatchJob:
    """One case of a batch (``runtime/batch.py:34-40``): a Fluent mesh file, the
    ``UnstructuredCase`` keyword arguments (``boundaries`` as dicts of ``BoundarySpec`` fields,
    overriding ``default_boundaries`` of the mesh), and an optional output folder."""
    name: str
    mesh: str
    case: dict = field(default_factory=dict)
    out_dir: str = ""
    solver: str = "pressure-based"


def _turbulent(kw: dict) -> bool:
    return str(kw.get("turbulence", "laminar")).lower() not in ("laminar", "none")


def _shared_key(kw: dict) -> str:
    """Case settings other than the boundary conditions: jobs that agree on them can share one
    solver (one viscosity, one set of relaxation factors and tolerances)."""
    rest = {k: v for k, v in kw.items() if k not in ("boundaries", "backend")}
    return json.dumps(rest, sort_keys=True, default=str)


def plan_batch(jobs: Sequence[BatchJob], max_cells: int = 50_000_000) -> dict:
    """Groups of jobs solved together (same shared settings, at most ``max_cells`` cells per
    launch, largest first) and each job's cell count; no solving (``batch.py:116-141``)."""
    from .fluent import read_any_mesh

    names = [j.name for j in jobs]
    if len(set(names)) != len(names):
        raise ValueError("job names must be unique")
    rows, meshes, errors = [], {}, {}
    by_name_case = {j.name: j.case for j in jobs}
    for j in jobs:
        try:
            if j.solver != "pressure-based":
                raise ValueError(f"solver {j.solver!r} is not available (pressure-based only)")
            meshes[j.name] = read_any_mesh(j.mesh)[0]
            key = _shared_key(j.case)
            if j.case.get("relax_p") is None:
                # automatic pressure relaxation depends on the mesh (simple.py:53-64): cases
                # share a solver only when they would pick the same value
                key += f"|relax_p={auto_relax_p(non_orthogonality_deg(meshes[j.name]))}"
            rows.append({"name": j.name, "cells": meshes[j.name].n_cells, "key": key})
        except Exception as exc:  # reported per job by run_batch
            errors[j.name] = f"{type(exc).__name__}: {exc}"
    groups: List[List[str]] = []
    by_key: Dict[str, List[dict]] = {}
    for r in rows:
        # a turbulent case runs alone: its wall distance must see its own walls only
        key = r["name"] if _turbulent(by_name_case[r["name"]]) else r["key"]
        by_key.setdefault(key, []).append(r)
    for key, rs in by_key.items():
        cur, size = [], 0
        for r in sorted(rs, key=lambda r: -r["cells"]):
            if cur and size + r["cells"] > max_cells:
                groups.append(cur)
                cur, size = [], 0
            cur.append(r["name"])
            size += r["cells"]
        if cur:
            groups.append(cur)
    return {"groups": groups, "jobs": [{k: v for k, v in r.items() if k != "key"} for r in rows],
            "errors": errors, "_meshes": meshes}


def run_batch(jobs: Sequence[BatchJob], precision: str = "fp64", device="cuda", use_graph: bool = True,
              max_cells: int = 50_000_000, log: Callable[[str], None] = print) -> dict:
    """Solve every job (``runtime/batch.py:144-222``). Instead of one process per case, each group
    of compatible jobs runs as one solver on the concatenation of its meshes (every kernel one
    launch over all of them, outer iterations replayed from a HIP graph); closed cases get a
    pressure reference cell each. Returns ``{"plan", "results" (job order), "wall_seconds"}``;
    a job that cannot be loaded or solved gets an ``error`` entry and the others go on."""
    from .case import BoundarySpec, UnstructuredCase, default_boundaries

    t0 = time.perf_counter()
    plan = plan_batch(jobs, max_cells)
    meshes = plan.pop("_meshes")
    by_name = {j.name: j for j in jobs}
    results: Dict[str, dict] = {n: {"name": n, "mesh": by_name[n].mesh, "error": e}
                                for n, e in plan["errors"].items()}
    log(f"Batch: {len(jobs)} cases in {len(plan['groups'])} launch group(s) on {device}")
    for gi, group in enumerate(plan["groups"]):
        tg = time.perf_counter()
        try:
            kw = {k: v for k, v in by_name[group[0]].case.items() if k not in ("boundaries", "backend")}
            bnd, closed = {}, []
            for k, name in enumerate(group):
                m = meshes[name]
                b = default_boundaries(m)
                for z, spec in (by_name[name].case.get("boundaries") or {}).items():
                    if z not in b:
                        raise ValueError(f"{name}: zone {z!r} is not a boundary zone of {by_name[name].mesh}")
                    b[z] = spec if isinstance(spec, BoundarySpec) else BoundarySpec(**spec)
                bnd.update({f"c{k}/{z}": v for z, v in b.items()})
                if not any(v.kind in ("outlet", "pressure") for v in b.values()):
                    closed.append(k)
            cat = concatenate_meshes([meshes[n] for n in group], [f"c{k}/" for k in range(len(group))])
            offs = np.concatenate([[0], np.cumsum([meshes[n].n_cells for n in group])])
            kw.setdefault("precision", precision)
            solver = SimpleSolver(cat, UnstructuredCase(boundaries=bnd, **kw), device)
            solver.use_graph = use_graph
            if closed:
                solver.set_reference_cells(offs[closed])
            if len(group) > 1:
                solver.set_case_partition(offs)
            t1 = time.perf_counter()
            while not solver.finished:
                solver.step()
            solve_s = time.perf_counter() - t1
            forces = solver.wall_forces()
            mb = solver.mass_balance()
            heat = solver.heat_flow() if solver.energy else {}
            fields = solver.fields()
            centres = solver.cell_centres()
            for k, name in enumerate(group):
                pre = f"c{k}/"
                strip = lambda d: {z[len(pre):]: v for z, v in d.items() if z.startswith(pre)}  # noqa: E731
                a, b = int(offs[k]), int(offs[k + 1])
                if len(group) > 1:  # the case's own convergence within the batch
                    at = int(solver.case_converged_at[k])
                    its, conv = (at if at >= 0 else solver.iteration), bool(solver.case_converged[k])
                    resid = dict(solver.case_residuals[k])
                else:
                    its, conv, resid = solver.iteration, bool(solver.converged), solver.residuals()["scaled"]
                res = {"name": name, "mesh": by_name[name].mesh, "error": None, "cells": b - a,
                       "device": f"{solver.device} (batched launch {gi}, {len(group)} cases)",
                       "iterations": its, "converged": conv, "solve_seconds": solve_s, "residuals": resid,
                       "mass_balance": strip(mb), "wall_forces": strip(forces), "heat_flow": strip(heat),
                       # the guard and the relaxation are shared by the launch group
                       "stability_rollbacks": len(solver.guard_events), "group_iterations": solver.iteration,
                       "group_stability_rollbacks": len(solver.guard_events)}
                job = by_name[name]
                if job.out_dir:
                    os.makedirs(job.out_dir, exist_ok=True)
                    out = {f: v[a:b] for f, v in fields.items() if v.shape[0] == solver.nc}
                    path = os.path.join(job.out_dir, name + ".npz")
                    np.savez_compressed(path, centres=centres[a:b], **out)
                    res["npz"] = path
                res["seconds"] = time.perf_counter() - tg
                results[name] = res
            log(f"  group {gi}: {len(group)} case(s), {solver.nc} cells, {solver.iteration} iterations, "
                f"converged={solver.converged}, {solve_s:.2f} s")
        except Exception as exc:  # one failing group must not stop the batch
            for name in group:
                results[name] = {"name": name, "mesh": by_name[name].mesh, "error": f"{type(exc).__name__}: {exc}",
                                 "seconds": time.perf_counter() - tg}
    return {"plan": plan, "results": [results[j.name] for j in jobs], "wall_seconds": time.perf_counter() - t0}


_TUNE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "autotune.json")


def _retune_hierarchies(solver: SimpleSolver):
    """Re-pick the threads per row of the solver's live AMG hierarchies with the current bias."""
    L = nat.lib()
    for amg in (solver._amg, getattr(solver, "_amg32", None)):
        h = getattr(amg, "handle", None) if amg is not None else None
        if h:
            nat.check(L.ftmi_amg_retune(h), "amg_retune")


def autotune(solver: SimpleSolver, caps=(1024, 2048, 4096), tpr_biases=(-1, 0, 1), replays: int = 3,
             cache: Optional[str] = _TUNE_FILE) -> dict:
    """Launch shapes that give the fastest outer iteration of ``solver``: the workgroup cap of
    the streaming kernels (``caps``) and the shift of the V-cycle's threads per row on every
    level (``tpr_biases``: steps of 1 / 4 / 16 / 64 lanes per row from the row-length rule).
    Each candidate is recorded in a graph of its own and ``replays`` replayed iterations are
    timed. The choice is cached per (cells, interior faces, precision, device name) and
    applied; returns {"cap", "tpr_bias", "ms_per_iteration"}. The solver advances by the
    iterations it times."""
    key = f"{solver.nc}:{solver.fv.n_if}:{solver.case.precision}:{torch.cuda.get_device_name(solver.device)}"
    L = nat.lib()
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            known = json.load(fh)
        hit = known.get(key)
        if hit and "tpr_bias" in hit and int(hit["cap"]) in caps and int(hit["tpr_bias"]) in tpr_biases:
            L.ftmi_set_grid_cap(int(hit["cap"]))
            L.ftmi_set_tpr_bias(int(hit["tpr_bias"]))
            _retune_hierarchies(solver)
            solver._graph = None
            return hit
    if not solver.use_graph:
        raise ValueError("autotuning times graph replays: set solver.use_graph")
    times = {}
    for bias in tpr_biases:
        for cap in caps:
            L.ftmi_set_grid_cap(int(cap))
            L.ftmi_set_tpr_bias(int(bias))
            _retune_hierarchies(solver)
            solver._graph = None  # launch shapes are recorded in the graph
            # step (rebuilding the hierarchy if due) until a graph is recorded with room for the timed replays
            for _ in range(2 * solver.amg_rebuild + 2):
                if solver.finished:
                    break
                room = solver._amg is not None and solver._has_lag and \
                    solver._amg_age + replays + 1 < solver.amg_rebuild
                solver.step()
                if room and solver._graph is not None:
                    break
            if solver.finished or solver._graph is None:
                break
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(replays):
                solver.step()
            torch.cuda.synchronize()
            times[(int(cap), int(bias))] = (time.perf_counter() - t) / replays * 1e3
    if not times:
        L.ftmi_set_grid_cap(int(caps[0]))
        L.ftmi_set_tpr_bias(0)
        _retune_hierarchies(solver)
        solver._graph = None
        return {"cap": int(caps[0]), "tpr_bias": 0, "ms_per_iteration": {}}
    cap, bias = min(times, key=times.get)
    L.ftmi_set_grid_cap(cap)
    L.ftmi_set_tpr_bias(bias)
    _retune_hierarchies(solver)
    solver._graph = None
    out = {"cap": cap, "tpr_bias": bias, "ms_per_iteration": {f"{c}/{b}": round(v, 3) for (c, b), v in times.items()}}
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[key] = out
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return out


def autotune_grid_cap(solver: SimpleSolver, candidates=(1024, 2048, 4096, 8192), replays: int = 3,
                      cache: Optional[str] = _TUNE_FILE) -> int:
    """``autotune`` over the workgroup cap alone (threads per row at the current bias)."""
    L = nat.lib()
    bias = int(L.ftmi_set_tpr_bias(0))
    L.ftmi_set_tpr_bias(bias)
    return int(autotune(solver, caps=tuple(candidates), tpr_biases=(bias,), replays=replays, cache=cache)["cap"])


__all__ = ["concatenate_meshes", "CavitySpec", "StepSpec", "sample_cavities", "sample_cases", "BatchSolver",
           "autotune", "autotune_grid_cap", "BatchJob", "plan_batch", "run_batch"]
