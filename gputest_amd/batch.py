"""Batched solves: many independent cases in one set of launches (no reference counterpart;
the reference batches at the process level, ``flowtrace/runtime/batch.py:34-151``).

The cases' meshes are concatenated into one disconnected mesh (node, cell and face arrays
appended with offsets, zones renamed per case), so every kernel of the outer iteration is one
launch over the cells and faces of all cases, and the linear operators are block diagonal.
``BatchSolver`` runs one ``SimpleSolver`` on that mesh, replaying each outer iteration from a
HIP graph (``SimpleSolver.use_graph``), with the streaming kernels' workgroup cap picked by
``autotune_grid_cap``. Each case keeps its own boundary conditions; Reynolds numbers differ by
lid speed at a shared viscosity. The Krylov reductions and the residual normalisation run over
the whole batch, so the iterates differ from separate solves, while each converged field is the
case's own steady solution (the batch is checked per case against the oracle). Closed cases fix
their pressure level at a reference cell each (``SimpleSolver.set_reference_cells``).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from .case import BoundarySpec, UnstructuredCase
from .mesh import UnstructuredMesh, Zone, cartesian_mesh_2d, sort_faces_by_zone
from .simple import SimpleSolver


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

    def mesh(self) -> UnstructuredMesh:
        return cartesian_mesh_2d(self.n, self.n, triangulate=self.triangles)

    @property
    def cells(self) -> int:
        return self.n * self.n * (2 if self.triangles else 1)


def sample_cavities(count: int = 256, cells=(16_000, 200_000), reynolds=(100.0, 1000.0),
                    seed: int = 0) -> List[CavitySpec]:
    """Cell counts log-uniform in ``cells``, triangles/quads 50/50, Re uniform in ``reynolds``
    (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        target = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        tri = bool(k % 2)
        n = max(4, int(round(np.sqrt(target / (2.0 if tri else 1.0)))))
        out.append(CavitySpec(n, tri, float(rng.uniform(*reynolds))))
    return out


class BatchSolver:
    """Independent lid-driven cavities solved together on one concatenated mesh."""

    def __init__(self, specs: Sequence[CavitySpec], viscosity: float = 0.01, precision: str = "mixed",
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
            lid = s.reynolds * viscosity  # unit side
            bnd[f"c{k}/top"] = BoundarySpec("wall", velocity=[lid, 0.0])
            for side in ("left", "right", "bottom"):
                bnd[f"c{k}/{side}"] = BoundarySpec("wall")
        self.case = UnstructuredCase(boundaries=bnd, viscosity=viscosity, precision=precision, **case_kw)
        self.solver = SimpleSolver(self.mesh, self.case, device)
        self.solver.use_graph = use_graph
        # every cavity is closed: one pressure reference cell per case
        self.solver.set_reference_cells(self.cell_offsets[:-1])
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
        """Case k's velocity components and mean-free pressure (host arrays)."""
        a, b = int(self.cell_offsets[k]), int(self.cell_offsets[k + 1])
        f = self.solver.fields()
        out = {name: v[a:b].copy() for name, v in f.items() if v.shape[0] == self.n_cells}
        out["pressure"] = out["pressure"] - out["pressure"].mean()
        return out


_TUNE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "autotune.json")


def autotune_grid_cap(solver: SimpleSolver, candidates=(1024, 2048, 4096, 8192), replays: int = 3,
                      cache: Optional[str] = _TUNE_FILE) -> int:
    """Workgroup cap of the streaming kernels that gives the fastest outer iteration of
    ``solver`` (``replays`` graph-replayed iterations timed per candidate, each candidate in a
    graph of its own). The choice is cached per (cells, interior faces, precision, device
    name). The solver advances by the iterations it times."""
    key = f"{solver.nc}:{solver.fv.n_if}:{solver.case.precision}:{torch.cuda.get_device_name(solver.device)}"
    L = nat.lib()
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            known = json.load(fh)
        if key in known:
            L.ftmi_set_grid_cap(int(known[key]["cap"]))
            return int(known[key]["cap"])
    if not solver.use_graph:
        raise ValueError("autotuning times graph replays: set solver.use_graph")
    times = {}
    for cap in candidates:
        L.ftmi_set_grid_cap(int(cap))
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
        times[int(cap)] = (time.perf_counter() - t) / replays * 1e3
    if not times:
        return int(L.ftmi_set_grid_cap(int(candidates[0])))
    best = min(times, key=times.get)
    L.ftmi_set_grid_cap(best)
    solver._graph = None
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[key] = {"cap": best, "ms_per_iteration": times}
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return best


__all__ = ["concatenate_meshes", "CavitySpec", "sample_cavities", "BatchSolver", "autotune_grid_cap"]
