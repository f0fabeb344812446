"""Pressure-based SIMPLE/SIMPLEC outer iteration on one MI355X.

``SimpleSolver`` keeps the reference's interface (``flowtrace/unstructured/simple.py:71-1072``:
``step``, ``finished``, ``fields``, ``residuals``, ``state_dict``/``load_state``,
``wall_forces``, ``heat_flow``, ``mass_balance``, ``describe``, ``timers``) and the same
arithmetic, but one outer iteration is a short sequence of fused HIP kernels on HBM-resident
fields:

1. boundary values (``ftmi_simplec_bvals``);
2. least-squares gradients of p and every velocity component in one pass, then the
   Green-Gauss pressure force (``ftmi_ls_gradients``, ``ftmi_gg_gradient_fused``);
3. momentum coefficients, right-hand sides, Fluent residual sums, relaxed system and the
   SIMPLEC d-coefficient in one pass (``ftmi_transport_assemble``);
4. one device-resident Jacobi-BiCGSTAB per component (``ftmi_bicgstab_ldu``);
5. Rhie-Chow fluxes, pressure-correction matrix and mass imbalance (``ftmi_rhie_chow``);
6. AMG-preconditioned PCG for p' (``ftmi_pcg_ldu``; hierarchy rebuilt every 10 iterations);
7. flux, pressure and velocity corrections with the p' gradient (``ftmi_correct``);
8. energy, when on (the same kernels for one scalar).

The host reads one block of reduced scalars per iteration (residual sums, continuity, flux
scale) and runs the reference's residual normalisation, convergence test and stability guard.

``case.precision == "mixed"`` runs steps 3-4 and the energy system in FP32 (assembly
arithmetic, matrices and BiCGSTAB, solving for FP32 corrections of the FP64 fields) and
step 6 as FP32 AMG-PCG corrections refined until the FP64 residual of the FP64 pressure
system meets the tolerance; steps 1-2, 5 and 7 and all state stay FP64.
"""
from __future__ import annotations

import ctypes
import gc
import time
from collections import defaultdict
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _native as nat
from .amg import SmoothedAggregationAMG
from .case import FLUENT_TO_KIND, BoundarySpec, UnstructuredCase
from .fvmesh import FVMesh
from .guard import Settings, StabilityGuard
from .krylov import KrylovInfo, Workspace, bicgstab_ldu, bicgstab_ldu_mixed, pcg_ldu, pcg_ldu_mixed, to_f32
from .mesh import UnstructuredMesh

WALL, INLET, OUTLET, PRESSURE, SYMMETRY = range(5)
T_ADIABATIC, T_FIXED, T_FLUX = range(3)
_KIND_CODE = {"wall": WALL, "inlet": INLET, "outlet": OUTLET, "pressure": PRESSURE, "symmetry": SYMMETRY}

# slots of the per-iteration scalar block (device doubles)
_S_MOM = 0  # 2 per component: sum |b - A u|, sum |diag u|
_S_CONT = 6
_S_SUMS = 7  # sum p_new, sum |F|, sum |Fb|
_S_EN = 10  # energy: sum |b - A T|, sum |diag T|
_S_TURB = 16  # SST: k then omega, each sum |r|, sum |diag phi|, max diag
_NSCAL = 22
# per-solve reports of a replayed graph (4 doubles each: ok, iterations, converged, residual),
# after the reduced scalars so that one host read returns both: momentum components,
# pressure, energy, SST k and omega
_MAX_SOLVES = 8


class _PendingSolve:
    """A Krylov solve still running on the energy stream: its report (ok, iterations,
    converged, residual, as the capture mode writes it) is copied to pinned host memory,
    readable once ``done`` has passed. ``targets`` are the iteration's last_linear and
    diagnostics capture, which receive the info."""

    def __init__(self, report_h: torch.Tensor, done):
        self.report_h, self.done, self.targets = report_h, done, None

    def result(self) -> KrylovInfo:
        self.done.synchronize()
        _, it, conv, res = (float(v) for v in self.report_h.tolist())
        return KrylovInfo(int(it), res, conv == 1.0)


def auto_relax_p(non_orthogonality_deg: float) -> float:
    """SIMPLEC pressure relaxation when none is given (``simple.py:53-64``)."""
    return 1.0 if non_orthogonality_deg <= 5.0 else 0.8


class SimpleSolver:
    def __init__(self, mesh: UnstructuredMesh, case: UnstructuredCase, device="cuda"):
        t0 = time.perf_counter()
        self.case = case
        if case.periodic and not mesh.periodic:
            from .mesh import make_periodic

            for za, zb in case.periodic:
                mesh = make_periodic(mesh, za, zb)
        self.mesh = mesh
        self.device = torch.device(device)
        nat.lib()  # fail now, not mid-iteration, when the HIP library is missing
        fv = self.fv = FVMesh(mesh, device)
        self.dim, self.nc = mesh.dim, mesh.n_cells
        self.energy = case.energy
        # "mixed": FP32 transport assembly and solves (FP32 corrections of the FP64 fields) and
        # FP32 AMG-PCG refined in FP64; the continuity residual, mass imbalance, pressure
        # system and the fields stay FP64
        self.mixed = case.precision == "mixed"
        # replay each outer iteration from a HIP graph (see _graph_step)
        self.use_graph = False
        self._graph = None
        self.graph_captures = 0
        # each recorded solve runs at most ceil(a * m) + b iterations, m the largest count of
        # that solve since the previous recording (graph_cap = (a, b)); a replay in which a
        # solve did not finish within its cap is redone eagerly (graph_redos counts them)
        self.graph_cap = (1.5, 8)
        self.graph_redos = 0
        self.graph_record_seconds = 0.0
        self._seen: list = []
        self._saved = None
        self._pin = None
        self.max_refine = 4
        self._amg32 = None
        self.timers: Dict[str, float] = defaultdict(float)
        self.iteration = 0
        self.step_count = 0
        self.converged = False
        self._res_max5: Dict[str, float] = {}
        self._amg = None
        self._amg_age = 10 ** 9
        self.amg_rebuild = 10
        self.amg_setup_seconds = 0.0
        self.amg_rebuilds = 0
        # mixed precision: numeric refreshes of the FP32 preconditioner between full set-ups
        self.mixed_amg_refreshes = 5
        self.amg_refreshes = 0
        self._amg_refreshed = 0
        # a refresh keeps the aggregates of an older matrix: after a guard rollback (new
        # relaxation, possibly SIMPLE instead of SIMPLEC) or a pressure solve that missed its
        # tolerance, the next rebuild is a full set-up instead
        self._amg_needs_setup = False
        self.amg_timing: Dict[str, float] = {}
        self.amg_levels_kept: List[int] = []  # per rebuild: levels whose aggregates stayed
        self._amg_kept = None  # (hierarchy, FP32 twin) that the last full set-up replaced
        self.amg_reuses = 0  # rebuilds that found their matrix in a kept hierarchy
        self.last_linear = {}
        # The energy step of an iteration runs on a stream of its own, beside the next
        # iteration's flow work, when T feeds nothing back into the flow (no turbulence model,
        # eager steps, one case): ``overlap_energy`` switches it off.
        self.overlap_energy = True
        self._estream = None
        self._energy_pending = None
        self.relax_p = float(case.relax_p) if case.relax_p is not None else auto_relax_p(fv.non_orthogonality_deg)
        self._build_boundaries()
        self._alloc()
        U0 = np.zeros((self.nc, self.dim))
        U0[:] = (list(case.init_velocity) + [0.0] * 3)[: self.dim]
        self.U.copy_(torch.as_tensor(U0.T.copy(), device=self.device))
        self.p.zero_()
        if self.energy:
            self.T.fill_(case.init_T)
        self._has_lag = False
        self._has_lagT = False
        self._init_flux()
        self.mom_src = None  # (dim, cells) volumetric momentum source times the volume
        self.en_src = None  # (cells,) volumetric heat source times the volume over cp
        self.turb = None
        if case.turbulent:
            from .turbulence import DeviceSST

            self.turb = DeviceSST(self)
        self.guard_enabled = case.relax_p is None
        self._capture = None
        self._parts = None  # per-case convergence of a batch (set_case_partition)
        self._reset_guard()
        self.setup_seconds = time.perf_counter() - t0

    # ------------------------------------------------------------------ set-up
    def _build_boundaries(self):
        fv, case, dim = self.fv, self.case, self.dim
        n = fv.n_bf
        kind = np.zeros(n, np.int32)
        ub = np.zeros((n, dim))
        pb = np.zeros(n)
        tk = np.zeros(n, np.int32)
        tv = np.zeros(n)
        self.zone_specs = {}
        for z in self.mesh.boundary_zones():
            spec = case.boundaries.get(z.name)
            if spec is None:
                spec = BoundarySpec(kind=FLUENT_TO_KIND.get(z.type, "wall"))
                if spec.kind == "inlet":
                    spec.normal_speed = 1.0
            self.zone_specs[z.name] = spec
            m = fv.bzone == z.id
            if spec.kind == "periodic" or z.type in ("periodic", "periodic-shadow"):
                if m.any():
                    raise ValueError(f"zone {z.name!r} is periodic but not paired")
                continue
            kind[m] = _KIND_CODE[spec.kind]
            if spec.kind in ("wall", "inlet"):
                if spec.profile is not None:
                    ub[m] = np.asarray(spec.profile(fv.xfb_h[m]), float).reshape(-1, dim)
                elif spec.velocity is not None:
                    ub[m] = (list(spec.velocity) + [0.0] * 3)[:dim]
                elif spec.normal_speed is not None:
                    ub[m] = -spec.normal_speed * fv.nb_h[m]
            pb[m] = spec.pressure
            if spec.kind == "wall":
                tk[m] = {"adiabatic": T_ADIABATIC, "temperature": T_FIXED, "flux": T_FLUX}[spec.thermal]
            elif spec.kind in ("inlet", "pressure"):
                tk[m] = T_FIXED
            tv[m] = spec.value
        self.bkind_h, self.ub_fixed_h, self.pb_fixed_h, self.tkind_h, self.tval_h = kind, ub, pb, tk, tv
        dev = self.device
        self.bkind = torch.as_tensor(kind, device=dev)
        self.tkind = torch.as_tensor(tk, device=dev)
        self.ub_fixed = torch.as_tensor(ub, device=dev)
        self.pb_fixed = torch.as_tensor(pb, device=dev)
        self.tval = torch.as_tensor(tv, device=dev)
        self.singular_p = not bool(((kind == OUTLET) | (kind == PRESSURE)).any())
        b = nat.FtmiBc()
        b.kind, b.tkind = nat.ptr(self.bkind), nat.ptr(self.tkind)
        b.ub_fixed, b.pb_fixed, b.tval = nat.ptr(self.ub_fixed), nat.ptr(self.pb_fixed), nat.ptr(self.tval)
        self.c_bc = b

    def _alloc(self):
        nc, nf, nb, dim, dev = self.nc, self.fv.n_if, self.fv.n_bf, self.dim, self.device
        z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=dev)  # noqa: E731
        self.U, self.p = z(dim, nc), z(nc)
        self.F, self.Fb = z(nf), z(nb)
        self._T = z(nc) if self.energy else None
        self.ub, self.pb = z(dim, nb), z(nb)
        self.gp_ls, self.gp, self.gU = z(nc, dim), z(nc, dim), z(dim, nc, dim)
        self.diag, self.diag_r, self.inv_r, self.off = z(nc), z(nc), z(nc), z(2 * nf)
        self.bmom, self.Ustar, self.dP = z(dim, nc), z(dim, nc), z(nc)
        self.Fst, self.cf, self.Fbst, self.cb = z(nf), z(nf), z(nb), z(nb)
        self.diag_p, self.offp, self.mass, self.rhs_p, self.pp = z(nc), z(2 * nf), z(nc), z(nc), z(nc)
        self.gpp = z(nc, dim)
        # the iteration-start fields, kept for the diagnostics of the latest iteration
        self.U_old, self.p_old = z(dim, nc), z(nc)
        # the transport operator in the slice layout (FP64 BiCGSTAB), shared by momentum and energy
        self.soff = z(self.fv.sl_size) if self.fv.sl_base is not None and not self.mixed else None
        # the pressure operator in the same layout (FP64 AMG-PCG)
        self.soffp = z(self.fv.sl_size) if self.soff is not None else None
        if self.energy:
            self.Tb, self.gT, self.Tnew, self.bT = z(nb), z(nc, dim), z(nc), z(nc)
            self.diagT, self.diagT_r, self.offT, self.T_old = z(nc), z(nc), z(2 * nf), z(nc)
        if self.mixed:
            f = lambda *s: torch.zeros(*s, dtype=torch.float32, device=dev)  # noqa: E731
            self.diag32, self.diag_r32, self.inv_r32, self.off32 = f(nc), f(nc), f(nc), f(2 * nf)
            self.ex32 = f(nc)
            self.b32 = f(dim, nc)
            self.offp32 = f(2 * nf)
            # the FP32 transport operator in the slice layout, re-sliced after each assembly
            self.soff32 = f(self.fv.sl_size) if self.fv.sl_base is not None else None
        self.scal = z(_NSCAL + 4 * _MAX_SOLVES)
        self.red_partials = z(2048 * 8)
        self.red_counter = torch.zeros(64, dtype=torch.int32, device=dev)
        r = nat.FtmiRed()
        r.partials, r.counter = nat.ptr(self.red_partials), nat.ptr(self.red_counter)
        self.c_red = r
        self.ws = Workspace(dev)

    def _init_flux(self):
        L, s = nat.lib(), nat.stream_ptr()
        fv = self.fv
        nat.check(L.ftmi_simplec_bvals(ctypes.byref(fv.c_mesh), ctypes.byref(self.c_bc), nat.ptr(self.U),
                                       nat.ptr(self.p), None, None, nat.ptr(self.ub), nat.ptr(self.pb), s), "bvals")
        nat.check(L.ftmi_init_flux(ctypes.byref(fv.c_mesh), ctypes.byref(self.c_bc), float(self.case.density),
                                   nat.ptr(self.U), nat.ptr(self.ub), nat.ptr(self.F), nat.ptr(self.Fb), s), "init_flux")

    def set_sources(self, momentum=None, energy=None) -> None:
        """Volumetric sources per unit volume (``simple.py:164-182``): momentum (cells, dim)
        [force/volume] and energy (cells,) [heat/volume]. They are kept on the device times the
        cell volume (energy also over cp) and added to the momentum source after the body force
        and to the energy equation's heat input after the flux walls, as the reference adds
        them. ``None`` removes a source."""
        vol = np.asarray(self.fv.vol_h, np.float64)
        if momentum is not None:
            if self.turb is not None:
                raise NotImplementedError("volumetric momentum sources with the SST model are not built")
            m = np.asarray(momentum, np.float64).reshape(self.nc, self.dim) * vol[:, None]
            self.mom_src = torch.as_tensor(np.ascontiguousarray(m.T), device=self.device)
        else:
            self.mom_src = None
        if energy is not None:
            if not self.energy:
                raise ValueError("energy source given but the energy equation is off")
            e = np.asarray(energy, np.float64).reshape(self.nc) * vol / self.case.cp
            self.en_src = torch.as_tensor(e, device=self.device)
        else:
            self.en_src = None
        self._graph = None  # a recorded graph holds the previous source pointers

    # ------------------------------------------------------------------ the outer iteration
    @property
    def finished(self) -> bool:
        return self.converged or self.iteration >= self.case.max_iter

    def _mark(self, name):
        if self.use_graph:
            return
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self._events.append((name, ev))

    def step(self) -> float:
        if self.finished:
            return 0.0
        self._events = []
        if self.use_graph:
            return self._graph_step()
        self._mark("start")
        out = self._launch(rebuild=True)
        self._amg_age += 1
        return self._finish(*out)

    def _graph_step(self) -> float:
        """One outer iteration replayed from a HIP graph of the device work. The graph is
        recorded (in capture mode: Krylov solves without host reads, each capped near its recent
        iteration counts) after each eager step that rebuilds the AMG hierarchy, and replayed
        until the next rebuild. Each replay leaves a report per solve beside the reduced
        scalars; when a solve did not finish within its cap, the iteration-start state that
        the graph saved is put back and the iteration is redone eagerly (and the graph is
        recorded again with the new counts). The host logic (residual normalisation,
        convergence, guard) runs after every iteration."""
        due = self._amg is None or self._amg_age >= self.amg_rebuild or not self._has_lag
        sc = None
        if due:
            self._graph = None
            out = self._launch(rebuild=True)
        else:
            if self._graph is None:
                self._record_graph()
            self._graph.replay()
            sc = self._read_scalars()  # the one host read of the iteration
            out = self._replay_infos(sc[_NSCAL:])
            if out is None:
                self._restore_saved()
                self._graph = None
                self.graph_redos += 1
                sc = None
                out = self._launch(rebuild=False)
        self._note_counts(*out)
        self._amg_age += 1
        return self._finish(*out, sc=sc)

    def _read_scalars(self) -> np.ndarray:
        """The reduced scalars, copied into pinned host memory on the current stream and read
        after that stream alone has reached them: work queued on the energy stream is not
        waited for (a copy into pageable memory can wait for the whole device)."""
        if getattr(self, "_scal_h", None) is None or self._scal_h.shape != self.scal.shape:
            self._scal_h = torch.empty(self.scal.shape, dtype=self.scal.dtype, pin_memory=True)
        self._scal_h.copy_(self.scal, non_blocking=True)
        torch.cuda.current_stream(self.device).synchronize()
        return self._scal_h.numpy().copy()

    def _iteration_state(self) -> Dict[str, torch.Tensor]:
        """The device arrays an outer iteration reads before it writes them: the fields, the
        lagged gradients of the boundary values, T and its gradient, the SST fields."""
        st = {"U": self.U, "p": self.p, "F": self.F, "Fb": self.Fb, "gp": self.gp, "gU": self.gU}
        if self.energy:
            st.update(T=self.T, gT=self.gT)
        if self.turb is not None:
            st.update({"turb." + k: getattr(self.turb, k) for k in ("k", "w", "mut")})
        return st

    def _restore_saved(self):
        for k, t in self._iteration_state().items():
            t.copy_(self._saved[k])

    def _record_graph(self):
        t0 = time.perf_counter()
        L = nat.lib()
        a, b = self.graph_cap
        caps = [min(int(np.ceil(a * m)) + int(b), 1 << 30) for m in self._seen]
        self._seen = []
        arr = (ctypes.c_int64 * max(len(caps), 1))(*caps)
        state = self._iteration_state()
        if self._saved is None or set(self._saved) != set(state):
            self._saved = {k: torch.empty_like(t) for k, t in state.items()}
        nat.check(L.ftmi_capture_plan(arr, len(caps), nat.ptr(self.scal[_NSCAL:]), _MAX_SOLVES), "capture plan")
        g = torch.cuda.CUDAGraph()
        prev = L.ftmi_capture_mode(1)
        # No garbage collection while the stream is captured: objects kept alive only by
        # reference cycles (a solver and its turbulence model point at each other; graphs,
        # events and pinned buffers of earlier solvers) would be destroyed inside the capture,
        # and their destructors' runtime calls are not allowed there (the process aborts);
        # such garbage waits until the capture has ended.
        gc_was_on = gc.isenabled()
        gc.disable()
        try:
            with torch.cuda.graph(g):
                for k, t in state.items():  # the iteration-start state, for a redo
                    self._saved[k].copy_(t)
                self._graph_out = self._launch(rebuild=False)
            used = L.ftmi_capture_slots_used()
        finally:
            if gc_was_on:
                gc.enable()
            L.ftmi_capture_mode(prev)
            L.ftmi_capture_plan(None, 0, None, 0)
        expect = len(self._solve_list(*self._graph_out))
        if used != expect or used > _MAX_SOLVES:
            raise RuntimeError(f"graph recording: {used} capped solves enqueued, {expect} expected")
        self._graph = g
        self.graph_captures += 1
        self.graph_record_seconds += time.perf_counter() - t0

    def _solve_list(self, infos, pinfo, tinfo) -> list:
        """The iteration's solves in the order they are enqueued (the capture plan's slots)."""
        out = list(infos) + [pinfo] + ([tinfo] if tinfo is not None else [])
        if self.turb is not None and self.turb.last_info:
            out += [self.turb.last_info[k] for k in ("k", "omega")]
        return out

    def _note_counts(self, infos, pinfo, tinfo):
        n = [int(i.iterations) for i in self._solve_list(infos, pinfo, tinfo)]
        self._seen = n if len(self._seen) != len(n) else [max(x, y) for x, y in zip(self._seen, n)]

    def _replay_infos(self, rep: np.ndarray):
        """The Krylov infos of a replay from its reports, or None when a solve did not end
        within its cap (or a report is missing)."""
        infos, pinfo, tinfo = self._graph_out
        nm = len(infos)
        out = []
        for s in range(len(self._solve_list(infos, pinfo, tinfo))):
            ok, it, conv, res = (float(v) for v in rep[4 * s:4 * s + 4])
            if ok != 1.0:
                return None
            it = int(it)
            rounds = 0
            if self.mixed and s == nm:  # FP32 corrections, as the eager refinement counts them
                rounds = 0 if self.max_refine < 1 or (it == 0 and conv == 1.0) else 1
            out.append(KrylovInfo(it, res, conv == 1.0, rounds))
        minfo, pinfo = out[:nm], out[nm]
        tinfo = out[nm + 1] if tinfo is not None else None
        if self.turb is not None and self.turb.last_info:
            self.turb.last_info = {"k": out[-2], "omega": out[-1]}
        return minfo, pinfo, tinfo

    def _launch(self, rebuild: bool):
        """Enqueue the device work of one outer iteration; returns the Krylov infos."""
        c, fv, dim = self.case, self.fv, self.dim
        L, s = nat.lib(), nat.stream_ptr()
        M, B, R = ctypes.byref(fv.c_mesh), ctypes.byref(self.c_bc), ctypes.byref(self.c_red)
        rho, mu = c.density, c.viscosity
        self.U_old.copy_(self.U)
        self.p_old.copy_(self.p)
        # 1-2. boundary values with the lagged gradients, then the new gradients
        lag = self._has_lag
        nat.check(L.ftmi_simplec_bvals(M, B, nat.ptr(self.U), nat.ptr(self.p), nat.ptr(self.gU) if lag else None,
                                       nat.ptr(self.gp) if lag else None, nat.ptr(self.ub), nat.ptr(self.pb), s),
                  "bvals")
        phis = [self.p] + [self.U[k] for k in range(dim)]
        phibs = [self.pb] + [self.ub[k] for k in range(dim)]
        outs = [self.gp_ls] + [self.gU[k] for k in range(dim)]
        arr = ctypes.c_void_p * len(phis)
        nat.check(L.ftmi_ls_gradients(M, len(phis), arr(*[t.data_ptr() for t in phis]),
                                      arr(*[t.data_ptr() for t in phibs]), arr(*[t.data_ptr() for t in outs]),
                                      None, s), "ls_gradients")
        if c.pressure_gradient == "green-gauss":
            nat.check(L.ftmi_gg_gradient_fused(M, nat.ptr(self.p), nat.ptr(self.pb), nat.ptr(self.gp_ls),
                                               nat.ptr(self.gp), s), "gg_gradient")
        else:
            self.gp.copy_(self.gp_ls)
        self._has_lag = True
        self._mark("gradients")
        # 3-4. momentum
        au = self.relax_u
        drive = np.asarray(list(c.body_force)[:dim], float) - np.asarray(
            (list(c.periodic_pressure_gradient) + [0.0] * 3)[:dim], float)
        a = nat.FtmiAssembleArgs()
        a.F, a.Fb, a.gamma, a.energy, a.nfields = nat.ptr(self.F), nat.ptr(self.Fb), float(mu), 0, dim
        for k in range(dim):
            a.phi[k], a.grad[k], a.phib[k] = nat.ptr(self.U[k]), nat.ptr(self.gU[k]), nat.ptr(self.ub[k])
            a.b[k], a.x0[k] = nat.ptr(self.bmom[k]), nat.ptr(self.Ustar[k])
            a.drive[k] = float(drive[k])
        a.gp, a.has_drive, a.relax = nat.ptr(self.gp), int(bool(np.any(drive))), float(au)
        a.second_order = int(c.scheme == "second-order")
        a.simplec = 1 if self.algo == "SIMPLEC" else 0
        a.cp = float(c.cp)
        a.diag, a.diag_r, a.off, a.dP = nat.ptr(self.diag), nat.ptr(self.diag_r), nat.ptr(self.off), nat.ptr(self.dP)
        a.resid, a.inv_r = nat.ptr(self.scal[_S_MOM:]), nat.ptr(self.inv_r)
        if self._parts is not None:
            a.cell_resid = nat.ptr(self._parts["mom"])
        if self.mixed:
            self._fp32_outputs(a, dim)
        if self.turb is not None:  # mu + mu_t, wall functions, transpose stress (simple.py:546-548)
            gm_f, gm_b, extra = self.turb.momentum_terms(L, M, B, s)
            a.gam_f, a.gam_b = nat.ptr(gm_f), nat.ptr(gm_b)
            for k in range(dim):
                a.extra[k] = nat.ptr(extra[k])
        elif self.mom_src is not None:  # simple.py:535-536
            for k in range(dim):
                a.extra[k] = nat.ptr(self.mom_src[k])
        nat.check(L.ftmi_transport_assemble(M, B, ctypes.byref(a), R, s), "momentum assembly")
        infos = []
        if self.soff is not None:
            fv.slice_values(self.off, self.soff)
        if self.mixed and self.soff32 is not None:
            fv.slice_values(self.off32, self.soff32)
        for k in range(dim):
            if self.mixed:
                infos.append(bicgstab_ldu_mixed(fv, self.ex32, self.off32, self.b32[k], self.inv_r32,
                                                self.Ustar[k], 0.1, 30, True, ws=self.ws, b_is_residual=True,
                                                sliced_off=self.soff32))
            else:
                infos.append(bicgstab_ldu(fv, self.diag_r, self.off, self.bmom[k], self.inv_r, self.Ustar[k], 0.1,
                                          30, True, ws=self.ws, sliced_off=self.soff))
        self._mark("momentum")
        # 5. Rhie-Chow fluxes and the pressure-correction system
        nat.check(L.ftmi_rhie_chow(M, B, float(rho), nat.ptr(self.Ustar), nat.ptr(self.p), nat.ptr(self.gp),
                                   nat.ptr(self.dP), nat.ptr(self.Fst), nat.ptr(self.cf), nat.ptr(self.Fbst),
                                   nat.ptr(self.cb), nat.ptr(self.diag_p), nat.ptr(self.offp), nat.ptr(self.mass),
                                   nat.ptr(self.rhs_p), nat.ptr(self.scal[_S_CONT:]), R, s), "rhie_chow")
        if self._pin is not None:  # reference cells of a batch of closed cases (set_reference_cells)
            pin, pin_he = self._pin
            self.rhs_p.index_fill_(0, pin, 0.0)
            self.offp.index_fill_(0, pin_he, 0.0)
        self._mark("assembly")
        # 6. pressure correction
        if rebuild and (self._amg is None or self._amg_age >= self.amg_rebuild):
            t0 = time.perf_counter()
            reused = self._reuse_hierarchy(fv) if self._amg is not None else None
            if reused is not None:  # the hierarchy of exactly this matrix is at hand
                self.amg_levels_kept.append(len(self._amg.levels))
                self.amg_reuses += 1
            elif (self.mixed and self._amg is not None and self._amg_refreshed < self.mixed_amg_refreshes
                  and not self._amg_needs_setup):
                # mixed precision: the hierarchy is only the FP32 preconditioner, so its values
                # are refreshed on the kept aggregates and patterns (no aggregation, no host
                # reads but the coarse matrix); a full set-up every few refreshes
                self._amg.refresh_ldu(fv, self.diag_p, self.offp)
                self._amg32.refresh()
                self._amg_refreshed += 1
                self.amg_refreshes += 1
            else:
                how = self._next_hierarchy(fv)
                self._amg_refreshed = 0
                self._amg_needs_setup = False
                for k, v in self._amg.timing.items():
                    self.amg_timing[k] = self.amg_timing.get(k, 0.0) + v
                self.amg_levels_kept.append(int(self._amg.timing["levels_kept"]) if how == "rebuilt" else 0)
                self.amg_rebuilds += 1
            self._amg_age = 0
            self.amg_setup_seconds += time.perf_counter() - t0
        if self.mixed:
            to_f32(self.offp, self.offp32)
            self.pp.zero_()
            pinfo = pcg_ldu_mixed(fv, self.diag_p, self.offp, self.offp32, self.rhs_p, self.pp, c.pressure_tol, 200,
                                  self._amg32, self.singular_p, self.max_refine, ws=self.ws)
            if pinfo.iterations >= 0 and not pinfo.converged:  # (capture mode reports later)
                self._amg_needs_setup = True
        else:
            if self.soffp is not None:
                fv.slice_values(self.offp, self.soffp)
            pinfo = pcg_ldu(fv, self.diag_p, self.offp, self.rhs_p, self.pp, c.pressure_tol, 200, self._amg,
                            self.singular_p, ws=self.ws, sliced_off=self.soffp)
        self._mark("pressure")
        # 7. corrections
        nat.check(L.ftmi_correct(M, B, nat.ptr(self.pp), nat.ptr(self.Ustar), nat.ptr(self.dP), nat.ptr(self.Fst),
                                 nat.ptr(self.cf), nat.ptr(self.Fbst), nat.ptr(self.cb), float(self.relax_p),
                                 int(self.singular_p), nat.ptr(self.U), nat.ptr(self.p), nat.ptr(self.F),
                                 nat.ptr(self.Fb), nat.ptr(self.gpp), nat.ptr(self.scal[_S_SUMS:]), R, s), "correct")
        # 8. energy
        tinfo = None
        if self.energy:
            if self._overlap_energy():
                tinfo = self._energy_step_overlapped(L, M, B)
            else:
                tinfo = self._energy_step(L, s, M, B, R)
        self._mark("correct+energy")
        if self.turb is not None:  # simple.py:662-669
            self.turb.solve(L, M, B, s, self.scal[_S_TURB:], self.scal[_S_TURB + 3:])
            self._mark("turbulence")
        return infos, pinfo, tinfo

    def _finish(self, infos, pinfo, tinfo, sc=None) -> float:
        """Host part of an outer iteration: one read of the reduced scalars (unless the caller
        read them), the residual normalisation, the convergence test and the stability guard."""
        c, dim = self.case, self.dim
        cs = self._case_sums() if self._parts is not None else None
        if sc is None:
            sc = self._read_scalars()  # the one host read of the iteration
        self._accumulate_timers()
        res = {}
        num = [sc[_S_MOM + 2 * k] for k in range(dim)]
        den = [sc[_S_MOM + 2 * k + 1] for k in range(dim)]
        floor = 1e-3 * max(max(den), 1e-300)
        for k, name in enumerate(("x-velocity", "y-velocity", "z-velocity")[:dim]):
            res[name] = num[k] / max(den[k], floor)
        res["continuity"] = float(sc[_S_CONT])
        if self.energy:
            r, sd = float(sc[_S_EN]), float(sc[_S_EN + 1])
            res["energy"] = r / sd if sd > 0 else r
        if self.turb is not None:
            res["k"] = float(sc[_S_TURB]) / max(float(sc[_S_TURB + 1]), 1e-300)
            res["omega"] = float(sc[_S_TURB + 3]) / max(float(sc[_S_TURB + 4]), 1e-300)
        self.iteration += 1
        self.step_count = self.iteration
        if self._energy_pending is not None:  # the previous iteration's solve (done by now)
            self._resolve_energy()
        lin = {"momentum_iters": [i.iterations for i in infos], "pressure_iters": pinfo.iterations,
               "pressure_residual": pinfo.residual}
        if self.mixed:
            lin["pressure_rounds"] = pinfo.rounds
        if self.turb is not None and self.turb.last_info:
            lin["turbulence_iters"] = [i.iterations for i in self.turb.last_info.values()]
        # what the diagnostics of this iteration need besides the device systems it left behind
        self._capture = {"mom_info": infos, "p_info": pinfo, "energy_info": tinfo, "iteration": self.iteration,
                         "alpha_u": float(self.relax_u), "alpha_T": float(c.relax_T), "rolled_back": False}
        if isinstance(tinfo, _PendingSolve):  # still running on the energy stream
            tinfo.targets = (lin, self._capture)
            self._energy_pending = tinfo
        elif tinfo is not None:
            lin["energy_iters"] = tinfo.iterations
        self._last_linear = lin
        prev = getattr(self, "_last_res", {}), getattr(self, "_last_raw", {})
        self._last_raw = dict(res)
        if self.iteration <= 5:
            self._res_max5["continuity"] = max(self._res_max5.get("continuity", 0.0), res["continuity"], 1e-300)
        flux_scale = float(sc[_S_SUMS + 1]) + float(sc[_S_SUMS + 2])
        res["continuity"] /= max(self._res_max5["continuity"], 1e-3 * flux_scale)
        self._last_res = res
        rolled_back = self._stability_guard(self._last_raw["continuity"], flux_scale)
        if rolled_back:
            self._capture["rolled_back"] = True
            self._graph = None  # relaxation and fields changed under the recorded graph
            self._last_res, self._last_raw = prev
            return 1.0
        if cs is not None:
            self._judge_cases(cs)
            self.converged = bool(self.case_converged.all())
        elif self.iteration > 5 and max(res.values()) < c.tol:
            self.converged = True
        return 1.0

    def _energy_assemble(self, L, s, M, B, R, buf):
        """The energy step up to its solve (``simple.py:778-813``): T's boundary values with the
        lagged gradient, T's gradient, and the relaxed energy system with its residual sums.
        ``buf`` holds the system buffers the momentum solves do not share (the overlapped step
        gives the energy solve buffers of its own)."""
        c = self.case
        T = self._T
        self.T_old.copy_(T)
        nat.check(L.ftmi_energy_bvals(M, B, nat.ptr(T), nat.ptr(self.gT) if self._has_lagT else None,
                                      nat.ptr(self.Fb), nat.ptr(self.Tb), s), "energy bvals")
        arr = ctypes.c_void_p * 1
        nat.check(L.ftmi_ls_gradients(M, 1, arr(T.data_ptr()), arr(self.Tb.data_ptr()),
                                      arr(self.gT.data_ptr()), None, s), "T gradient")
        self._has_lagT = True
        a = nat.FtmiAssembleArgs()
        a.F, a.Fb, a.gamma, a.energy, a.nfields = nat.ptr(self.F), nat.ptr(self.Fb), float(c.conductivity / c.cp), 1, 1
        a.phi[0], a.grad[0], a.phib[0] = nat.ptr(T), nat.ptr(self.gT), nat.ptr(self.Tb)
        a.b[0], a.x0[0] = nat.ptr(self.bT), nat.ptr(self.Tnew)
        a.relax, a.second_order, a.simplec, a.cp = float(c.relax_T), int(c.scheme == "second-order"), -1, float(c.cp)
        a.diag, a.diag_r, a.off = nat.ptr(self.diagT), nat.ptr(self.diagT_r), nat.ptr(self.offT)
        a.resid, a.inv_r = nat.ptr(self.scal[_S_EN:]), nat.ptr(buf["inv_r"])
        if self._parts is not None:
            a.cell_resid = nat.ptr(self._parts["en"])
        if self.mixed:
            a.out_fp32 = 1
            a.diag32, a.diag_r32, a.off32 = nat.ptr(buf["diag32"]), nat.ptr(buf["diag_r32"]), nat.ptr(buf["off32"])
            a.inv_r32, a.ex32, a.b32[0] = nat.ptr(buf["inv_r32"]), nat.ptr(buf["ex32"]), nat.ptr(buf["b32"])
        if self.turb is not None:  # k/cp + mu_t / Pr_t (simple.py:781-782)
            ge_f, ge_b = self.turb.energy_terms(L, M, B, s, c.conductivity / c.cp)
            a.gam_f, a.gam_b = nat.ptr(ge_f), nat.ptr(ge_b)
        if self.en_src is not None:  # simple.py:795-796
            a.extra[0] = nat.ptr(self.en_src)
        nat.check(L.ftmi_transport_assemble(M, B, ctypes.byref(a), R, s), "energy assembly")

    def _energy_solve(self, buf):
        """The energy system's Jacobi-BiCGSTAB (tol 1e-3, 60 iterations) into T."""
        if self.mixed:
            if buf["soff32"] is not None:
                self.fv.slice_values(buf["off32"], buf["soff32"])
            info = bicgstab_ldu_mixed(self.fv, buf["ex32"], buf["off32"], buf["b32"], buf["inv_r32"], self.Tnew,
                                      1e-3, 60, True, ws=buf["ws"], b_is_residual=True, sliced_off=buf["soff32"])
        else:
            if buf["soff"] is not None:
                self.fv.slice_values(self.offT, buf["soff"])
            info = bicgstab_ldu(self.fv, self.diagT_r, self.offT, self.bT, buf["inv_r"], self.Tnew, 1e-3, 60, True,
                                ws=buf["ws"], sliced_off=buf["soff"])
        self._T.copy_(self.Tnew)  # a copy, not a swap: recorded graphs keep their pointers
        return info

    def _energy_step(self, L, s, M, B, R):
        self._join_energy()  # an earlier overlapped step may still be on the energy stream
        buf = {"inv_r": self.inv_r, "ws": self.ws, "soff": self.soff}
        if self.mixed:
            buf.update(diag32=self.diag32, diag_r32=self.diag_r32, off32=self.off32, inv_r32=self.inv_r32,
                       ex32=self.ex32, b32=self.b32[0], soff32=self.soff32)
        self._energy_assemble(L, s, M, B, R, buf)
        return self._energy_solve(buf)

    # ------------------------------------------------------------------ energy beside the flow
    def _overlap_energy(self) -> bool:
        """Whether this iteration's energy step may run beside the next iteration's flow work:
        T must feed nothing back into the flow (no turbulence model, whose viscosity the
        energy diffusivity reads and the turbulence update rewrites), the step must be eager
        (graph recording captures one stream) and the solver must hold one case."""
        return (self.energy and self.overlap_energy and self.turb is None and not self.use_graph
                and self._parts is None and self._pin is None)

    def _energy_side(self):
        """The energy stream (default priority) and the energy solve's own buffers: inverse
        diagonal, slice-layout operator, the FP32 system in mixed precision, Krylov workspace and
        report."""
        if self._estream is None:
            dev, nc, nf = self.device, self.nc, self.fv.n_if
            es = torch.cuda.Stream(device=dev)
            z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=dev)  # noqa: E731
            f = lambda *s: torch.zeros(*s, dtype=torch.float32, device=dev)  # noqa: E731
            sl = self.fv.sl_size if self.fv.sl_base is not None else None
            # allocated on the energy stream, so that the caching allocator never hands their
            # memory to the main stream while energy kernels may still use it
            with torch.cuda.stream(es):
                buf = {"inv_r": z(nc), "ws": Workspace(dev),
                       "soff": z(sl) if sl is not None and not self.mixed else None,
                       "report": z(4)}
                if self.mixed:
                    buf.update(diag32=f(nc), diag_r32=f(nc), off32=f(2 * nf), inv_r32=f(nc), ex32=f(nc), b32=f(nc),
                               soff32=f(sl) if sl is not None else None)
            buf.update(report_h=torch.zeros(2, 4, dtype=torch.float64, pin_memory=True), slot=0)
            # the solve reads the energy system, the mesh's slice layout and Tnew, and writes T:
            # every device tensor of the solver and its mesh is marked as used by the energy
            # stream, so that memory freed with them (a solver dropped while its last solve
            # still runs) is not handed out again before that solve ends
            for t in [*vars(self.fv).values(), *vars(self).values()]:
                if isinstance(t, torch.Tensor) and t.is_cuda:
                    t.record_stream(es)
            self._ebuf = buf
            self._estream = es
        return self._estream, self._ebuf

    def _energy_step_overlapped(self, L, M, B):
        """The energy step with its solve on the energy stream. The assembly (boundary values,
        gradient, system, residual sums) runs on the main stream once the previous solve has
        finished with the energy buffers (a device-side wait, normally long passed: that solve
        ran beside this iteration's flow work); the solve, every iteration of it enqueued without
        host reads (the Krylov solvers' capture mode, one report slot), then runs on the energy
        stream while the host starts the next iteration. The main stream never waits for the
        energy stream within an iteration. Same kernels, same operands, same results as the
        serial step; its Krylov info is read when it is first asked for."""
        main = torch.cuda.current_stream(self.device)
        es, buf = self._energy_side()
        main.wait_stream(es)  # the previous solve is done with T, Tnew and the energy system
        self._energy_assemble(L, nat.stream_ptr(), M, B, ctypes.byref(self.c_red), buf)
        es.wait_stream(main)  # this iteration's system
        with torch.cuda.stream(es):
            self._energy_solve_unread(L, buf)
            # two host slots: the previous iteration's report is read (in that iteration's
            # successor's host part) while this one may still be on its way
            slot = buf["slot"]
            buf["slot"] ^= 1
            buf["report_h"][slot].copy_(buf["report"], non_blocking=True)
            done = es.record_event()
        return _PendingSolve(buf["report_h"][slot], done)

    def _energy_solve_unread(self, L, buf):
        """The energy solve with every iteration enqueued and no host read (the Krylov
        solvers' capture mode, one report slot)."""
        caps = (ctypes.c_int64 * 1)(0)  # 0: the full 60 iterations
        nat.check(L.ftmi_capture_plan(caps, 1, nat.ptr(buf["report"]), 1), "energy capture plan")
        prev = L.ftmi_capture_mode(1)
        try:
            self._energy_solve(buf)
        finally:
            L.ftmi_capture_mode(prev)
            L.ftmi_capture_plan(None, 0, None, 0)

    def _resolve_energy(self):
        """Wait for the pending energy solve and put its Krylov info where the serial step
        would have put it (the iteration's last_linear and diagnostics capture)."""
        p, self._energy_pending = self._energy_pending, None
        info = p.result()
        lin, cap = p.targets
        lin["energy_iters"] = info.iterations
        cap["energy_info"] = info

    def _join_energy(self):
        """Order the current stream after the energy stream's work (a device-side wait)."""
        if self._estream is not None:
            torch.cuda.current_stream(self.device).wait_stream(self._estream)

    @property
    def T(self):
        """The temperature field. With the energy step beside the flow, the current stream
        first waits for the energy stream, so every reader sees the finished solve."""
        self._join_energy()
        return self._T

    @T.setter
    def T(self, value):
        self._T = value

    @property
    def last_linear(self) -> dict:
        """Inner iteration counts of the latest iteration (the energy solve's once finished)."""
        if self._energy_pending is not None:
            self._resolve_energy()
        return self._last_linear

    @last_linear.setter
    def last_linear(self, value):
        self._last_linear = value

    @property
    def last_pressure_iters(self) -> int:
        """The latest iteration's pressure PCG count, without waiting for an energy solve that
        is still running beside the flow (``last_linear`` waits for it)."""
        return int(self._last_linear.get("pressure_iters") or 0)

    def set_reference_cells(self, cells):
        """Fix the pressure correction to zero at ``cells`` (one per closed case of a batch): their
        rows keep only the diagonal and a zero right-hand side, and the couplings of their
        neighbours to them are dropped. A closed case's pressure system is singular with the
        constant as null vector; the reference removes it by projecting out the mean
        (``simple.py:629-634``), which handles one closed case but not several side by side. With
        a reference cell per case the system is regular, and the flux and velocity corrections,
        which use only differences of p', are the same."""
        dev = self.device
        pin = torch.as_tensor(np.asarray(cells, np.int64), device=dev)
        is_pin = np.zeros(self.nc, bool)
        is_pin[np.asarray(cells)] = True
        fv = self.fv
        rows = np.repeat(np.arange(self.nc), np.diff(fv.cptr_h))
        he = np.flatnonzero(is_pin[rows] | is_pin[fv.cnbr_h])
        self._pin = (pin, torch.as_tensor(he, device=dev))
        self.singular_p = False
        self._graph = None

    def set_case_partition(self, cell_offsets):
        """Judge convergence case by case on a mesh that holds several independent cases side
        by side (a batch): case k owns cells ``cell_offsets[k]:cell_offsets[k+1]`` and, in the
        same order, contiguous runs of the interior and of the boundary faces. Each case then has
        the reference's residual normalisation of its own (``simple.py:675-703``: momentum and
        energy sums over its cells, continuity scaled by its largest value of the first five
        iterations and floored at 1e-3 of its own face flux), ``case_residuals`` holds them, and
        the solve is converged when every case is, ``case_converged_at`` recording the first
        iteration at which each case passed. The stability guard and the relaxation stay shared
        by the batch. The per-case sums come from per-cell residual terms the assembly writes
        and one deterministic segmented sum per quantity."""
        fv, dev = self.fv, self.device
        off = np.asarray(cell_offsets, np.int64)
        if off[0] != 0 or off[-1] != self.nc or np.any(np.diff(off) <= 0):
            raise ValueError("cell_offsets must rise from 0 to the cell count")
        ncase = off.size - 1
        case_of = np.repeat(np.arange(ncase), np.diff(off))

        def runs(owner):
            k = case_of[owner]
            if k.size and np.any(np.diff(k) < 0):
                raise ValueError("the faces of the cases are not grouped case by case")
            return np.searchsorted(k, np.arange(ncase + 1), side="left")

        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int64), device=dev)  # noqa: E731
        z = lambda *sh: torch.zeros(*sh, dtype=torch.float64, device=dev)  # noqa: E731
        self._parts = {"n": ncase, "cells": t(off), "faces": t(runs(fv.P_h[:fv.n_if])), "bfaces": t(runs(fv.Pb_h)),
                       "mom": z(2 * self.dim, self.nc), "en": z(2, self.nc) if self.energy else None,
                       "out": z(2 * self.dim + 5, ncase)}
        self.case_res_max5 = np.zeros(ncase)
        self.case_residuals = [dict() for _ in range(ncase)]
        self.case_converged = np.zeros(ncase, bool)
        self.case_converged_at = np.full(ncase, -1, np.int64)
        self._graph = None  # recorded launches must write the per-cell terms

    def _case_sums(self):
        """Per-case reductions of the iteration just enqueued, as a (rows, cases) host array:
        momentum |r| and |a_P u| per component, then sum |mass|, sum |F|, sum |Fb|, and the
        energy pair."""
        P, L, s = self._parts, nat.lib(), nat.stream_ptr()
        n, out, d2 = P["n"], P["out"], 2 * self.dim
        op = P["out"].data_ptr()
        row = lambda r: op + 8 * r * n  # noqa: E731
        nat.check(L.ftmi_segment_sums(n, nat.ptr(P["cells"]), d2, self.nc, nat.ptr(P["mom"]), 0, row(0), s), "case sums")
        nat.check(L.ftmi_segment_sums(n, nat.ptr(P["cells"]), 1, self.nc, nat.ptr(self.mass), 1, row(d2), s),
                  "case sums")
        nat.check(L.ftmi_segment_sums(n, nat.ptr(P["faces"]), 1, self.fv.n_if, nat.ptr(self.F), 1, row(d2 + 1), s),
                  "case sums")
        nat.check(L.ftmi_segment_sums(n, nat.ptr(P["bfaces"]), 1, self.fv.n_bf, nat.ptr(self.Fb), 1, row(d2 + 2), s),
                  "case sums")
        if self.energy:
            nat.check(L.ftmi_segment_sums(n, nat.ptr(P["cells"]), 2, self.nc, nat.ptr(P["en"]), 0, row(d2 + 3), s),
                      "case sums")
        return out.cpu().numpy()

    def _judge_cases(self, cs):
        """The reference's residual normalisation and convergence test, case by case."""
        c, dim, names = self.case, self.dim, ("x-velocity", "y-velocity", "z-velocity")
        cont = cs[2 * dim]
        if self.iteration <= 5:
            self.case_res_max5 = np.maximum(self.case_res_max5, np.maximum(cont, 1e-300))
        flux = cs[2 * dim + 1] + cs[2 * dim + 2]
        for k in range(self._parts["n"]):
            den = [cs[2 * q + 1, k] for q in range(dim)]
            floor = 1e-3 * max(max(den), 1e-300)
            r = {names[q]: cs[2 * q, k] / max(den[q], floor) for q in range(dim)}
            r["continuity"] = cont[k] / max(self.case_res_max5[k], 1e-3 * flux[k])
            if self.energy:
                rr, sd = cs[2 * dim + 3, k], cs[2 * dim + 4, k]
                r["energy"] = rr / sd if sd > 0 else rr
            self.case_residuals[k] = {q: float(v) for q, v in r.items()}
            ok = self.iteration > 5 and max(r.values()) < c.tol
            self.case_converged[k] = ok
            if ok and self.case_converged_at[k] < 0:
                self.case_converged_at[k] = self.iteration

    def _fp32_outputs(self, a, nfields):
        """Point a transport assembly at the FP32 system buffers (mixed precision)."""
        a.out_fp32 = 1
        a.diag32, a.diag_r32, a.off32 = nat.ptr(self.diag32), nat.ptr(self.diag_r32), nat.ptr(self.off32)
        a.inv_r32, a.ex32 = nat.ptr(self.inv_r32), nat.ptr(self.ex32)
        for k in range(nfields):
            a.b32[k] = nat.ptr(self.b32[k])

    def _accumulate_timers(self):
        evs = self._events
        for (name, e1), (_, e0) in zip(evs[1:], evs[:-1]):
            self.timers[name] += e0.elapsed_time(e1) * 1e-3

    # ------------------------------------------------------------------ stability guard (simple.py:705-776)
    def _reset_guard(self):
        """A fresh guard (``guard.StabilityGuard``) with the case's algorithm and momentum
        relaxation, holding a snapshot of the current state."""
        self.algo = self.case.algorithm
        self.relax_u = self.case.relax_u
        self.guard = StabilityGuard(self.guard_enabled, self._device_snapshot)

    @property
    def guard_events(self):
        return self.guard.events

    @property
    def guard_exhausted_at(self):
        return self.guard.exhausted_at

    def _device_snapshot(self):
        st = {"U": self.U.clone(), "p": self.p.clone(), "F": self.F.clone(), "Fb": self.Fb.clone(),
              "iteration": self.iteration, "res_continuity_norm": self._res_max5.get("continuity", 0.0)}
        if self.energy:
            st["T"] = self.T.clone()
        if self.turb is not None:
            st.update(self.turb.snapshot())
        return st

    def _restore_snapshot(self, st):
        self.U.copy_(st["U"])
        self.p.copy_(st["p"])
        self.F.copy_(st["F"])
        self.Fb.copy_(st["Fb"])
        if self.energy and "T" in st:
            self.T.copy_(st["T"])
        if self.turb is not None:
            self.turb.load(st)
        self.iteration = int(st["iteration"])
        self._has_lag = self._has_lagT = False
        if float(st.get("res_continuity_norm", 0.0)) > 0:
            self._res_max5["continuity"] = float(st["res_continuity_norm"])

    def _stability_guard(self, raw_continuity: float, flux_scale: float) -> bool:
        """Ask the guard about the iteration that just ended; on a rollback restore its snapshot
        (the iteration counter keeps counting), take the new relaxation and force an AMG
        rebuild. Returns True on a rollback."""
        self.guard.enabled = self.guard_enabled
        now = Settings(self.algo, self.relax_u, self.relax_p)
        rb = self.guard.observe(self.iteration, raw_continuity, flux_scale, now)
        if rb is None:
            return False
        at = self.iteration
        self._restore_snapshot(rb.snapshot)
        self.iteration = self.step_count = at
        self.algo, self.relax_u, self.relax_p = rb.settings.algorithm, rb.settings.relax_u, rb.settings.relax_p
        self._amg_age = 10 ** 9
        self._amg_needs_setup = True
        return True

    def _new_hierarchy(self, fv):
        c = self.case
        lv = {} if c.amg_levels is None else {"max_levels": int(c.amg_levels),
                                              "coarse_sweeps": int(c.amg_coarse_sweeps)}
        self._amg = SmoothedAggregationAMG.from_ldu(fv, self.diag_p, self.offp, singular=self.singular_p,
                                                   refreshable=True, **lv)
        self._amg32 = self._amg.to_fp32() if self.mixed else None

    def _reuse_hierarchy(self, fv) -> Optional[str]:
        """Two AMG hierarchies are kept: the current one and the one that the last full set-up
        replaced. One whose finest matrix is, bit for bit, the pressure matrix just assembled is
        what a set-up (or, in mixed precision, a refresh) of it would give again, and is used as
        it is: "current", or "kept" (the two are swapped). This happens when the stability guard
        restores a state whose first matrix was seen before. None when neither matches."""
        cur = (self._amg, self._amg32)
        if self._amg.matches_ldu(fv, self.diag_p, self.offp):
            return "current"
        kept = self._amg_kept
        if kept is not None and kept[0].matches_ldu(fv, self.diag_p, self.offp):
            self._amg_kept = cur
            self._amg, self._amg32 = kept
            return "kept"
        return None

    def _next_hierarchy(self, fv) -> str:
        """The AMG hierarchy of the pressure matrix just assembled, as a full set-up of it gives
        it (``amg.py:80-146`` re-aggregates at every rebuild), after ``_reuse_hierarchy`` found
        none: the current hierarchy is rebuilt on the levels whose strength masks stay
        ("rebuilt"), unless the finest mask
        changes: then a new hierarchy is set up and the current one is kept ("new")."""
        if self._amg is None:
            self._new_hierarchy(fv)
            return "new"
        if self._amg.rebuild_ldu(fv, self.diag_p, self.offp, keep=True):
            self._amg32 = self._amg.to_fp32() if self.mixed else None
            return "rebuilt"
        self._amg_kept = (self._amg, self._amg32)
        self._new_hierarchy(fv)
        return "new"

    # ------------------------------------------------------------------ state, fields, reports
    def diagnostics(self) -> dict:
        """Algebraic, conservation and health diagnostics of the latest outer iteration, from the
        systems it left on the device (``simple.py:diagnostics``, ``diagnostics.py:214-360``)."""
        from .diagnostics import evaluate

        if self._energy_pending is not None:
            self._resolve_energy()
        self._join_energy()  # the energy system and T_old of an overlapped energy step
        return evaluate(self)

    def residuals(self) -> dict:
        r = dict(getattr(self, "_last_res", {}))
        raw = dict(getattr(self, "_last_raw", r))
        return {"raw": raw, "scaled": r, "normalized": r}

    def fields(self, names=None) -> Dict[str, np.ndarray]:
        U = self.U.cpu().numpy().T
        out = {"x-velocity": U[:, 0].copy(), "y-velocity": U[:, 1].copy()}
        if self.dim == 3:
            out["z-velocity"] = U[:, 2].copy()
        out["velocity-magnitude"] = np.linalg.norm(U, axis=1)
        out["pressure"] = self.p.cpu().numpy()
        if self.energy:
            out["temperature"] = self.T.cpu().numpy()
        if self.turb is not None:
            out.update(self.turb.fields())
        if names:
            out = {k: v for k, v in out.items() if k in names}
        return out

    def cell_centres(self) -> np.ndarray:
        return self.fv.xc_h

    def state_dict(self) -> dict:
        st = {"U": self.U.cpu().numpy().T.copy(), "p": self.p.cpu().numpy(), "F": self.F.cpu().numpy(),
              "Fb": self.Fb.cpu().numpy(), "iteration": self.iteration, "time": 0.0, "time_step": 0,
              "res_continuity_norm": self._res_max5.get("continuity", 0.0)}
        if self.energy:
            st["T"] = self.T.cpu().numpy()
        if self.turb is not None:
            st.update(self.turb.state())
        return st

    def load_state(self, st: dict):
        dev = self.device
        self.U.copy_(torch.as_tensor(np.asarray(st["U"], float).T.copy(), device=dev))
        self.p.copy_(torch.as_tensor(np.asarray(st["p"], float), device=dev))
        self.F.copy_(torch.as_tensor(np.asarray(st["F"], float), device=dev))
        self.Fb.copy_(torch.as_tensor(np.asarray(st["Fb"], float), device=dev))
        if self.energy and "T" in st:
            self.T.copy_(torch.as_tensor(np.asarray(st["T"], float), device=dev))
        if self.turb is not None:
            self.turb.load(st)
        self.iteration = self.step_count = int(st.get("iteration", 0))
        self._has_lag = self._has_lagT = False
        self._capture = None
        if float(st.get("res_continuity_norm", 0.0)) > 0:
            self._res_max5["continuity"] = float(st["res_continuity_norm"])

    def set_lagged_gradients(self, gp, gU, gT=None):
        """Gradients of the previous iteration, used to reconstruct boundary values."""
        dev = self.device
        if gp is None:
            self._has_lag = False
        else:
            self.gp.copy_(torch.as_tensor(np.asarray(gp, float), device=dev))
            self.gU.copy_(torch.as_tensor(np.stack([np.asarray(g, float) for g in gU]), device=dev))
            self._has_lag = True
        if self.energy:
            if gT is None:
                self._has_lagT = False
            else:
                self.gT.copy_(torch.as_tensor(np.asarray(gT, float), device=dev))
                self._has_lagT = True

    def _pressure_bvals_h(self):
        fv = self.fv
        p = self.p.cpu().numpy()
        pP = p[fv.Pb_h]
        if self._has_lag:
            pP = pP + np.sum(self.gp.cpu().numpy()[fv.Pb_h] * fv.db.cpu().numpy(), axis=1)
        m_pres = ((self.bkind_h == OUTLET) | (self.bkind_h == PRESSURE)).astype(float)
        return m_pres * self.pb_fixed_h + (1.0 - m_pres) * pP

    def wall_forces(self) -> dict:
        """Pressure + viscous force on every wall zone (``simple.py:970-1001``); the viscous part
        uses the wall viscosity of the momentum equation (laminar, or the SST wall function)."""
        fv = self.fv
        pb = self._pressure_bvals_h()
        U = self.U.cpu().numpy().T
        du = U[fv.Pb_h] - self.ub_fixed_h
        du_t = du - np.einsum("fd,fd->f", du, fv.nb_h)[:, None] * fv.nb_h
        f_p = pb[:, None] * fv.Sb_h
        mu_b = self.turb.boundary_viscosity() if self.turb is not None else np.full(fv.n_bf, self.case.viscosity)
        f_v = (mu_b * fv.alpha_b_h)[:, None] * du_t
        out = {}
        for z in self.mesh.boundary_zones():
            m = (fv.bzone == z.id) & (self.bkind_h == WALL)
            if m.any():
                fp, fvis = f_p[m].sum(axis=0), f_v[m].sum(axis=0)
                out[z.name] = {"pressure": fp.tolist(), "viscous": fvis.tolist(), "total": (fp + fvis).tolist()}
        return out

    def heat_flow(self) -> dict:
        """Conduction + enthalpy flow into the domain per boundary zone (``simple.py:1003-1031``)."""
        if not self.energy:
            return {}
        c, fv = self.case, self.fv
        T = self.T.cpu().numpy()
        Fb = self.Fb.cpu().numpy()
        TP = T[fv.Pb_h]
        tk, tv, kind = self.tkind_h, self.tval_h, self.bkind_h
        Tb = np.where(tk == T_FIXED, tv, TP)
        Tb = np.where((kind == OUTLET) & (Fb < 0), tv, Tb)
        q = c.conductivity * fv.alpha_b_h * (Tb - TP) * (tk == T_FIXED)
        q = q + np.where(tk == T_FLUX, tv * fv.Sb_mag_h, 0.0) - c.cp * Fb * Tb
        return {z.name: float(q[fv.bzone == z.id].sum()) for z in self.mesh.boundary_zones()}

    def mass_balance(self) -> dict:
        Fb = self.Fb.cpu().numpy()
        return {z.name: float(Fb[self.fv.bzone == z.id].sum()) for z in self.mesh.boundary_zones()
                if (self.fv.bzone == z.id).any()}

    def describe(self) -> dict:
        return {"solver": f"pressure-based {self.algo}", "backend": "hip", "device": str(self.device),
                "precision": "mixed (FP32 solves, FP64 refinement)" if self.mixed else "float64", "cells": self.nc, "faces": self.mesh.n_faces, "dim": self.dim,
                "convection": self.case.scheme, "relax_p": self.relax_p,
                "stability_rollbacks": list(self.guard_events),
                "stability_guard_exhausted_at": self.guard_exhausted_at,
                "pressure_solver": "AMG-PCG (SA-AMG, V(2,2) Jacobi, HIP)",
                "max_non_orthogonality_deg": round(self.fv.non_orthogonality_deg, 2),
                "setup_seconds": round(self.setup_seconds, 3)}


__all__ = ["SimpleSolver", "auto_relax_p", "KrylovInfo"]
