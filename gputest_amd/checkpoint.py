"""Checkpoints of a running solve (``flowtrace/runtime/checkpoint.py:1-85``).

The file format is the reference's: one ``<tag>.npz`` per checkpoint holding the arrays of
``SimpleSolver.state_dict()`` (``U`` as (cells, dim), ``p``, ``F``, ``Fb``, ``T`` when the energy
equation is on, ``k``, ``w`` and ``mut`` with the SST model, and the scalars ``iteration``, ``time``, ``time_step``,
``res_continuity_norm``) plus a ``__meta__`` entry, the UTF-8 bytes of a JSON object. It is
written to a temporary name and renamed, so a reader never sees half a file, and only the
newest ``keep`` files of the directory are kept.

The GPU side differs from a host ``state_dict()``: ``snapshot(solver)`` enqueues the copies of
the state into pinned host buffers on the solver's stream (so they are ordered before the
next iteration's kernels) and records an event; the writer thread waits for that event and
writes the file while the solver keeps iterating. Two buffer sets alternate, and the queue
holds at most two pending snapshots, so a slow disk makes ``submit`` wait instead of
filling host memory.
"""
from __future__ import annotations

import glob
import json
import os
import queue
import threading
import time
from typing import Dict, Optional

import numpy as np
import torch

_STATE_ARRAYS = ("U", "p", "F", "Fb", "T")
_TURB_ARRAYS = ("k", "w", "mut")


class Snapshot:
    """State of a solver copied (asynchronously) into pinned host memory."""

    def __init__(self, arrays: Dict[str, torch.Tensor], scalars: dict, event: Optional[torch.cuda.Event],
                 release=None):
        self.arrays, self.scalars, self.event, self._release = arrays, scalars, event, release

    def state(self) -> dict:
        """The reference's state dictionary, after the copies have landed (the pinned buffers
        are handed back for reuse once copied out)."""
        try:
            if self.event is not None:
                self.event.synchronize()
            st = {k: (v.numpy().T.copy() if k == "U" else v.numpy().copy()) for k, v in self.arrays.items()}
        finally:
            if self._release is not None:
                self._release()
                self._release = None
        st.update(self.scalars)
        return st


def _device_state(solver) -> Dict[str, torch.Tensor]:
    """The device arrays of ``state_dict()``: the flow fields, T with the energy equation, and
    the SST fields k, w, mut with turbulence (``simple.py:1033-1046`` of the reference; the
    turbulence model's ``state()``)."""
    out = {k: getattr(solver, k) for k in _STATE_ARRAYS if k != "T" or solver.energy}
    turb = getattr(solver, "turb", None)
    if turb is not None:
        out.update({k: getattr(turb, k) for k in _TURB_ARRAYS})
    return out


class _Staging:
    """Pinned host buffers for one snapshot of a given solver layout, and whether they are free."""

    def __init__(self, solver):
        self.bufs = {k: torch.empty(v.shape, dtype=v.dtype, pin_memory=True) for k, v in _device_state(solver).items()}
        self.free = threading.Event()
        self.free.set()


def snapshot(solver, staging: Optional[_Staging] = None) -> Snapshot:
    """Enqueue device-to-host copies of the solver state on the current stream."""
    staging = staging or _Staging(solver)
    staging.free.wait()
    staging.free.clear()
    src = _device_state(solver)
    if set(src) != set(staging.bufs):
        staging.free.set()
        raise ValueError(f"checkpoint staging holds {sorted(staging.bufs)}, the solver has {sorted(src)}")
    for k, buf in staging.bufs.items():
        buf.copy_(src[k], non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    scalars = {"iteration": int(solver.iteration), "time": 0.0, "time_step": 0,
               "res_continuity_norm": float(solver._res_max5.get("continuity", 0.0))}
    return Snapshot(dict(staging.bufs), scalars, ev, staging.free.set)


class CheckpointWriter:
    """Background writer of solver checkpoints (``checkpoint.py:23-74``)."""

    def __init__(self, directory: str, keep: int = 3):
        self.directory, self.keep = directory, keep
        os.makedirs(directory, exist_ok=True)
        self._q: "queue.Queue" = queue.Queue(maxsize=2)
        self._staging = []  # two alternating pinned sets per solver layout
        self._next = 0
        self.written: list = []
        self.last_error: Optional[str] = None
        self.bytes_written = 0
        self.seconds = 0.0
        self._thread = threading.Thread(target=self._run, name="gputest-checkpoint", daemon=True)
        self._thread.start()

    def submit(self, tag: str, state, meta: dict) -> None:
        """Queue a checkpoint: ``state`` is a dict of host arrays, a ``Snapshot``, or a solver
        (snapshotted here into the next pinned buffer set)."""
        if hasattr(state, "state_dict") and hasattr(state, "U"):
            if not self._staging:
                self._staging = [_Staging(state), _Staging(state)]
            # a buffer set is reused two submissions later, once the writer has copied it out
            st = self._staging[self._next]
            self._next ^= 1
            state = snapshot(state, st)
        self._q.put((tag, state, dict(meta)))

    def _run(self):
        while True:
            item = self._q.get()
            if item is None:
                self._q.task_done()
                return
            tag, state, meta = item
            t0 = time.perf_counter()
            path = os.path.join(self.directory, f"{tag}.npz")
            tmp = path + ".tmp.npz"
            try:
                st = state.state() if isinstance(state, Snapshot) else state
                arrays = {k: np.asarray(v) for k, v in st.items()}
                bad = [k for k, v in arrays.items() if v.dtype == object]
                if bad:  # the files hold plain arrays only, never pickles
                    raise TypeError(f"not numeric arrays: {', '.join(bad)}")
                arrays["__meta__"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
                np.savez(tmp, **arrays)
                os.replace(tmp, path)
                self.written.append(path)
                self.bytes_written += os.path.getsize(path)
                self._prune()
            except Exception as exc:  # a full disk must not stop the solve
                self.last_error = f"{type(exc).__name__}: {exc}"
                if isinstance(state, Snapshot) and state._release is not None:
                    state._release()
            finally:
                self.seconds += time.perf_counter() - t0
                self._q.task_done()

    def _prune(self):
        files = sorted(glob.glob(os.path.join(self.directory, "*.npz")), key=os.path.getmtime)
        for f in files[:-self.keep] if self.keep > 0 else []:
            try:
                os.remove(f)
            except OSError:
                pass

    def flush(self):
        self._q.join()

    def close(self):
        self.flush()
        self._q.put(None)
        self._thread.join(timeout=10)


def load_checkpoint(path: str):
    """(state, meta) of a checkpoint file (``checkpoint.py:77-85``); nothing is unpickled."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["__meta__"]).decode())
        state = {k: z[k] for k in z.files if k != "__meta__"}
    for k in ("iteration", "time", "time_step", "res_continuity_norm", "t", "step", "dt_prev"):
        if k in state and state[k].ndim == 0:
            state[k] = state[k].item()
    return state, meta


__all__ = ["CheckpointWriter", "load_checkpoint", "snapshot", "Snapshot"]
