"""Divergence guard of the outer iteration: detect a diverging pressure-velocity coupling from
the raw continuity residual and prescribe a rollback with gentler relaxation.

Behaviour (``flowtrace/unstructured/simple.py:705-776``):

* **Trip condition.** After an iteration, the raw continuity ``r`` trips the guard when it is
  not finite, or, past iteration 10, when it exceeds 20 x its running minimum, the minimum being
  floored at 1e-6 x the total face mass flux (round-off of periodic or converged flows is not
  growth).
* **Remedies**, tried in the order of ``REMEDIES``; the first one whose precondition holds is
  applied, together with a restore of the last healthy snapshot:

  ======================  ====================================  ===============================
  name                    precondition                          new settings
  ======================  ====================================  ===============================
  ``lower_relax_p``       ``0.6 relax_p >= 0.1``                relax_p x 0.6
  ``simple_fallback``     algorithm is SIMPLEC                  SIMPLE, relax_u <= 0.3, relax_p 0.3
  ``lower_relax_u``       algorithm is SIMPLE and               relax_u x 0.7
                          ``0.7 relax_u >= 0.2``
  ======================  ====================================  ===============================

  When no remedy applies, the first such iteration is recorded as ``exhausted_at`` and the run
  goes on.
* **Bookkeeping** (only when no rollback was prescribed): past iteration 5 a finite ``r``
  lowers the running minimum, and every 20th iteration whose ``r`` is within 2 x that minimum
  becomes the new healthy snapshot. The snapshot taken at set-up has an infinite residual.
* After a rollback the running minimum becomes the restored snapshot's residual. The iteration
  counter is not rewound (discarded iterations count toward ``max_iter``).

The guard only decides. The solver owns the relaxation settings (it passes the current ones to
``observe``) and the device state: the guard holds a snapshot taken through the callable the
solver gives it, and hands it back for the solver to restore, so the same guard serves every
precision and layout.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple

GROWTH = 20.0
ROUNDOFF = 1e-6
SNAPSHOT_EVERY = 20
SNAPSHOT_HEALTH = 2.0
QUIET_UNTIL = 10  # growth is not judged before this iteration (start-up transient)
TRACK_FROM = 5  # the running minimum starts after this iteration


@dataclass
class Settings:
    """The relaxation the outer iteration runs with; remedies return a changed copy."""
    algorithm: str
    relax_u: float
    relax_p: float


def _lower_relax_p(s: Settings) -> Optional[Tuple[Settings, dict]]:
    if s.relax_p * 0.6 < 0.1:
        return None
    new = Settings(s.algorithm, s.relax_u, s.relax_p * 0.6)
    return new, {"relax_p": round(new.relax_p, 4)}


def _simple_fallback(s: Settings) -> Optional[Tuple[Settings, dict]]:
    if s.algorithm != "SIMPLEC":
        return None
    # a high-Reynolds flow on a skewed mesh that SIMPLEC cannot hold at any relax_p: the classical
    # robust setting is SIMPLE with momentum relaxation 0.3
    new = Settings("SIMPLE", min(s.relax_u, 0.3), 0.3)
    return new, {"algorithm": new.algorithm, "relax_p": round(new.relax_p, 4), "relax_u": round(new.relax_u, 4)}


def _lower_relax_u(s: Settings) -> Optional[Tuple[Settings, dict]]:
    if s.algorithm != "SIMPLE" or s.relax_u * 0.7 < 0.2 - 1e-12:
        return None
    new = Settings(s.algorithm, s.relax_u * 0.7, s.relax_p)
    return new, {"algorithm": new.algorithm, "relax_p": round(new.relax_p, 4), "relax_u": round(new.relax_u, 4)}


REMEDIES: List[Tuple[str, Callable[[Settings], Optional[Tuple[Settings, dict]]]]] = [
    ("lower_relax_p", _lower_relax_p),
    ("simple_fallback", _simple_fallback),
    ("lower_relax_u", _lower_relax_u),
]


@dataclass
class Rollback:
    """What the solver must do after a tripped iteration: restore ``snapshot`` and run on with
    ``settings``; ``event`` is the record appended to ``events``."""
    snapshot: object
    settings: Settings
    event: dict


@dataclass
class StabilityGuard:
    enabled: bool
    take_snapshot: Callable[[], dict]
    events: List[dict] = field(default_factory=list)
    exhausted_at: Optional[int] = None
    _floor: float = math.inf
    _snap: Optional[Tuple[dict, float]] = None

    def __post_init__(self):
        if self.enabled:
            self._snap = (self.take_snapshot(), math.inf)

    def _tripped(self, iteration: int, r: float, flux_scale: float) -> bool:
        if not math.isfinite(r):
            return True
        return iteration > QUIET_UNTIL and r > GROWTH * max(self._floor, ROUNDOFF * flux_scale)

    def observe(self, iteration: int, r: float, flux_scale: float, settings: Settings) -> Optional[Rollback]:
        """Judge the iteration that just ended (``iteration`` counts it), run with ``settings``.
        Returns the rollback to perform, or None to keep the iteration."""
        if not self.enabled:
            return None
        if self._tripped(iteration, r, flux_scale):
            for name, remedy in REMEDIES if self._snap is not None else ():
                out = remedy(settings)
                if out is None:
                    continue
                new, extra = out
                snap, snap_r = self._snap
                event = {"detected_at": iteration, "restored_to": int(snap["iteration"])}
                event.update(extra)
                event["continuity"] = r
                self.events.append(event)
                self._floor = snap_r
                return Rollback(snap, new, event)
            if self.exhausted_at is None:
                self.exhausted_at = iteration
        if math.isfinite(r) and iteration > TRACK_FROM:
            self._floor = min(self._floor, r)
            if iteration % SNAPSHOT_EVERY == 0 and r <= SNAPSHOT_HEALTH * self._floor:
                self._snap = (self.take_snapshot(), r)
        return None
