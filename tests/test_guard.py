"""The stability guard and the boundary heat flows, pinned on the CPU.

* The oracle follows the reference's own guard trajectory on the reference's guard test case
  (``tests/test_planner.py:93-113`` of the reference), event for event, with the reference's
  fields at iterations 40 and 100 and at convergence (``guard_sheared24.npz``).
* The product's guard (``gputest_amd.guard``, host logic) takes the same decisions as the
  oracle's restatement on synthetic residual sequences that walk every remedy.
* ``heat_flow`` of the oracle equals the reference's per zone on the reference's own states
  (``heat_flow_*.npz``), and meets the closed forms: flux walls bring in flux x wall length,
  and at steady state the zone values sum to zero (``simple.py:1003-1008``).
"""
import math

import numpy as np
import pytest

import oracle.simplec_oracle as O
from gputest_amd import guard as G
from tests.guard_cases import (events_match, heat_channel, heat_duct, heat_states, load, raw_continuity,
                               run_until_nonfinite, sheared_case, sheared_mesh)


def rel_l2(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_oracle_guard_follows_the_reference_on_the_sheared_mesh():
    g = load("guard_sheared24.npz")
    s = O.Simplec(sheared_mesh(), sheared_case(O, None))
    assert abs(s.fv.non_orthogonality_deg - float(g["non_orth_deg"])) < 1e-12
    assert s.fv.non_orthogonality_deg > 40
    snaps = {}
    while not s.finished and math.isfinite(raw_continuity(s)):
        s.step()
        if s.iteration in (40, 100) and s.iteration not in snaps:
            snaps[s.iteration] = (s.U.copy(), s.p.copy())
    # the reference's own assertions
    assert s.converged and s.guard_events and s.relax_p < 0.5
    # and its trajectory: every event, the iteration count, the fields on the way
    events_match(s.guard_events, g)
    assert s.iteration == int(g["iterations"]) and s.relax_p == float(g["relax_p_final"])
    for it, (U, p) in snaps.items():
        assert rel_l2(U, g[f"it{it}_U"]) < 1e-9, it
        assert rel_l2(p - p.mean(), g[f"it{it}_p"] - g[f"it{it}_p"].mean()) < 1e-9, it
    assert rel_l2(s.U, g["U"]) < 1e-9
    assert rel_l2(s.p - s.p.mean(), g["p"] - g["p"].mean()) < 1e-9


def test_oracle_user_relaxation_diverges_without_rollback():
    g = load("guard_sheared24.npz")
    s = run_until_nonfinite(O.Simplec(sheared_mesh(), sheared_case(O, 0.8)), raw_continuity)
    assert not s.converged and not s.guard_events
    assert s.iteration == int(g["fixed_iterations"]) and int(g["fixed_events"]) == 0


class _Stub:
    """The state the oracle's guard method reads and writes, for driving it with synthetic
    residuals."""
    GUARD_GROWTH, GUARD_ROUNDOFF, GUARD_EVERY = O.Simplec.GUARD_GROWTH, O.Simplec.GUARD_ROUNDOFF, O.Simplec.GUARD_EVERY
    GUARD_FACTOR, GUARD_MIN_RELAX = O.Simplec.GUARD_FACTOR, O.Simplec.GUARD_MIN_RELAX
    GUARD_MIN_RELAX_U = O.Simplec.GUARD_MIN_RELAX_U

    def __init__(self, algo, relax_u, relax_p):
        self.algo, self.relax_u, self.relax_p = algo, relax_u, relax_p
        self.iteration, self.guard_enabled, self._amg_age = 0, True, 0
        self.guard_events, self.guard_exhausted_at = [], None
        self._guard_min = float("inf")
        self._guard_snap = (self.state_dict(), float("inf"))

    def state_dict(self):
        return {"iteration": self.iteration}

    def load_state(self, st):
        self.iteration = st["iteration"]


def _sequence(seed, n):
    """Raw continuity values that decay with noise and occasionally blow up or go non-finite."""
    rng = np.random.default_rng(seed)
    r = np.exp(-np.arange(n) / 40.0) * (1.0 + 0.3 * rng.random(n))
    spikes = rng.random(n)
    r[spikes < 0.06] *= 100.0
    r[spikes > 0.985] = np.nan
    return r


@pytest.mark.parametrize("algo,relax_u,relax_p,seed", [("SIMPLEC", 0.7, 1.0, 0), ("SIMPLEC", 0.7, 0.8, 1),
                                                       ("SIMPLE", 0.7, 0.3, 2), ("SIMPLEC", 0.25, 0.8, 3)])
def test_product_guard_takes_the_oracles_decisions(algo, relax_u, relax_p, seed):
    ref = _Stub(algo, relax_u, relax_p)
    cur = G.Settings(algo, relax_u, relax_p)
    it = {"n": 0}
    guard = G.StabilityGuard(True, lambda: {"iteration": it["n"]})
    rollbacks = 0
    for r in _sequence(seed, 400):
        ref.iteration += 1
        it["n"] += 1
        flux = 1.0
        rb_ref = O.Simplec._guard(ref, float(r), flux)
        rb = guard.observe(it["n"], float(r), flux, cur)
        assert (rb is not None) == rb_ref
        if rb is not None:
            rollbacks += 1
            cur = rb.settings  # the counter keeps counting, as in the solver
        assert (cur.algorithm, cur.relax_u, cur.relax_p) == (ref.algo, ref.relax_u, ref.relax_p)
    assert repr(guard.events) == repr(ref.guard_events)  # repr: a non-finite trip records NaN
    assert guard.exhausted_at == ref.guard_exhausted_at
    assert rollbacks >= 3
    kinds = {("algorithm" in e) for e in guard.events}
    if algo == "SIMPLEC" and relax_p < 0.5:
        assert kinds == {False}


def test_guard_walks_every_remedy_and_exhausts():
    """A residual that trips every iteration after the quiet start: relax_p falls by 0.6 until its
    floor, then SIMPLE with relax_u 0.3, then relax_u falls by 0.7 to its floor, then the guard is
    exhausted and stops rolling back."""
    cur = G.Settings("SIMPLEC", 0.7, 1.0)
    guard = G.StabilityGuard(True, lambda: {"iteration": 0})
    seen = []
    for it in range(1, 60):
        rb = guard.observe(it, float("nan") if it > 11 else 1.0, 1.0, cur)
        if rb is not None:
            cur = rb.settings
            seen.append((cur.algorithm, round(cur.relax_u, 6), round(cur.relax_p, 6)))
    assert seen[:4] == [("SIMPLEC", 0.7, 0.6), ("SIMPLEC", 0.7, 0.36), ("SIMPLEC", 0.7, 0.216),
                        ("SIMPLEC", 0.7, 0.1296)]
    assert seen[4] == ("SIMPLE", 0.3, 0.3)
    assert seen[5:] == [("SIMPLE", 0.3, 0.18), ("SIMPLE", 0.3, 0.108), ("SIMPLE", 0.21, 0.108)]
    assert guard.exhausted_at == 20
    assert guard.observe(61, float("nan"), 1.0, cur) is None


# ---------------------------------------------------------------- heat flow
@pytest.mark.parametrize("name,builder", [("heat_flow_channel.npz", heat_channel), ("heat_flow_duct.npz", heat_duct)])
def test_oracle_heat_flow_equals_the_references(name, builder):
    g = load(name)
    mesh, case = builder(O)
    s = O.Simplec(mesh, case)
    for tag, st, q in heat_states(g):
        s.load_state(st)
        got = s.heat_flow()
        assert set(got) == set(q)
        scale = max(abs(v) for v in q.values())
        for z, v in q.items():
            assert abs(got[z] - v) <= 1e-12 * scale, (tag, z, got[z], v)


@pytest.mark.parametrize("name,builder", [("heat_flow_channel.npz", heat_channel), ("heat_flow_duct.npz", heat_duct)])
def test_oracle_heat_flow_closed_forms_at_steady_state(name, builder):
    """Run the oracle to the reference's convergence: the zone heat flows sum to zero, the flux
    walls bring in exactly flux x length, and the iteration count is the reference's."""
    g = load(name)
    mesh, case = builder(O)
    s = O.Simplec(mesh, case)
    while not s.finished:
        s.step()
    assert s.converged and s.iteration == int(g["iterations"])
    q = s.heat_flow()
    scale = max(abs(v) for v in q.values())
    assert abs(sum(q.values())) <= 1e-6 * scale, q
    ref = {k[len("it0_q_"):]: float(g[k]) for k in g.files if k.startswith("it0_q_")}
    for z in q:
        assert abs(q[z] - ref[z]) <= 1e-8 * scale, (z, q[z], ref[z])
    if name == "heat_flow_duct.npz":
        assert abs(q["bottom"] - 5000.0 * 4.0) < 1e-9 and abs(q["top"] - 5000.0 * 4.0) < 1e-9
        assert abs(q["inlet"] + q["outlet"] + 40000.0) <= 1e-6 * scale
    else:  # adiabatic top wall, no heat through it
        assert q["top"] == 0.0 and q["bottom"] > 0.0
