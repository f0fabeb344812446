"""The stability guard and the boundary heat flows of the HIP solver against the reference's own
trajectories (``guard_sheared24.npz``, ``heat_flow_*.npz``) and against the oracle.

The guard case is the reference's (``tests/test_planner.py:93-113``): SIMPLEC on 24 x 24 quads
sheared by 45 degrees diverges at the automatic relax_p; the guard rolls back to a healthy
snapshot and lowers relax_p until the iteration converges. The device solver must trip at the
same iterations, restore the same snapshots, pick the same relaxations and converge at the same
iteration to the same fields.
"""
import math

import numpy as np
import pytest

import oracle.simplec_oracle as O
from gputest_amd import case as C
from gputest_amd.simple import SimpleSolver
from tests.guard_cases import (events_match, heat_channel, heat_duct, heat_states, load, raw_continuity,
                               run_until_nonfinite, sheared_case, sheared_mesh)

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def host(s):
    return s.U.cpu().numpy().T.copy(), s.p.cpu().numpy().copy()


def test_guard_trajectory_matches_the_reference_event_for_event():
    g = load("guard_sheared24.npz")
    s = SimpleSolver(sheared_mesh(), sheared_case(C, None), "cuda")
    print("guard case: sheared 24x24, automatic relax_p", s.relax_p, flush=True)
    snaps = {}
    while not s.finished and math.isfinite(raw_continuity(s)):
        s.step()
        if s.iteration in (40, 100) and s.iteration not in snaps:
            snaps[s.iteration] = host(s)
    print("events", s.guard_events, "iterations", s.iteration, flush=True)
    assert s.converged and s.guard_events and s.relax_p < 0.5
    assert s.describe()["stability_rollbacks"] == s.guard_events
    events_match(s.guard_events, g, cont_rel=1e-7)
    assert s.iteration == int(g["iterations"]) and s.relax_p == float(g["relax_p_final"])
    for it, (U, p) in snaps.items():
        assert rel_l2(U, g[f"it{it}_U"]) < 1e-8, it
        assert rel_l2(p - p.mean(), g[f"it{it}_p"] - g[f"it{it}_p"].mean()) < 1e-8, it
    U, p = host(s)
    assert rel_l2(U, g["U"]) < 1e-8
    assert rel_l2(p - p.mean(), g["p"] - g["p"].mean()) < 1e-8
    assert np.isfinite(s.fields()["velocity-magnitude"]).all()


def test_guard_trajectory_matches_the_oracle():
    o = O.Simplec(sheared_mesh(), sheared_case(O, None))
    s = SimpleSolver(sheared_mesh(), sheared_case(C, None), "cuda")
    while not o.finished:
        o.step()
    while not s.finished:
        s.step()
    assert [(e["detected_at"], e["restored_to"], e["relax_p"]) for e in s.guard_events] == \
        [(e["detected_at"], e["restored_to"], e["relax_p"]) for e in o.guard_events]
    U, p = host(s)
    assert s.iteration == o.iteration
    assert rel_l2(U, o.U) < 1e-9 and rel_l2(p - p.mean(), o.p - o.p.mean()) < 1e-9


def test_user_relaxation_is_left_alone_and_diverges():
    g = load("guard_sheared24.npz")
    s = run_until_nonfinite(SimpleSolver(sheared_mesh(), sheared_case(C, 0.8), "cuda"), raw_continuity)
    assert not s.converged and not s.guard_events
    # the diverging trajectory amplifies round-off, so the iteration at which it overflows is not
    # a parity quantity (the oracle, with the reference's summation order, overflows at the
    # reference's iteration, 52); the reference asserts only that it diverges
    assert not math.isfinite(raw_continuity(s)) and s.iteration < 3000
    print("overflow at", s.iteration, "reference", int(g["fixed_iterations"]))


@pytest.mark.parametrize("name,builder", [("heat_flow_channel.npz", heat_channel), ("heat_flow_duct.npz", heat_duct)])
def test_heat_flow_equals_the_references_on_its_states(name, builder):
    g = load(name)
    mesh, case = builder(C)
    s = SimpleSolver(mesh, case, "cuda")
    for tag, st, q in heat_states(g):
        s.load_state(st)
        got = s.heat_flow()
        scale = max(abs(v) for v in q.values())
        for z, v in q.items():
            assert abs(got[z] - v) <= 1e-12 * scale, (tag, z, got[z], v)


@pytest.mark.parametrize("name,builder", [("heat_flow_channel.npz", heat_channel), ("heat_flow_duct.npz", heat_duct)])
def test_heat_flow_balances_at_steady_state(name, builder):
    g = load(name)
    mesh, case = builder(C)
    s = SimpleSolver(mesh, case, "cuda")
    while not s.finished:
        s.step()
    assert s.converged and s.iteration == int(g["iterations"])
    q = s.heat_flow()
    scale = max(abs(v) for v in q.values())
    assert abs(sum(q.values())) <= 1e-6 * scale, q
    for z in q:
        assert abs(q[z] - float(g["it0_q_" + z])) <= 1e-8 * scale, z
    if name == "heat_flow_duct.npz":
        assert abs(q["bottom"] - 20000.0) < 1e-9 and abs(q["inlet"] + q["outlet"] + 40000.0) <= 1e-6 * scale
