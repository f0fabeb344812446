"""Batched solves on the GPU: per-case parity with the oracle at convergence, graph replays
identical to eager outer iterations, the mixed-precision graph batch, and the autotuner."""
import dataclasses

import numpy as np
import pytest
import torch

import oracle.simplec_oracle as O
from gputest_amd.batch import BatchSolver, CavitySpec, StepSpec, autotune, autotune_grid_cap

pytestmark = pytest.mark.gpu
NU = 0.01
SPECS = [CavitySpec(12, False, 100.0), CavitySpec(9, True, 250.0), CavitySpec(14, False, 400.0)]
# cavities and backward-facing steps in one launch (closed and open cases side by side)
MIXED = [CavitySpec(10, False, 100.0), StepSpec(5, False, 100.0, 10.0), StepSpec(4, True, 100.0, 10.0),
         CavitySpec(8, True, 200.0)]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _oracle(spec, tol):
    bnd = {k: O.Boundary(**{f.name: getattr(b, f.name) for f in dataclasses.fields(O.Boundary)})
           for k, b in spec.boundaries("", NU).items()}
    s = O.Simplec(spec.mesh(), O.Case(boundaries=bnd, viscosity=NU, max_iter=20000, tol=tol))
    while not s.finished:
        s.step()
    assert s.converged
    return s


def _compare(batch, tol_err, specs=SPECS):
    for k, spec in enumerate(specs):
        ref = _oracle(spec, 1e-10)
        f = batch.fields(k)
        p = ref.p - ref.p.mean() if spec.closed else ref.p
        e = {"u": rel_l2(f["x-velocity"], ref.U[:, 0]), "v": rel_l2(f["y-velocity"], ref.U[:, 1]),
             "p": rel_l2(f["pressure"], p)}
        print(k, spec, e)
        assert max(e.values()) <= tol_err, (k, e)


def test_fp64_batch_matches_each_case_at_convergence():
    b = BatchSolver(SPECS, NU, precision="fp64", use_graph=False, max_iter=20000, tol=1e-9)
    while not b.solver.finished:
        b.step()
    assert b.solver.converged
    _compare(b, 1e-6)


def test_fp64_batch_of_cavities_and_steps_matches_each_case_at_convergence():
    b = BatchSolver(MIXED, NU, precision="fp64", use_graph=True, max_iter=20000, tol=1e-9)
    while not b.solver.finished:
        b.step()
    assert b.solver.converged
    print("cavities and steps: iterations", b.solver.iteration, "per case", b.solver.case_converged_at)
    _compare(b, 1e-6, MIXED)


def _replay_against_eager(precision, steps, cap=None, specs=SPECS):
    eager = BatchSolver(specs, NU, precision=precision, use_graph=False)
    graph = BatchSolver(specs, NU, precision=precision, use_graph=True)
    if cap is not None:
        graph.solver.graph_cap = cap
    for it in range(steps):  # spans hierarchy rebuilds and re-recordings
        eager.step()
        graph.step()
        # the solves' reports of a replay give the eager iteration counts and residuals
        assert graph.solver.last_linear == eager.solver.last_linear, (it, graph.solver.last_linear,
                                                                      eager.solver.last_linear)
        assert graph.solver.residuals()["raw"] == eager.solver.residuals()["raw"], it
    for name in ("U", "p", "F", "Fb"):
        a, b = getattr(eager.solver, name), getattr(graph.solver, name)
        assert torch.equal(a, b), name
    return graph.solver


@pytest.mark.parametrize("precision,cases", [("fp64", "cavities"), ("mixed", "cavities"), ("mixed", "mixed")])
def test_graph_replays_equal_eager_iterations(precision, cases):
    """Capped recordings (each solve at most 1.5 m + 8 iterations, m its recent count); a
    replay whose solve would have needed more, or (mixed) a second FP32 correction, is redone
    eagerly, so every iteration equals the eager one bit for bit."""
    g = _replay_against_eager(precision, 25, specs=SPECS if cases == "cavities" else MIXED)
    print(precision, "recordings", g.graph_captures, "redone", g.graph_redos)
    assert g.graph_captures >= 2
    assert g.graph_redos <= 3


def test_replays_past_their_caps_are_redone_eagerly():
    """Caps of one iteration per solve: the pressure solve never finishes within its cap, so
    every replay is rolled back to its saved start state and redone eagerly."""
    g = _replay_against_eager("fp64", 14, cap=(0.0, 1))
    print("recordings", g.graph_captures, "redone", g.graph_redos)
    assert g.graph_redos >= 10 and g.graph_captures == g.graph_redos


@pytest.mark.parametrize("cases,budget", [("cavities", 800), ("cavities and steps", 6000)])
def test_mixed_precision_graph_batch_reaches_each_case(cases, budget):
    """The mixed batch's continuity residual levels off near 1e-7 (scaled), so it runs a fixed
    budget of outer iterations (past the slowest case's FP64 convergence); each case's fields
    are then within 1e-6 of its converged FP64 oracle."""
    specs = SPECS if cases == "cavities" else MIXED
    b = BatchSolver(specs, NU, precision="mixed", use_graph=True, max_iter=budget, tol=1e-9)
    while not b.solver.finished:
        b.step()
    print("mixed batch residuals", b.solver.residuals()["scaled"], "iterations", b.solver.iteration,
          "redone", b.solver.graph_redos)
    _compare(b, 1e-6, specs)


def test_autotuner_picks_a_candidate_and_caches_it(tmp_path):
    b = BatchSolver(SPECS, NU, precision="mixed", use_graph=True)
    cache = str(tmp_path / "tune.json")
    cap = autotune_grid_cap(b.solver, candidates=(256, 1024), replays=2, cache=cache)
    assert cap in (256, 1024)
    assert autotune_grid_cap(b.solver, candidates=(256, 1024), replays=2, cache=cache) == cap
    from gputest_amd import _native as nat

    nat.lib().ftmi_set_grid_cap(2048)


def test_autotuner_over_cap_and_threads_per_row(tmp_path):
    """The joint search picks one of its candidates, applies it (the hierarchies' threads per
    row re-picked), caches it, and the tuned batch still converges each case."""
    from gputest_amd import _native as nat

    L = nat.lib()
    b = BatchSolver(SPECS, NU, precision="mixed", use_graph=True, max_iter=3000, tol=1e-6)
    cache = str(tmp_path / "tune.json")
    try:
        t = autotune(b.solver, caps=(256, 1024), tpr_biases=(0, 1), replays=2, cache=cache)
        assert t["cap"] in (256, 1024) and t["tpr_bias"] in (0, 1)
        assert len(t["ms_per_iteration"]) == 4
        assert autotune(b.solver, caps=(256, 1024), tpr_biases=(0, 1), replays=2, cache=cache) == t
        assert L.ftmi_set_tpr_bias(t["tpr_bias"]) == t["tpr_bias"]
        while not b.solver.finished:
            b.step()
        print("tuned", t, "iterations", b.solver.iteration, "redone", b.solver.graph_redos)
        assert b.solver.converged
    finally:
        L.ftmi_set_grid_cap(2048)
        L.ftmi_set_tpr_bias(0)
