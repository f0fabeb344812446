"""Checkpoint files and the batch planner on the CPU (``flowtrace/runtime/checkpoint.py``,
``flowtrace/runtime/batch.py``): the reference's .npz + ``__meta__`` format written by the
background writer and read back with nothing unpickled, pruning to the newest files, and the
grouping of a job list into batched launches (errors reported per job)."""
import json
import os

import numpy as np
import pytest

import oracle.simplec_oracle as O
from gputest_amd.batch import BatchJob, plan_batch
from gputest_amd.checkpoint import CheckpointWriter, load_checkpoint
from gputest_amd.fluent import write_fluent_mesh
from gputest_amd.mesh import cartesian_mesh_2d

REF_KEYS = {"U", "p", "F", "Fb", "iteration", "time", "time_step", "res_continuity_norm"}


def _oracle_state(steps=3):
    B = O.Boundary
    s = O.Simplec(cartesian_mesh_2d(8, 6), O.Case(boundaries={"top": B("wall", velocity=[1.0, 0.0])}, viscosity=0.05))
    for _ in range(steps):
        s.step()
    return s


def test_checkpoint_files_have_the_reference_layout(tmp_path):
    s = _oracle_state()
    st = s.state_dict()
    w = CheckpointWriter(str(tmp_path), keep=3)
    for k in range(5):
        st = dict(st, iteration=k)
        w.submit(f"it{k:04d}", st, {"iteration": k, "case": "cavity8x6"})
    w.close()
    assert w.last_error is None and len(w.written) == 5
    files = sorted(os.listdir(tmp_path))
    assert files == ["it0002.npz", "it0003.npz", "it0004.npz"]  # the newest three
    # plain NumPy reads it without pickles; __meta__ is UTF-8 JSON bytes
    with np.load(tmp_path / "it0004.npz", allow_pickle=False) as z:
        assert REF_KEYS <= set(z.files) and "__meta__" in z.files
        assert json.loads(bytes(z["__meta__"]).decode()) == {"iteration": 4, "case": "cavity8x6"}
        assert z["U"].shape == (s.nc, 2)
    state, meta = load_checkpoint(str(tmp_path / "it0004.npz"))
    assert meta["case"] == "cavity8x6" and state["iteration"] == 4 and isinstance(state["iteration"], int)
    for k in ("U", "p", "F", "Fb"):
        np.testing.assert_array_equal(state[k], st[k])
    # the file holds everything load_state uses: fresh solvers restored from the file and from
    # the in-memory state continue identically
    sa, sb = _oracle_state(0), _oracle_state(0)
    sa.load_state(st)
    sb.load_state(state)
    for _ in range(3):
        sa.step()
        sb.step()
    np.testing.assert_array_equal(sa.U, sb.U)
    np.testing.assert_array_equal(sa.p, sb.p)


def test_checkpoint_writer_reports_errors_without_raising(tmp_path):
    w = CheckpointWriter(str(tmp_path))
    w.submit("bad", {"U": object()}, {})  # not an array: the writer records the error
    w.close()
    assert w.last_error is not None and not w.written


def _msh(tmp_path, name, n):
    path = str(tmp_path / f"{name}.msh")
    write_fluent_mesh(cartesian_mesh_2d(n, n), path)
    return path


def test_plan_groups_compatible_jobs_largest_first(tmp_path):
    a, b, c = _msh(tmp_path, "a", 8), _msh(tmp_path, "b", 12), _msh(tmp_path, "c", 10)
    jobs = [BatchJob("a", a, {"viscosity": 0.01}), BatchJob("b", b, {"viscosity": 0.01, "boundaries": {}}),
            BatchJob("c", c, {"viscosity": 0.02}), BatchJob("d", c, {"viscosity": 0.01}),
            BatchJob("bad", str(tmp_path / "missing.msh")), BatchJob("dens", a, solver="density-based")]
    plan = plan_batch(jobs)
    plan.pop("_meshes")
    assert sorted(plan["groups"], key=len) == [["c"], ["b", "d", "a"]]
    assert {r["name"]: r["cells"] for r in plan["jobs"]} == {"a": 64, "b": 144, "c": 100, "d": 100}
    assert set(plan["errors"]) == {"bad", "dens"}
    # a launch holds at most max_cells cells
    small = plan_batch(jobs[:2] + jobs[3:4], max_cells=200)
    assert small["groups"] == [["b"], ["d", "a"]]
    with pytest.raises(ValueError):
        plan_batch([BatchJob("x", a), BatchJob("x", b)])


def test_turbulent_jobs_run_alone(tmp_path):
    """A k-omega SST job gets a launch of its own (its wall distance must see only its own walls),
    even beside identical settings."""
    a, b = _msh(tmp_path, "a", 8), _msh(tmp_path, "b", 12)
    sst = {"viscosity": 0.01, "turbulence": "k-omega-sst"}
    jobs = [BatchJob("a", a, dict(sst)), BatchJob("b", b, dict(sst)), BatchJob("c", a, {"viscosity": 0.01}),
            BatchJob("d", b, {"viscosity": 0.01})]
    plan = plan_batch(jobs)
    assert sorted(plan["groups"]) == [["a"], ["b"], ["d", "c"]]
