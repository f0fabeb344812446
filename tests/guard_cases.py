"""The stability-guard and heat-flow cases of ``tests/golden/make_guard_golden.py``, built with
this repository's generators for the oracle (``mod`` = ``oracle.simplec_oracle``) and for the
device solver (``mod`` = ``gputest_amd.case``)."""
import os
import re

import numpy as np

from gputest_amd.mesh import cartesian_mesh_2d, duct_mesh
from tests.golden_cases import duct_case
from tests.reference_pins import heated_channel, make_case

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def sheared_mesh():
    """The reference's guard mesh (``tests/test_planner.py:93-113``): 24 x 24 quads sheared by
    45 degrees on every face."""
    m = cartesian_mesh_2d(24, 24)
    m.nodes[:, 0] = m.nodes[:, 0] + m.nodes[:, 1]
    return m


def sheared_case(mod, relax_p):
    return make_case(mod, {"top": dict(kind="wall", velocity=[1.0, 0.0])}, viscosity=0.1, max_iter=3000, tol=1e-6,
                     relax_p=relax_p)


def run_until_nonfinite(s, continuity):
    """``test_planner.py:99-103``: step until finished or the raw continuity is not finite."""
    with np.errstate(all="ignore"):
        while not s.finished and np.isfinite(continuity(s)):
            s.step()
    return s


def raw_continuity(s):
    return s.residuals()["raw"].get("continuity", 0.0)


def events_match(events, g, cont_rel=1e-8):
    """The guard events against the fixture: iterations and relaxations exactly, the tripping
    continuity to ``cont_rel``."""
    assert [e["detected_at"] for e in events] == g["ev_detected_at"].tolist()
    assert [e["restored_to"] for e in events] == g["ev_restored_to"].tolist()
    assert [e["relax_p"] for e in events] == g["ev_relax_p"].tolist()
    for e, c in zip(events, g["ev_continuity"]):
        assert abs(e["continuity"] - c) <= cont_rel * abs(c), (e, c)


def heat_channel(mod):
    return heated_channel(mod)


def heat_duct(mod):
    case = duct_case(mod, pressure_tol=1e-8, max_iter=400)
    case.tol = 1e-8
    return duct_mesh(24, 8), case


def heat_states(g):
    """(tag, state dict, zone heat flows) of every state recorded in a heat-flow fixture."""
    out = []
    for tag in sorted({k.split("_")[0] for k in g.files if re.match(r"it\d+_", k)}):
        st = {k: g[f"{tag}_{k}"] for k in ("U", "p", "F", "Fb", "T")}
        st["iteration"] = int(tag[2:])
        q = {k[len(tag) + 3:]: float(g[k]) for k in g.files if k.startswith(tag + "_q_")}
        out.append((tag, st, q))
    return out
