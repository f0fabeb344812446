"""Case description of the SIMPLEC hot path: the reference's ``UnstructuredCase`` and
``BoundarySpec`` (``flowtrace/unstructured/case.py:31-177``), restricted to what the hot
path uses (laminar, optional energy, steady), plus one addition: ``BoundarySpec.profile``,
a per-face fixed velocity for walls and inlets (the workloads' parabolic inlet)."""
from __future__ import annotations

from dataclasses import asdict, dataclass, field
from typing import Callable, Dict, List, Optional

FLUENT_TO_KIND = {
    "wall": "wall", "velocity-inlet": "inlet", "mass-flow-inlet": "inlet", "pressure-inlet": "pressure",
    "pressure-outlet": "outlet", "outflow": "outlet", "pressure-far-field": "outlet", "outlet-vent": "outlet",
    "exhaust-fan": "outlet", "inlet-vent": "pressure", "intake-fan": "pressure", "symmetry": "symmetry",
    "axis": "symmetry", "periodic": "periodic", "periodic-shadow": "periodic",
}
KINDS = ("wall", "inlet", "outlet", "pressure", "symmetry", "periodic")


@dataclass
class BoundarySpec:
    """Condition of one face zone.

    wall: ``velocity`` (default at rest); ``thermal`` adiabatic | temperature | flux with ``value``.
    inlet: ``velocity`` vector or ``normal_speed`` (> 0 into the domain); T = ``value``.
    outlet: static ``pressure``; zero-gradient velocity and temperature (``value`` on backflow).
    pressure: opening at ``pressure``; T = ``value``. symmetry: slip.
    ``profile``: optional callable (face centres (n, dim) -> (n, dim)) giving a per-face fixed
    velocity on walls and inlets; it overrides ``velocity``.
    """
    kind: str = "wall"
    velocity: Optional[List[float]] = None
    normal_speed: Optional[float] = None
    pressure: float = 0.0
    thermal: str = "adiabatic"
    value: float = 0.0
    profile: Optional[Callable] = None

    def __post_init__(self):
        if self.kind not in KINDS:
            raise ValueError(f"unknown boundary kind {self.kind!r}")
        if self.thermal not in ("adiabatic", "temperature", "flux"):
            raise ValueError(f"unknown thermal condition {self.thermal!r}")


@dataclass
class UnstructuredCase:
    """Case parameters; defaults are the reference's (``case.py:69-177``)."""
    name: str = "unstructured case"
    boundaries: Dict[str, BoundarySpec] = field(default_factory=dict)
    density: float = 1.0
    viscosity: float = 0.01
    energy: bool = False
    conductivity: float = 0.01
    cp: float = 1.0
    body_force: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    algorithm: str = "SIMPLEC"
    relax_u: float = 0.7
    relax_p: Optional[float] = None
    relax_T: float = 1.0
    scheme: str = "second-order"
    pressure_gradient: str = "green-gauss"
    pressure_tol: float = 0.1
    max_iter: int = 2000
    tol: float = 1e-5
    init_velocity: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    init_T: float = 0.0
    precision: str = "fp64"
    periodic: List[List[str]] = field(default_factory=list)
    periodic_pressure_gradient: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])

    def __post_init__(self):
        self.boundaries = {k: (BoundarySpec(**v) if isinstance(v, dict) else v) for k, v in self.boundaries.items()}
        self.algorithm = self.algorithm.upper()
        if self.algorithm not in ("SIMPLE", "SIMPLEC"):
            raise ValueError("algorithm must be SIMPLE or SIMPLEC")
        if self.relax_p is None and self.algorithm == "SIMPLE":
            self.relax_p = 0.3
        if self.pressure_gradient not in ("green-gauss", "least-squares"):
            raise ValueError("pressure_gradient must be 'green-gauss' or 'least-squares'")
        if self.scheme not in ("upwind", "second-order"):
            raise ValueError("scheme must be 'upwind' or 'second-order'")
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed' (FP32 transport assembly and Krylov solves, "
                             "FP32 AMG-PCG refined in FP64)")
        self.periodic = [list(p) for p in self.periodic]

    def to_dict(self) -> dict:
        d = asdict(self)
        for b in d["boundaries"].values():
            b.pop("profile", None)
        return d


def default_boundaries(mesh) -> Dict[str, BoundarySpec]:
    """One spec per boundary zone from its Fluent type (``case.py:180-188``)."""
    out = {}
    for z in mesh.boundary_zones():
        kind = FLUENT_TO_KIND.get(z.type, "wall")
        out[z.name] = BoundarySpec(kind=kind, normal_speed=1.0 if kind == "inlet" else None)
    return out


# short names used by the tests for both this package and the oracle
Case = UnstructuredCase
Boundary = BoundarySpec
