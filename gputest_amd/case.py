"""Case description of the SIMPLEC hot path: the reference's ``UnstructuredCase`` and
``BoundarySpec`` (``flowtrace/unstructured/case.py:31-177``), restricted to what the hot
path uses (laminar or k-omega SST, optional energy, steady), plus one addition:
``BoundarySpec.profile``, a per-face fixed velocity for walls and inlets (the workloads'
parabolic inlet)."""
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
TURBULENCE_ALIASES = {"sst": "k-omega-sst", "kw-sst": "k-omega-sst", "k-omega": "k-omega-sst", "none": "laminar"}
TURBULENCE_MODELS = ("laminar", "k-omega-sst")


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
    # turbulence entering at this zone (None: the case defaults) (case.py:53-58)
    turb_intensity: Optional[float] = None
    turb_viscosity_ratio: Optional[float] = None
    k: Optional[float] = None
    omega: Optional[float] = None

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
    # pressure AMG (new; the reference fixes them, amg.py:80-112): at most amg_levels levels
    # counting the coarsest (None: the reference's 20, coarsening to <= 300 unknowns); when that
    # count leaves the coarsest system above 300 unknowns it is solved by amg_coarse_sweeps
    # weighted-Jacobi sweeps instead of a dense inverse
    amg_levels: Optional[int] = None
    amg_coarse_sweeps: int = 20
    periodic: List[List[str]] = field(default_factory=list)
    periodic_pressure_gradient: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    # turbulence (case.py:99-110): laminar | k-omega-sst
    turbulence: str = "laminar"
    turb_intensity: float = 0.05
    turb_viscosity_ratio: float = 10.0
    turb_reference_velocity: Optional[float] = None
    turb_init: Optional[Dict[str, float]] = None
    relax_k: Optional[float] = None
    relax_w: Optional[float] = None
    turb_scheme: str = "upwind"

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
        if self.amg_levels is not None and int(self.amg_levels) < 2:
            raise ValueError("amg_levels counts the coarsest level: at least 2")
        if int(self.amg_coarse_sweeps) < 2:
            raise ValueError("amg_coarse_sweeps must be at least 2")
        self.periodic = [list(p) for p in self.periodic]
        t = str(self.turbulence).lower()
        self.turbulence = TURBULENCE_ALIASES.get(t, t)
        if self.turbulence not in TURBULENCE_MODELS:
            raise ValueError(f"turbulence must be one of {TURBULENCE_MODELS}")
        if self.turbulence != "laminar" and self.turb_scheme != "upwind":
            raise ValueError("the SST transports are built for turb_scheme='upwind' (the reference's default)")

    @property
    def turbulent(self) -> bool:
        return self.turbulence != "laminar"

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
