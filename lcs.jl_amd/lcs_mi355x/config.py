"""Typed configuration mirroring the reference's YAML surface.

The reference deserialises ``.lcs-yaml`` into ``LCS.Config`` with PolySerde
``kind``/``params`` variants (src/config.jl:173-233, src/flows/config.jl,
src/particles/config.jl).  The same YAML loads here; ``backend: {kind: mi355x}``
selects this backend (the reference registers backends the same way,
ext/AMDGPUExt.jl:8-14).  New, builder-defined variants:

* ``flow.init``: ``{kind: spectrum, params: {seed, k0, uprime}}``
* ``particles[].init.size``: ``{kind: stokes, params: {st_min, st_max, eps, seed}}``
* ``particles[].init.velocity``: ``{kind: zero-slip}``
* ``particles[].init.position``: ``{kind: random, params: {seed}}`` (project RNG)
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import yaml

BACKENDS = ("mi355x",)


def _kind(d, what):
    if not isinstance(d, dict) or "kind" not in d:
        raise ValueError(f"{what}: expected a mapping with 'kind'")
    return d["kind"], (d.get("params") or {})


@dataclass(frozen=True)
class Grid:  # src/config.jl:56-60
    dims: Tuple[int, int, int]
    halo_size: int = 3


@dataclass(frozen=True)
class TimeStep:  # src/config.jl:127-131
    cfl: float
    dtmax: float


@dataclass(frozen=True)
class SimulateConfig:  # src/config.jl:144-151
    last_step: int
    interval_stat: int = 0
    interval_restart: int = 0
    interval_gc: int = 0
    num_restart_files: int = 1


@dataclass(frozen=True)
class FlowParams:  # src/flows/config.jl:11-18
    U0_m_s: float
    L0_m: float
    NU_m2_s: float

    @property
    def Re(self):
        return self.U0_m_s * self.L0_m / self.NU_m2_s


@dataclass(frozen=True)
class HSMAC:  # src/flows/config.jl:88-98
    time_blocking: bool
    overlap: bool
    coloring: str
    omega: float
    epsp: float
    itrmin: int
    itrmax: int

    @property
    def coloring_id(self):
        from ._lib import COLORING
        if self.coloring not in COLORING:
            raise ValueError(f"unsupported coloring {self.coloring!r}")
        return COLORING[self.coloring]


@dataclass(frozen=True)
class NoForcing:
    kind: str = "no"


@dataclass(frozen=True)
class ConstantPowerLF:  # src/flows/config.jl:132-136
    power: float
    kind: str = "linear-power"


@dataclass(frozen=True)
class EnergyPreserveRCF:  # src/flows/config.jl:162-175
    energy: float
    kmin: int
    kmax: int
    coarse_size: int
    kind: str = "rc-energy"

    def __post_init__(self):
        if self.kmin < 1:
            raise ValueError(f"kmin must be at least 1, got {self.kmin}")
        if self.kmax < self.kmin:
            raise ValueError(f"kmax must be at least kmin, got kmax={self.kmax} kmin={self.kmin}")


@dataclass(frozen=True)
class SpectrumInit:
    seed: int
    k0: float = 4.0
    uprime: float = 1.0
    kind: str = "spectrum"


@dataclass(frozen=True)
class FlowConfig:  # src/flows/config.jl:199-206
    params: FlowParams
    init: object
    couple: HSMAC
    forcing: object


@dataclass(frozen=True)
class Population:  # src/particles/config.jl:11-24
    valid: int
    invalid_ratio: float = 0.0
    exchange_ratio: float = 0.0

    @property
    def ntotal(self):
        import math
        return int(math.ceil(self.valid * (1 + self.invalid_ratio)))


@dataclass(frozen=True)
class ParticleParams:  # src/particles/config.jl:35-41
    gravity: float
    density_ratio: float


@dataclass(frozen=True)
class ParticleInit:
    seed: int = 0
    velocity: str = "rest"  # "rest" | "zero-slip"
    size: str = "const"  # "const" | "stokes"
    diam_m: float = 0.0
    st_min: float = 1.0
    st_max: float = 1.0
    eps: float = 0.5


@dataclass(frozen=True)
class ParticleConfig:  # src/particles/config.jl:272-282
    population: Population
    params: ParticleParams
    grid_ratio: int
    drag: str
    init: ParticleInit = field(default_factory=ParticleInit)

    @property
    def drag_id(self):
        from ._lib import DRAG
        return DRAG[self.drag]


@dataclass(frozen=True)
class Config:  # src/config.jl:173-194
    grid: Grid
    timestep: TimeStep
    simulate: SimulateConfig
    flow: FlowConfig
    particles: Tuple[ParticleConfig, ...] = ()
    mode: str = "flow-particle"
    backend: str = "mi355x"
    dtype: str = "f64"
    outdir: str = "output"


def _forcing(d):
    kind, p = _kind(d, "flow.forcing")
    if kind == "no":
        return NoForcing()
    if kind == "linear-power":
        return ConstantPowerLF(float(p["power"]))
    if kind == "rc-energy":
        return EnergyPreserveRCF(float(p["energy"]), int(p["kmin"]), int(p["kmax"]), int(p["coarse_size"]))
    raise ValueError(f"unknown forcing kind {kind!r}")


def _couple(d):
    kind, p = _kind(d, "flow.couple")
    if kind != "hsmac":
        raise ValueError(f"unknown couple kind {kind!r}")
    ck, _ = _kind(p["coloring"], "coloring")
    return HSMAC(bool(p["time_blocking"]), bool(p["overlap"]), ck, float(p["omega"]), float(p["epsp"]),
                 int(p["itrmin"]), int(p["itrmax"]))


def _flow_init(d):
    kind, p = _kind(d, "flow.init")
    if kind == "spectrum":
        return SpectrumInit(int(p.get("seed", 0)), float(p.get("k0", 4.0)), float(p.get("uprime", 1.0)))
    raise ValueError(f"flow.init kind {kind!r} is not available on the mi355x backend (use 'spectrum')")


def _particle(d, L0_m):
    pop = d["population"]
    prm = d["params"]
    init = d.get("init") or {}
    seed = 0
    pos_kind, pos_p = _kind(init.get("position", {"kind": "random"}), "position")
    if pos_kind != "random":
        raise ValueError(f"unsupported particle position init {pos_kind!r}")
    seed = int(pos_p.get("seed", 0))
    vel_kind, _ = _kind(init.get("velocity", {"kind": "rest"}), "velocity")
    size_kind, size_p = _kind(init.get("size", {"kind": "const", "params": {"diam_m": 0.0}}), "size")
    pinit = ParticleInit(seed=seed, velocity=vel_kind, size=size_kind, diam_m=float(size_p.get("diam_m", 0.0)),
                         st_min=float(size_p.get("st_min", 1.0)), st_max=float(size_p.get("st_max", 1.0)),
                         eps=float(size_p.get("eps", 0.5)))
    drag, _ = _kind(d.get("drag", {"kind": "linear"}), "drag")
    return ParticleConfig(Population(int(pop["valid"]), float(pop.get("invalid_ratio", 0.0)),
                                     float(pop.get("exchange_ratio", 0.0))),
                          ParticleParams(float(prm["gravity"]), float(prm["density_ratio"])),
                          int(d["cell"]["grid_ratio"]), drag, pinit)


def from_dict(x):
    """PolySerde.deserialize(LCS.Config, x) for the mi355x backend."""
    backend, _ = _kind(x.get("backend", {"kind": "mi355x"}), "backend")
    if backend not in BACKENDS:
        raise ValueError(f"backend kind {backend!r} is not served by this package (use 'mi355x')")
    mode, _ = _kind(x.get("mode", {"kind": "flow-particle"}), "mode")
    g = x["grid"]
    f = x["flow"]
    fp = FlowParams(float(f["params"]["U0_m_s"]), float(f["params"]["L0_m"]), float(f["params"]["NU_m2_s"]))
    flow = FlowConfig(fp, _flow_init(f["init"]), _couple(f["couple"]), _forcing(f["forcing"]))
    parts = ()
    if mode == "flow-particle" and x.get("particles"):
        parts = tuple(_particle(p, fp.L0_m) for p in x["particles"])
    s = x.get("simulate", {"last_step": 0})
    return Config(grid=Grid(tuple(int(v) for v in g["dims"]), int(g.get("halo_size", 3))),
                  timestep=TimeStep(float(x["timestep"]["cfl"]), float(x["timestep"]["dtmax"])),
                  simulate=SimulateConfig(int(s["last_step"]), int(s.get("interval_stat", 0)),
                                          int(s.get("interval_restart", 0)), int(s.get("interval_gc", 0)),
                                          int(s.get("num_restart_files", 1))),
                  flow=flow, particles=parts, mode=mode, backend=backend, dtype=str(x.get("dtype", "f64")),
                  outdir=str(x.get("outdir", "output")))


def readconfig(path):
    """LCSIO.readconfig (src/LCSIO/config.jl:6-31) without the patch expression."""
    with open(path) as fh:
        return from_dict(yaml.safe_load(fh))
