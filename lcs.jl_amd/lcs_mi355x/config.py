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
class RestartInit:  # LCS.Restart (src/config.jl:235-238), kind "restart" (src/flows/config.jl:27)
    file: str
    kind: str = "restart"


@dataclass(frozen=True)
class FlowStatConfig:  # Flows.Stat (src/flows/config.jl:182-185)
    integral_length: bool = False


@dataclass(frozen=True)
class FlowConfig:  # src/flows/config.jl:199-206
    params: FlowParams
    init: object
    couple: HSMAC
    forcing: object
    stat: FlowStatConfig = field(default_factory=FlowStatConfig)


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
    velocity: str = "rest"  # "rest" | "zero-slip" | "restart"
    size: str = "const"  # "const" | "stokes" | "restart"
    diam_m: float = 0.0
    st_min: float = 1.0
    st_max: float = 1.0
    eps: float = 0.5
    # Particles.ParticleRestart (src/particles/config.jl:150-154): id, position,
    # velocity and size all read from restart_file, profile restart_profile (1-based)
    restart_file: Optional[str] = None
    restart_profile: int = 1


@dataclass(frozen=True)
class ParticleStatConfig:  # Particles.Stat, RDF, Spectrum (src/particles/config.jl:66-108)
    contact_tolerance: float = 0.01
    bin_size: int = 16
    min_radius: float = 1.25e-4
    search_cell_size: int = 2
    max_wavenumber: int = 768
    nsamples: int = 64


@dataclass(frozen=True)
class ParticleConfig:  # src/particles/config.jl:272-282
    population: Population
    params: ParticleParams
    grid_ratio: int
    drag: str
    init: ParticleInit = field(default_factory=ParticleInit)
    stat: ParticleStatConfig = field(default_factory=ParticleStatConfig)

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
    resume: bool = False  # set by lcsio.init only (src/LCSIO/config.jl:44-77)


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
    if kind == "restart":
        return RestartInit(str(p["file"]))
    raise ValueError(f"flow.init kind {kind!r} is not available on the mi355x backend (use 'spectrum' or 'restart')")


def _particle_restart(init):
    """The ParticleRestart variants of one profile's init, or None.  This
    backend reads id, position, velocity and size from one file together."""
    kinds = {k: _kind(init[k], k) for k in ("id", "position", "velocity", "size") if k in init}
    rs = {k: p for k, (kind, p) in kinds.items() if kind == "restart"}
    if not rs:
        return None
    if len(rs) != 4:
        raise ValueError("particle init: id, position, velocity and size must all be 'restart' or none of them")
    srcs = {(str(p["file"]), int(p["iprofile"])) for p in rs.values()}
    if len(srcs) != 1:
        raise ValueError("particle init: every 'restart' variant must name the same file and iprofile")
    return srcs.pop()


def _particle_stat(d):
    if not d:
        return ParticleStatConfig()
    r = d.get("rdf") or {}
    sp = d.get("spectrum") or {}
    dflt = ParticleStatConfig()
    return ParticleStatConfig(float(r.get("contact_tolerance", dflt.contact_tolerance)),
                              int(r.get("bin_size", dflt.bin_size)), float(r.get("min_radius", dflt.min_radius)),
                              int(r.get("search_cell_size", dflt.search_cell_size)),
                              int(sp.get("max_wavenumber", dflt.max_wavenumber)),
                              int(sp.get("nsamples", dflt.nsamples)))


def _particle(d, L0_m):
    pop = d["population"]
    prm = d["params"]
    init = d.get("init") or {}
    restart = _particle_restart(init)
    if restart is not None:
        drag, _ = _kind(d.get("drag", {"kind": "linear"}), "drag")
        return ParticleConfig(Population(int(pop["valid"]), float(pop.get("invalid_ratio", 0.0)),
                                         float(pop.get("exchange_ratio", 0.0))),
                              ParticleParams(float(prm["gravity"]), float(prm["density_ratio"])),
                              int(d["cell"]["grid_ratio"]), drag,
                              ParticleInit(velocity="restart", size="restart", restart_file=restart[0],
                                           restart_profile=restart[1]),
                              _particle_stat(d.get("stat")))
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
                          int(d["cell"]["grid_ratio"]), drag, pinit, _particle_stat(d.get("stat")))


def from_dict(x):
    """PolySerde.deserialize(LCS.Config, x) for the mi355x backend."""
    backend, _ = _kind(x.get("backend", {"kind": "mi355x"}), "backend")
    if backend not in BACKENDS:
        raise ValueError(f"backend kind {backend!r} is not served by this package (use 'mi355x')")
    mode, _ = _kind(x.get("mode", {"kind": "flow-particle"}), "mode")
    g = x["grid"]
    f = x["flow"]
    fp = FlowParams(float(f["params"]["U0_m_s"]), float(f["params"]["L0_m"]), float(f["params"]["NU_m2_s"]))
    fstat = FlowStatConfig(bool((f.get("stat") or {}).get("integral_length", False)))
    flow = FlowConfig(fp, _flow_init(f["init"]), _couple(f["couple"]), _forcing(f["forcing"]), fstat)
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
                  outdir=str(x.get("outdir", "output")), resume=bool(x.get("resume", False)))


def _forcing_dict(f):
    if isinstance(f, NoForcing):
        return {"kind": "no"}
    if isinstance(f, ConstantPowerLF):
        return {"kind": f.kind, "params": {"power": f.power}}
    return {"kind": f.kind, "params": {"energy": f.energy, "kmin": f.kmin, "kmax": f.kmax,
                                       "coarse_size": f.coarse_size}}


def _particle_dict(p: ParticleConfig):
    i = p.init
    if i.restart_file is not None:
        r = {"kind": "restart", "params": {"file": i.restart_file, "iprofile": i.restart_profile}}
        init = {"id": r, "position": r, "velocity": r, "size": r}
    else:
        if i.size == "stokes":
            size = {"kind": "stokes", "params": {"st_min": i.st_min, "st_max": i.st_max, "eps": i.eps}}
        else:
            size = {"kind": "const", "params": {"diam_m": i.diam_m}}
        init = {"id": {"kind": "generate"}, "position": {"kind": "random", "params": {"seed": i.seed}},
                "velocity": {"kind": i.velocity}, "size": size}
    s = p.stat
    return {"population": {"valid": p.population.valid, "invalid_ratio": p.population.invalid_ratio,
                           "exchange_ratio": p.population.exchange_ratio},
            "params": {"gravity": p.params.gravity, "density_ratio": p.params.density_ratio},
            "cell": {"grid_ratio": p.grid_ratio},
            "stat": {"rdf": {"contact_tolerance": s.contact_tolerance, "bin_size": s.bin_size,
                             "min_radius": s.min_radius, "search_cell_size": s.search_cell_size},
                     "spectrum": {"max_wavenumber": s.max_wavenumber, "nsamples": s.nsamples}},
            "drag": {"kind": p.drag}, "init": init}


def to_dict(c: Config):
    """PolySerde.serialize(LCS.Config, c) for the mi355x backend: a mapping
    that from_dict reads back to an equal Config.  ``resume`` is left out, as
    LCSIO.init leaves it out of the saved config (src/LCSIO/config.jl:66-70)."""
    f = c.flow
    if isinstance(f.init, RestartInit):
        finit = {"kind": "restart", "params": {"file": f.init.file}}
    else:
        finit = {"kind": "spectrum", "params": {"seed": f.init.seed, "k0": f.init.k0, "uprime": f.init.uprime}}
    h = f.couple
    d = {"outdir": c.outdir, "mode": {"kind": c.mode}, "backend": {"kind": c.backend}, "dtype": c.dtype,
         "grid": {"dims": list(c.grid.dims), "halo_size": c.grid.halo_size},
         "timestep": {"cfl": c.timestep.cfl, "dtmax": c.timestep.dtmax},
         "simulate": {"last_step": c.simulate.last_step, "interval_stat": c.simulate.interval_stat,
                      "interval_restart": c.simulate.interval_restart, "interval_gc": c.simulate.interval_gc,
                      "num_restart_files": c.simulate.num_restart_files},
         "flow": {"params": {"U0_m_s": f.params.U0_m_s, "L0_m": f.params.L0_m, "NU_m2_s": f.params.NU_m2_s},
                  "init": finit,
                  "couple": {"kind": "hsmac", "params": {"time_blocking": h.time_blocking, "overlap": h.overlap,
                                                         "omega": h.omega, "epsp": h.epsp, "itrmin": h.itrmin,
                                                         "itrmax": h.itrmax, "coloring": {"kind": h.coloring}}},
                  "forcing": _forcing_dict(f.forcing),
                  "stat": {"integral_length": f.stat.integral_length}}}
    if c.particles:
        d["particles"] = [_particle_dict(p) for p in c.particles]
    return d


def readconfig(path, patch=None):
    """LCSIO.readconfig(file; patch) (src/LCSIO/config.jl:6-31): a YAML file
    that sets ``resume`` is refused, since resuming is chosen by the
    LCS_RESUME environment variable."""
    with open(path) as fh:
        d = yaml.safe_load(fh)
    if "resume" in d:
        raise ValueError("resume must not be set in configuration, resume mode is controlled by LCS_RESUME "
                         "environment variable")
    d["resume"] = False
    c = from_dict(d)
    return patch(c) if patch is not None else c
