"""Synthetic workloads of the benchmark and of the workload-sized tests.

Every input is generated from a seed by the package's own initialisers
(seeded spectrum, uniform particle positions, Stokes numbers, zero slip), so
the benchmark, the tests and the oracle start from identical states.

* ``single_box("c1" | "c2" | "c3" | "c6")``: one periodic box as a ``Config``;
* ``ensemble_configs(nbox)``: the batched ensemble of mixed 32^3-128^3 boxes;
* ``serve_requests(nreq)``: the request stream of the served workload, as
  keyword arguments of ``Server.submit``, and ``request_config`` giving the
  ``Config`` of the box a request is served in.
"""
import numpy as np

from . import config as C

# the HSMAC parameters of every workload: cfgs/N128.lcs-yaml:36-46
HSMAC_DEFAULT = C.HSMAC(True, True, "red-black-fast", 1.7, 1e-3, 6, 300)


def _particles(npart, seed, st, drag="linear", invalid_ratio=0.0):
    if not npart:
        return ()
    return (C.ParticleConfig(C.Population(npart, invalid_ratio), C.ParticleParams(0.0, 843.0), 4, drag,
                             C.ParticleInit(seed=seed, velocity="zero-slip", size="stokes", st_min=st[0],
                                            st_max=st[1], eps=0.5)),)


def single_box(name, n_override=None, particles_override=None, dtype=None):
    """One periodic 2 pi box: returns (Config, description).

    c1: 64^3 FP64, linear-power P = 0.5, Re = 53.3, 1e4 particles St = 1.
    c2: 256^3 FP32, rc-energy (energy 0.649, kmin 1, kmax 2, coarse 8),
        Re = 360, 2e6 particles St log-uniform [0.1, 10].
    c3: 1024^3 FP32, linear-power P = 0.5, Re = 1333, 6.4e7 particles St = 1.
    c6: 750^3 FP64, rc-energy (energy 0.66, kmin 1, kmax 2, coarse 10),
        Re = 2220, 325^3 particles St = 1 with nonlinear drag and half the
        slots spare: the single-device case of the reference's weak-scaling
        runs (data/gpu-weak.csv:2 with the settings of cfgs/N1500.lcs-yaml;
        coarse 12 there does not divide 750, 10 is the nearest divisor).
    All start from the seeded spectrum E(k) ~ k^4 exp(-2 (k/4)^2), u' = 1."""
    drag, spare = "linear", 0.0
    if name == "c6":
        n, npart, dt, Re, forcing, st = 750, 325**3, "f64", 2220.0, C.EnergyPreserveRCF(0.66, 1, 2, 10), (1.0, 1.0)
        drag, spare = "nonlinear", 0.5
        desc = ("single periodic box 750^3 FP64, rc-energy (coarse 10), Re=2220, 325^3 particles St=1, "
                "nonlinear drag")
    elif name == "c3":
        n, npart, dt, Re, forcing, st = 1024, 64_000_000, "f32", 1333.3, C.ConstantPowerLF(0.5), (1.0, 1.0)
        desc = "single periodic box 1024^3 FP32, linear-power P=0.5, Re=1333, 6.4e7 particles St=1"
    elif name == "c2":
        n, npart, dt, Re, forcing, st = 256, 2_000_000, "f32", 360.0, C.EnergyPreserveRCF(0.649, 1, 2, 8), (0.1, 10.0)
        desc = "single periodic box 256^3 FP32, rc-energy (coarse 8), Re=360, 2e6 particles St log-U[0.1,10]"
    elif name == "c1":
        n, npart, dt, Re, forcing, st = 64, 10_000, "f64", 53.3, C.ConstantPowerLF(0.5), (1.0, 1.0)
        desc = "single periodic box 64^3 FP64, linear-power P=0.5, Re=53.3, 1e4 particles St=1"
    else:
        raise ValueError(f"unknown workload {name}")
    if n_override:
        npart = int(round(npart * (n_override / n) ** 3))
        n = n_override
        if isinstance(forcing, C.EnergyPreserveRCF) and n % forcing.coarse_size:
            # the coarse size must divide the grid: the largest divisor not
            # above the workload's own that still holds the forced shells
            fit = [c for c in range(2 * forcing.kmax, n + 1) if n % c == 0]
            below = [c for c in fit if c <= forcing.coarse_size]
            c = below[-1] if below else fit[0]
            forcing = C.EnergyPreserveRCF(forcing.energy, forcing.kmin, forcing.kmax, c)
            desc += f" (sampled at {n}^3 with {npart} particles, coarse {c})"
        else:
            desc += f" (sampled at {n}^3 with {npart} particles)"
    if particles_override is not None:
        npart = particles_override
    if dtype is not None:
        dt = dtype
    cfg = C.Config(grid=C.Grid((n, n, n), 3), timestep=C.TimeStep(0.3, 0.01), simulate=C.SimulateConfig(0),
                   flow=C.FlowConfig(C.FlowParams(1.0, 1.0, 1.0 / Re), C.SpectrumInit(1, 4.0, 1.0), HSMAC_DEFAULT,
                                     forcing),
                   particles=_particles(npart, 7, st, drag, spare), dtype=dt)
    return cfg, desc


def ensemble_configs(nbox, seed=12345):
    """Batched-ensemble workload: nbox FP32 periodic boxes, N uniform over
    {32, 48, 64, 96, 128}, particle counts log-uniform in [1e3, 1e6] with Stokes
    numbers log-uniform in [0.1, 10], seeded spectra with k0 from {2, 4, 6},
    linear-power forcing P = 0.5 (Re scaled with N as Re = 53.3 (N/64)^(4/3))."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nbox):
        n = int(rng.choice([32, 48, 64, 96, 128]))
        npart = int(round(10 ** rng.uniform(3.0, 6.0)))
        k0 = float(rng.choice([2.0, 4.0, 6.0]))
        Re = 53.3 * (n / 64.0) ** (4.0 / 3.0)
        out.append(C.Config(grid=C.Grid((n, n, n), 3), timestep=C.TimeStep(0.3, 0.01), simulate=C.SimulateConfig(0),
                            flow=C.FlowConfig(C.FlowParams(1.0, 1.0, 1.0 / Re), C.SpectrumInit(1000 + i, k0, 1.0),
                                              HSMAC_DEFAULT, C.ConstantPowerLF(0.5)),
                            particles=_particles(npart, 2000 + i, (0.1, 10.0)), dtype="f32"))
    return out


def serve_requests(nreq, seed=4242):
    """Served-stream workload: N in {32, 48, ..., 128}, particle count
    log-uniform in [1e3, 2e6], one Stokes number per run log-uniform in
    [0.1, 10], forcing linear-power (P = 0.5) or rc-energy (energy 0.649),
    step count uniform in [50, 500], a seed per request."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nreq):
        n = int(rng.choice(np.arange(32, 129, 16)))
        st = float(10 ** rng.uniform(-1.0, 1.0))
        rc = bool(rng.integers(0, 2))
        out.append(dict(n=n, particles=int(round(10 ** rng.uniform(3.0, np.log10(2e6)))), st=(st, st),
                        forcing="rc-energy" if rc else "linear-power", forcing_param=0.649 if rc else 0.5,
                        steps=int(rng.integers(50, 501)), seed=10_000 + i, k0=4.0,
                        Re=53.3 * (n / 64.0) ** (4.0 / 3.0), dtype="f32"))
    return out


def request_config(r, hsmac=HSMAC_DEFAULT, cfl=0.3, dtmax=0.01):
    """The Config of the box a request is served in, for stepping the same run
    outside the server (its initial state comes from the server's own
    initialisers, see ``request_box``)."""
    forcing = {"linear-power": lambda: C.ConstantPowerLF(r.get("forcing_param", 0.5)),
               "rc-energy": lambda: C.EnergyPreserveRCF(r.get("forcing_param", 0.649), 1, 2, 8),
               "none": C.NoForcing}[r.get("forcing", "linear-power")]()
    npart = int(r.get("particles", 0))
    st = r.get("st", (1.0, 1.0))
    return C.Config(grid=C.Grid((r["n"],) * 3, 3), timestep=C.TimeStep(cfl, dtmax), simulate=C.SimulateConfig(0),
                    flow=C.FlowConfig(C.FlowParams(1.0, 1.0, 1.0 / r.get("Re", 53.3)),
                                      C.SpectrumInit(r["seed"], r.get("k0", 4.0), 1.0), hsmac, forcing),
                    particles=_particles(npart, r["seed"], st), dtype=r.get("dtype", "f32"))


# the particle seed the server derives from a request's seed (lcs_serve.hip)
SERVE_PARTICLE_SEED_XOR = 0x9E3779B97F4A7C15


def request_box(r):
    """A standalone box holding the initial state the server gives request r:
    the seeded spectrum with u' = 1, periodic ghosts, and one profile of
    r["particles"] particles (grid ratio 4, density ratio 843, linear drag)
    placed and sized from the request's seed with zero slip."""
    from . import core
    cfg = request_config(r)
    box = core.Box(cfg.grid.dims, 3, cfg.flow.params, cfg.dtype)
    box.init_spectrum(r["seed"], r.get("k0", 4.0), 1.0)
    core.fill_ghost(box, ("U", "V", "W", "P"), 3, True)
    if r.get("particles", 0):
        box.add_particles(r["particles"], r["particles"], 0.0, 843.0, "linear", 4)
        st = r.get("st", (1.0, 1.0))
        box.init_particles(0, (r["seed"] ^ SERVE_PARTICLE_SEED_XOR) & (2 ** 64 - 1), st[0], st[1], 0.5, True)
    return box, cfg
