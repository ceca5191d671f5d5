"""``LCS.simulate(config)`` for the mi355x backend (src/simulate.jl:13-90).

Builds the box (``LCS.Buffer``), runs the builder-defined initialisers, then
loops ``timestep!`` + ``evolve!`` up to ``simulate.last_step`` and prints the
per-step log the way ``printstate``/``printlog`` do (src/stat.jl:48-85).
Statistics and restart IO (src/LCSIO) are outside this backend's scope.
"""
import sys

from . import core
from .config import Config, SpectrumInit, readconfig


def build(config: Config, device=0):
    """LCS.Buffer + LCS.init! (src/buffer.jl:27-34, src/init.jl:6-26)."""
    g = config.grid
    box = core.Box(g.dims, g.halo_size, config.flow.params, config.dtype, device)
    init = config.flow.init
    if not isinstance(init, SpectrumInit):
        raise ValueError("the mi355x backend initialises flows with kind 'spectrum'")
    box.init_spectrum(init.seed, init.k0, init.uprime)
    core.fill_ghost(box, ("U", "V", "W", "P"), g.halo_size, True)
    for prof in config.particles:
        i = box.add_particles(prof.population.valid, prof.population.ntotal, prof.params.gravity,
                              prof.params.density_ratio, prof.drag, prof.grid_ratio)
        pi = prof.init
        if pi.size == "stokes":
            st_min, st_max, eps = pi.st_min, pi.st_max, pi.eps
        elif pi.size == "const":
            # ConstSize: diam = diam_m / L0 (src/particles/config.jl:233); expressed as a Stokes number
            # through the same taup/taueta relation so one initialiser serves both kinds
            Re = config.flow.params.Re
            d = pi.diam_m / config.flow.params.L0_m
            eps = pi.eps
            taueta = (1.0 / (eps * Re)) ** 0.5
            st_min = st_max = (d * d * prof.params.density_ratio * Re / 18.0) / taueta
        else:
            raise ValueError(f"unsupported particle size init {pi.size!r}")
        box.init_particles(i, pi.seed, st_min, st_max, eps, zero_slip=(pi.velocity == "zero-slip"))
    return box


def simulate(config, callbacks=(), quiet=False, device=0):
    if isinstance(config, str):
        config = readconfig(config)
    box = build(config, device)
    scfg = core.step_config(config)
    logs = []
    for _ in range(config.simulate.last_step):
        log = core.step(box, scfg)
        logs.append(log)
        if not quiet:
            f = log["flow"]
            print(f"step {log['step']}  t {log['t']:.6g}  dt {log['dt']:.6g}  "
                  f"couple1 itr {f['couple1']['itr']} div {f['couple1']['div']:.4g}  "
                  f"couple2 itr {f['couple2']['itr']} div {f['couple2']['div']:.4g}  p {f['forcing']['p']:.4g}",
                  file=sys.stdout, flush=True)
        for cb in callbacks:
            cb(box, log)
    return box, logs
