"""``LCS.simulate(config)`` for the mi355x backend (src/simulate.jl:13-90).

LCSIO.init (output directory, resume policy), the box (``LCS.Buffer``), the
initialisers (``LCS.init!``: builder-defined spectrum and Stokes-number
initial conditions, or a restart file), then ``timestep!`` + ``evolve!`` from
state.step through ``simulate.last_step`` with the per-step log printed the
way ``printstate``/``printlog`` do (src/stat.jl:48-85), statistics saved every
``interval_stat`` steps and restart files every ``interval_restart`` steps.
"""
import sys

from . import core, lcsio
from .config import Config, RestartInit, SpectrumInit, readconfig


def build(config: Config, device=0):
    """LCS.Buffer + LCS.init! (src/buffer.jl:27-34, src/init.jl:6-26)."""
    g = config.grid
    box = core.Box(g.dims, g.halo_size, config.flow.params, config.dtype, device)
    init = config.flow.init
    if isinstance(init, RestartInit):
        lcsio.load_flow_fields(init.file, box)  # src/flows/inits/init.jl:16-19
    elif isinstance(init, SpectrumInit):
        box.init_spectrum(init.seed, init.k0, init.uprime)
        core.fill_ghost(box, ("U", "V", "W", "P"), g.halo_size, True)
    else:
        raise ValueError("the mi355x backend initialises flows with kind 'spectrum' or 'restart'")
    for prof in config.particles:
        box.add_particles(prof.population.valid, prof.population.ntotal, prof.params.gravity,
                          prof.params.density_ratio, prof.drag, prof.grid_ratio)
    for i, prof in enumerate(config.particles):
        pi = prof.init
        if pi.restart_file is not None:
            # ParticleRestart(file, iprofile) fills profile iprofile (src/particles/init.jl:22-31,61-70)
            lcsio.load_particles(pi.restart_file, box, pi.restart_profile - 1)
            continue
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


def stat(box, config: Config):
    """LCS.stat (src/stat.jl:92-105): flow statistics, then per profile the
    particle statistics with that flow's tau_eta.  Returns (flow, particles)."""
    flow = core.flow_stat(box, config.flow.stat.integral_length)
    parts = []
    for i, p in enumerate(config.particles):
        s = p.stat
        parts.append(core.particle_stat(box, i, flow["tau_eta"], s.contact_tolerance, s.bin_size, s.min_radius,
                                        s.search_cell_size, s.max_wavenumber, s.nsamples))
    return flow, parts


def _print_log(log):
    f = log["flow"]
    print(f"step {log['step']}  t {log['t']:.6g}  dt {log['dt']:.6g}  "
          f"couple1 itr {f['couple1']['itr']} div {f['couple1']['div']:.4g}  "
          f"couple2 itr {f['couple2']['itr']} div {f['couple2']['div']:.4g}  p {f['forcing']['p']:.4g}",
          file=sys.stdout, flush=True)


def simulate(config, callbacks=(), quiet=False, device=0):
    """The loop of src/simulate.jl:54-88: steps state.step..last_step (so a
    fresh run takes last_step + 1 steps, as the reference's does)."""
    if isinstance(config, str):
        config = readconfig(config)
    config = lcsio.init(config)
    box = build(config, device)
    if config.resume:
        step, t = lcsio.load_resume_state(config)  # LCS.init_state! (src/init.jl:11-15)
        core.set_state(box, step, t)
    scfg = core.step_config(config)
    sim = config.simulate
    logs = []
    first = core.read_log(box)["step"]
    for _ in range(first, sim.last_step + 1):
        log = core.step(box, scfg)
        logs.append(log)
        if not quiet:
            _print_log(log)
        s = log["step"]
        if lcsio.oninterval(s, sim.interval_stat):
            flow, parts = stat(box, config)
            state = lcsio.state_tree(box, log)
            if not quiet:
                print(f"stat {s}  k {flow['k']:.6g}  eps {flow['eps']:.6g}  Re_lambda {flow['Re_lambda']:.6g}",
                      file=sys.stdout, flush=True)
            lcsio.savestat(config, state, lcsio.stat_tree(state, flow, parts),
                           lcsio.log_tree(log, config.flow.forcing.kind, [p["nvalid"] for p in box.profiles]))
        if lcsio.oninterval(s, sim.interval_restart):
            lcsio.saverestart(box, config, s)
            lcsio.cleanrestart(s, config)
        for cb in callbacks:
            cb(box, log)
    return box, logs
