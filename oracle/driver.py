"""Whole-step CPU oracle (test infrastructure only).

Composes the C restatement's phases in the reference's per-step order:
``LCS.timestep!`` (``src/timestep.jl:14-27``: flow CFL dt, particle dt, min)
followed by ``LCS.evolve!(::FlowParticleMode)`` (``src/evolve.jl:63-146``).
The config argument is duck-typed: it needs the attributes of the reference's
``LCS.Config`` (grid, timestep, flow.params/couple/forcing, particles).
"""
import numpy as np

from . import oracle as O


def _Re(params):
    # src/flows/config.jl:18
    return params.U0_m_s * params.L0_m / params.NU_m2_s


class ParticleSet:
    """Particles.Properties + ParticleState for one profile (src/particles/buffer.jl:1-30)."""

    def __init__(self, xss, uss, diams, ids, ntotal=None):
        n = len(xss[0])
        ntotal = n if ntotal is None else ntotal
        dt = xss[0].dtype

        def pad(a, fill=0):
            out = np.full(ntotal, fill, dtype=a.dtype)
            out[:n] = a
            return out

        self.nvalid = n
        self.ntotal = ntotal
        self.xss = [pad(a) for a in xss]
        self.xss2 = [np.zeros(ntotal, dt) for _ in range(3)]
        self.uss = [pad(a) for a in uss]
        self.uss2 = [np.zeros(ntotal, dt) for _ in range(3)]
        self.dudtss = [np.zeros(ntotal, dt) for _ in range(3)]
        self.diams = pad(diams)
        self.ids = pad(ids, O.INVALID)


class OracleSim:
    def __init__(self, config, dtype, Us, P, particles=()):
        self.config = config
        self.dtype = np.dtype(dtype)
        self.h = config.grid.halo_size
        self.dims = tuple(config.grid.dims)
        self.Us = [np.array(a, dtype=dtype, order="F") for a in Us]
        self.P = np.array(P, dtype=dtype, order="F")
        self.Us2 = [np.zeros_like(a) for a in self.Us]
        self.particles = list(particles)
        self.step = 0
        self.t = 0.0
        self.dt = 0.0
        self.Re = _Re(config.flow.params)

    # src/timestep.jl:14-27
    def timestep(self):
        c = self.config
        dt, _ = O.flow_timestep(*self.Us, self.h, c.timestep.cfl, c.timestep.dtmax, self.Re)
        if self.particles:
            dtp = min(O.p_timestep(ps.diams, self.Re, prof.params.density_ratio, c.timestep.dtmax)
                      for ps, prof in zip(self.particles, c.particles))
            dt = min(dt, dtp)
        self.dt = dt
        return dt

    def _couple(self, Us, dt):
        cp = self.config.flow.couple
        return O.hsmac(*Us, self.P, self.h, dt, cp.time_blocking, cp.overlap, cp.coloring_id, cp.omega, cp.epsp,
                       cp.itrmin, cp.itrmax)

    def _forcing(self, dt):
        f = self.config.flow.forcing
        if f.kind == "linear-power":
            p, _ = O.force_linear(*self.Us, self.h, f.power, dt)
            return {"p": p}
        if f.kind == "rc-energy":
            p, e10 = O.force_rcf(*self.Us, self.h, f.energy, f.kmin, f.kmax, f.coarse_size, dt)
            return {"p": p, "e10": e10}
        return {}

    def _particles_stage(self, stage, dt_stage):
        c = self.config
        for ps, prof in zip(self.particles, c.particles):
            n = ps.nvalid
            if stage == 1:
                O.p_boundary(ps.xss, ps.xss2, n)
                O.p_motion(self.Us, self.h, self.Re, prof.params.gravity, prof.params.density_ratio, prof.drag_id,
                           ps.xss, ps.uss, ps.diams, ps.dudtss, n)
                O.p_update(ps.xss2, ps.xss, ps.uss2, ps.uss, ps.dudtss, n, dt_stage)
            else:
                O.p_boundary(ps.xss2, ps.xss, n)
                O.p_motion(self.Us2, self.h, self.Re, prof.params.gravity, prof.params.density_ratio, prof.drag_id,
                           ps.xss2, ps.uss2, ps.diams, ps.dudtss, n)
                O.p_update(ps.xss, ps.xss, ps.uss, ps.uss, ps.dudtss, n, dt_stage)

    # src/evolve.jl:63-146
    def evolve(self):
        h = self.h
        dt1, dt2 = 0.5 * self.dt, self.dt  # src/state.jl:37-38
        for A in (*self.Us, self.P):
            O.ghost_fill(A, h, h, True)
        self._particles_stage(1, dt1)
        dUs = O.rhs(*self.Us, self.P, h, self.Re)
        O.update(self.Us2, self.Us, dUs, h, dt1)
        couple1 = self._couple(self.Us2, dt1)
        for A in (*self.Us2, self.P):
            O.ghost_fill(A, h, h, True)
        self._particles_stage(2, dt2)
        dUs = O.rhs(*self.Us2, self.P, h, self.Re)
        O.update(self.Us, self.Us, dUs, h, dt2)
        forcing = self._forcing(dt2)
        couple2 = self._couple(self.Us, dt2)
        self.step += 1
        self.t += self.dt
        return {"flow": {"couple1": {"itr": couple1[0], "div": couple1[1]}, "forcing": forcing,
                         "couple2": {"itr": couple2[0], "div": couple2[1]}}}
