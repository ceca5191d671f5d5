"""Writes the committed golden fixtures of this directory (FP64, small).

Every fixture holds seeded inputs and the CPU oracle's outputs for them; the
GPU tests (tests/test_gpu_golden.py) run the HIP library on the stored inputs
and compare with the stored outputs without needing the oracle, and the CPU
tests (tests/test_golden.py) check that the oracle still reproduces them and
that they satisfy the closed-form invariants of each phase.

* flow16_h3.npz -- 16^3 cells, halo 3, per-phase input/output pairs:
  rhs! (src/flows/rhs.jl:9-81), update! fused with it (src/flows/update.jl:12-54),
  couple!(::HSMAC) with itrmin = itrmax = 6 as test/test_flows.jl:82-101 runs it,
  time-blocked + overlapped and naive (src/flows/couples/hsmac.jl:52-201),
  the flow CFL step (src/flows/timestep.jl:18-65), the linear-power forcing
  (src/flows/forcings/lf.jl:6-106) and the rc-energy forcing with coarse 4
  (src/flows/forcings/rcf.jl:21-191).
* particles16.npz -- 1,000 particles in the same 16^3 fields: boundary!,
  motion! (linear and nonlinear drag) and update! of stage 1
  (src/particles/{boundary,motion,update}.jl), the particle dt
  (src/particles/timestep.jl:6-25), and makeindex! of 10 and 1,000 particles
  (src/particles/makeindex.jl:29-67).
* step24_precompile.npz -- one whole step (src/timestep.jl:14-27 then
  src/evolve.jl:63-146) of the precompile.lcs-yaml shape: 24^3, halo 3,
  rc-energy forcing (kmin 1, kmax 2, coarse 6, energy 0.649), HSMAC with
  itrmin = itrmax = 6, 432 particles with linear drag and density ratio 843.
  Inputs are interiors: the step fills every ghost before it reads one.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "lcs.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import oracle as O  # noqa: E402
from oracle.driver import OracleSim, ParticleSet  # noqa: E402

F64 = np.float64

# phase parameters (the GPU and CPU tests read them back from the files)
FLOW = dict(dims=(16, 16, 16), h=3, rhs_Re=(2.0, 3.0, 0.5), update_dt=0.007, hsmac_dt=0.01, cfl=0.3, dtmax=1.0,
            cfl_Re=(1.0, 1.0, 0.01), lf_power=1.0, lf_dt=0.01, rcf=(1.0, 1, 2, 4), rcf_dt=0.01)
PART = dict(dims=(16, 16, 16), h=3, n=1000, Re=(1.0, 1.0, 0.02), gravity=9.8, density_ratio=5.0, grid_ratio=4,
            dt=0.003, dtmax=10.0, index_sizes=(10, 1000))
STEP_YAML = """
mode: {kind: flow-particle}
backend: {kind: mi355x}
grid: {dims: [24, 24, 24], halo_size: 3}
timestep: {cfl: 0.3, dtmax: 0.01}
simulate: {last_step: 1}
flow:
  params: {U0_m_s: 0.429, L0_m: 0.005, NU_m2_s: 1.5e-5}
  init: {kind: spectrum, params: {seed: 3, k0: 4, uprime: 1}}
  couple:
    kind: hsmac
    params: {time_blocking: true, overlap: true, omega: 1.7, epsp: 0.001, itrmin: 6, itrmax: 6,
             coloring: {kind: red-black-fast}}
  forcing: {kind: rc-energy, params: {kmin: 1, kmax: 2, coarse_size: 6, energy: 0.649}}
particles:
  - population: {valid: 432, invalid_ratio: 0.02, exchange_ratio: 0.02}
    params: {gravity: 0.0, density_ratio: 843}
    cell: {grid_ratio: 4}
    drag: {kind: linear}
    init:
      id: {kind: generate}
      position: {kind: random}
      velocity: {kind: rest}
      size: {kind: const, params: {diam_m: 1.14e-5}}
"""


def _Re(t):
    return t[0] * t[1] / t[2]


def rand_halo(rng, dims, h):
    return np.asfortranarray(rng.random(tuple(n + 2 * h for n in dims)))


def interior(A, h):
    return np.ascontiguousarray(A[h:-h, h:-h, h:-h])


def flow_fixture():
    f = FLOW
    dims, h = f["dims"], f["h"]
    rng = np.random.default_rng(20240601)
    U, V, W, P = (rand_halo(rng, dims, h) for _ in range(4))
    out = dict(U=U, V=V, W=W, P=P)
    cp = lambda a: a.copy(order="F")  # noqa: E731
    # rhs!
    dU = O.rhs(U, V, W, P, h, _Re(f["rhs_Re"]))
    out.update(rhs_dU=dU[0], rhs_dV=dU[1], rhs_dW=dU[2])
    # rhs! + update! of stage 1: U2 = U + dt rhs(U, P)
    Us2 = [np.zeros_like(U) for _ in range(3)]
    O.update(Us2, [U, V, W], dU, h, f["update_dt"])
    out.update(upd_U2=interior(Us2[0], h), upd_V2=interior(Us2[1], h), upd_W2=interior(Us2[2], h))
    # couple! on (U2, V2, W2, P) <- (U, V, W, P)
    for tag, tb, ov in (("tbov", True, True), ("naive", False, False)):
        A = [cp(a) for a in (U, V, W, P)]
        itr, div = O.hsmac(*A, h, f["hsmac_dt"], tb, ov, 0, 1.7, 1e-3, 6, 6)
        out.update({f"hs_{tag}_{k}": interior(a, h) for k, a in zip("UVWP", A)})
        out[f"hs_{tag}_itr"] = np.int64(itr)
        out[f"hs_{tag}_div"] = F64(div)
    # CFL on (U - 1/2) * 3 (signed velocities of order one)
    Us = [np.asfortranarray((a - 0.5) * 3.0) for a in (U, V, W)]
    dt, um = O.flow_timestep(*[cp(a) for a in Us], h, f["cfl"], f["dtmax"], _Re(f["cfl_Re"]))
    out.update(cfl_dt=F64(dt), cfl_umax=um)
    # linear-power forcing
    A = [cp(a) for a in (U, V, W)]
    p, _ = O.force_linear(*A, h, f["lf_power"], f["lf_dt"])
    out.update(lf_p=F64(p), lf_U=interior(A[0], h), lf_V=interior(A[1], h), lf_W=interior(A[2], h))
    # rc-energy forcing
    A = [cp(a) for a in (U, V, W)]
    e, kmin, kmax, c = f["rcf"]
    p, _ = O.force_rcf(*A, h, e, kmin, kmax, c, f["rcf_dt"])
    out.update(rcf_p=F64(p), rcf_U=interior(A[0], h), rcf_V=interior(A[1], h), rcf_W=interior(A[2], h))
    return out


PROPS_F = ("x", "y", "z", "u", "v", "w", "diams")


def particle_fixture():
    f = PART
    dims, h, n = f["dims"], f["h"], f["n"]
    rng = np.random.default_rng(20240602)
    Us = [rand_halo(rng, dims, h) for _ in range(3)]
    out = dict(U=Us[0], V=Us[1], W=Us[2])
    L = O.DOMAIN_LENGTH
    x = [rng.uniform(-0.3, L + 0.3, n) for _ in range(3)]
    u = [rng.standard_normal(n) for _ in range(3)]
    diams = rng.uniform(0.005, 0.05, n)
    ids = (rng.permutation(n) + 1).astype(np.int64)
    out.update(x=x[0], y=x[1], z=x[2], u=u[0], v=u[1], w=u[2], diams=diams, ids=ids)
    Re = _Re(f["Re"])
    for drag in (0, 1):
        xs = [a.copy() for a in x]
        xs2 = [np.zeros(n) for _ in range(3)]
        us = [a.copy() for a in u]
        us2 = [np.zeros(n) for _ in range(3)]
        du = [np.zeros(n) for _ in range(3)]
        # stage 1 as the driver runs it: wrap (x, and x2 by the same shift),
        # drag at the wrapped x, then x2 = x + dt u2 and u2 from dudt
        O.p_boundary(xs, xs2, n)
        O.p_motion(Us, h, Re, f["gravity"], f["density_ratio"], drag, xs, us, diams, du, n)
        O.p_update(xs2, xs, us2, us, du, n, f["dt"])
        tag = ("linear", "nonlinear")[drag]
        out.update({f"{tag}_{k}": a for k, a in zip(("xw", "yw", "zw"), xs)})
        out.update({f"{tag}_{k}": a for k, a in zip(("x2", "y2", "z2"), xs2)})
        out.update({f"{tag}_{k}": a for k, a in zip(("u2", "v2", "w2"), us2)})
        out.update({f"{tag}_{k}": a for k, a in zip(("dudt", "dvdt", "dwdt"), du)})
    out["pdt"] = F64(O.p_timestep(diams, Re, f["density_ratio"], f["dtmax"]))
    for m in f["index_sizes"]:
        xm = [a[:m].copy() for a in x]
        hh, perm, st, sp = O.makeindex(dims, f["grid_ratio"], [a.copy() for a in xm], ids[:m].copy(), [], m)
        out.update({f"mi{m}_hashes": hh, f"mi{m}_perm": np.asarray(perm, np.int64), f"mi{m}_starts": st,
                    f"mi{m}_stops": sp})
    return out


def step_config():
    import yaml

    from lcs_mi355x import config as C
    return C.from_dict(yaml.safe_load(STEP_YAML))


def step_fixture():
    cfg = step_config()
    dims, h = cfg.grid.dims, cfg.grid.halo_size
    init = cfg.flow.init
    (U, V, W), P = O.init_spectrum(dims, h, init.seed, init.k0, init.uprime, F64)
    pc = cfg.particles[0]
    n, ntotal = pc.population.valid, pc.population.ntotal
    rng = np.random.default_rng(20240603)
    x = [rng.uniform(0.0, O.DOMAIN_LENGTH, n) for _ in range(3)]
    u = [np.zeros(n) for _ in range(3)]
    diams = np.full(n, pc.init.diam_m / cfg.flow.params.L0_m)
    ids = (rng.permutation(n) + 1).astype(np.int64)
    out = dict(U=interior(U, h), V=interior(V, h), W=interior(W, h), P=interior(P, h), x=x[0], y=x[1], z=x[2],
               u=u[0], v=u[1], w=u[2], diams=diams, ids=ids, ntotal=np.int64(ntotal))
    sim = OracleSim(cfg, F64, [U, V, W], P, [ParticleSet(x, u, diams, ids, ntotal)])
    dt = sim.timestep()
    log = sim.evolve()
    out.update(dt=F64(dt), itr1=np.int64(log["flow"]["couple1"]["itr"]), div1=F64(log["flow"]["couple1"]["div"]),
               itr2=np.int64(log["flow"]["couple2"]["itr"]), div2=F64(log["flow"]["couple2"]["div"]),
               p=F64(log["flow"]["forcing"]["p"]))
    out.update({f"out_{k}": interior(a, h) for k, a in zip("UVW", sim.Us)})
    out["out_P"] = interior(sim.P, h)
    ps = sim.particles[0]
    out.update({f"out_{k}": a[:n].copy() for k, a in zip(("x", "y", "z"), ps.xss)})
    out.update({f"out_{k}": a[:n].copy() for k, a in zip(("u", "v", "w"), ps.uss)})
    out["out_ids"] = ps.ids[:n].copy()
    return out


def main():
    O.set_threads(1)
    for name, make in (("flow16_h3.npz", flow_fixture), ("particles16.npz", particle_fixture),
                       ("step24_precompile.npz", step_fixture)):
        data = make()
        path = os.path.join(HERE, name)
        np.savez_compressed(path, **data)
        print(f"{name}: {len(data)} arrays, {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
