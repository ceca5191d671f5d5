"""Whole-step parity: LCS.timestep! + LCS.evolve! on the device (one C call per
step) against the oracle's phase-by-phase restatement, from identical initial
states (the device initialises; the oracle starts from the downloaded copy).

Bar: after 100 steps at 64^3 in FP64, U, V, W, P agree
within relative L2 1e-5; HSMAC sweep counts agree step by step; the device's
cell indices of all 10^4 particles equal the oracle's makeindex of the same
positions bit for bit (so they agree wherever positions agree).
"""
import numpy as np
import pytest

from conftest import interior, rel_l2
from oracle import oracle as O
from oracle.driver import OracleSim, ParticleSet

pytestmark = pytest.mark.gpu


def c1_config(dims=(64, 64, 64), nparticles=10000, forcing=None, drag="linear", dtype="f64", profiles=1):
    from lcs_mi355x import config as C
    forcing = forcing or C.ConstantPowerLF(0.5)
    part = tuple(C.ParticleConfig(C.Population(nparticles, 0.02), C.ParticleParams(0.0, 843.0), 4, drag,
                                  C.ParticleInit(seed=100 + q, velocity="zero-slip", size="stokes", st_min=1.0,
                                                 st_max=1.0 if q == 0 else 3.0, eps=0.5))
                 for q in range(profiles))
    # Re_lambda = u'^2 sqrt(15 Re / eps) = 40 with u' = 1, eps = P = 0.5  ->  Re = 53.3
    return C.Config(grid=C.Grid(dims, 3), timestep=C.TimeStep(0.3, 0.01), simulate=C.SimulateConfig(100),
                    flow=C.FlowConfig(C.FlowParams(1.0, 1.0, 1.0 / 53.3), C.SpectrumInit(1, 4.0, 1.0),
                                      C.HSMAC(True, True, "red-black-fast", 1.7, 1e-3, 6, 300), forcing),
                    particles=part, dtype=dtype)


def oracle_from_device(cfg, box):
    dt = np.float64 if cfg.dtype == "f64" else np.float32
    Us = [box.download(n) for n in ("U", "V", "W")]
    P = box.download("P")
    sets = []
    for q, prof in enumerate(box.profiles):
        nv = prof["nvalid"]
        xs = [box.download_prop(q, k)[:nv] for k in ("x", "y", "z")]
        us = [box.download_prop(q, k)[:nv] for k in ("u", "v", "w")]
        sets.append(ParticleSet(xs, us, box.download_prop(q, "diams")[:nv], box.download_prop(q, "ids")[:nv],
                                prof["ntotal"]))
    return OracleSim(cfg, dt, Us, P, sets)


def run_both(cfg, nsteps):
    import lcs_mi355x as L
    box = L.build(cfg)
    sim = oracle_from_device(cfg, box)
    scfg = L.step_config(cfg)
    logs = []
    for s in range(nsteps):
        dlog = L.step(box, scfg)
        sim.timestep()
        olog = sim.evolve()
        logs.append((dlog, olog, sim.dt))
    return box, sim, logs


def test_c1_64cubed_fp64_100_steps():
    cfg = c1_config()
    box, sim, logs = run_both(cfg, 100)
    for s, (d, o, odt) in enumerate(logs):
        assert d["flow"]["couple1"]["itr"] == o["flow"]["couple1"]["itr"], s
        assert d["flow"]["couple2"]["itr"] == o["flow"]["couple2"]["itr"], s
        assert np.isclose(d["dt"], odt, rtol=1e-9), s
        assert np.isclose(d["flow"]["forcing"]["p"], o["flow"]["forcing"]["p"], rtol=1e-7), s
    h = 3
    errs = {}
    for name, ref in zip(("U", "V", "W", "P"), (*sim.Us, sim.P)):
        errs[name] = rel_l2(interior(box.download(name), h), interior(ref, h))
    print("rel L2 after 100 steps:", errs, flush=True)
    assert all(e <= 1e-5 for e in errs.values()), errs
    assert box.profiles and logs[-1][0]["step"] == 100
    # particles: positions compared by id; then the device re-binning of the
    # device positions against the oracle's makeindex of the same positions
    ps = sim.particles[0]
    nv = ps.nvalid
    assert nv == 10000
    ids = box.download_prop(0, "ids")[:nv]
    order = np.argsort(ids)
    gx = [box.download_prop(0, k)[:nv][order] for k in ("x", "y", "z")]
    oo = np.argsort(ps.ids[:nv])
    ox = [a[:nv][oo] for a in ps.xss]
    for a, b in zip(gx, ox):
        assert rel_l2(a, b) < 1e-6
    same = np.all([a == b for a, b in zip(gx, ox)], axis=0)
    import lcs_mi355x as L
    L.makeindex(box, 0, 1)
    dh, dstarts, dstops = box.cellindex(0)
    dx = [box.download_prop(0, k)[:nv] for k in ("x", "y", "z")]
    dids = box.download_prop(0, "ids")[:nv]
    # every particle: the device hash of its device position is the oracle's
    # hash of that position, bit for bit, and the cell index agrees
    oh, _, ostarts, ostops = O.makeindex(cfg.grid.dims, 4, [a.copy() for a in dx], dids.copy(), [], nv)
    assert len(dh) == nv
    np.testing.assert_array_equal(dh, oh)
    np.testing.assert_array_equal(dstarts, ostarts)
    np.testing.assert_array_equal(dstops, ostops)
    # where the two runs left a particle at the same position, its cell is the
    # oracle's cell for its own run's position (the same value, so the same hash)
    pos = {int(i): k for k, i in enumerate(dids)}
    sel = np.sort(ids)[same]
    ohash_same = O.hash_cells(cfg.grid.dims, 4, [a[same].copy() for a in ox], int(same.sum()))
    np.testing.assert_array_equal(np.array([dh[pos[int(i)]] for i in sel], dtype=np.int64), ohash_same)
    print("particles re-binned:", nv, "; with bit-identical positions after 100 steps:", int(same.sum()), flush=True)


def test_precompile_like_rc_energy_two_profiles():
    """precompile.lcs-yaml shape: 24^3, rc-energy (coarse 6), nonlinear drag."""
    from lcs_mi355x import config as C
    cfg = c1_config(dims=(24, 24, 24), nparticles=432, forcing=C.EnergyPreserveRCF(0.649, 1, 2, 6),
                    drag="nonlinear", profiles=2)
    box, sim, logs = run_both(cfg, 20)
    for s, (d, o, odt) in enumerate(logs):
        assert d["flow"]["couple1"]["itr"] == o["flow"]["couple1"]["itr"], s
        assert d["flow"]["couple2"]["itr"] == o["flow"]["couple2"]["itr"], s
        assert np.isclose(d["flow"]["forcing"]["p"], o["flow"]["forcing"]["p"], rtol=1e-7), s
    for name, ref in zip(("U", "V", "W", "P"), (*sim.Us, sim.P)):
        assert rel_l2(interior(box.download(name), 3), interior(ref, 3)) <= 1e-8


def test_flow_mode_without_particles_fp32():
    cfg = c1_config(dims=(32, 32, 32), dtype="f32", profiles=0)
    box, sim, logs = run_both(cfg, 10)
    for s, (d, o, odt) in enumerate(logs):
        assert d["flow"]["couple1"]["itr"] == o["flow"]["couple1"]["itr"], s
    for name, ref in zip(("U", "V", "W", "P"), (*sim.Us, sim.P)):
        assert rel_l2(interior(box.download(name), 3), interior(ref, 3)) <= 1e-4
