"""The HIP library (through the C ABI) on the committed golden fixtures:
stored FP64 inputs in, stored oracle outputs out, no oracle at run time.

Bars, as in the per-phase parity tests: positions after the wrap, hashes,
cell starts/stops and permutations bit for bit; floating-point results
within relative L2 1e-12 per phase (the kernels contract multiply-adds and
reduce in another order than the oracle); the whole 24^3 step within 1e-10,
with its sweep counts equal.
"""
import os

import numpy as np
import pytest

from conftest import interior, rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12
INVALID = 2**53 - 1


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def gen():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def lcs():
    import lcs_mi355x
    lcs_mi355x.load()
    return lcs_mi355x


def box_of(dims, h, Re):
    L = lcs()
    return L.Box(dims, h, L.FlowParams(*Re), np.float64)


@pytest.fixture(scope="module")
def flow():
    return load("flow16_h3.npz"), gen().FLOW


def test_golden_rhs_and_update(flow):
    d, f = flow
    L = lcs()
    h = f["h"]
    box = box_of(f["dims"], h, f["rhs_Re"])
    for c in "UVWP":
        box.upload(c, d[c])
    L.rhs(box, 1)
    for c in "UVW":
        assert rel_l2(box.download(f"d{c}"), d[f"rhs_d{c}"]) < TOL, c
    L.rhs_update(box, 1, f["update_dt"])
    for c in "UVW":
        assert rel_l2(interior(box.download(f"{c}2"), h), d[f"upd_{c}2"]) < TOL, c


@pytest.mark.parametrize("tag,tb,ov", [("tbov", True, True), ("naive", False, False)])
def test_golden_hsmac(flow, tag, tb, ov):
    d, f = flow
    L = lcs()
    h = f["h"]
    box = box_of(f["dims"], h, (1.0, 1.0, 1.0))
    for c, name in zip("UVWP", ("U2", "V2", "W2", "P")):
        box.upload(name, d[c])
    res = L.couple(box, L.HSMAC(tb, ov, "red-black-fast", 1.7, 1e-3, 6, 6), 1, f["hsmac_dt"])
    assert res["itr"] == int(d[f"hs_{tag}_itr"])
    assert np.isclose(res["div"], float(d[f"hs_{tag}_div"]), rtol=1e-10)
    for c, name in zip("UVWP", ("U2", "V2", "W2", "P")):
        assert rel_l2(interior(box.download(name), h), d[f"hs_{tag}_{c}"]) < TOL, c


def test_golden_cfl(flow):
    d, f = flow
    L = lcs()
    box = box_of(f["dims"], f["h"], f["cfl_Re"])
    for c in "UVW":
        box.upload(c, (d[c] - 0.5) * 3.0)
    dt, um = L.flow_timestep(box, f["cfl"], f["dtmax"])
    assert np.isclose(dt, float(d["cfl_dt"]), rtol=TOL)
    assert np.allclose(um, d["cfl_umax"], rtol=TOL)


def test_golden_forcings(flow):
    d, f = flow
    L = lcs()
    h = f["h"]
    e, kmin, kmax, c = f["rcf"]
    for tag, force, dt in (("lf", L.ConstantPowerLF(f["lf_power"]), f["lf_dt"]),
                           ("rcf", L.EnergyPreserveRCF(e, kmin, kmax, c), f["rcf_dt"])):
        box = box_of(f["dims"], h, (1.0, 1.0, 1.0))
        for k in "UVW":
            box.upload(k, d[k])
        r = L.forcing(box, force, dt)
        assert np.isclose(r["p"], float(d[f"{tag}_p"]), rtol=1e-10), tag
        for k in "UVW":
            assert rel_l2(interior(box.download(k), h), d[f"{tag}_{k}"]) < TOL, (tag, k)


@pytest.mark.parametrize("drag", ["linear", "nonlinear"])
def test_golden_particle_stage(drag):
    d = load("particles16.npz")
    f = gen().PART
    L = lcs()
    n = f["n"]
    box = box_of(f["dims"], f["h"], f["Re"])
    for c in "UVW":
        box.upload(c, d[c])
    prof = box.add_particles(n, n, gravity=f["gravity"], density_ratio=f["density_ratio"], drag=drag,
                             grid_ratio=f["grid_ratio"])
    for name in L.core._lib.PROPS:
        a = d[name] if name in ("x", "y", "z", "u", "v", "w", "diams", "ids") else np.zeros(n)
        box.upload_prop(prof, name, a)
    L.boundary(box, prof, 1)
    for k in "xyz":
        assert np.array_equal(box.download_prop(prof, k), d[f"{drag}_{k}w"]), k
    L.motion(box, prof, 1)
    for k in ("dudt", "dvdt", "dwdt"):
        assert rel_l2(box.download_prop(prof, k), d[f"{drag}_{k}"]) < TOL, k
    L.particle_update(box, prof, 1, f["dt"])
    for k in ("x2", "y2", "z2", "u2", "v2", "w2"):
        assert rel_l2(box.download_prop(prof, k), d[f"{drag}_{k}"]) < TOL, k
    assert np.isclose(L.particle_timestep(box, prof, f["dtmax"]), float(d["pdt"]), rtol=TOL)


@pytest.mark.parametrize("m", [10, 1000])
def test_golden_makeindex(m):
    d = load("particles16.npz")
    f = gen().PART
    L = lcs()
    box = box_of(f["dims"], f["h"], (1.0, 1.0, 1.0))
    prof = box.add_particles(m, m, grid_ratio=f["grid_ratio"])
    for k in "xyz":
        box.upload_prop(prof, k, d[k][:m])
    L.makeindex(box, prof, 1)
    hashes, starts, stops = box.cellindex(prof)
    assert np.array_equal(hashes, d[f"mi{m}_hashes"])
    assert np.array_equal(starts, d[f"mi{m}_starts"]) and np.array_equal(stops, d[f"mi{m}_stops"])
    perm = d[f"mi{m}_perm"] - 1
    for k in "xyz":
        assert np.array_equal(box.download_prop(prof, k), d[k][:m][perm]), k


def test_golden_precompile_step():
    d = load("step24_precompile.npz")
    G = gen()
    cfg = G.step_config()
    L = lcs()
    g = cfg.grid
    h = g.halo_size
    box = L.Box(g.dims, h, cfg.flow.params, np.float64)
    for c in "UVWP":
        box.upload(c, np.pad(d[c], h, mode="wrap"))
    pc = cfg.particles[0]
    n, nt = pc.population.valid, int(d["ntotal"])
    prof = box.add_particles(n, nt, pc.params.gravity, pc.params.density_ratio, pc.drag, pc.grid_ratio)
    for name in L.core._lib.PROPS:
        if name == "ids":
            a = np.full(nt, INVALID, np.int64)
        else:
            a = np.zeros(nt)
        if name in ("x", "y", "z", "u", "v", "w", "diams", "ids"):
            a[:n] = d[name]
        box.upload_prop(prof, name, a)
    log = L.step(box, L.step_config(cfg))
    assert np.isclose(log["dt"], float(d["dt"]), rtol=TOL)
    fl = log["flow"]
    assert fl["couple1"]["itr"] == int(d["itr1"]) and fl["couple2"]["itr"] == int(d["itr2"])
    assert np.isclose(fl["couple1"]["div"], float(d["div1"]), rtol=1e-9)
    assert np.isclose(fl["couple2"]["div"], float(d["div2"]), rtol=1e-9)
    assert np.isclose(fl["forcing"]["p"], float(d["p"]), rtol=1e-9)
    for c in "UVWP":
        assert rel_l2(interior(box.download(c), h), d[f"out_{c}"]) < 1e-10, c
    ids = box.download_prop(prof, "ids")[:n]
    order = np.argsort(ids)
    ref = np.argsort(d["out_ids"])
    assert np.array_equal(ids[order], d["out_ids"][ref])
    for k in ("x", "y", "z", "u", "v", "w"):
        assert rel_l2(box.download_prop(prof, k)[:n][order], d[f"out_{k}"][ref]) < 1e-10, k
