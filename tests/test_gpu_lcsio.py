"""Restart files of device boxes (lcs_save_restart, lcs_load_flow_fields,
lcs_load_particles: src/LCSIO/save.jl:40-167, load.jl:23-118) and the
simulate loop's stat/restart/resume cadence (src/simulate.jl:54-88), after
test/test_integration_mpi.jl:68-237.

* a saved restart file holds the box interior and the first nvalid particles
  bit for bit, in the reference's layout;
* loading it into a fresh box reproduces fields (ghosts included) and
  particles bit for bit, and one more step from both boxes gives identical
  logs, fields and particles;
* a file of FP64 host arrays loads into an FP32 box as float32(value);
* size, process-count and particle-count mismatches are errors;
* simulate() with interval_stat = interval_restart = 1 writes config, stat and
  rotated restart files, and a run stopped and resumed with LCS_RESUME=1
  ends bit for bit where the uninterrupted run ends.
"""
import dataclasses
import os

import numpy as np
import pytest

from test_gpu_step import c1_config

pytestmark = pytest.mark.gpu

H = 3
PROPS = (("id", "ids"), ("x", "x"), ("y", "y"), ("z", "z"), ("u", "u"), ("v", "v"), ("w", "w"), ("diam", "diams"))


def small(dtype="f64", dims=(16, 12, 20), n=300, forcing=None):
    return c1_config(dims=dims, nparticles=n, dtype=dtype, forcing=forcing)


def interior_of(box, name):
    return box.download(name)[H:-H, H:-H, H:-H]


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_saved_restart_holds_the_box(tmp_path, dtype):
    import lcs_mi355x as L
    from lcs_mi355x import lcsio
    cfg = small(dtype)
    box = L.build(cfg)
    L.step(box, L.step_config(cfg))
    f = lcsio.saverestart(box, str(tmp_path / "restart.h5.1"))
    assert lcsio.keys(f) == ["flow", "particles"]
    assert lcsio.keys(f, "flow") == ["P", "U", "V", "W"]
    assert set(lcsio.keys(f, "particles/1")) == {"counts", "id", "x", "y", "z", "u", "v", "w", "diam"}
    assert lcsio.h5_info(f, "flow/U")[1:] == (lcsio.H5_F64 if dtype == "f64" else lcsio.H5_F32, (16, 12, 20))
    for name in ("U", "V", "W", "P"):
        assert np.array_equal(lcsio.h5_read(f, f"flow/{name}").T, interior_of(box, name)), name
    nv = box.profiles[0]["nvalid"]
    assert lcsio.h5_read(f, "particles/1/counts").tolist() == [nv]
    for key, prop in PROPS:
        assert np.array_equal(lcsio.h5_read(f, f"particles/1/{key}"), box.download_prop(0, prop)[:nv]), key
    with pytest.raises(L.LCSError, match="already holds group flow"):
        lcsio.saverestart(box, f)
    box.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_reload_reproduces_and_continues(tmp_path, dtype):
    import lcs_mi355x as L
    from lcs_mi355x import lcsio
    cfg = small(dtype, forcing=None)
    scfg = L.step_config(cfg)
    a = L.build(cfg)
    L.step(a, scfg)
    log_a = L.read_log(a)
    f = lcsio.saverestart(a, str(tmp_path / "restart.h5.1"))
    L.fill_ghost(a, ("U", "V", "W", "P"), H, True)
    b = L.build(dataclasses.replace(cfg, flow=dataclasses.replace(cfg.flow, init=L.RestartInit(f))))
    lcsio.load_particles(f, b, 0)
    L.set_state(b, log_a["step"], log_a["t"])
    nv = a.profiles[0]["nvalid"]
    for name in ("U", "V", "W", "P"):  # interior and ghosts
        assert np.array_equal(b.download(name), a.download(name)), name
    for _, prop in PROPS:
        assert np.array_equal(b.download_prop(0, prop)[:nv], a.download_prop(0, prop)[:nv]), prop
    la, lb = L.step(a, scfg), L.step(b, scfg)
    assert la == lb
    for name in ("U", "V", "W", "P"):
        assert np.array_equal(interior_of(b, name), interior_of(a, name)), name
    for _, prop in PROPS:
        assert np.array_equal(b.download_prop(0, prop)[:nv], a.download_prop(0, prop)[:nv]), prop
    a.close()
    b.close()


def test_fp64_file_into_fp32_box(tmp_path):
    import lcs_mi355x as L
    from lcs_mi355x import lcsio
    rng = np.random.default_rng(3)
    nx, ny, nz = 16, 12, 20
    flow = {n: rng.standard_normal((nz, ny, nx)) for n in ("U", "V", "W", "P")}
    f = str(tmp_path / "restart.h5.0")
    lcsio.h5_write(f, {"flow": flow})
    box = L.build(small("f32"))
    lcsio.load_flow_fields(f, box)
    for name in ("U", "V", "W", "P"):
        assert np.array_equal(interior_of(box, name), flow[name].T.astype(np.float32)), name
    box.close()


def test_mismatches_are_errors(tmp_path):
    import lcs_mi355x as L
    from lcs_mi355x import lcsio
    box = L.build(small("f64"))
    nv = box.profiles[0]["nvalid"]
    f = str(tmp_path / "restart.h5.0")
    lcsio.h5_write(f, {"flow": {n: np.zeros((8, 8, 8)) for n in ("U", "V", "W", "P")},
                       "particles": {"1": {"counts": np.array([nv, 1]), "id": np.arange(nv + 1)},
                                     "2": {"counts": np.array([nv + 1])}}})
    with pytest.raises(L.LCSArgumentError, match=r"dataset U size mismatch, expected \(16, 12, 20\) got \(8, 8, 8\)"):
        lcsio.load_flow_fields(f, box)
    with pytest.raises(L.LCSArgumentError, match="process count mismatch, expected 1 got 2"):
        lcsio.load_particles(f, box, 0)
    with pytest.raises(L.LCSError, match="no such particle profile"):
        lcsio.load_particles(f, box, 1)
    g = str(tmp_path / "restart.h5.1")
    lcsio.h5_write(g, {"particles": {"1": {"counts": np.array([nv + 1])}}})
    with pytest.raises(L.LCSArgumentError, match="particle count mismatch"):
        lcsio.load_particles(g, box, 0)
    box.close()


def run_config(outdir, last_step):
    from lcs_mi355x import config as C
    cfg = small("f64", dims=(16, 16, 16), n=200, forcing=C.EnergyPreserveRCF(0.649, 1, 2, 4))
    return dataclasses.replace(cfg, outdir=str(outdir), simulate=C.SimulateConfig(last_step, 1, 1, 0, 1),
                               flow=dataclasses.replace(cfg.flow, stat=C.FlowStatConfig(True)))


def test_simulate_writes_and_resumes(tmp_path, monkeypatch):
    import lcs_mi355x as L
    from lcs_mi355x import lcsio
    monkeypatch.delenv("LCS_RESUME", raising=False)
    # uninterrupted: steps 1..3 (last_step 2 runs state.step..last_step = 0..2)
    full, _ = L.simulate(run_config(tmp_path / "full", 2), quiet=True)
    # stopped after step 1 (last_step 0), then resumed up to last_step 2
    out = tmp_path / "part"
    box1, logs1 = L.simulate(run_config(out, 0), quiet=True)
    assert [g["step"] for g in logs1] == [1]
    files = sorted(os.listdir(out))
    assert files == ["config.lcs-yaml", "restart.h5.1", "stat.h5"]
    st = str(out / "stat.h5")
    assert lcsio.keys(st) == ["1"] and lcsio.keys(st, "1") == ["log", "stat", "state"]
    assert lcsio.deepread(st, "1/state")["step"] == 1
    saved = lcsio.deepread(st, "1/stat")
    assert set(saved) == {"state", "flow", "particles"} and saved["state"]["step"] == 1
    assert saved["flow"]["L"] > 0 and saved["particles"]["1"]["rdf"]["npairs"].dtype == np.int64
    assert lcsio.h5_read(str(out / "restart.h5.1"), "particles/1/counts").sum() == 200
    box1.close()
    with pytest.raises(ValueError, match="already exists"):
        L.simulate(run_config(out, 2), quiet=True)
    monkeypatch.setenv("LCS_RESUME", "1")
    box2, logs2 = L.simulate(L.readconfig(str(out / "config.lcs-yaml"),
                                          patch=lambda c: dataclasses.replace(
                                              c, simulate=dataclasses.replace(c.simulate, last_step=2))),
                             quiet=True)
    assert [g["step"] for g in logs2] == [2, 3]
    assert sorted(os.listdir(out)) == ["config.lcs-yaml", "restart.h5.3", "stat.h5"]  # one file kept
    assert lcsio.keys(st) == ["1", "2", "3"]
    assert L.read_log(box2) == L.read_log(full)
    for name in ("U", "V", "W", "P"):
        assert np.array_equal(interior_of(box2, name), interior_of(full, name)), name
    nv = full.profiles[0]["nvalid"]
    for _, prop in PROPS:
        assert np.array_equal(box2.download_prop(0, prop)[:nv], full.download_prop(0, prop)[:nv]), prop
    box2.close()
    full.close()
