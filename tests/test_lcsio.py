"""CPU checks of the restart/statistics files (LCSIO) and the output-directory
policy.  No GPU call: statistics files and host-array restart files are
written through liblcs_io.so, and the layout is checked both with this
package's reader and with HDF5's own ``h5dump`` (an independent reader),
against the assertions of the reference's tests (test/test_LCSIO.jl:22-128,
test/test_integration_mpi.jl:68-129,161-237)."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
import yaml

from conftest import ROOT

IOHEADER = os.path.join(ROOT, "include", "lcs_io.h")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(
    not os.path.exists(os.path.join(ROOT, "lcs.jl_amd", "lcs_mi355x", "liblcs_io.so")),
    reason="liblcs_io.so is built only where the HDF5 headers are installed")


def io_functions():
    txt = re.sub(r"/\*.*?\*/", "", open(IOHEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lcs_[a-z_0-9]+)\s*\(", txt, flags=re.M)))


def test_io_library_exports_every_declared_symbol():
    from lcs_mi355x import lcsio
    L = lcsio.load_io()
    names = io_functions()
    assert "lcs_save_restart" in names and "lcs_load_flow_fields" in names and "lcs_save_stat" in names
    assert not [n for n in names if not hasattr(L, n)]
    assert set(lcsio.IO_SIGNATURES) == set(names)


def test_h5_entry_layout_matches_header(tmp_path):
    from lcs_mi355x import lcsio
    src = tmp_path / "t.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lcs_io.h"\n'
                   'int main(void){printf("%zu %zu %zu\\n", sizeof(lcs_h5_entry), offsetof(lcs_h5_entry, dims),'
                   ' offsetof(lcs_h5_entry, data)); return 0;}\n')
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out == [C.sizeof(lcsio.H5Entry), lcsio.H5Entry.dims.offset, lcsio.H5Entry.data.offset]


# ---- statistics file --------------------------------------------------------
def sample_stat(step):
    from lcs_mi355x import lcsio
    state = {"step": step, "t": 0.125 * step, "dt": 0.125, "particles": [{"nvalid": 432}]}
    summ = {"mean": 0.5, "max": 1.0, "min": -1.0}
    flow = {"u": summ, "v": summ, "w": summ, "k": 1.5, "eps": 0.25, "tau_eta": 0.3, "l_mu": 0.4, "l_eta": 0.01,
            "Re_lambda": 40.0, "urms": 1.0, "k_eta": 1.2, "skew": -0.5, "flat": 4.0, "L": 1.1, "C_eps": 0.6}
    part = {c: summ for c in ("x", "y", "z", "u", "v", "w", "diam")}
    part.update(taup=0.02, St=0.07, rdf={"gr_contact": 1.5, "npairs_contact": 7, "edges": np.linspace(0, 1, 11),
                                         "gr": np.ones(10), "npairs": np.arange(10)},
                spectrum={"wavenumbers": np.array([1.0, 2.0, 3.0]), "Enp": np.array([0.1, 0.2, 0.3])})
    stat = lcsio.stat_tree(state, flow, [part])
    log = lcsio.log_tree({"flow": {"couple1": {"itr": 6, "div": 1e-3}, "couple2": {"itr": 7, "div": 2e-3},
                                   "forcing": {"p": 0.5, "e10": 0.9}}}, "rc-energy", [432])
    return state, stat, log


def test_savestat_layout(tmp_path):
    from lcs_mi355x import lcsio
    f = str(tmp_path / "stat.h5")
    for step in (1, 2):
        lcsio.savestat(f, *sample_stat(step))
    # test_integration_mpi.jl:82-96 and :219-227
    assert lcsio.keys(f) == ["1", "2"]
    assert lcsio.keys(f, "1") == ["log", "stat", "state"]
    saved_state = lcsio.deepread(f, "1/state")
    assert saved_state["step"] == 1 and saved_state["particles"]["1"]["nvalid"] == 432
    saved_stat = lcsio.deepread(f, "2/stat")
    assert set(saved_stat) == {"state", "flow", "particles"}
    assert saved_stat["state"]["step"] == 2
    # the reference's FlowStat / ParticleStat field names (src/flows/stat.jl:23-41, particles/stat.jl:64-78)
    assert set(saved_stat["flow"]) == {"u", "v", "w", "k", "ϵ", "τη", "lμ", "lη", "Reλ", "urms", "kη", "skew",
                                       "flat", "L", "Cϵ"}
    p1 = saved_stat["particles"]["1"]
    assert p1["τp"] == 0.02 and p1["rdf"]["npairs_contact"] == 7
    assert np.array_equal(p1["rdf"]["npairs"], np.arange(10)) and p1["rdf"]["npairs"].dtype == np.int64
    assert np.array_equal(p1["spectrum"]["Enp"], [0.1, 0.2, 0.3])
    log = lcsio.deepread(f, "1/log")
    assert log["flow"]["couple1"] == {"itr": 6, "div": 1e-3}
    assert log["flow"]["forcing"] == {"p": 0.5, "e10": 0.9}
    assert log["particles"]["boundary1"]["1"]["nvalid"] == {"mean": 432, "min": 432, "max": 432}
    # savestat replaces a step's group (save.jl:24-26)
    st, sa, lg = sample_stat(2)
    st["t"] = 9.0
    lcsio.savestat(f, st, sa, lg)
    assert lcsio.keys(f) == ["1", "2"] and lcsio.h5_read(f, "2/state/t") == 9.0


def test_savestat_skips_empty_and_none(tmp_path):
    from lcs_mi355x import lcsio
    f = str(tmp_path / "stat.h5")
    lcsio.savestat(f, {"step": 3, "t": 1.0, "dt": 0.1, "flow": {}, "particles": None},
                   {"state": {"step": 3}, "flow": {"k": 1.0}, "particles": []}, {})
    assert lcsio.keys(f, "3") == ["stat", "state"]
    assert lcsio.keys(f, "3/state") == ["dt", "step", "t"]
    assert lcsio.keys(f, "3/stat") == ["flow", "state"]
    with pytest.raises(TypeError, match="cannot write value of type bool"):
        lcsio.savestat(f, {"step": 4, "flag": True}, {}, {})


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
def test_savestat_read_by_h5dump(tmp_path):
    from lcs_mi355x import lcsio
    f = str(tmp_path / "stat.h5")
    lcsio.savestat(f, *sample_stat(5))
    out = subprocess.run([H5DUMP, "-d", "/5/stat/flow/τη", f], capture_output=True, text=True, check=True).stdout
    assert "H5T_IEEE_F64LE" in out and "SCALAR" in out and "0.3" in out
    out = subprocess.run([H5DUMP, "-d", "/5/stat/particles/1/rdf/npairs", f], capture_output=True, text=True,
                         check=True).stdout
    assert "H5T_STD_I64LE" in out and "( 10 )" in out


# ---- restart file of host arrays: the column-major layout --------------------
def host_restart(tmp_path, dtype=np.float64, nvalid=5, dims=(6, 5, 4)):
    from lcs_mi355x import lcsio
    rng = np.random.default_rng(1)
    nx, ny, nz = dims
    flow = {n: rng.random((nz, ny, nx)).astype(dtype) for n in ("U", "V", "W", "P")}  # C order of (nx, ny, nz)
    parts = {"counts": np.array([nvalid], np.int64), "id": np.arange(1, nvalid + 1, dtype=np.int64)}
    parts.update({k: rng.random(nvalid).astype(dtype) for k in ("x", "y", "z", "u", "v", "w", "diam")})
    f = str(tmp_path / "restart.h5.7")
    lcsio.h5_write(f, {"flow": flow, "particles": {"1": parts}})
    return f, flow, parts


def test_restart_layout_from_host_arrays(tmp_path):
    from lcs_mi355x import lcsio
    f, flow, parts = host_restart(tmp_path)
    # test_integration_mpi.jl:101-129
    assert lcsio.keys(f) == ["flow", "particles"]
    assert lcsio.keys(f, "flow") == ["P", "U", "V", "W"]
    assert set(lcsio.keys(f, "particles/1")) == {"counts", "id", "x", "y", "z", "u", "v", "w", "diam"}
    kind, typ, dims = lcsio.h5_info(f, "flow/U")
    assert kind == lcsio.H5_DATASET and typ == lcsio.H5_F64 and dims == (6, 5, 4)  # Julia size (nx, ny, nz)
    assert np.array_equal(lcsio.h5_read(f, "flow/U"), flow["U"])
    assert lcsio.h5_read(f, "particles/1/counts").tolist() == [5]
    assert np.array_equal(lcsio.h5_read(f, "particles/1/diam"), parts["diam"])
    with pytest.raises(Exception, match="already exists"):
        lcsio.h5_write(f, {"flow": {"U": flow["U"]}})


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
def test_restart_bytes_are_column_major(tmp_path):
    """h5dump's raw dump of flow/U is the field in column-major order, x
    fastest: the bytes HDF5.jl reads back as a (nx, ny, nz) Julia array."""
    f, flow, _ = host_restart(tmp_path)
    hdr = subprocess.run([H5DUMP, "-H", "-d", "/flow/U", f], capture_output=True, text=True, check=True).stdout
    assert "( 4, 5, 6 )" in hdr  # HDF5 (C) order of the Julia size (6, 5, 4)
    raw = tmp_path / "U.bin"
    subprocess.run([H5DUMP, "-d", "/flow/U", "-b", "LE", "-o", str(raw), f], capture_output=True, check=True)
    got = np.fromfile(raw, np.float64)
    U = flow["U"]  # U[k, j, i] is the reference's U[i+1, j+1, k+1]
    col_major = np.array([U[k, j, i] for k in range(4) for j in range(5) for i in range(6)])
    assert np.array_equal(got, col_major)


def test_h5_read_errors(tmp_path):
    from lcs_mi355x import lcsio
    from lcs_mi355x._lib import LCSArgumentError
    f, _, _ = host_restart(tmp_path)
    with pytest.raises(KeyError):
        lcsio.h5_read(f, "flow/Q")
    with pytest.raises(LCSArgumentError, match="does not exist"):
        lcsio.h5_read(str(tmp_path / "none.h5"), "flow/U")
    assert not lcsio.haskey(f, "flow/Q") and not lcsio.haskey(f, "a/b/c") and lcsio.haskey(f, "flow")


# ---- output directory and resume policy (test/test_LCSIO.jl) -----------------
def write_yaml(tmp_path, extra="", name="run.lcs-yaml"):
    from test_abi import YAML
    p = tmp_path / name
    p.write_text(YAML + extra)
    return str(p)


def test_readconfig_resume_flag_and_patch(tmp_path):
    import dataclasses

    from lcs_mi355x import config as Cfg
    src = write_yaml(tmp_path)
    c = Cfg.readconfig(src)
    assert not c.resume
    with pytest.raises(ValueError, match="resume mode is controlled by LCS_RESUME environment variable"):
        Cfg.readconfig(write_yaml(tmp_path, "\nresume: true\n", "resume.lcs-yaml"))
    c2 = Cfg.readconfig(src, patch=lambda c: dataclasses.replace(
        c, timestep=dataclasses.replace(c.timestep, cfl=c.timestep.cfl * 2)))
    assert c2.timestep.cfl == 2 * c.timestep.cfl
    c3 = Cfg.readconfig(src, patch=lambda c: dataclasses.replace(c, grid=Cfg.Grid((32, 32, 32), 3)))
    assert c3.grid.dims == (32, 32, 32)


def test_config_yaml_round_trip():
    """YAML -> Config -> YAML -> Config (test/test_base.jl:9-27)."""
    from lcs_mi355x import config as Cfg
    from test_abi import YAML
    c = Cfg.from_dict(yaml.safe_load(YAML))
    d = Cfg.to_dict(c)
    assert Cfg.from_dict(yaml.safe_load(yaml.safe_dump(d))) == c
    assert Cfg.to_dict(Cfg.from_dict(d)) == d


def test_isresumable(tmp_path):
    import dataclasses

    from lcs_mi355x import config as Cfg, lcsio
    cfg = dataclasses.replace(Cfg.readconfig(write_yaml(tmp_path)), outdir=str(tmp_path / "out"))
    assert not lcsio.isresumable(cfg)
    os.makedirs(cfg.outdir)
    assert not lcsio.isresumable(cfg)
    open(os.path.join(cfg.outdir, lcsio.FILE_CONFIG), "w").close()
    assert not lcsio.isresumable(cfg)
    open(os.path.join(cfg.outdir, "restart.h5.1"), "w").close()
    assert lcsio.isresumable(cfg)


def test_init_policy(tmp_path, monkeypatch):
    import dataclasses

    from lcs_mi355x import config as Cfg, lcsio
    base = Cfg.readconfig(write_yaml(tmp_path))
    # outdir exists without LCS_RESUME
    d = tmp_path / "a"
    d.mkdir()
    monkeypatch.delenv("LCS_RESUME", raising=False)
    with pytest.raises(ValueError, match="already exists.*LCS_RESUME=1"):
        lcsio.init(dataclasses.replace(base, outdir=str(d)))
    # outdir exists with LCS_RESUME but not resumable
    monkeypatch.setenv("LCS_RESUME", "1")
    with pytest.raises(ValueError, match="not resumable"):
        lcsio.init(dataclasses.replace(base, outdir=str(d)))
    # resumable: flow and particles restart from the latest restart file
    open(d / "config.lcs-yaml", "w").close()
    for s in (2, 10, 9):
        open(d / f"restart.h5.{s}", "w").close()
    r = lcsio.init(dataclasses.replace(base, outdir=str(d)))
    assert r.resume
    assert r.flow.init.file == str(d / "restart.h5.10")
    assert r.particles[0].init.restart_file == str(d / "restart.h5.10") and r.particles[0].init.restart_profile == 1
    # outdir does not exist with LCS_RESUME: a new run, config saved
    e = tmp_path / "b"
    r = lcsio.init(dataclasses.replace(base, outdir=str(e)))
    assert not r.resume and os.path.isdir(e)
    saved = Cfg.readconfig(str(e / "config.lcs-yaml"))
    assert saved == dataclasses.replace(base, outdir=str(e))


def test_restart_file_names_and_rotation(tmp_path):
    import dataclasses

    from lcs_mi355x import config as Cfg, lcsio
    assert lcsio.extract_step("/x/restart.h5.12") == "12"
    with pytest.raises(ValueError, match="invalid restart file name"):
        lcsio.extract_step("/x/restart.h5")
    with pytest.raises(ValueError, match="no restart files"):
        lcsio.find_latest_restart_file(str(tmp_path))
    cfg = dataclasses.replace(Cfg.readconfig(write_yaml(tmp_path)), outdir=str(tmp_path),
                              simulate=Cfg.SimulateConfig(10, 1, 2, 0, 2))
    for s in (2, 4, 6):
        open(tmp_path / f"restart.h5.{s}", "w").close()
    lcsio.cleanrestart(6, cfg)  # removes step 6 - 2*2 = 2 (src/LCSIO/utils.jl:13-22)
    assert sorted(os.path.basename(f) for f in lcsio.find_restart_files(str(tmp_path))) == ["restart.h5.4",
                                                                                          "restart.h5.6"]
    assert lcsio.oninterval(6, 2) and not lcsio.oninterval(5, 2) and not lcsio.oninterval(5, 0)
