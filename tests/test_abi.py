"""CPU-side checks of the C ABI boundary and the host mirror: the in-tree
library loads and exports every function include/lcs_mi355x.h declares; the
ctypes struct layouts match the header; configs parse like the reference's
YAML surface.  No compute call is made here (no GPU in this container)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lcs_mi355x.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lcs_[a-z_0-9]+)\s*\(", txt, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("lcs_box_create", "lcs_rhs", "lcs_hsmac", "lcs_force_linear", "lcs_force_rcf",
                 "lcs_particles_motion", "lcs_particles_makeindex", "lcs_step", "lcs_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from lcs_mi355x import _lib
    L = C.CDLL(_lib.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers the header exactly
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_struct_layouts_match_header():
    from lcs_mi355x import _lib
    # offsets the C compiler assigns for these plain structs (x86-64 SysV)
    assert C.sizeof(_lib.BoxDesc) == 5 * 4 + 4 + 3 * 8
    assert _lib.BoxDesc.U0_m_s.offset == 24
    assert C.sizeof(_lib.HsmacParams) == 40 and _lib.HsmacParams.omega.offset == 16
    assert C.sizeof(_lib.ParticleDesc) == 40 and _lib.ParticleDesc.drag.offset == 32
    assert C.sizeof(_lib.Forcing) == 40
    assert C.sizeof(_lib.StepConfig) == 16 + 40 + 40 + 8  # + localize_interval, padded
    assert _lib.StepConfig.localize_interval.offset == 96
    assert C.sizeof(_lib.StepLog) == 8 + 8 + 8 + 4 + 4 + 4 * 8


def test_c_struct_sizes_agree_with_ctypes(tmp_path):
    """Compile a tiny C program against the header and compare sizeof/offsetof."""
    import subprocess
    from lcs_mi355x import _lib
    src = tmp_path / "t.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lcs_mi355x.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu ", sizeof(lcs_box_desc),'
                   ' sizeof(lcs_hsmac_params), sizeof(lcs_particle_desc), sizeof(lcs_forcing),'
                   ' sizeof(lcs_step_config), sizeof(lcs_step_log), offsetof(lcs_step_config, forcing),'
                   ' offsetof(lcs_step_log, div1)); printf("%zu %zu %zu %zu %zu %zu\\n",'
                   ' offsetof(lcs_step_config, localize_interval), sizeof(lcs_request), offsetof(lcs_request, seed),'
                   ' sizeof(lcs_server_config), offsetof(lcs_server_config, localize_interval),'
                   ' sizeof(lcs_run_status)); printf("%zu %zu %zu\\n", sizeof(lcs_flow_stats),'
                   ' offsetof(lcs_flow_stats, C_eps), sizeof(lcs_tile_shapes));'
                   ' printf("%zu %zu %zu %zu\\n", offsetof(lcs_tile_shapes, sweep2_ty), sizeof(lcs_particle_stat_params),'
                   ' sizeof(lcs_particle_stats), offsetof(lcs_particle_stats, npairs_contact)); return 0;}\n')
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out == [C.sizeof(_lib.BoxDesc), C.sizeof(_lib.HsmacParams), C.sizeof(_lib.ParticleDesc),
                   C.sizeof(_lib.Forcing), C.sizeof(_lib.StepConfig), C.sizeof(_lib.StepLog),
                   _lib.StepConfig.forcing.offset, _lib.StepLog.div1.offset, _lib.StepConfig.localize_interval.offset,
                   C.sizeof(_lib.Request), _lib.Request.seed.offset, C.sizeof(_lib.ServerConfig),
                   _lib.ServerConfig.localize_interval.offset, C.sizeof(_lib.RunStatus), C.sizeof(_lib.FlowStat),
                   _lib.FlowStat.C_eps.offset, C.sizeof(_lib.TileShapes),
                   _lib.TileShapes.sweep2_ty.offset, C.sizeof(_lib.ParticleStatParams), C.sizeof(_lib.ParticleStats),
                   _lib.ParticleStats.npairs_contact.offset]


YAML = """
outdir: out
mode: {kind: flow-particle}
backend: {kind: mi355x}
topo: {proc_dims: [0, 0, 0], reorder: false}
grid: {dims: [24, 24, 24], halo_size: 3}
timestep: {cfl: 0.3, dtmax: 0.01}
simulate: {last_step: 3, interval_stat: 1, interval_restart: 1, interval_gc: 1, num_restart_files: 1}
flow:
  params: {U0_m_s: 0.429, L0_m: 0.005, NU_m2_s: 1.5e-5}
  init: {kind: spectrum, params: {seed: 1, k0: 4, uprime: 1}}
  couple:
    kind: hsmac
    params: {time_blocking: true, overlap: true, omega: 1.7, epsp: 0.001, itrmin: 6, itrmax: 6,
             coloring: {kind: red-black-fast}}
  forcing: {kind: rc-energy, params: {kmin: 1, kmax: 2, coarse_size: 6, energy: 0.649}}
  stat: {integral_length: true}
particles:
  - population: {valid: 432, invalid_ratio: 0.02, exchange_ratio: 0.02}
    params: {gravity: 0.0, density_ratio: 843}
    cell: {grid_ratio: 4}
    drag: {kind: nonlinear}
    init:
      id: {kind: generate}
      position: {kind: random}
      velocity: {kind: rest}
      size: {kind: const, params: {diam_m: 1.0e-5}}
"""


def test_config_parses_reference_yaml_surface():
    import yaml
    from lcs_mi355x import config
    c = config.from_dict(yaml.safe_load(YAML))
    assert c.grid.dims == (24, 24, 24) and c.grid.halo_size == 3
    assert c.flow.couple.time_blocking and c.flow.couple.coloring_id == 0
    assert c.flow.forcing.kind == "rc-energy" and c.flow.forcing.coarse_size == 6
    assert c.particles[0].population.ntotal == 441  # ceil(432 * 1.02)
    assert abs(c.flow.params.Re - 0.429 * 0.005 / 1.5e-5) < 1e-12


def test_config_errors_mirror_reference():
    import yaml
    from lcs_mi355x import config
    d = yaml.safe_load(YAML)
    d["flow"]["forcing"]["params"]["kmin"] = 0
    with pytest.raises(ValueError, match="kmin must be at least 1"):
        config.from_dict(d)
    d = yaml.safe_load(YAML)
    d["backend"]["kind"] = "cuda"
    with pytest.raises(ValueError):
        config.from_dict(d)
