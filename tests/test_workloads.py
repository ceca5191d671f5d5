"""Host logic of the synthetic workloads: the parameters each one is quoted
with, and the coarse size of the rc-energy forcing on sampled grids."""
import pytest

from lcs_mi355x import config as C
from lcs_mi355x.workloads import single_box


def test_c6_is_the_single_device_weak_scaling_case():
    # cfgs/N1500.lcs-yaml:35-63,67-92 at one device's share (750^3 cells,
    # 325^3 particles: data/gpu-weak.csv:2)
    cfg, _ = single_box("c6")
    assert cfg.grid.dims == (750, 750, 750) and cfg.grid.halo_size == 3 and cfg.dtype == "f64"
    assert cfg.flow.params.Re == pytest.approx(6.66 * 0.005 / 1.5e-5)
    f = cfg.flow.forcing
    assert (f.energy, f.kmin, f.kmax) == (0.66, 1, 2) and 750 % f.coarse_size == 0
    (p,) = cfg.particles
    assert p.population.valid == 325**3 and p.population.ntotal == int(1.5 * 325**3 + 0.5)
    assert p.drag == "nonlinear" and p.params.density_ratio == 843.0 and p.grid_ratio == 4
    h = cfg.flow.couple
    assert h.time_blocking and (h.omega, h.epsp, h.itrmin, h.itrmax) == (1.7, 1e-3, 6, 300)


@pytest.mark.parametrize("n,want", [(750, 10), (256, 8), (128, 8), (100, 10), (96, 8), (36, 9), (5, 5)])
def test_sampled_grid_coarse_size_divides(n, want):
    cfg, desc = single_box("c6", n_override=n)
    f = cfg.flow.forcing
    assert f.coarse_size == want and n % f.coarse_size == 0 and f.kmax <= f.coarse_size // 2
    if want != 10:
        assert f"coarse {want}" in desc


def test_sampled_linear_forcing_unchanged():
    cfg, _ = single_box("c3", n_override=64)
    assert isinstance(cfg.flow.forcing, C.ConstantPowerLF) and cfg.grid.dims == (64, 64, 64)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; asked for more it says so and exits 2
    before touching a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr and r.stdout == ""
