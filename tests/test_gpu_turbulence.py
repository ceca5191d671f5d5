"""Forced isotropic turbulence against the reference's published statistics.

data/flow-stats.csv (rows N = 128 and 256, kept in tests/golden/
flow_stats_reference.csv) lists the stationary statistics of the reference's
own runs: u_rms, Re_lambda, k_max eta, L, -S (velocity-derivative skewness)
and F (flatness), as Flows.stat computes them (src/flows/stat.jl:66-182).
The runs here use the reference's settings for those sizes
(cfgs/N256.lcs-yaml: rc-energy forcing with kmin 1, kmax 2, energy 0.649,
coarse size 8; cfl 0.3, dtmax 0.01; HSMAC as cfgs/N128.lcs-yaml:36-46;
U0 0.429 m/s at N = 128 and 1.08 m/s at N = 256 with L0 5 mm and nu
1.5e-5 m^2/s, i.e. Re = 143 and 360), FP32, from this build's seeded
spectrum (the reference's random-flow RNG is Julia's).  The statistics are
averaged every 50 steps over the second half of the run (steps 3000-6000 at
N = 128, 4000-8000 at N = 256, past the transient of the seeded start) and
compared with the row: u_rms within 5 %, Re_lambda 8 %, k_max eta 5 %,
F 10 %, -S 15 %, and L / 2 pi 25 % (the CSV's L column matches the integral
length divided by the domain length 2 pi; its header does not say so).  The
reference's N128 config file forces with constant power instead; the
CSV's N = 128 row (epsilon about 0.3, not the 0.5 that constant power 0.5
would set) matches the rc-energy runs.
"""
import csv
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {"u_rms": 0.05, "Re_lambda": 0.08, "kmax_leta": 0.05, "F": 0.10, "neg_S": 0.15, "L": 0.25}


def reference_row(n):
    with open(os.path.join(GOLDEN, "flow_stats_reference.csv")) as fh:
        for r in csv.DictReader(fh):
            if int(r["N"]) == n:
                return {k: float(v) for k, v in r.items() if k != "N"}
    raise KeyError(n)


def stationary_stats(n, U0, nsteps, every=50):
    import lcs_mi355x as L
    from lcs_mi355x import config as C
    cfg = C.Config(grid=C.Grid((n, n, n), 3), timestep=C.TimeStep(0.3, 0.01), simulate=C.SimulateConfig(0),
                   flow=C.FlowConfig(C.FlowParams(U0, 0.005, 1.5e-5), C.SpectrumInit(7, 4.0, 1.0),
                                     C.HSMAC(True, True, "red-black-fast", 1.7, 1e-3, 6, 300),
                                     C.EnergyPreserveRCF(0.649, 1, 2, 8), C.FlowStatConfig(True)),
                   particles=(), mode="flow", dtype="f32")
    box = L.build(cfg)
    sc = L.step_config(cfg)
    rows = []
    for s in range(1, nsteps + 1):
        L.step(box, sc, want_log=False)
        if s > nsteps // 2 and s % every == 0:
            st = L.flow_stat(box, True)
            rows.append([st["urms"], st["Re_lambda"], st["k_eta"], st["flat"], st["skew"], st["L"] / (2 * math.pi)])
    box.close()
    return dict(zip(["u_rms", "Re_lambda", "kmax_leta", "F", "neg_S", "L"], np.mean(rows, axis=0)))


@pytest.mark.parametrize("n,U0,nsteps", [(128, 0.429, 6000), (256, 1.08, 8000)])
def test_stationary_statistics_match_the_reference(n, U0, nsteps):
    ref = reference_row(n)
    got = stationary_stats(n, U0, nsteps)
    print(f"N={n}: " + ", ".join(f"{k} {got[k]:.4g} (reference {ref[k]:.4g})" for k in TOL), flush=True)
    for k, tol in TOL.items():
        assert abs(got[k] - ref[k]) <= tol * abs(ref[k]), (k, got[k], ref[k])
