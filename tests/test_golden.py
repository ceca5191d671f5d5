"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) against the oracle and against the closed forms
each phase obeys.

The oracle must reproduce every stored output from the stored inputs bit for
bit: the fixtures are the oracle's record, so a change to the restatement
that moves any result shows up here before it reaches a GPU comparison.  The
invariants pin the stored numbers independently of the oracle's code:
update! is U + dt dU/dt, HSMAC conserves each velocity component's mean and
leaves RMS divergence below its input's, the linear forcing scales every
fluctuation by one common factor, the wrap keeps positions in [0, L), and the
cell index counts every particle once with nondecreasing hashes.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle.driver import OracleSim, ParticleSet

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def gen():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.fixture(scope="module")
def G():
    O.set_threads(1)
    return gen()


def interior(A, h):
    return A[h:-h, h:-h, h:-h]


def wrap(A, h):
    return np.asfortranarray(np.pad(A, h, mode="wrap"))


@pytest.mark.parametrize("name,maker", [("flow16_h3.npz", "flow_fixture"), ("particles16.npz", "particle_fixture"),
                                        ("step24_precompile.npz", "step_fixture")])
def test_oracle_reproduces_fixture(G, name, maker):
    stored = load(name)
    fresh = getattr(G, maker)()
    assert sorted(stored) == sorted(fresh)
    for k in stored:
        assert np.array_equal(stored[k], np.asarray(fresh[k])), k


def test_flow_fixture_invariants(G):
    d = load("flow16_h3.npz")
    h = G.FLOW["h"]
    dt = G.FLOW["update_dt"]
    # update! (src/flows/update.jl:23-54): U2 = U + dt dU
    for c in "UVW":
        np.testing.assert_array_equal(d[f"upd_{c}2"], interior(d[c], h) + dt * d[f"rhs_d{c}"])
    # HSMAC: six sweeps cut the RMS of the periodic divergence of the
    # staggered field, (U[i] - U[i-1]) / dx + ..., by more than 5x (1.78 ->
    # 0.22 and 0.27 here); the opposite staggering would leave it near 1.2
    dx = O.DOMAIN_LENGTH / G.FLOW["dims"][0]

    def rms_div(U, V, W):
        D = (U - np.roll(U, 1, 0)) + (V - np.roll(V, 1, 1)) + (W - np.roll(W, 1, 2))
        return np.sqrt(np.mean((D / dx) ** 2))

    before = rms_div(*(interior(d[c], h) for c in "UVW"))
    for tag in ("tbov", "naive"):
        assert d[f"hs_{tag}_itr"] == 6
        assert rms_div(*(d[f"hs_{tag}_{c}"] for c in "UVW")) < before / 5, tag
    # linear forcing (src/flows/forcings/lf.jl:70-106): u' -> (1 + A dt) u'
    # with one A for all three components, and the means kept
    facs = []
    for c in "UVW":
        a = interior(d[c], h)
        b = d[f"lf_{c}"]
        assert abs(b.mean() - a.mean()) < 1e-14
        r = (b - b.mean()) / (a - a.mean())
        facs.append(r)
    facs = np.concatenate([f.ravel() for f in facs])
    assert np.ptp(facs) < 1e-10 and facs[0] > 1.0
    # rc-energy: injected power positive (the forcing pushes energy in)
    assert d["rcf_p"] > 0
    # CFL: dt below dtmax and below the advective bound cfl * dx / max|u|
    dx = O.DOMAIN_LENGTH / G.FLOW["dims"][0]
    assert 0 < d["cfl_dt"] <= G.FLOW["dtmax"]
    assert d["cfl_dt"] <= G.FLOW["cfl"] * dx / np.max(d["cfl_umax"]) * (1 + 1e-12)


def test_particle_fixture_invariants(G):
    d = load("particles16.npz")
    L = O.DOMAIN_LENGTH
    dt = G.PART["dt"]
    for tag in ("linear", "nonlinear"):
        for k in "xyz":
            # boundary!: a whole period added or taken where x left [0, L)
            x, xw = d[k], d[f"{tag}_{k}w"]
            assert np.all((xw >= 0) & (xw < L))
            assert np.all((xw == x) | (xw == x + L) | (xw == x - L))
            assert np.any(xw != x)
        for k, u, du in zip("xyz", "uvw", ("dudt", "dvdt", "dwdt")):
            # update! (src/particles/update.jl): x2 = x + dt u, u2 = u + dt du/dt
            np.testing.assert_array_equal(d[f"{tag}_{k}2"], d[f"{tag}_{k}w"] + dt * d[u])
            np.testing.assert_array_equal(d[f"{tag}_{u}2"], d[u] + dt * d[f"{tag}_{du}"])
    # the same wrapped positions and velocities under the two drag laws
    assert not np.array_equal(d["linear_dudt"], d["nonlinear_dudt"])
    for m in G.PART["index_sizes"]:
        hh, st, sp, perm = d[f"mi{m}_hashes"], d[f"mi{m}_starts"], d[f"mi{m}_stops"], d[f"mi{m}_perm"]
        assert len(hh) == m and np.all(np.diff(hh) >= 0)
        assert int(np.sum(sp - st + 1)) == m
        assert sorted(perm.tolist()) == list(range(1, m + 1))
        # hashes are the cells of the permuted positions
        xs = [d[k][:m][perm - 1].copy() for k in "xyz"]
        np.testing.assert_array_equal(O.hash_cells(G.PART["dims"], G.PART["grid_ratio"], xs, m), hh)
        # positions wrapped to [0, L) by boundary! land in the hashed cells' range
        assert np.all((hh >= 0))
    assert 0 < d["pdt"] <= G.PART["dtmax"]


def test_step_fixture_invariants(G):
    d = load("step24_precompile.npz")
    assert d["itr1"] == 6 and d["itr2"] == 6  # itrmin = itrmax = 6
    assert 0 < d["dt"] <= 0.01 and d["p"] > 0
    n = len(d["ids"])
    assert n == 432 and sorted(d["out_ids"].tolist()) == sorted(d["ids"].tolist())
    for k in "xyz":
        assert np.all((d[f"out_{k}"] >= 0) & (d[f"out_{k}"] < O.DOMAIN_LENGTH + 1))
    # the step moved the field: a whole step is not the identity
    assert not np.array_equal(d["out_U"], d["U"])


def test_step_fixture_from_interiors(G):
    """The step fills every ghost before reading one: stored interiors with
    periodic ghosts reproduce the step."""
    d = load("step24_precompile.npz")
    cfg = G.step_config()
    h = cfg.grid.halo_size
    Us = [wrap(d[c], h) for c in "UVW"]
    P = wrap(d["P"], h)
    xs = [d[k] for k in "xyz"]
    us = [d[k] for k in "uvw"]
    sim = OracleSim(cfg, np.float64, Us, P, [ParticleSet(xs, us, d["diams"], d["ids"], int(d["ntotal"]))])
    sim.timestep()
    sim.evolve()
    for c, A in zip("UVW", sim.Us):
        np.testing.assert_array_equal(interior(A, h), d[f"out_{c}"])
    np.testing.assert_array_equal(interior(sim.P, h), d["out_P"])
