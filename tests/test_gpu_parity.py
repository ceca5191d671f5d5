"""Parity of the HIP path (through the C ABI) with the CPU oracle, phase by
phase, on identical seeded inputs.  Like the reference's per-phase tests
(test/test_flows.jl, test/test_particles.jl) the inputs are random whole
arrays, ghosts included.

Tolerances: integer/index/copy work is compared bit-exactly; floating point
uses a relative L2 bound stated per test (FP64 ~1e-12, FP32 ~1e-5): the HIP
kernels contract multiply-adds and reduce in a different order than the
oracle, so results differ at rounding level only.
"""
import numpy as np
import pytest

from conftest import interior, rand_halo, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 2e-5}
DTYPES = [np.float64, np.float32]


def lcs():
    import lcs_mi355x
    lcs_mi355x.load()
    return lcs_mi355x


def make_box(dims, h=3, dtype=np.float64, Re=(1.0, 1.0, 1.0)):
    L = lcs()
    return L.Box(dims, h, L.FlowParams(*Re), dtype)


def upload_all(box, arrays):
    for name, A in arrays.items():
        box.upload(name, A)


def tol(dtype):
    return TOL[np.dtype(dtype).type]


# --------------------------------------------------------------------------
def test_library_is_in_tree_native_code():
    L = lcs()
    import os
    assert os.path.dirname(L.LIBPATH).endswith(os.path.join("lcs.jl_amd", "lcs_mi355x"))
    assert b"gfx950" in L.load().lcs_version()


@pytest.mark.parametrize("full", [True, False])
@pytest.mark.parametrize("width", [1, 3])
def test_ghost_fill_bit_exact(rng, full, width):
    dims, h = (9, 10, 11), 3
    A = rand_halo(rng, dims, h)
    box = make_box(dims, h)
    box.upload("V", A)
    lcs().fill_ghost(box, ["V"], width, full)
    O.ghost_fill(A, h, width, full)
    assert np.array_equal(box.download("V"), A)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("dims", [(16, 16, 16), (12, 14, 10)])
def test_rhs(rng, dtype, dims):
    h = 3
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(4)]
    box = make_box(dims, h, dtype, (2.0, 3.0, 0.5))
    upload_all(box, dict(zip(("U", "V", "W", "P"), Us)))
    lcs().rhs(box, 1)
    ref = O.rhs(*Us, h, 12.0)
    for name, r in zip(("dU", "dV", "dW"), ref):
        assert rel_l2(box.download(name), r) < tol(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_rhs_stage2_reads_us2_and_update(rng, dtype):
    dims, h, dt = (12, 12, 12), 3, 0.013
    L = lcs()
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    Us2 = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    P = rand_halo(rng, dims, h, dtype)
    box = make_box(dims, h, dtype)
    upload_all(box, {"U": Us[0], "V": Us[1], "W": Us[2], "U2": Us2[0], "V2": Us2[1], "W2": Us2[2], "P": P})
    L.rhs(box, 2)
    dUs = O.rhs(*Us2, P, h, 1.0)
    for name, r in zip(("dU", "dV", "dW"), dUs):
        assert rel_l2(box.download(name), r) < tol(dtype)
    L.update(box, 2, dt)
    exp = [a.copy(order="F") for a in Us]
    O.update(exp, Us, [box.download(n) for n in ("dU", "dV", "dW")], h, dt)
    for name, e in zip(("U", "V", "W"), exp):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < tol(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("stage", [1, 2])
def test_fused_rhs_update(rng, dtype, stage):
    dims, h, dt = (16, 12, 20), 3, 0.007
    L = lcs()
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    Us2 = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    P = rand_halo(rng, dims, h, dtype)
    box = make_box(dims, h, dtype, (1.0, 2.0, 0.1))
    upload_all(box, {"U": Us[0], "V": Us[1], "W": Us[2], "U2": Us2[0], "V2": Us2[1], "W2": Us2[2], "P": P})
    L.rhs_update(box, stage, dt)
    src = Us if stage == 1 else Us2
    dUs = O.rhs(*src, P, h, 20.0)
    out = [a.copy(order="F") for a in (Us2 if stage == 1 else Us)]
    O.update(out, Us, dUs, h, dt)
    names = ("U2", "V2", "W2") if stage == 1 else ("U", "V", "W")
    for name, e in zip(names, out):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < tol(dtype)


HSMAC_CASES = [(tb, ov) for tb in (True, False) for ov in (True, False)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("tb,ov", HSMAC_CASES)
def test_hsmac_reference_case(rng, dtype, tb, ov):
    """test/test_flows.jl:82-101: random U,V,W,P, HSMAC(tb, ov, RedBlackFast(), 1.7, 1e-3, 6, 6)."""
    dims, h, dt = (16, 16, 16), 3, 0.01
    L = lcs()
    A = [rand_halo(rng, dims, h, dtype) for _ in range(4)]
    box = make_box(dims, h, dtype)
    upload_all(box, {"U2": A[0], "V2": A[1], "W2": A[2], "P": A[3]})
    hs = L.HSMAC(tb, ov, "red-black-fast", 1.7, 1e-3, 6, 6)
    res = L.couple(box, hs, 1, dt)
    itr, div = O.hsmac(*A, h, dt, tb, ov, 0, 1.7, 1e-3, 6, 6)
    assert res["itr"] == itr == 6
    assert abs(res["div"] - div) <= tol(dtype) * 10 * div
    for name, e in zip(("U2", "V2", "W2", "P"), A):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < tol(dtype)


@pytest.mark.parametrize("tb,ov", HSMAC_CASES)
@pytest.mark.parametrize("coloring", ["red-black-fast", "red-black"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_hsmac_multi_tile_non_cubic(rng, tb, ov, coloring, dtype):
    """Several x tiles, y tiles and z chunks of the fused sweep kernel, odd
    extents and both colourings (red-black's truncating % skips cells)."""
    dims, h, dt = (70, 21, 40), 3, 0.02
    L = lcs()
    A = [rand_halo(rng, dims, h, dtype) for _ in range(4)]
    box = make_box(dims, h, dtype)
    upload_all(box, {"U": A[0], "V": A[1], "W": A[2], "P": A[3]})
    hs = L.HSMAC(tb, ov, coloring, 1.7, 1e-3, 5, 5)
    res = L.couple(box, hs, 2, dt)
    itr, div = O.hsmac(*A, h, dt, tb, ov, 0 if coloring == "red-black-fast" else 1, 1.7, 1e-3, 5, 5)
    assert res["itr"] == itr
    assert abs(res["div"] - div) <= tol(dtype) * 10 * div
    for name, e in zip(("U", "V", "W", "P"), A):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < tol(dtype)


@pytest.mark.parametrize("tb", [True, False])
@pytest.mark.parametrize("coloring", ["red-black-fast", "red-black"])
@pytest.mark.parametrize("dtype", DTYPES)
def test_hsmac_x_slab_pair(rng, tb, coloring, dtype):
    """A box alone with at least 2^15 rows per x face slab runs both colours
    over its x slabs as one patch launch plus a commit (overlap order,
    halo.jl:349-366): 189 x 179 slab rows make ragged 32- and 16-row
    patches."""
    dims, h, dt = (6, 190, 180), 3, 0.02
    L = lcs()
    A = [rand_halo(rng, dims, h, dtype) for _ in range(4)]
    box = make_box(dims, h, dtype)
    upload_all(box, {"U": A[0], "V": A[1], "W": A[2], "P": A[3]})
    hs = L.HSMAC(tb, True, coloring, 1.7, 1e-3, 4, 4)
    res = L.couple(box, hs, 2, dt)
    itr, div = O.hsmac(*A, h, dt, tb, True, 0 if coloring == "red-black-fast" else 1, 1.7, 1e-3, 4, 4)
    assert res["itr"] == itr
    assert abs(res["div"] - div) <= tol(dtype) * 10 * div
    for name, e in zip(("U", "V", "W", "P"), A):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < tol(dtype)


@pytest.mark.parametrize("tb,ov", HSMAC_CASES)
@pytest.mark.parametrize("coloring", ["red-black-fast", "red-black"])
def test_hsmac_converging_smooth_field(tb, ov, coloring):
    """Convergence path (itrmin 6, itrmax 300, epsp 1e-3): same exit sweep."""
    dims, h, dt = (24, 24, 24), 3, 0.005
    L = lcs()
    (U, V, W), P = O.init_spectrum(dims, h, 21, 3.0, 1.0, np.float64)
    for a in (U, V, W, P):
        O.ghost_fill(a, h, h, True)
    box = make_box(dims, h)
    upload_all(box, {"U": U, "V": V, "W": W, "P": P})
    hs = L.HSMAC(tb, ov, coloring, 1.7, 1e-3, 6, 300)
    res = L.couple(box, hs, 2, dt)
    cid = 0 if coloring == "red-black-fast" else 1
    itr, div = O.hsmac(U, V, W, P, h, dt, tb, ov, cid, 1.7, 1e-3, 6, 300)
    assert res["itr"] == itr and 6 <= itr < 300
    assert np.isclose(res["div"], div, rtol=1e-10)
    for name, e in zip(("U", "V", "W", "P"), (U, V, W, P)):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < 1e-11


@pytest.mark.parametrize("dtype", DTYPES)
def test_flow_timestep(rng, dtype):
    dims, h = (16, 20, 12), 3
    Us = [(rand_halo(rng, dims, h, dtype) - dtype(0.5)) * dtype(3) for _ in range(3)]
    box = make_box(dims, h, dtype, (1.0, 1.0, 0.01))
    upload_all(box, dict(zip(("U", "V", "W"), Us)))
    dt, um = lcs().flow_timestep(box, 0.3, 0.01 if dtype == np.float64 else 1.0)
    dto, umo = O.flow_timestep(*Us, h, 0.3, 0.01 if dtype == np.float64 else 1.0, 100.0)
    assert np.isclose(dt, dto, rtol=tol(dtype))
    assert np.allclose(um, umo, rtol=tol(dtype))
    for name, e in zip(("U", "V", "W"), Us):  # face ghosts were filled in place
        assert np.array_equal(box.download(name), e)


@pytest.mark.parametrize("dtype", DTYPES)
def test_linear_forcing(rng, dtype):
    dims, h, dt = (16, 16, 16), 3, 0.01
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    box = make_box(dims, h, dtype)
    upload_all(box, dict(zip(("U", "V", "W"), Us)))
    r = lcs().forcing(box, lcs().ConstantPowerLF(1.0), dt)
    p, _ = O.force_linear(*Us, h, 1.0, dt)
    assert np.isclose(r["p"], p, rtol=tol(dtype) * 10)
    for name, e in zip(("U", "V", "W"), Us):
        assert rel_l2(box.download(name), e) < tol(dtype)


def test_boxmean_kat_on_device():
    dims, h = (8, 8, 8), 3
    A = np.zeros(tuple(n + 2 * h for n in dims), order="F")
    ref = np.arange(1, 9, dtype=np.float64).reshape((2, 2, 2), order="F")
    A[h:-h, h:-h, h:-h] = np.kron(ref, np.ones((4, 4, 4)))
    box = make_box(dims, h)
    box.upload("U", A)
    c = lcs().boxmean(box, 2)
    assert np.array_equal(c[0], ref)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("dims,c", [((16, 16, 16), 4), ((24, 24, 24), 6), ((32, 32, 32), 8),
                                    ((1040, 10, 20), 10), ((30, 20, 10), 10)])
def test_rc_energy_forcing(rng, dtype, dims, c):
    """test/test_flows.jl:150-164 uses 16^3 with coarse 4 (aliased Nyquist shell).
    The apply takes x-rows in groups of up to 1024 cells: 1040-cell rows take
    two passes each, 30-cell rows go six to a group with a ragged last group."""
    h, dt = 3, 0.01
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    box = make_box(dims, h, dtype)
    upload_all(box, dict(zip(("U", "V", "W"), Us)))
    f = lcs().EnergyPreserveRCF(1.0, 1, 2, c)
    r = lcs().forcing(box, f, dt)
    p, _ = O.force_rcf(*Us, h, 1.0, 1, 2, c, dt)
    assert np.isclose(r["p"], p, rtol=max(tol(dtype) * 10, 1e-10))
    for name, e in zip(("U", "V", "W"), Us):
        assert rel_l2(box.download(name), e) < tol(dtype)


# ---------------------------------------------------------------------------
# particles
# ---------------------------------------------------------------------------
def particle_box(rng, dtype, n, dims=(16, 16, 16), h=3, drag="linear", ntotal=None):
    L = lcs()
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(6)]
    box = make_box(dims, h, dtype, (1.0, 1.0, 0.02))
    upload_all(box, dict(zip(("U", "V", "W", "U2", "V2", "W2"), Us)))
    ntotal = n if ntotal is None else ntotal
    prof = box.add_particles(n, ntotal, gravity=9.8, density_ratio=5.0, drag=drag, grid_ratio=4)
    props = {}
    for name in L.core._lib.PROPS:
        if name == "ids":
            a = np.full(ntotal, O.INVALID, np.int64)
            a[:n] = rng.permutation(n) + 1
        elif name == "diams":
            a = np.zeros(ntotal, dtype)
            a[:n] = rng.uniform(0.005, 0.05, n)
        elif name in ("x", "y", "z", "x2", "y2", "z2"):
            a = np.zeros(ntotal, dtype)
            a[:n] = rng.uniform(-0.3, O.DOMAIN_LENGTH + 0.3, n)
        else:
            a = np.zeros(ntotal, dtype)
            a[:n] = rng.standard_normal(n)
        props[name] = a.astype(np.int64 if name == "ids" else dtype)
        box.upload_prop(prof, name, props[name])
    return box, prof, Us, props


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("drag", ["linear", "nonlinear"])
@pytest.mark.parametrize("stage", [1, 2])
def test_particle_phases(rng, dtype, drag, stage):
    n, h, dt = 3000, 3, 0.003
    L = lcs()
    box, prof, Us, p = particle_box(rng, dtype, n, drag=drag)
    x, x2 = ("x", "y", "z"), ("x2", "y2", "z2")
    u, u2 = ("u", "v", "w"), ("u2", "v2", "w2")
    dudt = ("dudt", "dvdt", "dwdt")
    # boundary!
    L.boundary(box, prof, stage)
    a, s = (x, x2) if stage == 1 else (x2, x)
    O.p_boundary([p[k] for k in a], [p[k] for k in s], n)
    for k in x + x2:
        assert np.array_equal(box.download_prop(prof, k), p[k])
    # motion!
    L.motion(box, prof, stage)
    fl = Us[:3] if stage == 1 else Us[3:]
    O.p_motion(fl, h, 50.0, 9.8, 5.0, 0 if drag == "linear" else 1, [p[k] for k in (x if stage == 1 else x2)],
               [p[k] for k in (u if stage == 1 else u2)], p["diams"], [p[k] for k in dudt], n)
    for k in dudt:
        assert rel_l2(box.download_prop(prof, k), p[k]) < tol(dtype) * 10
    # update!
    L.particle_update(box, prof, stage, dt)
    dst_x, dst_u = (x2, u2) if stage == 1 else (x, u)
    O.p_update([p[k] for k in dst_x], [p[k] for k in x], [p[k] for k in dst_u], [p[k] for k in u],
               [box.download_prop(prof, k) for k in dudt], n, dt)
    for k in dst_x + dst_u:
        assert rel_l2(box.download_prop(prof, k), p[k]) < tol(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_particle_cell_indices_bit_exact_through_motion_positions(rng, dtype):
    """Cell indices depend only on positions: with identical positions the
    device hashes, starts/stops and permutation are bit-identical."""
    n = 20000
    L = lcs()
    box, prof, Us, p = particle_box(rng, dtype, n, ntotal=n + 77)
    L.makeindex(box, prof, 1)
    hashes, starts, stops = box.cellindex(prof)
    xs = [p[k][:n].copy() for k in ("x", "y", "z")]
    ids = p["ids"][:n].copy()
    others = [p[k][:n].copy() for k in L.core._lib.PROPS[1:] if k not in ("x", "y", "z")]
    hh, perm, st, sp = O.makeindex((16, 16, 16), 4, xs, ids, others, n)
    assert np.array_equal(hashes, hh)
    assert np.array_equal(starts, st) and np.array_equal(stops, sp)
    assert np.array_equal(box.download_prop(prof, "ids")[:n], ids)
    assert np.array_equal(box.download_prop(prof, "ids")[n:], p["ids"][n:])
    assert np.array_equal(box.download_prop(prof, "x")[:n], xs[0])
    assert int(np.sum(sp - st + 1)) == n


@pytest.mark.parametrize("n", [0, 1, 7, 300000])
def test_makeindex_sizes(rng, n):
    L = lcs()
    dims = (32, 32, 32)
    box = make_box(dims, 3)
    prof = box.add_particles(n, n + 5, grid_ratio=4)
    xs = [rng.uniform(-0.1, O.DOMAIN_LENGTH + 0.1, n + 5) for _ in range(3)]
    for k, a in zip(("x", "y", "z"), xs):
        box.upload_prop(prof, k, a)
    L.makeindex(box, prof, 1)
    hashes, starts, stops = box.cellindex(prof)
    h0 = O.hash_cells(dims, 4, [a[:n].copy() for a in xs], n)
    order = np.argsort(h0, kind="stable")
    assert np.array_equal(hashes, h0[order])
    assert int(np.sum(stops - starts + 1)) == n
    assert np.array_equal(box.download_prop(prof, "x")[:n], xs[0][:n][order])


@pytest.mark.parametrize("dtype", DTYPES)
def test_particle_timestep(rng, dtype):
    L = lcs()
    box, prof, Us, p = particle_box(rng, dtype, 1000, ntotal=1200)
    dt = L.particle_timestep(box, prof, 10.0)
    assert np.isclose(dt, O.p_timestep(p["diams"], 50.0, 5.0, 10.0), rtol=tol(dtype))


# ---------------------------------------------------------------------------
# builder-defined initialisers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", DTYPES)
def test_init_spectrum_matches_oracle(dtype):
    dims, h = (32, 24, 16), 3
    box = make_box(dims, h, dtype)
    box.init_spectrum(5, 4.0, 1.0)
    (U, V, W), P = O.init_spectrum(dims, h, 5, 4.0, 1.0, dtype)
    for name, e in zip(("U", "V", "W", "P"), (U, V, W, P)):
        assert rel_l2(interior(box.download(name), h), interior(e, h)) < max(tol(dtype), 1e-12)


@pytest.mark.parametrize("dtype", DTYPES)
def test_init_particles_matches_oracle(dtype):
    n = 5000
    box = make_box((16, 16, 16), 3, dtype, (1.0, 1.0, 0.01))
    prof = box.add_particles(n, n + 10, density_ratio=843.0)
    box.init_particles(prof, 17, 0.1, 10.0, 0.5, zero_slip=False)
    xs, diams, ids = O.init_particles(17, n, 0.1, 10.0, 0.5, 100.0, 843.0, dtype)
    for k, e in zip(("x", "y", "z"), xs):
        assert np.array_equal(box.download_prop(prof, k)[:n], e)
    assert np.allclose(box.download_prop(prof, "diams")[:n], diams, rtol=1e-14 if dtype == np.float64 else 1e-6)
    got = box.download_prop(prof, "ids")
    assert np.array_equal(got[:n], ids) and np.all(got[n:] == O.INVALID)
