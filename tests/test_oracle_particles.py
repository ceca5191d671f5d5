"""Pins the CPU oracle's particle phases (no GPU needed): interpolation
exactness, drag KATs (test/test_particles.jl:96-97), the RK2 drag-relaxation
closed form, wrap invariants (test/test_particles_mpi.jl:46-61), hasher and
binary-search KATs (test/test_utils.jl:77-127) and makeindex against a
stable sort plus the reference's invariant (test/test_particles.jl:85)."""
import bisect

import numpy as np
import pytest

from conftest import rand_halo
from test_oracle_flow import field_from
from oracle import oracle as O

TWO_PI = O.DOMAIN_LENGTH


def linear_fields(dims, h, coefs):
    return [field_from(lambda x, y, z, c=c: c[0] + c[1] * x + c[2] * y + c[3] * z, dims, h, axis)
            for axis, c in enumerate(coefs)]


def test_trilinear_exact_for_linear_fields(rng):
    dims, h = (8, 12, 16), 1
    coefs = [(0.1, 0.5, -0.2, 0.3), (-0.4, 0.2, 0.7, -0.1), (0.9, -0.3, 0.05, 0.6)]
    Us = linear_fields(dims, h, coefs)
    n = 500
    xs = [rng.random(n) * TWO_PI for _ in range(3)]
    uf = O.p_fluid_velocity(Us, h, xs, n)
    for a, c in enumerate(coefs):
        expect = c[0] + c[1] * xs[0] + c[2] * xs[1] + c[3] * xs[2]
        assert np.allclose(uf[a], expect, rtol=1e-12, atol=1e-12)


def test_interpolation_hits_node_values(rng):
    dims, h = (8, 8, 8), 1
    U = rand_halo(rng, dims, h)
    Z = np.zeros_like(U)
    d = TWO_PI / 8
    # U node (i, j, k) sits at (i dx, (j-1/2) dy, (k-1/2) dz)
    pts = [(3, 4, 5), (1, 1, 1), (7, 2, 8)]
    xs = [np.array([p[0] * d for p in pts]), np.array([(p[1] - .5) * d for p in pts]),
          np.array([(p[2] - .5) * d for p in pts])]
    uf = O.p_fluid_velocity([U, Z, Z], h, xs, len(pts))
    for q, (i, j, k) in enumerate(pts):
        assert np.isclose(uf[0][q], U[i - 1 + h, j - 1 + h, k - 1 + h], rtol=1e-13)


@pytest.mark.parametrize("drag,f", [(0, 1.0), (1, 1.15)])
def test_drag_coefficient_kats(drag, f):
    """coeff(LinearDrag(), 1) == 1.0 and coeff(NonlinearDrag(), 1) == 1.15:
    a particle at rest in a uniform stream with Re_p = d |u| Re = 1."""
    dims, h = (8, 8, 8), 1
    Re, rho, uf = 20.0, 3.0, 0.5
    diam = 1.0 / (uf * Re)
    U = np.full(tuple(n + 2 * h for n in dims), uf, order="F")
    Z = np.zeros_like(U)
    xs = [np.array([1.0]), np.array([2.0]), np.array([3.0])]
    us = [np.zeros(1) for _ in range(3)]
    du = [np.zeros(1) for _ in range(3)]
    O.p_motion([U, Z, Z], h, Re, 0.0, rho, drag, xs, us, np.array([diam]), du, 1)
    taup = 1 / 18 * rho * Re * diam ** 2
    assert np.isclose(du[0][0], f * uf / taup, rtol=1e-13)
    assert du[1][0] == 0 and du[2][0] == 0


def test_rk2_drag_relaxation_closed_form():
    """Linear drag in a uniform stream, two stages (src/particles/update.jl):
    u* = dt/2 (U - u0)/τ, u1 = u0 + dt (U - u*)/τ, and the position uses the
    stage-start velocity: x1 = x0 + dt u0."""
    from oracle.driver import ParticleSet
    dims, h = (8, 8, 8), 3
    Re, rho, g = 50.0, 800.0, 0.3
    Uf, dt = 0.7, 0.004
    U = np.full(tuple(n + 2 * h for n in dims), Uf, order="F")
    Z = np.zeros_like(U)
    diam = 0.01
    taup = 1 / 18 * rho * Re * diam ** 2
    u0 = np.array([0.2])
    ps = ParticleSet([np.array([1.0]), np.array([2.0]), np.array([3.0])], [u0, np.zeros(1), np.zeros(1)],
                     np.array([diam]), np.array([1]))
    n = 1
    O.p_boundary(ps.xss, ps.xss2, n)
    O.p_motion([U, Z, Z], h, Re, g, rho, 0, ps.xss, ps.uss, ps.diams, ps.dudtss, n)
    O.p_update(ps.xss2, ps.xss, ps.uss2, ps.uss, ps.dudtss, n, dt / 2)
    O.p_boundary(ps.xss2, ps.xss, n)
    O.p_motion([U, Z, Z], h, Re, g, rho, 0, ps.xss2, ps.uss2, ps.diams, ps.dudtss, n)
    O.p_update(ps.xss, ps.xss, ps.uss, ps.uss, ps.dudtss, n, dt)
    ustar = u0[0] + dt / 2 * ((Uf - u0[0]) / taup - g)
    u1 = u0[0] + dt * ((Uf - ustar) / taup - g)
    assert np.isclose(ps.uss[0][0], u1, rtol=1e-13)
    assert np.isclose(ps.xss[0][0], 1.0 + dt * u0[0], rtol=1e-14)


def test_periodic_wrap_invariant(rng):
    n = 1000
    xs = [rng.uniform(-1.0, TWO_PI + 1.0, n) for _ in range(3)]
    shadow = [x + rng.uniform(-0.1, 0.1, n) for x in xs]
    dxs = [s - x for s, x in zip(shadow, xs)]
    X = [x.copy() for x in xs]
    S = [s.copy() for s in shadow]
    O.p_boundary(X, S, n)
    for a in range(3):
        assert np.all((X[a] >= 0) & (X[a] < TWO_PI))
        assert np.allclose(S[a] - X[a], dxs[a], atol=1e-12)
        assert np.allclose(np.mod(X[a] - xs[a], TWO_PI) * (np.abs(np.mod(X[a] - xs[a], TWO_PI) - TWO_PI) > 1e-9),
                           0, atol=1e-9)


def test_particle_update_closed_form(rng):
    n = 50
    x0 = [rng.random(n) for _ in range(3)]
    u0 = [rng.random(n) for _ in range(3)]
    f = [rng.random(n) for _ in range(3)]
    x = [np.zeros(n) for _ in range(3)]
    u = [np.zeros(n) for _ in range(3)]
    O.p_update(x, x0, u, u0, f, n, 0.01)
    for a in range(3):
        assert np.array_equal(x[a], x0[a] + 0.01 * u0[a])
        assert np.array_equal(u[a], u0[a] + 0.01 * f[a])


def test_particle_timestep_uses_all_slots():
    d = np.array([0.0, 0.01, 0.03, 0.0])
    Re, rho, dtmax = 100.0, 843.0, 10.0
    dt = O.p_timestep(d, Re, rho, dtmax)
    assert np.isclose(dt, 0.5 * (1 / 18 * rho * Re * 0.03 ** 2), rtol=1e-14)
    assert O.p_timestep(d, Re, rho, 1e-9) == 1e-9


def test_hasher_encode_kat():
    """Hasher((0:5, 0:5, 0:5)) encodes like LinearIndices((6,6,6)) (test/test_utils.jl:118-127):
    n = 16, grid_ratio 4 -> 4 cells per axis plus the ghost ring."""
    dims, ratio = (16, 16, 16), 4
    dxc = TWO_PI / 16 * 4
    c = np.arange(-1, 5)  # ci - 1 over 0:5
    I, J, K = np.meshgrid(c, c, c, indexing="ij")
    xs = [np.ascontiguousarray(((A.ravel(order="F") + 0.5) * dxc)) for A in (I, J, K)]
    h = O.hash_cells(dims, ratio, xs, len(xs[0]))
    lin = np.arange(1, 6 ** 3 + 1).reshape((6, 6, 6), order="F")
    assert np.array_equal(h, lin.ravel(order="F"))


SEARCH_SAMPLES = [([], 5), ([1], 0), ([1], 1), ([1], 2), ([1, 2, 2, 3, 5], 2), ([1, 2, 2, 3, 5], 4),
                  ([1, 2, 2, 3, 5], 5), ([1, 2, 2, 3, 5], 6), ([1, 1, 1, 1, 1], 0), ([1, 1, 1, 1, 1], 1),
                  ([1, 1, 1, 1, 1], 2)]


@pytest.mark.parametrize("v,x", SEARCH_SAMPLES)
def test_searchsorted_reference_samples(v, x):
    assert O.searchsortedfirst(v, x) == bisect.bisect_left(v, x) + 1
    assert O.searchsortedlast(v, x) == bisect.bisect_right(v, x)


@pytest.mark.parametrize("n", [0, 1, 10, 1000])
def test_makeindex_matches_stable_sort(rng, n):
    dims, ratio = (16, 16, 16), 4
    xs = [rng.uniform(-0.2, TWO_PI + 0.2, n) for _ in range(3)]
    ids = np.arange(1, n + 1, dtype=np.int64)
    other = [rng.random(n) for _ in range(14)]
    before_x = [x.copy() for x in xs]
    before_o = [o.copy() for o in other]
    hashes0 = O.hash_cells(dims, ratio, xs, n)
    hashes, perm, starts, stops = O.makeindex(dims, ratio, xs, ids, other + xs, n)
    order = np.argsort(hashes0, kind="stable")
    assert np.array_equal(perm - 1, order)
    assert np.array_equal(hashes, hashes0[order])
    assert np.array_equal(ids, (order + 1))
    for a, b in zip(other, before_o):
        assert np.array_equal(a, b[order])
    for a, b in zip(xs, before_x):
        assert np.array_equal(a, b[order])
    assert np.sum(stops - starts + 1) == n  # test/test_particles.jl:85
    ncell = 6 ** 3
    cnt = np.bincount(hashes0, minlength=ncell + 1)[1:]
    assert np.array_equal(stops - starts + 1, cnt)
    for c in np.nonzero(cnt)[0][:20]:
        assert np.all(hashes[starts[c] - 1:stops[c]] == c + 1)


def test_particle_init_ranges():
    n = 2000
    xs, diams, ids = O.init_particles(9, n, 0.1, 10.0, 0.5, 100.0, 843.0, np.float64)
    for x in xs:
        assert np.all((x >= 0) & (x < TWO_PI))
    taueta = np.sqrt(1 / (0.5 * 100.0))
    st = (1 / 18 * 843.0 * 100.0 * diams ** 2) / taueta
    assert st.min() >= 0.1 * (1 - 1e-12) and st.max() <= 10.0 * (1 + 1e-12)
    assert 0.3 < np.median(np.log10(st)) + 0.5 < 0.7  # log-uniform around 10^0
    assert np.array_equal(ids, np.arange(1, n + 1))
