"""The oracle's flow statistics (Flows.stat, src/flows/stat.jl:67-182), pinned
by what the mathematics fixes:

* polynomial fields of degree <= 3 on a non-periodic patch (ghosts set from the
  same polynomial, no ghost fill): diff1_o4_at_half (src/diff.jl:13-15) is
  exact there, so every strain component equals the analytic derivative at
  the point the reference's stencil is centred on -- cell centres for the
  diagonal, the edges (i, j-1/2... ) of each staggered pair for the
  off-diagonal terms (dudy/dvdx at x = i dx, y = (j-1) dy; dudz/dwdx at
  x = i dx, z = (k-1) dz; dvdz/dwdy at y = (j-1) dy, z = k dz).  A wrong
  index offset, sign or factor 2 moves the result;
* one sine mode in x: <dudx^2> = A^2 G^2 / 2 with G the stencil's discrete
  symbol, flatness 3/2, skewness 0, s2 = 2 <dudx^2>, k = A^2 / 4, and the
  integral length from the closed-form autocorrelation (n/2) cos(theta r)
  summed to the first negative lag;
* the derived quantities agree with the textbook definitions (tau_eta =
  sqrt(nu/eps), eta = (nu^3/eps)^(1/4), Re_lambda = u' lambda / nu).
"""
import numpy as np
import pytest

from oracle import oracle as O

TWO_PI = 2 * np.pi


def positions(n, h, offset):
    """Physical coordinate of array index a (0-based) for a quantity sitting
    at (d - offset) * dx, domain index d = a + 1 - h."""
    d = np.arange(n + 2 * h) + 1 - h
    return (d - offset) * (TWO_PI / n)


def poly_fields(dims, h, cU, cV, cW):
    """U at (i, j-1/2, k-1/2), V at (i-1/2, j, k-1/2), W at (i-1/2, j-1/2, k)
    (src/particles/interp.jl:1-51 staggering), each a polynomial given as a
    list of (coef, px, py, pz) terms."""
    nx, ny, nz = dims
    stag = {"U": (0.0, 0.5, 0.5), "V": (0.5, 0.0, 0.5), "W": (0.5, 0.5, 0.0)}
    out = []
    for name, terms in (("U", cU), ("V", cV), ("W", cW)):
        ox, oy, oz = stag[name]
        X = positions(nx, h, ox)[:, None, None]
        Y = positions(ny, h, oy)[None, :, None]
        Z = positions(nz, h, oz)[None, None, :]
        A = np.zeros((nx + 2 * h, ny + 2 * h, nz + 2 * h))
        for c, px, py, pz in terms:
            A = A + c * X**px * Y**py * Z**pz
        out.append(np.asfortranarray(A))
    return out


def dpoly(terms, axis, X, Y, Z):
    r = 0.0
    for c, px, py, pz in terms:
        p = [px, py, pz]
        if p[axis] == 0:
            continue
        q = list(p)
        q[axis] -= 1
        r = r + c * p[axis] * X ** q[0] * Y ** q[1] * Z ** q[2]
    return r + 0 * X * Y * Z


def expected_sums(dims, cU, cV, cW):
    nx, ny, nz = dims
    dx, dy, dz = (TWO_PI / n for n in dims)
    i = np.arange(1, nx + 1)[:, None, None]
    j = np.arange(1, ny + 1)[None, :, None]
    k = np.arange(1, nz + 1)[None, None, :]

    def at(xo, yo, zo):
        return (i - xo) * dx, (j - yo) * dy, (k - zo) * dz

    c = at(0.5, 0.5, 0.5)
    dudx = dpoly(cU, 0, *c)
    dvdy = dpoly(cV, 1, *c)
    dwdz = dpoly(cW, 2, *c)
    e12 = at(0.0, 1.0, 0.5)
    dudy, dvdx = dpoly(cU, 1, *e12), dpoly(cV, 0, *e12)
    e13 = at(0.0, 0.5, 1.0)
    dudz, dwdx = dpoly(cU, 2, *e13), dpoly(cW, 0, *e13)
    e23 = at(0.5, 1.0, 0.0)
    dvdz, dwdy = dpoly(cV, 2, *e23), dpoly(cW, 1, *e23)
    s12, s13, s23 = 0.5 * (dudy + dvdx), 0.5 * (dudz + dwdx), 0.5 * (dvdz + dwdy)
    s2 = 2 * (dudx**2 + dvdy**2 + dwdz**2 + 2 * (s12**2 + s13**2 + s23**2))
    return s2.mean(), (dudx**2).mean(), (dudx**3).mean(), (dudx**4).mean()


CASES = {
    "s12": ([(0.3, 0, 2, 0), (0.2, 1, 1, 0)], [(-0.7, 2, 0, 0), (0.4, 1, 1, 0)], []),
    "s13": ([(0.5, 0, 0, 2), (-0.3, 1, 0, 1)], [], [(0.6, 2, 0, 0), (0.25, 1, 0, 1)]),
    "s23": ([], [(0.45, 0, 0, 2), (0.2, 0, 1, 1)], [(-0.35, 0, 2, 0), (0.3, 0, 1, 1)]),
    "diagonal": ([(0.2, 2, 0, 0), (-0.05, 3, 0, 0)], [(0.3, 0, 2, 0), (0.02, 0, 3, 0)],
                 [(-0.25, 0, 0, 2), (0.03, 0, 0, 3)]),
    "all": ([(0.1, 1, 2, 0), (0.2, 2, 0, 1), (-0.04, 3, 0, 0), (0.05, 0, 1, 2)],
            [(0.15, 2, 1, 0), (-0.1, 0, 2, 1), (0.03, 1, 0, 2), (0.02, 0, 3, 0)],
            [(-0.2, 1, 1, 1), (0.07, 2, 0, 1), (0.04, 0, 2, 1), (-0.01, 0, 0, 3)]),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_strain_and_moments_are_exact_for_cubic_fields(case):
    dims, h = (12, 10, 9), 3
    cU, cV, cW = CASES[case]
    U, V, W = poly_fields(dims, h, cU, cV, cW)
    got = O.flow_stat_sums(U, V, W, h)
    s2, d2, d3, d4 = expected_sums(dims, cU, cV, cW)
    for name, g, e in (("s2", got[10], s2), ("d2", got[11], d2), ("d3", got[12], d3), ("d4", got[13], d4)):
        assert g == pytest.approx(e, rel=1e-10, abs=1e-10), name
    # u2 and the summaries are taken from the raw interior values
    inner = [a[h:-h, h:-h, h:-h] for a in (U, V, W)]
    assert got[9] == pytest.approx(sum((a**2) for a in inner).mean(), rel=1e-13)
    for q, a in enumerate(inner):
        assert got[3 * q] == pytest.approx(a.mean(), rel=1e-12, abs=1e-14)
        assert got[3 * q + 1] == a.max() and got[3 * q + 2] == a.min()


@pytest.mark.parametrize("n,m", [(30, 4), (24, 5), (40, 3)])
def test_single_mode_closed_forms(n, m):
    h, A = 3, 0.8
    dims = (n, 8, 6)
    U = O.halo_zeros(dims, h, np.float64)
    V = O.halo_zeros(dims, h, np.float64)
    W = O.halo_zeros(dims, h, np.float64)
    dx = TWO_PI / n
    theta = m * dx
    x_face = (np.arange(n + 2 * h) + 1 - h) * dx  # U[i] at x = i dx
    U[...] = (A * np.sin(m * x_face))[:, None, None]
    st = O.flow_stat(U, V, W, h, Re=50.0, integral_length=True)
    sums = O.flow_stat_sums(U, V, W, h, integral_length=True)
    G = (1 / dx) * (27 * 2 * np.sin(theta / 2) - 2 * np.sin(3 * theta / 2)) / 24
    assert sums[11] == pytest.approx(A**2 * G**2 / 2, rel=1e-12)
    assert sums[10] == pytest.approx(A**2 * G**2, rel=1e-12)
    assert st["flat"] == pytest.approx(1.5, rel=1e-12)
    assert abs(st["skew"]) < 1e-12
    assert st["k"] == pytest.approx(A**2 / 4, rel=1e-12)
    # autocorrelation (n/2) A^2 cos(theta r), summed to the first negative lag
    lags = np.cos(theta * np.arange(n // 2 + 1))
    stop = np.argmax(lags < 0) if (lags < 0).any() else len(lags)
    assert st["L"] == pytest.approx(dx * lags[:stop].sum(), rel=1e-12)


def test_derived_quantities_follow_the_textbook_definitions():
    rng = np.random.default_rng(3)
    dims, h, Re = (16, 16, 16), 3, 80.0
    U, V, W = (np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims))) for _ in range(3))
    st = O.flow_stat(U, V, W, h, Re)
    nu = 1 / Re
    assert st["tau_eta"] == pytest.approx(np.sqrt(nu / st["eps"]), rel=1e-13)
    assert st["l_eta"] == pytest.approx((nu**3 / st["eps"]) ** 0.25, rel=1e-13)
    assert st["l_mu"] == pytest.approx(np.sqrt(15 * nu * st["urms"] ** 2 / st["eps"]), rel=1e-13)
    assert st["Re_lambda"] == pytest.approx(st["urms"] * st["l_mu"] / nu, rel=1e-13)
    assert st["k_eta"] == pytest.approx(8 * st["l_eta"], rel=1e-13)
    # k = (3/2) u'^2 when the means vanish; here it is the fluctuation part
    m2 = st["u"]["mean"] ** 2 + st["v"]["mean"] ** 2 + st["w"]["mean"] ** 2
    assert st["k"] == pytest.approx(1.5 * st["urms"] ** 2 - 0.5 * m2, rel=1e-13)
