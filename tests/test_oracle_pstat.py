"""The oracle's particle statistics on the cell index (src/particles/rdf.jl,
src/particles/stat.jl:171-231, src/particles/spectrum.jl), pinned without the
restatement's own loops:

* RDF pair counts per log-spaced bin (edges[k] < d <= edges[k+1]) and contact
  pairs equal a brute-force count over all pairs id1 < id2 of particles in
  domain cells (the search over +-ncell coarse cells covers every pair within
  rmax, so nothing else can enter a bin) -- integers, exactly;
* SPH density equals a brute-force sum of the cubic spline over all particles
  within 2h (self included), and is zero for particles alone;
* the number-density spectrum of N coincident particles is
  (1 - 1/N) 4 pi k^2 in every shell, and 0 for one particle; the Fibonacci
  directions are unit vectors running from the +y pole to the -y pole.
"""
import numpy as np
import pytest

from oracle import oracle as O

TWO_PI = 2 * np.pi


def indexed(dims, n, grid_ratio, dtype, seed, cluster=None):
    rng = np.random.default_rng(seed)
    if cluster is None:
        xs = [rng.uniform(0, TWO_PI, n).astype(dtype) for _ in range(3)]
    else:  # clustered, so close pairs and contacts occur
        c = rng.uniform(0.5, TWO_PI - 0.5, (n // 4, 3))
        pts = np.repeat(c, 4, axis=0) + rng.normal(0, cluster, (n // 4 * 4, 3))
        xs = [np.mod(pts[:, a], TWO_PI).astype(dtype) for a in range(3)]
    ids = rng.permutation(n).astype(np.int64) + 1
    hashes, perm, starts, stops = O.makeindex(dims, grid_ratio, xs, ids, [], len(xs[0]))
    return xs, ids, hashes, starts, stops


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ncell", [1, 2])
def test_rdf_equals_brute_force(dtype, ncell):
    dims, gr = (32, 32, 32), 4
    xs, ids, hashes, starts, stops = indexed(dims, 1200, gr, dtype, 7, cluster=0.02)
    dx = TWO_PI / 32
    rmax = ncell * dx
    edges = O.rdf_edges(1e-3, rmax, 12)
    diam, tol = 0.02, 0.05
    npairs, ncont = O.rdf_counts(dims, gr, xs, ids, hashes, starts, stops, diam, tol, edges, ncell)
    P = np.stack(xs, 1)
    D = P[:, None, :] - P[None, :, :]
    d = np.sqrt(((D[..., 0] * D[..., 0] + D[..., 1] * D[..., 1]) + D[..., 2] * D[..., 2])).astype(np.float64)
    pair = ids[:, None] < ids[None, :]
    k = np.searchsorted(edges, d, side="left")  # first edge >= d
    ref = np.array([np.sum(pair & (k == b)) for b in range(1, len(edges))])
    assert np.array_equal(npairs, ref)
    assert ncont == np.sum(pair & (np.abs(d / diam - 1) < tol))  # contacts lie well inside the search
    assert npairs.sum() > 50 and ncont > 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sph_density_equals_brute_force(dtype):
    dims, gr = (32, 32, 32), 4
    xs, ids, hashes, starts, stops = indexed(dims, 1600, gr, dtype, 3, cluster=0.08)
    dens = O.sph_density(dims, gr, xs, hashes, starts, stops)
    h = TWO_PI / 32
    P = np.stack(xs, 1)
    D = P[:, None, :] - P[None, :, :]
    d = np.sqrt(((D[..., 0] * D[..., 0] + D[..., 1] * D[..., 1]) + D[..., 2] * D[..., 2])).astype(np.float64)
    q = d / h
    w = np.where(q < 1, 1 - 1.5 * q**2 + 0.75 * q**3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0)) / (np.pi * h**3)
    np.testing.assert_allclose(dens, w.sum(1), rtol=1e-12)
    assert dens.min() >= 1 / (np.pi * h**3) * (1 - 1e-12)  # every particle counts itself


@pytest.mark.parametrize("nsamples", [8, 64])
def test_spectrum_closed_forms(nsamples):
    for N in (1, 5):
        pt = np.full(N, 1.234)
        ks, enp = O.pspectrum([pt, pt * 0.5, pt * 2.0], 48, nsamples)
        np.testing.assert_allclose(enp, (1 - 1 / N) * 4 * np.pi * ks**2, rtol=1e-11, atol=1e-11)
    assert list(ks) == [1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48]
    dirs = np.array([O.fib_dir(i, nsamples) for i in range(nsamples)])
    np.testing.assert_allclose(np.linalg.norm(dirs, axis=1), 1.0, rtol=1e-14)
    assert dirs[0][1] == 1.0 and dirs[-1][1] == -1.0
