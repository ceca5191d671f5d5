"""Particle statistics on the device (lcs_particle_stat, lcs_particle_density:
Particles.stat / Particles.density, src/particles/stat.jl:80-231) against the
oracle run on the device's own re-binned particles:

* RDF pair counts per bin and contact pairs are equal exactly (integers), for
  clustered particles where close pairs and contacts occur;
* SPH densities are equal bit for bit (FP64 and FP32 positions);
* the summaries equal numpy's over the downloaded values (extrema exactly);
  taup = (1/18) rho_p Re d^2 of the mean diameter and St = taup / tau_eta;
* the number-density spectrum equals the oracle's to rounding, and N
  coincident particles give (1 - 1/N) 4 pi k^2 in every shell.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_step import c1_config

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def clustered_box(dtype, n=4000, coincident=False):
    import lcs_mi355x as L
    cfg = c1_config(dims=(32, 32, 32), dtype=dtype, nparticles=n, profiles=1)
    box = L.build(cfg)
    nv, nt = box.profiles[0]["nvalid"], box.profiles[0]["ntotal"]
    rng = np.random.default_rng(11)
    if coincident:
        pts = np.tile([1.1, 2.2, 3.3], (nv, 1))
    else:
        c = rng.uniform(0.5, TWO_PI - 0.5, (nv // 4 + 1, 3))
        pts = np.mod(np.repeat(c, 4, axis=0)[:nv] + rng.normal(0, 0.02, (nv, 3)), TWO_PI)
    for a, k in enumerate("xyz"):
        full = box.download_prop(0, k)
        full[:nv] = pts[:, a]
        box.upload_prop(0, k, full)
    d = box.download_prop(0, "diams")
    d[:nv] = 0.02
    box.upload_prop(0, "diams", d)
    return cfg, box, nv


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_particle_stat_matches_oracle(dtype):
    import lcs_mi355x as L
    cfg, box, nv = clustered_box(dtype)
    prm = dict(contact_tolerance=0.05, bin_size=12, min_radius=1e-3, search_cell_size=2, max_wavenumber=48,
               nsamples=16)
    st = L.particle_stat(box, 0, tau_eta=0.5, **prm)
    xs = [box.download_prop(0, k)[:nv] for k in "xyz"]
    ids = box.download_prop(0, "ids")[:nv]
    hashes, starts, stops = box.cellindex(0)
    edges = st["rdf"]["edges"]
    assert edges[0] == 1e-3 and edges[-1] == pytest.approx(2 * TWO_PI / 32, rel=1e-15)
    npairs, ncont = O.rdf_counts((32, 32, 32), 4, xs, ids, hashes, starts, stops, st["diam"]["mean"],
                                 prm["contact_tolerance"], edges, prm["search_cell_size"])
    assert np.array_equal(st["rdf"]["npairs"], npairs)
    assert st["rdf"]["npairs_contact"] == ncont and ncont > 0 and npairs.sum() > 100
    for c, k in zip("xyz", xs):
        assert st[c]["max"] == k.max() and st[c]["min"] == k.min()
        assert st[c]["mean"] == pytest.approx(k.astype(np.float64).mean(), rel=1e-12)
    Re = 53.3
    assert st["taup"] == pytest.approx(1 / 18 * 843.0 * Re * st["diam"]["mean"] ** 2, rel=1e-12)
    assert st["St"] == pytest.approx(st["taup"] / 0.5, rel=1e-15)
    ks, enp = O.pspectrum(xs, prm["max_wavenumber"], prm["nsamples"])
    assert np.array_equal(st["spectrum"]["wavenumbers"], ks)
    err = np.abs(st["spectrum"]["Enp"] - enp)
    assert np.all(err <= 1e-6 * np.abs(enp) + 1e-9 * 4 * np.pi * ks**2), (st["spectrum"]["Enp"], enp)
    box.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_sph_density_matches_oracle_bitwise(dtype):
    import lcs_mi355x as L
    cfg, box, nv = clustered_box(dtype, n=5000)
    dens = L.particle_density(box, 0)
    xs = [box.download_prop(0, k)[:nv] for k in "xyz"]
    hashes, starts, stops = box.cellindex(0)
    ref = O.sph_density((32, 32, 32), 4, xs, hashes, starts, stops)
    np.testing.assert_array_equal(dens, ref)
    assert dens.min() > 0
    box.close()


def test_spectrum_of_coincident_particles():
    import lcs_mi355x as L
    cfg, box, nv = clustered_box("f64", n=300, coincident=True)
    st = L.particle_stat(box, 0, tau_eta=1.0, max_wavenumber=24, nsamples=8)
    ks = st["spectrum"]["wavenumbers"]
    np.testing.assert_allclose(st["spectrum"]["Enp"], (1 - 1 / nv) * 4 * np.pi * ks**2, rtol=1e-9)
    box.close()
