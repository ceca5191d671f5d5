"""ctypes bindings for the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module; it is the checker, never the product path.  The C
restatement lives in ``lcs_oracle.c`` (built by ``make -C oracle``); this module
only marshals numpy arrays and composes the phases in the order the reference
calls them (``src/evolve.jl:63-146``, ``src/timestep.jl:14-27``).

Arrays use the reference layout: Fortran order (x fastest) with halo ``h`` on
every side, shape ``(nx+2h, ny+2h, nz+2h)``; dU/dt arrays are interior-only.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

DOMAIN_LENGTH = 6.283185307179586
INVALID = 2**53 - 1  # src/particles/Particles.jl:18


def build():
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def lib(dtype):
    dtype = np.dtype(dtype)
    key = "f64" if dtype == np.float64 else "f32"
    if key not in _libs:
        path = os.path.join(_BUILD, f"liboracle_{key}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        _declare(L)
        _libs[key] = L
    return _libs[key]


def set_threads(n):
    """Sets the OpenMP thread count of both builds (0 keeps it); returns it."""
    return min(lib(dt).o_set_threads(int(n)) for dt in (np.float64, np.float32))


P = C.c_void_p
I = C.c_int
D = C.c_double
Lg = C.c_long
U64 = C.c_uint64


def _declare(L):
    sig = {
        "o_set_threads": (I, [I]),
        "o_ghost_fill": (None, [P, I, I, I, I, I, I]),
        "o_rhs": (None, [I, I, I, I, D, P, P, P, P, P, P, P]),
        "o_update": (None, [I, I, I, I, D, P, P, P, P, P, P, P, P, P]),
        "o_hsmac": (None, [I, I, I, I, D, I, I, I, D, D, I, I, P, P, P, P, P, P]),
        "o_flow_timestep": (D, [I, I, I, I, D, D, D, P, P, P, P]),
        "o_force_linear": (D, [I, I, I, I, D, D, P, P, P, P]),
        "o_boxmean": (None, [I, I, I, I, I, I, I, P, P]),
        "o_colour_mask": (None, [P, P, I, I, P]),
        "o_hsmac_pass": (None, [I, I, I, I, P, P, I, I, D, D, P, P, P, P, P]),
        "o_kcount": (I, [I]),
        "o_kset": (I, [I, P, I]),
        "o_enespe0": (D, [I, D]),
        "o_rcf_spectrum": (I, [I, I, I, D, P, P, P, P, P]),
        "o_rcf_apply": (D, [I, I, I, I, I, I, I, D, P, P, P, P]),
        "o_p_boundary": (None, [Lg, P, P, P, P, P, P]),
        "o_p_fluid_velocity": (None, [I, I, I, I, P, P, P, Lg, P, P, P, P, P, P]),
        "o_p_motion": (None, [I, I, I, I, D, D, D, I, P, P, P, Lg, P, P, P, P, P, P, P, P, P, P]),
        "o_p_update": (None, [Lg, D] + [P] * 15),
        "o_p_timestep": (D, [Lg, P, D, D, D]),
        "o_hash": (None, [I, I, I, I, Lg, P, P, P, P]),
        "o_makeindex": (None, [I, I, I, I, Lg, P, P, P, P, P, I, P, P, P, P]),
        "o_searchsortedfirst": (Lg, [P, Lg, C.c_int64]),
        "o_searchsortedlast": (Lg, [P, Lg, C.c_int64]),
        "o_rand_u01": (D, [U64, U64, U64]),
        "o_spectrum_kcut": (I, [I, D]),
        "o_spectrum_modes": (I, [U64, D, I, P, P, I]),
        "o_init_spectrum": (None, [I, I, I, I, U64, D, D, P, P, P, P]),
        "o_init_particles": (None, [U64, Lg, D, D, D, D, D, P, P, P, P, P]),
        "o_flow_stat": (None, [I, I, I, I, P, P, P, I, P]),
        "o_rdf": (None, [I, I, I, I, Lg, P, P, P, P, P, P, P, D, D, I, P, I, P, P]),
        "o_sph_density": (None, [I, I, I, I, Lg, P, P, P, P, P, P, P]),
        "o_fib_dir": (None, [I, I, P]),
        "o_pspectrum": (None, [Lg, P, P, P, I, P, I, D, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    if a is None:
        return None
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1, "oracle arrays must be Fortran-ordered"
    return a.ctypes.data


def dims_of(A, h):
    return tuple(int(s) - 2 * h for s in A.shape)


# --------------------------------------------------------------------------
# Flow phases
# --------------------------------------------------------------------------
def halo_zeros(dims, h, dtype):
    return np.zeros(tuple(n + 2 * h for n in dims), dtype=dtype, order="F")


def ghost_fill(A, h, width, full):
    nx, ny, nz = dims_of(A, h)
    lib(A.dtype).o_ghost_fill(_p(A), nx, ny, nz, h, width, int(full))


def rhs(U, V, W, P_, h, Re):
    """Flows.rhs!: src/flows/rhs.jl:59-81. Returns interior dU, dV, dW."""
    nx, ny, nz = dims_of(U, h)
    out = [np.zeros((nx, ny, nz), dtype=U.dtype, order="F") for _ in range(3)]
    lib(U.dtype).o_rhs(nx, ny, nz, h, Re, _p(U), _p(V), _p(W), _p(P_), *[_p(a) for a in out])
    return out


def update(Us, Us0, dUs, h, dt):
    """Flows.update!: src/flows/update.jl:23-54 (in place on Us)."""
    nx, ny, nz = dims_of(Us[0], h)
    lib(Us[0].dtype).o_update(nx, ny, nz, h, dt, *[_p(a) for a in Us], *[_p(a) for a in Us0], *[_p(a) for a in dUs])


def hsmac(U, V, W, P_, h, dt, time_blocking=True, overlap=True, coloring=0, omega=1.7, epsp=1e-3, itrmin=6, itrmax=300):
    """couple!(::HSMAC): src/flows/couples/hsmac.jl:52-201 (in place). Returns (itr, div)."""
    nx, ny, nz = dims_of(U, h)
    itr = C.c_int(0)
    div = C.c_double(0)
    lib(U.dtype).o_hsmac(nx, ny, nz, h, dt, int(time_blocking), int(overlap), int(coloring), omega, epsp, itrmin, itrmax,
                         _p(U), _p(V), _p(W), _p(P_), C.addressof(itr), C.addressof(div))
    return itr.value, div.value


def colour_mask(lo, hi, color, coloring=0):
    lo = np.asarray(lo, dtype=np.int32)
    hi = np.asarray(hi, dtype=np.int32)
    shape = tuple(int(b - a + 1) for a, b in zip(lo, hi))
    m = np.zeros(shape, dtype=np.int32, order="F")
    lib(np.float64).o_colour_mask(_p(lo), _p(hi), color, coloring, _p(m))
    return m


def hsmac_pass(U, V, W, P_, Div2, h, lo, hi, color, dt, omega=1.7, coloring=0):
    """One colour pass of hsmac_cellwise! (hsmac.jl:1-35) over the box [lo, hi]."""
    nx, ny, nz = dims_of(U, h)
    lo = np.asarray(lo, dtype=np.int32)
    hi = np.asarray(hi, dtype=np.int32)
    lib(U.dtype).o_hsmac_pass(nx, ny, nz, h, _p(lo), _p(hi), color, coloring, dt, omega, _p(U), _p(V), _p(W),
                              _p(P_), _p(Div2))


def flow_timestep(U, V, W, h, cfl, dtmax, Re):
    """Flows.timestep: src/flows/timestep.jl:18-65 (fills face ghosts in place)."""
    nx, ny, nz = dims_of(U, h)
    m = np.zeros(3)
    dt = lib(U.dtype).o_flow_timestep(nx, ny, nz, h, cfl, dtmax, Re, _p(U), _p(V), _p(W), _p(m))
    return dt, m


def force_linear(U, V, W, h, power, dt):
    """forcing!(::ConstantPowerLF): src/flows/forcings/lf.jl:27-106. Returns p."""
    nx, ny, nz = dims_of(U, h)
    m = np.zeros(5)
    p = lib(U.dtype).o_force_linear(nx, ny, nz, h, power, dt, _p(U), _p(V), _p(W), _p(m))
    return p, m


def flow_stat_sums(U, V, W, h, integral_length=False):
    """Raw interior statistics of Flows.stat (src/flows/stat.jl:67-130), read
    with the ghosts as they are: (means, max, min of U, V, W), the means of
    u2, s2, dudx^2, dudx^3, dudx^4, and the integral length (0 unless asked)."""
    nx, ny, nz = dims_of(U, h)
    out = np.zeros(15)
    lib(U.dtype).o_flow_stat(nx, ny, nz, h, _p(U), _p(V), _p(W), int(bool(integral_length)), out.ctypes.data)
    return out


def flow_stat_derived(sums, dims, Re):
    """The derived quantities of Flows.stat: src/flows/stat.jl:132-154."""
    um, vm, wm = sums[0], sums[3], sums[6]
    u2, s2, d2, d3, d4, L = sums[9:15]
    urms = np.sqrt(u2 / 3)
    k = 0.5 * (u2 - um**2 - vm**2 - wm**2)
    eps = s2 / Re
    tau_eta = np.sqrt(1 / eps / Re)
    l_mu = np.sqrt(15 * urms**2 / eps / Re)
    l_eta = (1 / eps / Re**3) ** 0.25
    return {"u": {"mean": um, "max": sums[1], "min": sums[2]}, "v": {"mean": vm, "max": sums[4], "min": sums[5]},
            "w": {"mean": wm, "max": sums[7], "min": sums[8]}, "k": k, "eps": eps, "tau_eta": tau_eta,
            "l_mu": l_mu, "l_eta": l_eta, "Re_lambda": l_mu * urms * Re, "urms": urms,
            "k_eta": (min(dims) / 2) * l_eta, "skew": -d3 / d2 ** 1.5, "flat": d4 / d2**2, "L": L,
            "C_eps": eps * L / urms**3}


def flow_stat(U, V, W, h, Re, integral_length=False):
    """Flows.stat (src/flows/stat.jl:67-155) on copies: FullHalo(h) on U, V, W, then the statistics."""
    Us = [np.array(a, order="F") for a in (U, V, W)]
    for a in Us:
        ghost_fill(a, h, h, True)
    return flow_stat_derived(flow_stat_sums(*Us, h, integral_length), dims_of(U, h), Re)


# particle statistics (src/particles/rdf.jl, spectrum.jl, stat.jl)
SPECTRUM_WAVENUMBERS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768)


def rdf_edges(rmin, rmax, nbins):
    """logrange(rmin, rmax, nbins + 1): log-spaced, ends exact."""
    e = np.exp(np.linspace(np.log(rmin), np.log(rmax), nbins + 1))
    e[0], e[-1] = rmin, rmax
    return e


def rdf_counts(dims, grid_ratio, xss, ids, hashes, starts, stops, diam, tol, edges, ncell):
    """rdf_assume_ci pair counts (per bin) and contact pairs, on makeindex!'s output."""
    n = len(hashes)
    nb = len(edges) - 1
    npairs = np.zeros(nb, np.int64)
    ncont = C.c_int64(0)
    xs = [np.ascontiguousarray(a) for a in xss]
    e = np.ascontiguousarray(edges, dtype=np.float64)
    ids, hashes, starts, stops = (np.ascontiguousarray(a, dtype=np.int64) for a in (ids, hashes, starts, stops))
    lib(xs[0].dtype).o_rdf(*dims, grid_ratio, n, *[_p(a) for a in xs], _p(ids), _p(hashes), _p(starts), _p(stops),
                           float(diam), float(tol), nb, _p(e), int(ncell), _p(npairs), C.addressof(ncont))
    return npairs, ncont.value


def sph_density(dims, grid_ratio, xss, hashes, starts, stops):
    n = len(hashes)
    out = np.zeros(n)
    xs = [np.ascontiguousarray(a) for a in xss]
    hashes, starts, stops = (np.ascontiguousarray(a, dtype=np.int64) for a in (hashes, starts, stops))
    lib(xs[0].dtype).o_sph_density(*dims, grid_ratio, n, *[_p(a) for a in xs], _p(hashes), _p(starts), _p(stops),
                                   _p(out))
    return out


def fib_dir(i, nsamples):
    out = np.zeros(3)
    lib(np.float64).o_fib_dir(int(i), int(nsamples), _p(out))
    return out


def pspectrum(xss, max_wavenumber, nsamples, nvalid_g=None):
    ks = [k for k in SPECTRUM_WAVENUMBERS if k <= max_wavenumber]
    xs = [np.ascontiguousarray(a) for a in xss]
    n = len(xs[0])
    wn = np.asarray(ks, dtype=np.int32)
    out = np.zeros(len(ks))
    lib(xs[0].dtype).o_pspectrum(n, *[_p(a) for a in xs], len(ks), _p(wn), int(nsamples),
                                 float(n if nvalid_g is None else nvalid_g), _p(out))
    return np.asarray(ks, dtype=np.float64), out


def boxmean(U, h, coarse):
    nx, ny, nz = dims_of(U, h)
    cx, cy, cz = coarse
    out = np.zeros((cx, cy, cz), dtype=U.dtype, order="F")
    lib(U.dtype).o_boxmean(nx, ny, nz, h, cx, cy, cz, _p(U), _p(out))
    return out


def kcount(k0):
    return lib(np.float64).o_kcount(k0)


def kset(k0):
    cap = 4 * (2 * k0 + 1) ** 3
    buf = np.zeros(3 * cap, dtype=np.int32)
    n = lib(np.float64).o_kset(k0, _p(buf), cap)
    return [tuple(int(v) for v in buf[3 * m:3 * m + 3]) for m in range(n)]


def enespe0(k0, e10):
    return lib(np.float64).o_enespe0(k0, e10)


def rcf_spectrum(coarse3, kmin, kmax, e10):
    """Spectral part of forcing!(::EnergyPreserveRCF): rcf.jl:43-67."""
    c = coarse3[0].shape[0]
    nk = 2 * kmax + 1
    ncrop = (kmax + 1) * nk * nk
    crop = np.zeros(2 * 3 * ncrop)
    en = np.zeros(kmax - kmin + 1)
    dt = coarse3[0].dtype
    keep = [np.asfortranarray(a, dtype=dt) for a in coarse3]
    rc = lib(dt).o_rcf_spectrum(c, kmin, kmax, e10, *[_p(a) for a in keep], _p(crop), _p(en))
    if rc != 0:
        raise ValueError("kmax must satisfy 1 <= kmin <= kmax <= coarse_size/2")
    return crop, en


def rcf_apply(U, V, W, h, coarse_size, kmin, kmax, dt, crop):
    nx, ny, nz = dims_of(U, h)
    return lib(U.dtype).o_rcf_apply(nx, ny, nz, h, coarse_size, kmin, kmax, dt, _p(crop), _p(U), _p(V), _p(W))


def force_rcf(U, V, W, h, energy, kmin, kmax, coarse_size, dt):
    """forcing!(::EnergyPreserveRCF): src/flows/forcings/rcf.jl:21-100. Returns (p, e10)."""
    coarse = [boxmean(A, h, (coarse_size,) * 3) for A in (U, V, W)]
    crop, _ = rcf_spectrum(coarse, kmin, kmax, energy)
    p = rcf_apply(U, V, W, h, coarse_size, kmin, kmax, dt, crop)
    return p, energy


# --------------------------------------------------------------------------
# Particle phases (SoA vectors of length >= n)
# --------------------------------------------------------------------------
def p_boundary(xss, xss2, n):
    lib(xss[0].dtype).o_p_boundary(n, *[_p(a) for a in xss], *[_p(a) for a in xss2])


def p_fluid_velocity(Us, h, xss, n):
    U = Us[0]
    nx, ny, nz = dims_of(U, h)
    out = [np.zeros(n, dtype=U.dtype) for _ in range(3)]
    lib(U.dtype).o_p_fluid_velocity(nx, ny, nz, h, *[_p(a) for a in Us], n, *[_p(a) for a in xss], *[_p(a) for a in out])
    return out


def p_motion(Us, h, Re, gravity, density_ratio, drag, xss, uss, diams, dudtss, n):
    U = Us[0]
    nx, ny, nz = dims_of(U, h)
    lib(U.dtype).o_p_motion(nx, ny, nz, h, Re, gravity, density_ratio, int(drag), *[_p(a) for a in Us], n,
                            *[_p(a) for a in xss], *[_p(a) for a in uss], _p(diams), *[_p(a) for a in dudtss])


def p_update(xss, xss0, uss, uss0, dudtss, n, dt):
    lib(xss[0].dtype).o_p_update(n, dt, *[_p(a) for a in xss], *[_p(a) for a in xss0], *[_p(a) for a in uss],
                                 *[_p(a) for a in uss0], *[_p(a) for a in dudtss])


def p_timestep(diams, Re, density_ratio, dtmax):
    return lib(diams.dtype).o_p_timestep(len(diams), _p(diams), Re, density_ratio, dtmax)


def hash_cells(dims, grid_ratio, xss, n):
    out = np.zeros(n, dtype=np.int64)
    lib(xss[0].dtype).o_hash(*dims, grid_ratio, n, *[_p(a) for a in xss], _p(out))
    return out


def makeindex(dims, grid_ratio, xss, ids, props, n):
    """makeindex!: src/particles/makeindex.jl:29-67 (in place on xss, ids, props).

    ``props`` are the other real SoA properties; they are permuted together with
    xss (which are permuted too).
    Returns (hashes, perm, starts, stops); perm/starts/stops are 1-based."""
    nc = [d // grid_ratio + 2 for d in dims]
    ncell = nc[0] * nc[1] * nc[2]
    hashes = np.zeros(max(n, 1), dtype=np.int64)
    perm = np.zeros(max(n, 1), dtype=np.int64)
    starts = np.zeros(ncell, dtype=np.int64)
    stops = np.zeros(ncell, dtype=np.int64)
    # all real properties are permuted (src/particles/buffer.jl:39-43), the
    # hashing positions included exactly once
    allp = list(xss) + [a for a in props if all(a is not x for x in xss)]
    arr = (C.c_void_p * max(len(allp), 1))(*[_p(a) for a in allp])
    lib(xss[0].dtype).o_makeindex(*dims, grid_ratio, n, *[_p(a) for a in xss], _p(ids), arr, len(allp),
                                  _p(hashes), _p(perm), _p(starts), _p(stops))
    return hashes[:n], perm[:n], starts, stops


def searchsortedfirst(v, x):
    v = np.ascontiguousarray(v, dtype=np.int64)
    return lib(np.float64).o_searchsortedfirst(_p(v) if len(v) else None, len(v), int(x))


def searchsortedlast(v, x):
    v = np.ascontiguousarray(v, dtype=np.int64)
    return lib(np.float64).o_searchsortedlast(_p(v) if len(v) else None, len(v), int(x))


# --------------------------------------------------------------------------
# Builder-defined initial conditions
# --------------------------------------------------------------------------
def rand_u01(seed, stream, idx):
    return lib(np.float64).o_rand_u01(seed, stream, idx)


def spectrum_modes(seed, k0, K):
    cap = (K + 1) * (2 * K + 1) ** 2
    kt = np.zeros(3 * cap, dtype=np.int32)
    at = np.zeros(6 * cap)
    n = lib(np.float64).o_spectrum_modes(seed, k0, K, _p(kt), _p(at), cap)
    return kt[:3 * n].reshape(n, 3), at[:6 * n].reshape(n, 6)


def spectrum_kcut(nmin, k0):
    return lib(np.float64).o_spectrum_kcut(nmin, k0)


def init_spectrum(dims, h, seed, k0, uprime, dtype):
    Us = [halo_zeros(dims, h, dtype) for _ in range(4)]
    lib(dtype).o_init_spectrum(*dims, h, seed, k0, uprime, *[_p(a) for a in Us])
    return Us[:3], Us[3]


def init_particles(seed, n, st_min, st_max, eps, Re, density_ratio, dtype):
    xs = [np.zeros(n, dtype=dtype) for _ in range(3)]
    diams = np.zeros(n, dtype=dtype)
    ids = np.zeros(n, dtype=np.int64)
    lib(dtype).o_init_particles(seed, n, st_min, st_max, eps, Re, density_ratio, *[_p(a) for a in xs], _p(diams), _p(ids))
    return xs, diams, ids
