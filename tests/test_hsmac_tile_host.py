"""The fused HSMAC sweep kernel's own phase functions, executed on the host.

``tests/native/hsmac_tile_emul.cpp`` runs every workgroup of the launch grid
of ``k_hsmac_sweep`` thread by thread and phase by phase, in the kernel's
barrier order, over the same ``HsTile`` code the kernel compiles
(lcs_hsmac_tile.h).  One fused sweep over a box R must equal the oracle's
colour-0 pass followed by its colour-1 pass over R (hsmac.jl:1-35,
lib/parallel/src/foraxes.jl:73-107) bit for bit -- every element of U, V, W,
P, ghosts included -- and its div^2 sum over (1:n+1)^3 must equal the
oracle's.  Boxes: the three time-blocking domains d1, d2, d3 and the overlap
core (2:n)^3; tiles 128, 64 and 32 wide (128 x 16: two cells per thread),
ragged tiles in x and y, several
z-chunkings, both colourings, both element types.
"""
import ctypes as C
import os
import subprocess
import zlib

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    if not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("HIP headers not installed")
    out = tmp_path_factory.mktemp("hsmac_tile") / "libhsmac_tile_emul.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "lcs.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "hsmac_tile_emul.cpp"), "-o", str(out)], check=True)
    return C.CDLL(str(out))


def _box(name, n):
    n = np.asarray(n)
    lo, hi = {"d1": (-np.ones(3), n + 3), "d2": (np.zeros(3), n + 2), "d3": (np.ones(3), n + 1),
              "core": (2 * np.ones(3), n)}[name]
    return np.asarray(lo, dtype=np.int32), np.asarray(hi, dtype=np.int32)


CASES = [((8, 8, 8), 4, 64, 14, "d1"), ((16, 12, 10), 4, 64, 5, "d2"), ((70, 9, 7), 8, 64, 3, "d3"),
         ((17, 5, 9), 8, 64, 4, "core"), ((20, 13, 11), 4, 64, 1, "d1"), ((66, 20, 6), 8, 64, 12, "d3"),
         ((5, 3, 4), 4, 64, 2, "d1"), ((40, 10, 9), 8, 32, 4, "d1"), ((27, 9, 6), 8, 32, 2, "core"),
         ((33, 7, 5), 4, 32, 3, "d2"), ((70, 20, 6), 16, 64, 3, "d1"), ((12, 35, 5), 16, 64, 2, "core"),
         ((126, 9, 7), 8, 128, 3, "d3"), ((130, 20, 6), 16, 128, 3, "d1"), ((20, 35, 5), 16, 128, 2, "core")]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("coloring", [0, 1])
@pytest.mark.parametrize("dims,ty,tx,zc,box", CASES)
def test_fused_sweep_equals_two_colour_passes(emul, dims, ty, tx, zc, box, coloring, dtype):
    h, dt, omega = 3, 0.01, 1.7
    rng = np.random.default_rng(zlib.crc32(repr((dims, ty, tx, zc, box, coloring)).encode()))
    src = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    lo, hi = _box(box, dims)
    ref = [a.copy(order="F") for a in src]
    div2 = np.zeros_like(ref[0])
    O.hsmac_pass(*ref, div2, h, lo, hi, 0, dt, omega, coloring)
    O.hsmac_pass(*ref, div2, h, lo, hi, 1, dt, omega, coloring)
    nx, ny, nz = dims
    s_ref = div2[h:h + nx + 1, h:h + ny + 1, h:h + nz + 1].astype(np.float64).sum()

    dst = [np.full_like(a, np.nan) for a in src]
    sp = (C.c_void_p * 4)(*[a.ctypes.data for a in src])
    dp = (C.c_void_p * 4)(*[a.ctypes.data for a in dst])
    s = C.c_double(0)
    rc = emul.hsmac_sweep_emul(ty, tx, np.dtype(dtype).itemsize, coloring, nx, ny, nz, h,
                               lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), C.c_double(dt),
                               C.c_double(omega), sp, dp, zc, C.byref(s))
    assert rc == 0
    for name, a, b in zip("UVWP", dst, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert s.value == pytest.approx(s_ref, rel=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,ty,tx,zc,box", [((20, 13, 11), 8, 64, 4, "d1"), ((40, 10, 9), 8, 32, 3, "core"),
                                               ((70, 20, 6), 16, 64, 2, "d3"), ((130, 20, 6), 16, 128, 2, "d3")])
def test_fused_sweep_with_linear_forcing_on_load(emul, dims, ty, tx, zc, box, dtype):
    """The first sweep of couple! after the linear forcing applies it as it
    loads U, V, W: the result equals x <- x + (A (x - mean)) dt on every
    element (lf.jl:88-100; ghosts are copies, so transforming them equals
    copying the transformed interior) followed by the two colour passes, and
    the power sum is that of the cells (1:n)^3 (lf.jl:96,103)."""
    h, dt, omega = 3, 0.01, 1.7
    rng = np.random.default_rng(7)
    src = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    A, means, fdt = 0.37, [0.11, -0.07, 0.02], 0.004
    T = np.dtype(dtype).type
    ref = [a.copy(order="F") for a in src]
    nx, ny, nz = dims
    inner = (slice(h, h + nx), slice(h, h + ny), slice(h, h + nz))
    power = 0.0
    srcs = []
    for c in range(3):
        usrc = T(A) * (ref[c] - T(means[c]))
        srcs.append(usrc)
        ref[c] = ref[c] + usrc * T(fdt)
    pw = (src[0] * srcs[0] + src[1] * srcs[1]) + src[2] * srcs[2]
    power = pw[inner].astype(np.float64).sum()
    lo, hi = _box(box, dims)
    div2 = np.zeros_like(ref[0])
    O.hsmac_pass(*ref, div2, h, lo, hi, 0, dt, omega, 0)
    O.hsmac_pass(*ref, div2, h, lo, hi, 1, dt, omega, 0)
    dst = [np.full_like(a, np.nan) for a in src]
    sp = (C.c_void_p * 4)(*[a.ctypes.data for a in src])
    dp = (C.c_void_p * 4)(*[a.ctypes.data for a in dst])
    force = (C.c_double * 5)(A, *means, fdt)
    s, pw_out = C.c_double(0), C.c_double(0)
    rc = emul.hsmac_sweep_forced_emul(ty, tx, np.dtype(dtype).itemsize, 0, nx, ny, nz, h,
                                      lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), C.c_double(dt),
                                      C.c_double(omega), sp, dp, zc, force, C.byref(s), C.byref(pw_out))
    assert rc == 0
    for name, a, b in zip("UVWP", dst, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert pw_out.value == pytest.approx(power, rel=1e-10)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("forced", [False, True])
@pytest.mark.parametrize("dims,ty,tx,zc,box,coloring", [((20, 13, 11), 8, 64, 4, "d1", 0),
                                                        ((16, 12, 10), 4, 64, 5, "core", 1),
                                                        ((5, 3, 4), 4, 64, 2, "d1", 0),
                                                        ((40, 10, 9), 8, 32, 3, "d1", 1),
                                                        ((70, 20, 6), 16, 64, 2, "core", 0),
                                                        ((130, 20, 6), 16, 128, 2, "d1", 1)])
def test_fused_sweep_with_fill_on_load(emul, dims, ty, tx, zc, box, coloring, forced, dtype):
    """The first sweep of a time-blocked group loads U, V, W of ghost cells
    from their periodic images: equal bit for bit, ghosts included, to the
    FullHalo fill of width h of U, V, W that precedes the group
    (hsmac.jl:131-201, halo.jl:138-159) followed by the two colour passes.
    P's ghosts are used as they are (the fill does not touch P).  With the
    linear forcing on load as well, the forcing applies to the filled
    values."""
    h, dt, omega = 3, 0.01, 1.7
    rng = np.random.default_rng(zlib.crc32(repr((dims, ty, tx, zc, box, coloring, forced)).encode()))
    src = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    T = np.dtype(dtype).type
    A, means, fdt = 0.37, [0.11, -0.07, 0.02], 0.004
    ref = [a.copy(order="F") for a in src]
    for c in range(3):
        O.ghost_fill(ref[c], h, h, True)
    nx, ny, nz = dims
    inner = (slice(h, h + nx), slice(h, h + ny), slice(h, h + nz))
    power = 0.0
    if forced:
        srcs = [T(A) * (ref[c] - T(means[c])) for c in range(3)]
        power = ((ref[0] * srcs[0] + ref[1] * srcs[1]) + ref[2] * srcs[2])[inner].astype(np.float64).sum()
        for c in range(3):
            ref[c] = ref[c] + srcs[c] * T(fdt)
    lo, hi = _box(box, dims)
    div2 = np.zeros_like(ref[0])
    O.hsmac_pass(*ref, div2, h, lo, hi, 0, dt, omega, coloring)
    O.hsmac_pass(*ref, div2, h, lo, hi, 1, dt, omega, coloring)
    s_ref = div2[h:h + nx + 1, h:h + ny + 1, h:h + nz + 1].astype(np.float64).sum()
    dst = [np.full_like(a, np.nan) for a in src]
    sp = (C.c_void_p * 4)(*[a.ctypes.data for a in src])
    dp = (C.c_void_p * 4)(*[a.ctypes.data for a in dst])
    force = (C.c_double * 5)(A, *means, fdt) if forced else None
    s, pw_out = C.c_double(0), C.c_double(0)
    rc = emul.hsmac_sweep_wrap_emul(ty, tx, np.dtype(dtype).itemsize, coloring, nx, ny, nz, h,
                                    lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), C.c_double(dt),
                                    C.c_double(omega), sp, dp, zc, force, C.byref(s), C.byref(pw_out))
    assert rc == 0
    for name, a, b in zip("UVWP", dst, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert s.value == pytest.approx(s_ref, rel=1e-12)
    if forced:
        assert pw_out.value == pytest.approx(power, rel=1e-10)


def _x_slabs(box, n):
    """The x- and x+ face slabs of region `box` in the overlap order
    (lib/topologies/src/halo.jl:349-366): rows (y, z) in [2, n]."""
    lo, hi = _box(box, n)
    nx, ny, nz = n
    return (np.array([lo[0], 2, 2, nx + 1, 2, 2], dtype=np.int32),
            np.array([1, ny, nz, hi[0], ny, nz], dtype=np.int32))


XCASES = [((16, 12, 20), 16, "d1"), ((40, 37, 35), 16, "d1"), ((9, 13, 10), 4, "d1"), ((9, 13, 10), 4, "d2"),
          ((11, 6, 14), 4, "d3"), ((2, 5, 7), 4, "d1"), ((3, 2, 2), 4, "d1"), ((20, 18, 17), 16, "d3"),
          ((6, 9, 4), 4, "d2"), ((40, 70, 45), 32, "d1"), ((9, 33, 36), 32, "d3")]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("coloring", [0, 1])
@pytest.mark.parametrize("dims,patch,box", XCASES)
def test_x_slab_pair_equals_four_colour_passes(emul, dims, patch, box, coloring, dtype):
    """Both colours over the x- and x+ slabs by patches of slab rows, final
    values through a compact buffer: equal bit for bit, ghosts included, to
    the reference's order x- red, x- black, x+ red, x+ black
    (hsmac.jl:1-35, foraxes.jl:73-107), and the div^2 sum over the slab
    cells in (1:n+1)^3 equal to the oracle's."""
    h, dt, omega = 3, 0.01, 1.7
    rng = np.random.default_rng(zlib.crc32(repr((dims, patch, box, coloring)).encode()))
    src = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    lo, hi = _x_slabs(box, dims)
    ref = [a.copy(order="F") for a in src]
    div2 = np.zeros_like(ref[0])
    for s in range(2):
        for color in range(2):
            O.hsmac_pass(*ref, div2, h, lo[3 * s:3 * s + 3].copy(), hi[3 * s:3 * s + 3].copy(), color, dt, omega,
                         coloring)
    nx, ny, nz = dims
    s_ref = div2[h:h + nx + 1, h:h + ny + 1, h:h + nz + 1].astype(np.float64).sum()
    out = [a.copy(order="F") for a in src]
    fp = (C.c_void_p * 4)(*[a.ctypes.data for a in out])
    s = C.c_double(0)
    rc = emul.hsmac_xface_emul(patch, np.dtype(dtype).itemsize, coloring, nx, ny, nz, h,
                               lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), C.c_double(dt),
                               C.c_double(omega), fp, C.byref(s))
    assert rc == 0
    for name, a, b in zip("UVWP", out, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert s.value == pytest.approx(s_ref, rel=1e-12)


def _slabs(ti, box, n):
    """The two face slabs across axis ti (0 x, 1 y, 2 z) of region `box` in
    the overlap order (lib/topologies/src/halo.jl:349-366)."""
    lo, hi = _box(box, n)
    nx, ny, nz = n
    if ti == 0:
        return _x_slabs(box, n)
    if ti == 1:
        return (np.array([lo[0], lo[1], 2, lo[0], ny + 1, 2], dtype=np.int32),
                np.array([hi[0], 1, nz, hi[0], hi[1], nz], dtype=np.int32))
    return (np.array([lo[0], lo[1], lo[2], lo[0], lo[1], nz + 1], dtype=np.int32),
            np.array([hi[0], hi[1], 1, hi[0], hi[1], hi[2]], dtype=np.int32))


YZCASES = [(1, (16, 12, 20), 16, "d1"), (1, (9, 13, 10), 4, "d2"), (1, (11, 6, 14), 4, "d3"),
           (1, (20, 18, 17), 16, "d3"), (2, (16, 12, 20), 16, "d1"), (2, (9, 13, 10), 4, "d2"),
           (2, (11, 6, 14), 4, "d3"), (2, (20, 18, 17), 16, "d1"), (1, (2, 5, 7), 4, "d1"),
           (2, (5, 3, 2), 4, "d1"), (0, (16, 12, 20), 4, "d2")]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("coloring", [0, 1])
@pytest.mark.parametrize("ti,dims,patch,box", YZCASES)
def test_y_and_z_slab_pairs_equal_four_colour_passes(emul, ti, dims, patch, box, coloring, dtype):
    """The same patch pair across y and z: rows (x, z) and (x, y) that reach
    into the ghosts; equal bit for bit to the reference's lower red, lower
    black, upper red, upper black passes, and the div^2 sum to the oracle's."""
    h, dt, omega = 3, 0.01, 1.7
    rng = np.random.default_rng(zlib.crc32(repr((ti, dims, patch, box, coloring)).encode()))
    src = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    lo, hi = _slabs(ti, box, dims)
    ref = [a.copy(order="F") for a in src]
    div2 = np.zeros_like(ref[0])
    for s in range(2):
        for color in range(2):
            O.hsmac_pass(*ref, div2, h, lo[3 * s:3 * s + 3].copy(), hi[3 * s:3 * s + 3].copy(), color, dt, omega,
                         coloring)
    nx, ny, nz = dims
    s_ref = div2[h:h + nx + 1, h:h + ny + 1, h:h + nz + 1].astype(np.float64).sum()
    out = [a.copy(order="F") for a in src]
    fp = (C.c_void_p * 4)(*[a.ctypes.data for a in out])
    s = C.c_double(0)
    rc = emul.hsmac_slab_pair_emul(ti, patch, np.dtype(dtype).itemsize, coloring, nx, ny, nz, h,
                                   lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), C.c_double(dt),
                                   C.c_double(omega), fp, C.byref(s))
    assert rc == 0
    for name, a, b in zip("UVWP", out, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert s.value == pytest.approx(s_ref, rel=1e-12)
