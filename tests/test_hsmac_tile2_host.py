"""The two-sweep HSMAC kernel's own phase functions, executed on the host.

``tests/native/hsmac_tile2_emul.cpp`` runs every workgroup of the launch grid
of ``k_hsmac_sweep2`` thread by thread and phase by phase, in the kernel's
barrier order, over the same ``HsTile2`` code the kernel compiles
(lcs_hsmac_tile2.h).  One pass must equal the oracle's colour-0 and colour-1
passes over box RA followed by its colour-0 and colour-1 passes over box RB
(src/flows/couples/hsmac.jl:1-35, lib/parallel/src/foraxes.jl:73-107) bit for
bit -- every element of U, V, W, P, ghosts included -- and its two div^2 sums
over (1:n+1)^3 must equal the oracle's after each sweep.  Boxes: (d2, d3) as
the time-blocked group uses them, and other pairs; ragged tiles in x and y,
several z-chunkings, three tile shapes, both colourings, both element types.
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
    out = tmp_path_factory.mktemp("hsmac_tile2") / "libhsmac_tile2_emul.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "lcs.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "hsmac_tile2_emul.cpp"), "-o", str(out)], check=True)
    return C.CDLL(str(out))


def _box(name, n):
    n = np.asarray(n)
    lo, hi = {"d1": (-np.ones(3), n + 3), "d2": (np.zeros(3), n + 2), "d3": (np.ones(3), n + 1),
              "core": (2 * np.ones(3), n)}[name]
    return np.asarray(lo, dtype=np.int32), np.asarray(hi, dtype=np.int32)


CASES = [((8, 8, 8), 24, 32, 14, "d2", "d3"), ((16, 12, 10), 24, 32, 5, "d2", "d3"),
         ((40, 30, 7), 24, 32, 3, "d2", "d3"), ((17, 5, 9), 12, 64, 4, "d1", "d2"),
         ((70, 13, 11), 12, 64, 6, "d2", "d3"), ((5, 3, 4), 16, 32, 2, "d3", "core"),
         ((36, 21, 9), 16, 32, 16, "d2", "d3"), ((33, 27, 12), 24, 32, 1, "core", "d1")]


def two_sweeps_oracle(src, h, RA, RB, dt, omega, coloring):
    ref = [a.copy(order="F") for a in src]
    sums = []
    nx, ny, nz = (s - 2 * h for s in src[0].shape)
    for lo, hi in (RA, RB):
        div2 = np.zeros_like(ref[0])
        O.hsmac_pass(*ref, div2, h, lo, hi, 0, dt, omega, coloring)
        O.hsmac_pass(*ref, div2, h, lo, hi, 1, dt, omega, coloring)
        sums.append(div2[h:h + nx + 1, h:h + ny + 1, h:h + nz + 1].astype(np.float64).sum())
    return ref, sums


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("coloring", [0, 1])
@pytest.mark.parametrize("dims,ty,tx,zc,ra,rb", CASES)
def test_two_sweeps_equal_four_colour_passes(emul, dims, ty, tx, zc, ra, rb, coloring, dtype):
    h, dt, omega = 3, 0.01, 1.7
    rng = np.random.default_rng(zlib.crc32(repr((dims, ty, tx, zc, ra, rb, coloring)).encode()))
    src = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    RA, RB = _box(ra, dims), _box(rb, dims)
    ref, sums = two_sweeps_oracle(src, h, RA, RB, dt, omega, coloring)
    dst = [np.full_like(a, np.nan) for a in src]
    sp = (C.c_void_p * 4)(*[a.ctypes.data for a in src])
    dp = (C.c_void_p * 4)(*[a.ctypes.data for a in dst])
    sa, sb = C.c_double(0), C.c_double(0)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = emul.hsmac_sweep2_emul(ty, tx, np.dtype(dtype).itemsize, coloring, *dims, h, p(RA[0]), p(RA[1]), p(RB[0]),
                                p(RB[1]), C.c_double(dt), C.c_double(omega), sp, dp, zc, C.byref(sa), C.byref(sb))
    assert rc == 0
    for name, a, b in zip("UVWP", dst, ref):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert sa.value == pytest.approx(sums[0], rel=1e-12)
    assert sb.value == pytest.approx(sums[1], rel=1e-12)
