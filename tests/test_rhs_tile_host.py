"""The tiled RHS kernel's own phase functions, executed on the host.

``tests/native/rhs_tile_emul.cpp`` runs every workgroup of the launch grid of
``k_rhs_tile`` thread by thread and phase by phase, in the kernel's barrier
order, over the same ``RhsTile`` code the kernel compiles (lcs_rhs_tile.h).
Without FMA contraction the result must equal the oracle's ``rhs``
(src/flows/rhs.jl:59-81) bit for bit, for full and ragged tiles, several
z-chunkings and both element types; the fused form must equal
``base + dUdt*dt`` (src/flows/update.jl:37-51).
"""
import ctypes as C
import os
import subprocess
import zlib

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHAPES = {0: (64, 4), 1: (8, 4), 2: (16, 2), 3: (64, 16), 4: (32, 8)}


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    out = tmp_path_factory.mktemp("rhs_tile") / "librhs_tile_emul.so"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "lcs.jl_amd", "csrc"), os.path.join(ROOT, "tests", "native", "rhs_tile_emul.cpp"),
           "-o", str(out)]
    if not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("HIP headers not installed")
    subprocess.run(cmd, check=True)
    return C.CDLL(str(out))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _run(lib, shape, arrays, h, Re, out, base=None, dt=0.0, zc=None, fused=None):
    nx, ny, nz = (s - 2 * h for s in arrays[0].shape)
    b = base if base is not None else (None, None, None)
    fused = base is not None if fused is None else fused
    rc = lib.rhs_tile_emul(shape, arrays[0].dtype.itemsize, nx, ny, nz, h, C.c_double(Re),
                           *[_p(a) for a in arrays], *[_p(o) for o in out], *[_p(x) for x in b],
                           int(fused), C.c_double(dt), zc or nz)
    assert rc == 0


CASES = [((16, 16, 16), 1, 16), ((20, 13, 11), 1, 4), ((70, 9, 7), 0, 3), ((17, 5, 9), 2, 5),
         ((8, 8, 8), 0, 8), ((33, 3, 12), 2, 1), ((5, 7, 6), 1, 2), ((70, 20, 6), 3, 4), ((40, 17, 5), 4, 2)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,shape,zc", CASES)
def test_tile_rhs_equals_oracle(emul, dims, shape, zc, dtype):
    h = 3
    rng = np.random.default_rng(zlib.crc32(repr((dims, shape, zc)).encode()))
    A = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    for a in A:
        O.ghost_fill(a, h, h, True)
    Re = 37.0
    ref = O.rhs(*A, h, Re)
    out = [np.full(dims, np.nan, dtype=dtype, order="F") for _ in range(3)]
    _run(emul, shape, A, h, Re, out, zc=zc)
    for o, r in zip(out, ref):
        np.testing.assert_array_equal(o, r)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tile_rhs_fused_update(emul, dtype):
    h, dims, dt = 3, (19, 10, 9), 0.0137
    rng = np.random.default_rng(5)
    A = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    for a in A:
        O.ghost_fill(a, h, h, True)
    base = [np.asfortranarray(rng.standard_normal(A[0].shape).astype(dtype)) for _ in range(3)]
    out = [b.copy(order="F") for b in base]
    _run(emul, 1, A, h, 11.0, out, base=base, dt=dt, zc=4)
    dU = O.rhs(*A, h, 11.0)
    for o, b, d in zip(out, base, dU):
        inner = o[h:-h, h:-h, h:-h]
        np.testing.assert_array_equal(inner, b[h:-h, h:-h, h:-h] + d * dtype(dt))
        # nothing outside the interior is written
        o2 = o.copy()
        o2[h:-h, h:-h, h:-h] = b[h:-h, h:-h, h:-h]
        np.testing.assert_array_equal(o2, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tile_rhs_fused_update_of_its_own_input(emul, dtype):
    """Stage 1 updates the array it reads (base == input): with no base array
    the kernel takes the base from its register copy of the input."""
    h, dims, dt = 3, (21, 9, 10), 0.011
    rng = np.random.default_rng(6)
    A = [np.asfortranarray(rng.standard_normal(tuple(n + 2 * h for n in dims)).astype(dtype)) for _ in range(4)]
    for a in A:
        O.ghost_fill(a, h, h, True)
    out = [np.zeros_like(A[0]) for _ in range(3)]
    _run(emul, 1, A, h, 13.0, out, dt=dt, zc=3, fused=True)
    dU = O.rhs(*A, h, 13.0)
    for o, u, d in zip(out, A[:3], dU):
        np.testing.assert_array_equal(o[h:-h, h:-h, h:-h], u[h:-h, h:-h, h:-h] + d * dtype(dt))
