"""lcs_force_rcf_noalloc, the rc-energy forcing that allocates nothing, against
lcs_force_rcf on the same inputs, called through the C ABI.

The two entry points run the same kernels; only the home of the coarse grids,
spectra and crop differs (a fixed per-device array against the box's own
buffer), so fields and power must agree bit for bit.  Two calls in a row, the
second on a smaller box with a smaller coarse grid, check that the shared
array's leftovers from the first call never reach the second.  A request
larger than the fixed array is refused with LCS_ERR_CAPACITY and leaves the
fields alone.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rand_halo

pytestmark = pytest.mark.gpu


def lcs():
    import lcs_mi355x
    lcs_mi355x.load()
    return lcs_mi355x


def twin_boxes(rng, dims, dtype):
    L = lcs()
    h = 3
    Us = [rand_halo(rng, dims, h, dtype) for _ in range(3)]
    boxes = []
    for _ in range(2):
        b = L.Box(dims, h, L.FlowParams(1.0, 1.0, 1.0), dtype)
        for name, A in zip(("U", "V", "W"), Us):
            b.upload(name, A)
        boxes.append(b)
    return boxes


def force(box, fn, energy, kmin, kmax, c, dt):
    from lcs_mi355x._lib import check
    p = C.c_double()
    check(getattr(box.L, fn)(box.handle, float(energy), int(kmin), int(kmax), int(c), float(dt), C.byref(p)))
    return p.value


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_rcf_equals_lcs_force_rcf_on_two_calls(dtype):
    rng = np.random.default_rng(11)
    # (dims, coarse_size, kmin, kmax, energy): the second call is smaller in
    # every respect, so a stale tail of the first call's scratch would show
    cases = [((48, 40, 32), 8, 1, 3, 0.649), ((16, 16, 16), 4, 1, 2, 1.0)]
    for dims, c, kmin, kmax, e in cases:
        print(f"# noalloc rcf {np.dtype(dtype).name} dims {dims} c {c} k {kmin}..{kmax}", flush=True)
        ref, new = twin_boxes(rng, dims, dtype)
        p_ref = force(ref, "lcs_force_rcf", e, kmin, kmax, c, 0.01)
        p_new = force(new, "lcs_force_rcf_noalloc", e, kmin, kmax, c, 0.01)
        assert p_new == p_ref, (dims, p_new, p_ref)
        for name in ("U", "V", "W"):
            a, b = new.download(name), ref.download(name)
            assert np.array_equal(a, b), (dims, name)
        ref.close()
        new.close()


def test_noalloc_rcf_refuses_sizes_beyond_its_array():
    from lcs_mi355x._lib import LCS_ERR_CAPACITY, LCSError
    rng = np.random.default_rng(12)
    (box, _) = twin_boxes(rng, (64, 64, 64), np.float64)
    before = box.download("U")
    with pytest.raises(LCSError) as ei:
        force(box, "lcs_force_rcf_noalloc", 1.0, 1, 2, 64, 0.01)  # 3*64^3 doubles > 2^17
    assert ei.value.code == LCS_ERR_CAPACITY and "scratch" in str(ei.value)
    assert np.array_equal(box.download("U"), before)
    # the same box still forces with a size that fits
    force(box, "lcs_force_rcf_noalloc", 1.0, 1, 2, 8, 0.01)
    box.close()
