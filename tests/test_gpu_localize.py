"""Particle storage order (lcs_particles_localize): the device may keep a
profile's particles in any slot order; every entry point presents the logical
order (the reference's array order), and no result changes.

* localize, then download: every property equals its pre-localize download,
  bit for bit; uploads after localize land on the right particles;
* whole steps with localize every few steps equal steps without it, bit for
  bit, fields and all 17 properties;
* makeindex after localize equals makeindex without it (the stable sort sees
  the logical order, src/particles/makeindex.jl:29-67).
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import lcs, make_box, particle_box
from test_gpu_step import c1_config

pytestmark = pytest.mark.gpu


def all_props(box, prof):
    L = lcs()
    return {k: box.download_prop(prof, k) for k in L.core._lib.PROPS}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_localize_keeps_logical_order(rng, dtype):
    L = lcs()
    box, prof, Us, props = particle_box(rng, dtype, 3000, dims=(32, 24, 16), ntotal=3100)
    before = all_props(box, prof)
    L.localize(box, prof)
    after = all_props(box, prof)
    for k in before:
        np.testing.assert_array_equal(after[k], before[k], err_msg=k)
    # an upload after localize reaches the same logical particles
    u = rng.standard_normal(3100).astype(dtype)
    box.upload_prop(prof, "u", u)
    L.localize(box, prof)  # compose a second reordering with the first
    np.testing.assert_array_equal(box.download_prop(prof, "u"), u)
    np.testing.assert_array_equal(box.download_prop(prof, "ids"), before["ids"])


def test_steps_with_localize_equal_steps_without():
    import lcs_mi355x as L
    cfg = c1_config(dims=(32, 32, 32), nparticles=4000, profiles=2)
    boxes = [L.build(cfg), L.build(cfg)]
    cfgs = [L.step_config(cfg), L.step_config(cfg, localize_interval=3)]
    for s in range(8):
        logs = [L.step(b, c) for b, c in zip(boxes, cfgs)]
        assert logs[0]["flow"]["couple1"]["itr"] == logs[1]["flow"]["couple1"]["itr"], s
        assert logs[0]["dt"] == logs[1]["dt"], s
    for name in ("U", "V", "W", "P"):
        np.testing.assert_array_equal(boxes[1].download(name), boxes[0].download(name), err_msg=name)
    for q in range(2):
        a, b = all_props(boxes[0], q), all_props(boxes[1], q)
        for k in a:
            np.testing.assert_array_equal(b[k], a[k], err_msg=f"profile {q} {k}")


def test_makeindex_after_localize(rng):
    L = lcs()
    dims, n = (32, 32, 32), 5000
    boxes = [make_box(dims, 3), make_box(dims, 3)]
    xs = [rng.uniform(-0.1, O.DOMAIN_LENGTH + 0.1, n + 7) for _ in range(3)]
    ids = np.arange(1, n + 8, dtype=np.int64)
    for box in boxes:
        prof = box.add_particles(n, n + 7, grid_ratio=4)
        for k, a in zip(("x", "y", "z"), xs):
            box.upload_prop(prof, k, a)
        box.upload_prop(prof, "ids", ids)
    L.localize(boxes[1], 0)
    for box in boxes:
        L.makeindex(box, 0, 1)
    c0, c1 = boxes[0].cellindex(0), boxes[1].cellindex(0)
    for a, b in zip(c0, c1):
        np.testing.assert_array_equal(b, a)
    for k in ("ids", "x", "y", "z"):
        np.testing.assert_array_equal(boxes[1].download_prop(0, k), boxes[0].download_prop(0, k), err_msg=k)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_repeated_reorders_keep_buffers_consistent(rng, dtype):
    """makeindex and localize permute every property through a spare buffer of
    the property's own element size; repeated calls in any order (FP32 has
    4-byte properties beside the 8-byte ids) must leave every download intact."""
    L = lcs()
    box, prof, Us, props = particle_box(rng, dtype, 2000, dims=(32, 32, 32), ntotal=2100)
    L.makeindex(box, prof, 1)
    once = all_props(box, prof)
    L.makeindex(box, prof, 1)  # stable sort of sorted data: unchanged
    for k, v in all_props(box, prof).items():
        np.testing.assert_array_equal(v, once[k], err_msg=k)
    for _ in range(2):
        L.localize(box, prof)
        for k, v in all_props(box, prof).items():
            np.testing.assert_array_equal(v, once[k], err_msg=k)
    L.makeindex(box, prof, 1)
    for k, v in all_props(box, prof).items():
        np.testing.assert_array_equal(v, once[k], err_msg=k)
