"""Tile shapes of the fused kernels (lcs_autotune / lcs_set_tiles).

* every candidate shape of the HSMAC sweep, of the two-sweep pass (FP32) and
  of the fused RHS gives the same box, bit for bit, after full steps (FP32 and
  FP64; ragged tiles in x and y), so the tuner can never change a result;
* lcs_autotune returns one of the candidates, keeps it for the box, and
  answers a second box of the same size from its cache with the same choice;
* shapes that are not candidates are rejected.
"""
import numpy as np
import pytest

from test_gpu_step import c1_config

pytestmark = pytest.mark.gpu

SWEEP = {"f32": [(64, 16), (64, 8), (32, 16), (32, 8), (128, 8), (128, 16)], "f64": [(64, 16), (64, 8), (32, 16), (32, 8)]}
RHS = {"f32": [(64, 16), (64, 8), (32, 16), (32, 8)], "f64": [(64, 8), (32, 8), (32, 16), (32, 4)]}
SWEEP2 = [(32, 24), (64, 12), (32, 16)]


def run(cfg, sweep=None, rhs=None, sweep2=(0, 0), steps=3):
    import lcs_mi355x as L
    box = L.build(cfg)
    if sweep or rhs:
        L.set_tiles(box, sweep=sweep, rhs=rhs, sweep2=sweep2)
    sc = L.step_config(cfg)
    itrs = [L.step(box, sc)["flow"]["couple1"]["itr"] for _ in range(steps)]
    out = {k: box.download(k) for k in ("U", "V", "W", "P")}
    out["x"] = box.download_prop(0, "x")
    box.close()
    return itrs, out


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_every_shape_gives_the_same_box(dtype):
    cfg = c1_config(dims=(44, 27, 30), dtype=dtype, nparticles=2000)
    ref_itrs, ref = run(cfg, sweep=(64, 8), rhs=RHS[dtype][-1])
    cases = [(s, RHS[dtype][-1], (0, 0)) for s in SWEEP[dtype]] + [((64, 8), r, (0, 0)) for r in RHS[dtype]]
    if dtype == "f32":  # the two-sweep pass for the d2 and d3 sweeps before itrmin
        cases += [((64, 8), RHS[dtype][-1], s2) for s2 in SWEEP2]
    for sweep, rhs, sweep2 in cases:
        itrs, out = run(cfg, sweep=sweep, rhs=rhs, sweep2=sweep2)
        assert itrs == ref_itrs, (sweep, rhs, sweep2)
        for k in ref:
            np.testing.assert_array_equal(out[k], ref[k], err_msg=f"{k} sweep {sweep} rhs {rhs} sweep2 {sweep2}")


def test_autotune_picks_a_candidate_and_caches_it():
    import lcs_mi355x as L
    cfg = c1_config(dims=(40, 40, 40), dtype="f32", nparticles=1000)
    a, b = L.build(cfg), L.build(cfg)
    t1 = L.autotune(a)
    assert tuple(t1["sweep"]) in SWEEP["f32"] and tuple(t1["rhs"]) in RHS["f32"]
    assert tuple(t1["sweep2"]) in SWEEP2 + [(0, 0)]
    assert L.get_tiles(a) == t1
    assert L.autotune(b) == t1  # cached by size
    a.close()
    b.close()


def test_unsupported_shapes_are_rejected():
    import lcs_mi355x as L
    box = L.build(c1_config(dims=(16, 16, 16), dtype="f64", nparticles=100))
    with pytest.raises(L.LCSArgumentError):
        L.set_tiles(box, sweep=(128, 8))  # too much LDS in FP64
    with pytest.raises(L.LCSArgumentError):
        L.set_tiles(box, rhs=(64, 16))  # too much LDS in FP64
    L.set_tiles(box)
    box.close()
