"""Batched ensembles (lcs_batch_*): boxes of uneven size, particle counts and
forcing power advanced with one launch per phase.

* every box of a mixed batch (linear-power and rc-energy forcing, with and
  without particles) follows the oracle's phase-by-phase restatement of
  timestep! + evolve! for that box alone (same bar as the single-box step
  test: sweep counts equal per step, fields within relative L2 1e-8 in FP64);
* a box's result does not depend on the mix: alone in a batch and among other
  boxes it is the same bit for bit (per-box reductions use per-box partials);
* adding, removing and destroying boxes keeps the batch consistent.
"""
import numpy as np
import pytest

from conftest import interior, rel_l2
from test_gpu_step import c1_config, oracle_from_device

pytestmark = pytest.mark.gpu


def mixed_configs(dtype="f64", hsmac=None, rcf=True):
    """Four boxes; the last has rc-energy forcing (linear-power with rcf=False)."""
    import dataclasses

    from lcs_mi355x import config as C
    cfgs = []
    last = C.EnergyPreserveRCF(0.649, 1, 2, 6) if rcf else C.ConstantPowerLF(0.4)
    for dims, npart, forcing, profiles in [((24, 24, 24), 500, C.ConstantPowerLF(0.5), 1),
                                           ((32, 16, 24), 1200, C.ConstantPowerLF(0.3), 2),
                                           ((16, 16, 16), 0, C.ConstantPowerLF(0.7), 0),
                                           ((24, 24, 24), 400, last, 1)]:
        cfg = c1_config(dims=dims, nparticles=max(npart, 1), forcing=forcing, dtype=dtype, profiles=profiles)
        if hsmac is not None:
            cfg = dataclasses.replace(cfg, flow=dataclasses.replace(cfg.flow, couple=hsmac))
        cfgs.append(cfg)
    return cfgs


def test_batched_boxes_follow_the_oracle():
    import lcs_mi355x as L
    cfgs = mixed_configs()
    boxes = [L.build(c) for c in cfgs]
    sims = [oracle_from_device(c, b) for c, b in zip(cfgs, boxes)]
    batch = L.Batch()
    for c, b in zip(cfgs, boxes):
        batch.add(b, L.step_config(c, localize_interval=3))
    for s in range(6):
        logs = batch.step()
        for q, (sim, log) in enumerate(zip(sims, logs)):
            sim.timestep()
            olog = sim.evolve()
            assert log["flow"]["couple1"]["itr"] == olog["flow"]["couple1"]["itr"], (s, q)
            assert log["flow"]["couple2"]["itr"] == olog["flow"]["couple2"]["itr"], (s, q)
            assert np.isclose(log["dt"], sim.dt, rtol=1e-12), (s, q)
            assert np.isclose(log["flow"]["forcing"]["p"], olog["flow"]["forcing"]["p"], rtol=1e-7), (s, q)
    for q, (sim, b) in enumerate(zip(sims, boxes)):
        for name, ref in zip(("U", "V", "W", "P"), (*sim.Us, sim.P)):
            assert rel_l2(interior(b.download(name), 3), interior(ref, 3)) <= 1e-8, (q, name)
    batch.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_box_result_does_not_depend_on_the_mix(dtype):
    import lcs_mi355x as L
    cfgs = mixed_configs(dtype)
    alone = L.build(cfgs[0])
    mix = [L.build(c) for c in cfgs]
    b1, b2 = L.Batch(), L.Batch()
    b1.add(alone, L.step_config(cfgs[0], localize_interval=2))
    for c, b in zip(cfgs, mix):
        b2.add(b, L.step_config(c, localize_interval=2))
    for s in range(5):
        la = b1.step()[0]
        lm = b2.step()[0]
        assert la == lm, s
    for name in ("U", "V", "W", "P"):
        np.testing.assert_array_equal(mix[0].download(name), alone.download(name), err_msg=name)
    for k in ("ids", "x", "y", "z", "u", "v", "w"):
        np.testing.assert_array_equal(mix[0].download_prop(0, k), alone.download_prop(0, k), err_msg=k)
    b1.close()
    b2.close()


def test_batch_membership_changes():
    import lcs_mi355x as L
    from lcs_mi355x._lib import LCSError
    cfgs = mixed_configs()
    boxes = [L.build(c) for c in cfgs]
    batch = L.Batch()
    for c, b in zip(cfgs, boxes):
        batch.add(b, L.step_config(c))
    assert len(batch) == 4
    batch.step()
    with pytest.raises(LCSError):  # a batched box is stepped by its batch only
        L.step(boxes[1], L.step_config(cfgs[1]))
    # leave, step alone, come back: states keep advancing consistently
    batch.remove(boxes[1])
    assert len(batch) == 3
    log = L.step(boxes[1], L.step_config(cfgs[1]))
    assert log["step"] == 2
    batch.add(boxes[1], L.step_config(cfgs[1]))
    logs = batch.step()
    assert [l["step"] for l in logs] == [2, 2, 2, 3]
    boxes[0].close()  # destroying a box detaches it
    assert len(batch) == 3
    batch.step()
    batch.close()


@pytest.mark.parametrize("mode,rcf", [(1, False), (2, False), (3, False), (1, True), (3, True)])
@pytest.mark.parametrize("dtype,tb,itrmin,itrmax", [("f32", True, 6, 300), ("f64", True, 6, 300),
                                                    ("f64", True, 2, 300), ("f32", True, 0, 300),
                                                    ("f64", False, 0, 300), ("f64", False, 3, 300),
                                                    ("f64", True, 6, 6), ("f32", False, 1, 4)])
def test_graph_replays_match_eager_steps(dtype, tb, itrmin, itrmax, mode, rcf):
    """Every replayed step equals the eager step bit for bit, with eager work
    (downloads, a membership change that drops and re-captures the graphs)
    between replays.  Small itrmin puts the group that applies the deferred
    linear forcing right before the first convergence poll; that group must
    not be the one the later polled groups replay.  Mode 1 runs whole-step
    graphs from the third step after a membership change, mode 2 the
    segments between polls only, mode 3 whole-step graphs that stop at their
    gate whenever a stage needs a group past itrmin, so the step goes on from
    the gate on the host, with and without an rc-energy box (its forcing is
    batched from the device table like every other phase).  With itrmax =
    itrmin every group is in the graph and there is no gate."""
    import lcs_mi355x as L
    from lcs_mi355x import config as C
    cfgs = mixed_configs(dtype, C.HSMAC(tb, True, "red-black-fast", 1.7, 1e-3, itrmin, itrmax), rcf=rcf)
    ga = [L.build(c) for c in cfgs]
    ea = [L.build(c) for c in cfgs]
    bg, be = L.Batch(), L.Batch()
    bg.set_graphs(mode)
    be.set_graphs(False)
    for c, b in zip(cfgs, ga):
        bg.add(b, L.step_config(c, localize_interval=2))
    for c, b in zip(cfgs, ea):
        be.add(b, L.step_config(c, localize_interval=2))
    for s in range(7):
        if s == 4:  # membership change: graphs are dropped and captured again
            bg.remove(ga[2])
            be.remove(ea[2])
        lg, le = bg.step(), be.step()
        assert lg == le, s
        for q, (a, b) in enumerate(zip(ga, ea)):
            if s >= 4 and q == 2:
                continue
            for name in ("U", "V", "W", "P"):
                np.testing.assert_array_equal(a.download(name), b.download(name), err_msg=f"step {s} box {q} {name}")
            if a.profiles:
                np.testing.assert_array_equal(a.download_prop(0, "x"), b.download_prop(0, "x"))
    st = bg.graph_stats()
    print("steps by path:", st, flush=True)
    assert sum(st.values()) == 7
    if mode == 2:
        assert st["whole"] == st["gated"] == 0
    else:  # steps 2, 3 and 6 (the two after each membership change are eager and segmented)
        assert st["whole"] + st["gated"] == 3
    bg.close()
    be.close()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_persistent_sweeps_change_no_result(dtype):
    """Persistent sweep grids give the one-workgroup-per-tile results bit for
    bit, logs included: a tile's arithmetic and its partial-sum slot do not
    depend on the workgroup that runs it.  First the grid one CU holds (fewer
    than 8 workgroups: each walks the table by stride), then the grid of two
    CUs (8 or more: workgroups draw tiles from the per-XCD queues, many tiles
    each).  Eight steps; steps 2-4 and 7 replay step graphs, so the queues'
    reset to zero at the end of every launch is exercised by replays too."""
    import lcs_mi355x as L
    cfgs = mixed_configs(dtype)
    ta = [L.build(c) for c in cfgs]
    da = [L.build(c) for c in cfgs]
    bt, bd = L.Batch(), L.Batch()
    for c, b in zip(cfgs, ta):
        bt.add(b, L.step_config(c, localize_interval=2))
    for c, b in zip(cfgs, da):
        bd.add(b, L.step_config(c, localize_interval=2))
    bt.set_persistent(True, cus=1)
    for s in range(8):
        if s == 5:  # (drops the captured graphs: steps 5 and 6 run eagerly)
            bt.set_persistent(True, cus=2)
        la, lb = bt.step(), bd.step()
        for a, b in zip(la, lb):
            assert a == b, (s, a, b)
        for q, (a, b) in enumerate(zip(ta, da)):
            for n in ("U", "V", "W", "P"):
                np.testing.assert_array_equal(a.download(n), b.download(n), err_msg=f"step {s} box {q} {n}")
    st = bt.graph_stats()
    assert sum(st.values()) == 8 and st["whole"] + st["gated"] == 4, st


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_tiling_changes_no_result(dtype):
    """The autotuner's timing runs (sweeps, and stage-1 RHS into the spare
    arrays) leave every box as it was, and the per-class sweep and RHS tilings
    it (or set_tiles, set_rhs_tiles) picks give the default tiling's fields
    and sweep counts bit for bit.  The logged divergence and forcing
    power are sums of per-tile partials: their last bits follow the tile
    count, so they are compared to rounding."""
    import lcs_mi355x as L
    cfgs = mixed_configs(dtype)
    ta = [L.build(c) for c in cfgs]
    da = [L.build(c) for c in cfgs]
    bt, bd = L.Batch(), L.Batch()
    for c, b in zip(cfgs, ta):
        bt.add(b, L.step_config(c, localize_interval=2))
    for c, b in zip(cfgs, da):
        bd.add(b, L.step_config(c, localize_interval=2))
    rtol = 1e-12 if dtype == "f64" else 1e-5

    def same_logs(la, lb, s):
        for a, b in zip(la, lb):
            for c in ("couple1", "couple2"):
                assert a["flow"][c]["itr"] == b["flow"][c]["itr"], s
                assert np.isclose(a["flow"][c]["div"], b["flow"][c]["div"], rtol=rtol, atol=0), s
            assert np.isclose(a["flow"]["forcing"]["p"], b["flow"]["forcing"]["p"], rtol=rtol, atol=0), s
            assert (a["step"], a["t"], a["dt"]) == (b["step"], b["t"], b["dt"]), s

    assert bt.tiles() == bd.tiles() and all(t[:2] == (32, 8) for t in bt.tiles())
    rdef = (32, 4) if dtype == "f64" else (32, 8)
    assert bt.rhs_tiles() == bd.rhs_tiles() == [rdef] * len(ta)
    same_logs(bt.step(), bd.step(), -1)
    names = ("U", "V", "W", "P", "U2", "V2", "W2")
    before = [[b.download(n) for n in names] for b in ta]
    tiles = bt.autotune()
    print("autotuned tiles:", tiles, "RHS tiles:", bt.rhs_tiles(), flush=True)
    for b, ref in zip(ta, before):  # the timing runs changed no box, stage arrays included
        for n, a in zip(names, ref):
            np.testing.assert_array_equal(b.download(n), a, err_msg=n)
    # boxes of one size class share a tiling
    assert tiles[0] == tiles[3] and bt.rhs_tiles()[0] == bt.rhs_tiles()[3]
    for s in range(3):
        if s == 1:  # an explicit tiling of the 32 x 16 x 24 class
            bt.set_tiles(1, 64, 16, 8)
            assert bt.tiles()[1][:2] == (64, 16)
            rset = (64, 8) if dtype == "f64" else (64, 16)
            bt.set_rhs_tiles(1, *rset)
            assert bt.rhs_tiles()[1] == rset
        same_logs(bt.step(), bd.step(), s)
        for q, (a, b) in enumerate(zip(ta, da)):
            for n in ("U", "V", "W", "P"):
                np.testing.assert_array_equal(a.download(n), b.download(n), err_msg=f"step {s} box {q} {n}")
    from lcs_mi355x._lib import LCSError
    with pytest.raises(LCSError):
        bt.set_tiles(0, 128, 8, 28)
    with pytest.raises(LCSError):
        bt.set_rhs_tiles(0, 128, 8)
    bt.close()
    bd.close()
