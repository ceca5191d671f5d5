"""The benchmark workloads themselves on the GPU, at their full sizes.

* C2 (256^3, rc-energy forcing, 2e6 particles, St log-uniform in [0.1, 10]):
  three steps in FP64 (the reference's float type) against the FP64 oracle
  within 1e-9, three in FP32 against the FP32 oracle (HSMAC sweep counts per
  stage equal, velocities within relative L2 1e-5, pressure within 1e-3),
  particle positions by id, cell indices of all device positions equal to the
  oracle's bit for bit; then 200 steps in FP32 against the same run in FP64:
  kinetic energy and dissipation within 0.5 %.
* C3 (1024^3 FP32, linear-power, 6.4e7 particles): size-independent
  properties after two steps: velocity means conserved, the RMS divergence at
  HSMAC exit below epsp/dxmin (from the log and recomputed on the host from
  the downloaded fields), ghost cells equal to their periodic images after a
  fill, cell indices of all 6.4e7 device positions equal to the oracle's hash
  of those positions.
* C4 (96 mixed boxes in one batch): every box equals, bit for bit, the same
  box stepped in a batch of its own, and the first 32^3 box follows the FP32
  oracle.
* C5 (served stream): 60 requests from the workload's own distribution,
  arriving over time; every run completes with its step count, and three
  sampled runs equal the same runs served alone, bit for bit.
* C6 (750^3 FP64, the reference's one-device weak-scaling case) sampled at
  100^3: three steps against the FP64 oracle within 1e-9, particles by id,
  every cell index.
* A served FP64 run, admitted while other runs are on the GPU, follows the
  oracle stepped from the same seeded initial state (fields within relative
  L2 1e-8), for both forcing kinds.
"""
import numpy as np
import pytest

from conftest import interior, rel_l2
from oracle import oracle as O
from test_gpu_step import oracle_from_device

pytestmark = pytest.mark.gpu


def say(*a):
    print("#", *a, flush=True)


def compare_particles_by_id(box, sim, q, tol):
    ps = sim.particles[q]
    nv = ps.nvalid
    ids = box.download_prop(q, "ids")[:nv]
    order = np.argsort(ids)
    oo = np.argsort(ps.ids[:nv])
    np.testing.assert_array_equal(ids[order], ps.ids[:nv][oo])
    for k, ref in zip(("x", "y", "z", "u", "v", "w"), (*ps.xss, *ps.uss)):
        e = rel_l2(box.download_prop(q, k)[:nv][order], ref[:nv][oo])
        assert e < tol, (k, e)


def check_cellindex_against_oracle(box, q, dims, grid_ratio, full_oracle=True):
    """Re-bin on the device, then compare the device hashes of the device
    positions with the oracle's hash of the same positions, every particle."""
    import lcs_mi355x as L
    nv = box.profiles[q]["nvalid"]
    L.makeindex(box, q, 1)
    dh, dstarts, dstops = box.cellindex(q)
    xs = [box.download_prop(q, k)[:nv] for k in ("x", "y", "z")]
    assert len(dh) == nv
    if full_oracle:
        ids = box.download_prop(q, "ids")[:nv]
        oh, _, ostarts, ostops = O.makeindex(dims, grid_ratio, [a.copy() for a in xs], ids.copy(), [], nv)
        np.testing.assert_array_equal(dh, oh)
        np.testing.assert_array_equal(dstarts, ostarts)
        np.testing.assert_array_equal(dstops, ostops)
    else:
        oh = O.hash_cells(dims, grid_ratio, xs, nv)
        np.testing.assert_array_equal(dh, oh)
        # starts/stops: searchsortedfirst/last over the sorted hashes (makeindex.jl:55-64), 1-based
        cells = np.arange(1, len(dstarts) + 1, dtype=np.int64)
        np.testing.assert_array_equal(dstarts, np.searchsorted(dh, cells, "left") + 1)
        np.testing.assert_array_equal(dstops, np.searchsorted(dh, cells, "right"))
    assert np.all(np.diff(dh) >= 0)
    return nv


# ---------------------------------------------------------------------------
# C2: 256^3, rc-energy, 2e6 particles
# ---------------------------------------------------------------------------
def _step_against(box, sim, scfg, nsteps, tol_uvw, tol_p, tol_power, tag):
    """Steps box and oracle side by side; per step: equal HSMAC sweep counts,
    equal dt, fields and forcing power within the given bounds."""
    import lcs_mi355x as L
    p_scale = None
    for s in range(nsteps):
        d = L.step(box, scfg)
        sim.timestep()
        o = sim.evolve()
        dp, op = d["flow"]["forcing"]["p"], o["flow"]["forcing"]["p"]
        errs = {n: rel_l2(interior(box.download(n), 3), interior(r, 3)) for n, r in zip("UVWP", (*sim.Us, sim.P))}
        say(f"{tag} step {s}: itr {d['flow']['couple1']['itr']}/{d['flow']['couple2']['itr']} oracle "
            f"{o['flow']['couple1']['itr']}/{o['flow']['couple2']['itr']} p {dp:.9g} oracle {op:.9g}; rel L2 {errs}")
        assert d["flow"]["couple1"]["itr"] == o["flow"]["couple1"]["itr"], s
        assert d["flow"]["couple2"]["itr"] == o["flow"]["couple2"]["itr"], s
        assert np.isclose(d["dt"], sim.dt, rtol=tol_uvw), s
        assert all(errs[n] <= tol_uvw for n in "UVW"), (s, errs)
        assert errs["P"] <= tol_p, (s, errs)
        # p = <u.src>/dt is a mean of products of both signs.  The first step
        # injects the most power (the shells start far below en0); later steps
        # inject a small difference of large terms, so field differences reach
        # p through that cancellation.  The bound is relative to the forcing's
        # own scale, the first step's power.
        p_scale = abs(op) if p_scale is None else p_scale
        assert abs(dp - op) <= tol_power * p_scale, (s, dp, op, p_scale)


def test_c2_256cubed_rc_energy_against_oracle_then_fp64():
    """C2 at its full size.  In FP64, the reference's float type, three steps
    follow the FP64 oracle to 1e-9 (fields, particles, all 2e6 cell indices
    bit for bit): this pins the algorithm on the workload itself, including
    its first step, whose rc-energy forcing injects far more energy than
    HSMAC removes in itrmax = 300 sweeps.  In FP32 the same three steps follow
    the FP32 oracle with equal sweep counts and velocities within 1e-5; the
    pressure sums the corrections of every sweep (336 in the first step, 300
    of them without convergence) and the device contracts multiply-adds the
    oracle keeps apart, so its bound is 1e-3.  Then 200 steps in all, FP32
    against FP64: kinetic energy and dissipation within 0.5 %."""
    import lcs_mi355x as L
    from lcs_mi355x.workloads import single_box
    cfg64, _ = single_box("c2", dtype="f64")
    cfg, desc = single_box("c2")
    say("C2:", desc)
    b64 = L.build(cfg64)
    s64 = L.step_config(cfg64, localize_interval=30)
    sim = oracle_from_device(cfg64, b64)
    _step_against(b64, sim, s64, 3, 1e-9, 1e-9, 1e-9, "C2 FP64")
    compare_particles_by_id(b64, sim, 0, 1e-9)
    nv = check_cellindex_against_oracle(b64, 0, cfg64.grid.dims, 4)
    say("C2 FP64 cell indices equal the oracle's for", nv, "particles")
    del sim
    box = L.build(cfg)
    scfg = L.step_config(cfg, localize_interval=30)
    sim = oracle_from_device(cfg, box)
    _step_against(box, sim, scfg, 3, 1e-5, 1e-3, 1e-5, "C2 FP32")
    compare_particles_by_id(box, sim, 0, 1e-5)
    nv = check_cellindex_against_oracle(box, 0, cfg.grid.dims, 4)
    say("C2 FP32 cell indices equal the oracle's for", nv, "particles")
    del sim
    # 200 steps in all: FP32 against the same workload in FP64
    for _ in range(197):
        L.step(box, scfg, want_log=False)
        L.step(b64, s64, want_log=False)
    l32, l64 = L.read_log(box), L.read_log(b64)
    assert l32["step"] == l64["step"] == 200
    f32, f64 = L.flow_stat(box), L.flow_stat(b64)
    say(f"C2 after 200 steps: k {f32['k']:.6g} (FP64 {f64['k']:.6g}), eps {f32['eps']:.6g} (FP64 {f64['eps']:.6g})")
    assert abs(f32["k"] / f64["k"] - 1) < 5e-3
    assert abs(f32["eps"] / f64["eps"] - 1) < 5e-3
    box.close()
    b64.close()


def test_c6_reference_case_sampled_against_oracle():
    """The 750^3 FP64 reference case (cfgs/N1500.lcs-yaml settings at one
    device's share: rc-energy, Re 2220, nonlinear drag, half the particle
    slots spare) sampled at 100^3 with the same particle density: three steps
    follow the FP64 oracle within 1e-9 (fields, forcing power, sweep counts),
    particles by id, and every cell index equals the oracle's."""
    import lcs_mi355x as L
    from lcs_mi355x.workloads import single_box
    cfg, desc = single_box("c6", n_override=100)
    say("C6 sampled:", desc)
    p = cfg.particles[0]
    assert p.drag == "nonlinear" and p.population.ntotal > p.population.valid
    box = L.build(cfg)
    scfg = L.step_config(cfg, localize_interval=30)
    sim = oracle_from_device(cfg, box)
    _step_against(box, sim, scfg, 3, 1e-9, 1e-9, 1e-9, "C6 sampled")
    compare_particles_by_id(box, sim, 0, 1e-9)
    nv = check_cellindex_against_oracle(box, 0, cfg.grid.dims, 4)
    say("C6 sampled cell indices equal the oracle's for", nv, "particles")
    box.close()


# ---------------------------------------------------------------------------
# C3: 1024^3 FP32 properties
# ---------------------------------------------------------------------------
def _host_rms_div(box, n, h):
    """RMS of the discrete divergence (hsmac.jl:19-21) over the interior cells
    of the box's current U, V, W (ghosts filled first), in z-slabs."""
    import lcs_mi355x as L
    L.fill_ghost(box, ("U", "V", "W"), 1, True)
    U, V, W = (box.download(k) for k in "UVW")
    di = n / (2 * np.pi)
    acc = 0.0
    for z0 in range(h, n + h, 64):
        z1 = min(z0 + 64, n + h)
        sl = np.s_[h:n + h, h:n + h, z0:z1]
        d = (U[h:n + h, h:n + h, z0:z1].astype(np.float64) - U[h - 1:n + h - 1, h:n + h, z0:z1]) * di
        d += (V[h:n + h, h:n + h, z0:z1].astype(np.float64) - V[h:n + h, h - 1:n + h - 1, z0:z1]) * di
        d += (W[sl].astype(np.float64) - W[h:n + h, h:n + h, z0 - 1:z1 - 1]) * di
        acc += float(np.sum(d * d))
    return np.sqrt(acc / n ** 3)


def test_c3_1024cubed_properties():
    import lcs_mi355x as L
    from lcs_mi355x.workloads import single_box
    cfg, desc = single_box("c3")
    say("C3:", desc)
    box = L.build(cfg)
    n, h = 1024, 3
    st0 = L.flow_stat(box)
    scfg = L.step_config(cfg, localize_interval=30)
    hs = cfg.flow.couple
    dxmin = 2 * np.pi / n
    for s in range(2):
        log = L.step(box, scfg)
        say(f"C3 step {s}: itr {log['flow']['couple1']['itr']}/{log['flow']['couple2']['itr']} "
            f"div {log['flow']['couple1']['div']:.4g}/{log['flow']['couple2']['div']:.4g} dt {log['dt']:.4g}")
        for c in ("couple1", "couple2"):
            itr, div = log["flow"][c]["itr"], log["flow"][c]["div"]
            assert hs.itrmin <= itr < hs.itrmax
            assert div * dxmin < hs.epsp  # the exit test of hsmac.jl:128,193
    st1 = L.flow_stat(box)
    for c in "uvw":  # HSMAC, the flux-form RHS and the linear forcing conserve the means
        assert abs(st1[c]["mean"] - st0[c]["mean"]) < 1e-6, (c, st0[c], st1[c])
    div = _host_rms_div(box, n, h)
    say(f"C3 host RMS div of the final fields {div:.4g} (log {log['flow']['couple2']['div']:.4g})")
    assert div * dxmin < hs.epsp
    # ghost cells after a FullHalo fill are the periodic images A[mod1(i, n)]
    L.fill_ghost(box, ("P",), h, True)
    P = box.download("P")
    for ax in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = slice(0, h), slice(n, n + h)
        np.testing.assert_array_equal(P[tuple(lo)], P[tuple(hi)])
        lo[ax], hi[ax] = slice(n + h, n + 2 * h), slice(h, 2 * h)
        np.testing.assert_array_equal(P[tuple(lo)], P[tuple(hi)])
    del P
    nv = check_cellindex_against_oracle(box, 0, cfg.grid.dims, 4, full_oracle=False)
    assert nv == 64_000_000
    say("C3 cell indices equal the oracle's hash for", nv, "particles")
    box.close()


# ---------------------------------------------------------------------------
# C4: 96 mixed boxes in one batch
# ---------------------------------------------------------------------------
def _snapshot(box):
    out = {k: box.download(k) for k in "UVWP"}
    if box.profiles:
        for k in ("ids", "x", "y", "z", "u", "v", "w"):
            out[k] = box.download_prop(0, k)
    return out


def test_c4_96_box_batch_each_box_as_alone_and_one_against_oracle():
    import lcs_mi355x as L
    from lcs_mi355x.workloads import ensemble_configs
    cfgs = ensemble_configs(96)
    say("C4: 96 boxes, N", sorted({c.grid.dims[0] for c in cfgs}), "cells",
        sum(int(np.prod(c.grid.dims)) for c in cfgs))
    boxes = [L.build(c) for c in cfgs]
    pick = next(i for i, c in enumerate(cfgs) if c.grid.dims[0] == 32)
    sim = oracle_from_device(cfgs[pick], boxes[pick])
    batch = L.Batch()
    for c, b in zip(cfgs, boxes):
        batch.add(b, L.step_config(c, localize_interval=2))
    nsteps = 3
    logs = [batch.step() for _ in range(nsteps)]
    for s in range(nsteps):
        sim.timestep()
        o = sim.evolve()
        d = logs[s][pick]
        assert d["flow"]["couple1"]["itr"] == o["flow"]["couple1"]["itr"], s
        assert d["flow"]["couple2"]["itr"] == o["flow"]["couple2"]["itr"], s
    errs = {n: rel_l2(interior(boxes[pick].download(n), 3), interior(r, 3)) for n, r in zip("UVWP", (*sim.Us, sim.P))}
    say(f"C4 box {pick} (32^3, {cfgs[pick].particles[0].population.valid} particles) vs oracle:", errs)
    assert all(e <= 1e-5 for e in errs.values()), errs
    compare_particles_by_id(boxes[pick], sim, 0, 1e-5)
    mixed = [_snapshot(b) for b in boxes]
    batch.close()
    for b in boxes:
        b.close()
    # each box again, alone in a batch of its own
    one = L.Batch()
    for i, c in enumerate(cfgs):
        b = L.build(c)
        one.add(b, L.step_config(c, localize_interval=2))
        alone_logs = [one.step()[0] for _ in range(nsteps)]
        assert alone_logs == [lg[i] for lg in logs], i
        snap = _snapshot(b)
        for k in snap:
            np.testing.assert_array_equal(snap[k], mixed[i][k], err_msg=f"box {i} {k}")
        one.remove(b)
        b.close()
    one.close()
    say("C4: all 96 boxes equal their runs alone, bit for bit")


# ---------------------------------------------------------------------------
# C5: a served stream
# ---------------------------------------------------------------------------
def _serve(reqs, arrivals, max_live=128, keep=()):
    """Submit reqs[i] at server step arrivals[i]; run until all finish.
    Returns (ids, snapshots of the kept request indices)."""
    import lcs_mi355x as L
    S = L.Server(max_live=max_live, localize_interval=10)
    ids = [None] * len(reqs)
    k, step = 0, 0
    order = np.argsort(arrivals, kind="stable")
    while True:
        while k < len(reqs) and arrivals[order[k]] <= step:
            ids[order[k]] = S.submit(**reqs[order[k]])
            k += 1
        S.step()
        step += 1
        if k == len(reqs) and all(S.status(i)["status"] in ("done", "failed") for i in ids):
            break
        assert step < 5000
    snaps = {}
    for i in keep:
        b = S.box(ids[i])
        snaps[i] = _snapshot(b)
    stats = [S.status(i) for i in ids]
    assert S.counts()["memory_used"] == 0
    S.close()
    return stats, snaps


def test_c5_served_stream_completes_and_runs_equal_their_runs_alone():
    from lcs_mi355x.workloads import serve_requests
    reqs = serve_requests(60)
    rng = np.random.default_rng(3)
    arrivals = np.floor(np.cumsum(rng.exponential(3.0, len(reqs)))).astype(int)  # one request per 3 steps
    lin = [i for i, r in enumerate(reqs) if r["forcing"] == "linear-power"]
    rcf = [i for i, r in enumerate(reqs) if r["forcing"] == "rc-energy"]
    sample = sorted({min(lin, key=lambda i: reqs[i]["steps"]), min(rcf, key=lambda i: reqs[i]["steps"]),
                     max(range(len(reqs)), key=lambda i: reqs[i]["particles"] / reqs[i]["steps"])})
    say("C5: 60 requests over", int(arrivals[-1]), "server steps; sampled", sample)
    stats, mixed = _serve(reqs, arrivals, keep=sample)
    for r, st in zip(reqs, stats):
        assert st["status"] == "done" and st["steps_done"] == r["steps"]
        assert st["submit_s"] <= st["start_s"] <= st["finish_s"]
    for i in sample:
        _, alone = _serve([reqs[i]], np.zeros(1, int), keep=[0])
        for k in alone[0]:
            np.testing.assert_array_equal(alone[0][k], mixed[i][k], err_msg=f"request {i} {k}")
    say("C5: sampled runs equal their runs alone")


@pytest.mark.parametrize("forcing", ["linear-power", "rc-energy"])
def test_served_fp64_run_follows_the_oracle(forcing):
    import lcs_mi355x as L
    from lcs_mi355x.workloads import request_box
    target = dict(n=32, particles=3000, st=(0.1, 10.0), forcing=forcing,
                  forcing_param=0.5 if forcing == "linear-power" else 0.649, steps=12, seed=321, k0=4.0, Re=40.0,
                  dtype="f64")
    others = [dict(n=48, particles=20000, st=(0.5, 2.0), forcing="rc-energy", forcing_param=0.649, steps=5, seed=5,
                   k0=2.0, Re=60.0, dtype="f64"),
              dict(n=16, particles=0, forcing="linear-power", forcing_param=0.3, steps=20, seed=6, k0=6.0, Re=20.0,
                   dtype="f64")]
    S = L.Server(max_live=8, localize_interval=3)
    for o in others:
        S.submit(**o)
    for _ in range(2):
        S.step()
    rid = S.submit(**target)
    for _ in range(400):
        if S.status(rid)["status"] == "done":
            break
        S.step()
    assert S.status(rid)["status"] == "done"
    served = _snapshot(S.box(rid))
    S.close()
    box, cfg = request_box(target)
    sim = oracle_from_device(cfg, box)
    for _ in range(target["steps"]):
        sim.timestep()
        sim.evolve()
    errs = {n: rel_l2(interior(served[n], 3), interior(r, 3)) for n, r in zip("UVWP", (*sim.Us, sim.P))}
    say(f"served {forcing} run vs oracle after {target['steps']} steps:", errs)
    assert all(e <= 1e-8 for e in errs.values()), errs
    ps = sim.particles[0]
    nv = ps.nvalid
    order = np.argsort(served["ids"][:nv])
    oo = np.argsort(ps.ids[:nv])
    np.testing.assert_array_equal(served["ids"][:nv][order], ps.ids[:nv][oo])
    for k, ref in zip(("x", "y", "z", "u", "v", "w"), (*ps.xss, *ps.uss)):
        assert rel_l2(served[k][:nv][order], ref[:nv][oo]) < 1e-10, k
    box.close()
