"""Serving (lcs_server_*): continuous batching of submitted runs.

* a run's result is the same, bit for bit, when it is served alone and when
  it arrives while other runs (other sizes, particle counts, forcings and
  step counts) are already on the GPU and finishing around it;
* admission respects the slot limit and the device-memory budget, queued
  runs start as slots and memory free up, and every run finishes with
  submit <= start <= finish;
* invalid requests are rejected with an argument error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TARGET = dict(n=32, steps=9, seed=77, particles=3000, st=(0.1, 10.0), forcing="linear-power", forcing_param=0.5,
              k0=4.0, Re=40.0)
OTHERS = [dict(n=48, steps=4, seed=5, particles=20000, st=(0.5, 2.0), forcing="rc-energy", forcing_param=0.649,
               k0=2.0, Re=60.0),
          dict(n=16, steps=14, seed=6, particles=0, forcing="linear-power", forcing_param=0.3, k0=6.0, Re=20.0),
          dict(n=40, steps=6, seed=7, particles=500, st=(1.0, 1.0), forcing="linear-power", forcing_param=0.8,
               k0=4.0, Re=45.0),
          dict(n=24, steps=20, seed=8, particles=7000, st=(0.1, 10.0), forcing="rc-energy", forcing_param=0.5,
               k0=2.0, Re=30.0)]


def run_to_end(S, rid, limit=400):
    for _ in range(limit):
        if S.status(rid)["status"] in ("done", "failed"):
            return
        S.step()
    raise AssertionError("run did not finish")


def snapshot(S, rid):
    b = S.box(rid)
    out = {k: b.download(k) for k in ("U", "V", "W", "P")}
    for k in ("ids", "x", "y", "z", "u", "v", "w", "diams"):
        out["p_" + k] = b.download_prop(0, k)
    return out


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_run_is_the_same_alone_and_in_the_mix(dtype):
    import lcs_mi355x as L
    S1 = L.Server(max_live=8, localize_interval=3)
    a = S1.submit(dtype=dtype, **TARGET)
    run_to_end(S1, a)
    alone = snapshot(S1, a)
    S1.close()

    S2 = L.Server(max_live=8, localize_interval=3)
    for o in OTHERS[:2]:
        S2.submit(dtype=dtype, **o)
    for _ in range(3):  # the target arrives while others are on the GPU
        S2.step()
    b = S2.submit(dtype=dtype, **TARGET)
    for o in OTHERS[2:]:
        S2.submit(dtype=dtype, **o)
    run_to_end(S2, b)
    st = S2.status(b)
    assert st["status"] == "done" and st["steps_done"] == TARGET["steps"]
    mixed = snapshot(S2, b)
    for k in alone:
        np.testing.assert_array_equal(mixed[k], alone[k], err_msg=k)
    S2.close()


def test_admission_by_slots_and_memory():
    import lcs_mi355x as L
    S = L.Server(max_live=2, localize_interval=0)
    ids = [S.submit(dtype="f32", **o) for o in OTHERS] + [S.submit(dtype="f32", **TARGET)]
    seen_live = 0
    for _ in range(200):
        live, _ = S.step()
        seen_live = max(seen_live, live)
        assert live <= 2
        c = S.counts()
        assert c["live"] <= 2
        if all(S.status(i)["status"] == "done" for i in ids):
            break
    assert seen_live == 2
    for i, r in zip(ids, OTHERS + [TARGET]):
        st = S.status(i)
        assert st["status"] == "done" and st["steps_done"] == r["steps"]
        assert st["submit_s"] <= st["start_s"] <= st["finish_s"]
    assert S.counts()["memory_used"] == 0
    for i in ids:
        S.release(i)
    S.close()
    # a budget that holds one 48^3 run (about 13 MB) but not two: they go one by one
    S = L.Server(max_live=8, memory_budget=20 * 2**20, localize_interval=0)
    ids = [S.submit(dtype="f32", **OTHERS[0]), S.submit(dtype="f32", **dict(OTHERS[0], seed=9))]
    live, _ = S.step()
    assert live == 1
    for i in ids:
        run_to_end(S, i)
    assert S.status(ids[0])["finish_s"] <= S.status(ids[1])["start_s"]
    S.close()


def test_requests_that_fit_pass_one_waiting_for_memory():
    """A request waiting for memory lets a later one that fits go first; one
    larger than the whole budget is refused at submit (LCS_ERR_CAPACITY)."""
    import lcs_mi355x as L
    from lcs_mi355x._lib import LCS_ERR_CAPACITY, LCSError
    big = dict(OTHERS[0], steps=6)
    small = dict(OTHERS[1], steps=3)
    S = L.Server(max_live=8, memory_budget=20 * 2**20, localize_interval=0)
    ids = [S.submit(dtype="f32", **big), S.submit(dtype="f32", **dict(big, seed=9)), S.submit(dtype="f32", **small)]
    live, _ = S.step()
    assert live == 2  # the first big run and the small one; the second big run waits
    assert S.status(ids[1])["status"] == "queued" and S.status(ids[2])["status"] == "live"
    for i in ids:
        run_to_end(S, i)
    assert S.status(ids[0])["finish_s"] <= S.status(ids[1])["start_s"]
    assert S.status(ids[2])["start_s"] <= S.status(ids[1])["start_s"]
    S.close()
    S = L.Server(max_live=8, memory_budget=4 * 2**20, localize_interval=0)
    with pytest.raises(LCSError) as e:
        S.submit(dtype="f32", **big)
    assert e.value.code == LCS_ERR_CAPACITY
    S.close()
    S = L.Server(max_live=8)  # the default budget: nine tenths of the free device memory
    S.submit(dtype="f32", **small)
    assert S.step()[0] == 1
    S.close()


def test_invalid_requests_are_rejected():
    import lcs_mi355x as L
    S = L.Server(max_live=2)
    with pytest.raises(L.LCSArgumentError):
        S.submit(n=4, steps=3, seed=1)
    with pytest.raises(L.LCSArgumentError):
        S.submit(n=32, steps=0, seed=1)
    with pytest.raises(L.LCSArgumentError):
        S.submit(n=36, steps=3, seed=1, forcing="rc-energy")  # coarse size 8 must divide n
    S.close()
