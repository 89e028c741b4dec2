"""Serving loop on the CPU: admission by footprint, slot reuse, eviction,
rejection and failure handling, with a fake backend standing in for the
device (the GPU tests run the real one)."""
import math
import os
import sys
import threading
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gevolution-1.2_amd"))

import serving as sv  # noqa: E402


class FakeBackend:
    """Boxes cost N^3 bytes; a run sleeps in proportion to N and returns
    which box it ran on."""

    def __init__(self, unit=0.004, fail=()):
        self.unit = unit
        self.fail = set(fail)
        self.lock = threading.Lock()
        self.live = 0
        self.peak_live = 0
        self.boxes = set()
        self.next_id = 0

    def key(self, r):
        return (r.ngrid, r.form)

    def footprint(self, r):
        return r.ngrid ** 3

    def create(self, r):
        with self.lock:
            self.next_id += 1
            box = (self.next_id, r.ngrid)
            self.boxes.add(box)
        return box

    def destroy(self, box):
        with self.lock:
            self.boxes.remove(box)

    def run(self, box, r):
        assert box in self.boxes and box[1] == r.ngrid
        with self.lock:
            self.live += 1
            self.peak_live = max(self.peak_live, self.live)
        time.sleep(self.unit * r.ngrid / 16)
        with self.lock:
            self.live -= 1
        if r.rid in self.fail:
            raise RuntimeError("injected failure")
        return {"particle_steps": r.ngrid ** 3 * 10, "box": box}


def reqs(grids, arrivals=None):
    arrivals = arrivals or [0.0] * len(grids)
    return [sv.Request(i, g, arrival=t) for i, (g, t) in enumerate(zip(grids, arrivals))]


def test_all_complete_within_slots_and_budget():
    be = FakeBackend()
    budget = 3 * 32 ** 3
    with sv.Server(be, budget, slots=3) as s:
        out = s.serve(reqs([16, 32, 16, 32, 16, 32, 16, 16]))
        assert [o.rid for o in out] == list(range(8))
        assert all(o.status == "ok" for o in out)
        assert s.peak_running <= 3 and be.peak_live <= 3
        assert s.peak_in_use <= budget
        assert all(o.admitted >= o.arrival and o.finished >= o.admitted for o in out)
    assert s.in_use == 0 and not be.boxes  # close() freed the pool


def test_slot_reuse_keeps_one_box_per_shape():
    be = FakeBackend()
    with sv.Server(be, 10 ** 9, slots=1) as s:
        out = s.serve(reqs([32, 32, 32, 32]))
        assert s.created == 1
        assert [o.reused for o in out] == [False, True, True, True]
        assert len({o.result["box"] for o in out}) == 1
    with sv.Server(be, 10 ** 9, slots=1, reuse=False) as s:
        out = s.serve(reqs([32, 32, 32]))
        assert s.created == 3 and not any(o.reused for o in out)
        assert not be.boxes  # plain runs free their box at once


def test_idle_boxes_are_evicted_to_admit_another_shape():
    be = FakeBackend()
    budget = 32 ** 3 + 16 ** 3 - 1  # a 32 box and a 16 box never fit together
    with sv.Server(be, budget, slots=2) as s:
        out = s.serve(reqs([16, 32, 16]))
        assert all(o.status == "ok" for o in out)
        assert s.evicted >= 1 and s.peak_in_use <= budget
        assert be.peak_live == 1


def test_oversized_request_is_rejected_without_blocking_the_queue():
    be = FakeBackend()
    with sv.Server(be, 32 ** 3, slots=2) as s:
        out = s.serve(reqs([16, 64, 32, 16]))
    assert [o.status for o in out] == ["ok", "rejected", "ok", "ok"]
    assert "exceeds the budget" in out[1].error


def test_failure_is_reported_and_its_box_not_reused():
    be = FakeBackend(fail={1})
    with sv.Server(be, 10 ** 9, slots=1) as s:
        out = s.serve(reqs([16, 16, 16]))
        assert [o.status for o in out] == ["ok", "failed", "ok"]
        assert "injected" in out[1].error
        assert out[2].result["box"] != out[1].result.get("box")
        assert s.created == 2  # the failed run's box was destroyed, not pooled
    assert s.in_use == 0 and not be.boxes


def test_arrivals_are_waited_for_and_admitted_in_order():
    be = FakeBackend(unit=0.002)
    arr = [0.0, 0.05, 0.06, 0.2]
    with sv.Server(be, 10 ** 9, slots=1) as s:
        out = s.serve(reqs([16, 16, 16, 16], arr))
    for o, t in zip(out, arr):
        assert o.admitted >= t - 1e-9
    adm = [o.admitted for o in out]
    assert adm == sorted(adm)
    summ = sv.summarize(out)
    assert summ["completed"] == 4 and summ["makespan_s"] >= 0.2
    assert summ["particle_steps"] == 4 * 16 ** 3 * 10


def test_overlap_beats_back_to_back_on_a_loaded_trace():
    trace = reqs([32] * 8)
    t0 = time.perf_counter()
    with sv.Server(FakeBackend(unit=0.01), 10 ** 9, slots=4) as s:
        mix = sv.summarize(s.serve(trace))
    t_mix = time.perf_counter() - t0
    t0 = time.perf_counter()
    with sv.Server(FakeBackend(unit=0.01), 10 ** 9, slots=1, reuse=False) as s:
        plain = sv.summarize(s.serve(trace))
    t_plain = time.perf_counter() - t0
    assert t_mix < t_plain
    assert mix["p99_s"] < plain["p99_s"]


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert sv.percentile(v, 50) == 50 and sv.percentile(v, 99) == 99 and sv.percentile(v, 100) == 100
    assert sv.percentile([3.0], 99) == 3.0
    assert math.isnan(sv.percentile([], 50))
    assert sv.percentile([5, 1, 4, 2, 3], 50) == 3


def test_poisson_trace():
    a = sv.poisson_trace(4000, 8.0, 3)
    b = sv.poisson_trace(4000, 8.0, 3)
    assert [(r.ngrid, r.gr, r.arrival) for r in a] == [(r.ngrid, r.gr, r.arrival) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 8.0) < 0.1 / 8.0
    grids = np.array([r.ngrid for r in a])
    for g, w in zip((64, 96, 128, 192, 256), (0.3, 0.25, 0.2, 0.15, 0.1)):
        assert abs(np.mean(grids == g) - w) < 0.03
    assert abs(np.mean([r.gr for r in a]) - 0.5) < 0.03
    om = np.array([r.Omega_m for r in a])
    assert om.min() >= 0.25 and om.max() <= 0.35
    assert [r.seed for r in a[:5]] == [0, 1, 2, 3, 4]
    assert a[0].box_mpc == 5.0 * a[0].ngrid


def test_default_template_is_the_sc1_crystal():
    gev = pytest.importorskip("gevpm")
    if not os.path.exists(gev.LIB_PATH):
        pytest.skip("library not built")
    tpl = gev.load_homogeneous_template(os.path.join(ROOT, "tests", "golden", "sc1_crystal.dat"))
    be = sv.GevBackend()
    assert be.tile == 4
    # the same particles (the file lists them in another order, which only
    # changes which particle gets which ID)
    key = lambda t: t[np.lexsort(t.T[::-1])]
    np.testing.assert_array_equal(key(np.asarray(tpl, dtype=np.float32).reshape(-1, 3)), key(be.template))


class RecordingBox:
    """Stands in for gevpm.Box: records the calls GevBackend.run makes."""

    def __init__(self):
        self.calls = []

    def generateIC_basic(self, sim, ic, tpl, k, dt, tt, dc, tc):
        self.calls.append(("ic", sim.gr_flag, ic.seed, ic.numtile, ic.boxsize, ic.A_s, len(tpl), float(k[0])))
        return 1e-3, 0.1

    def run_to_outputs(self, sim, st, z_pk, numbins, boxsize):
        self.calls.append(("outputs", list(z_pk), numbins, boxsize))
        st.cycle += 117
        st.a = 1.001
        z = np.zeros((1, numbins))
        return {"k": z, "pk": z + 1.0, "sigma_k": z, "sigma_pk": z, "occupation": np.ones((1, numbins), dtype=np.int32),
                "z": np.zeros(1), "cycle": np.array([117]), "interpolated": np.array([True])}

    def field_sum(self, f):
        return 0.5


@pytest.mark.parametrize("gr", [True, False])
def test_backend_runs_the_reference_sequence(gr):
    """IC generation with the request's seed, tiling N/4 and a 5 Mpc/h-per-cell
    box, then the main loop with the delta spectrum at z = 0 in N/2 bins
    (gev_run_to_outputs); the particle steps are the completed updates."""
    gev = pytest.importorskip("gevpm")
    if not os.path.exists(gev.LIB_PATH):
        pytest.skip("library not built")
    r = sv.Request(3, 32, gr=gr, Omega_m=0.3, A_s=2.0e-9, seed=11)
    sim, st = sv.GevBackend.setup(r)
    box = RecordingBox()
    res = sv.GevBackend().run(box, r)
    ic = box.calls[0]
    assert ic[:6] == ("ic", 1 if gr else 0, 11, 8, 160.0, 2.0e-9) and ic[6] == 64
    assert abs(ic[7] - 1e-5 * 160.0) < 1e-18  # k in units of 1/box (loadTransferFunctions)
    assert box.calls[1:] == [("outputs", [0.0], 16, 160.0)]
    assert res["cycles"] == 117 and res["particle_steps"] == 32 ** 3 * 117 and res["z"] == 0.0
    assert sim.fourpiG == 1.5 * 160.0 ** 2 / sv.C_LIGHT ** 2
    assert abs(sim.cosmo.Omega_cdm - 0.25) < 1e-15 and sim.cosmo.Omega_b == 0.05


def test_waiting_for_running_boxes_keeps_the_pool():
    """A head request that cannot fit beside the running boxes, even with
    the pool emptied, waits without evicting pooled boxes."""
    be = FakeBackend()
    s = sv.Server(be, 3 * 16 ** 3, slots=4)
    big, small = sv.Request(0, 16), sv.Request(1, 12)
    got = s._take(big, 2 * 16 ** 3)           # a running box
    assert got is not None
    s._give_back(small, s._take(small, 12 ** 3)[0], 12 ** 3, True)  # a pooled box
    assert s._take(sv.Request(2, 16), 2 * 16 ** 3) is None
    assert len(s._pool) == 1 and s.evicted == 0
    s.close()


def test_a_request_that_cannot_fit_beside_held_memory_is_rejected_not_waited_for():
    be = FakeBackend()
    with sv.Server(be, 32 ** 3, slots=1) as s:
        s.in_use = 32 ** 3  # memory held outside the pool (accounting the server cannot free)
        out = s.serve(reqs([16]))
        s.in_use = 0
    assert out[0].status == "rejected" and "does not fit" in out[0].error


class FakeBatchBackend:
    """Batches whose slots finish a request after N // 8 steps."""

    def __init__(self, fail_key=None, step_s=0.005):
        self.live = set()
        self.fail_key = fail_key
        self.step_s = step_s
        self.next = 0
        self.starts = []

    def key(self, r):
        return (r.ngrid, 1 if r.gr else 0, 0)

    def slot_footprint(self, r):
        return r.ngrid ** 3

    def create(self, key, slots):
        self.next += 1
        b = {"id": self.next, "key": key, "left": {}, "slots": slots}
        self.live.add(self.next)
        return b

    def destroy(self, b):
        self.live.discard(b["id"])

    def start(self, b, slot, r):
        assert slot not in b["left"] and 0 <= slot < b["slots"]
        b["left"][slot] = r.ngrid // 8
        self.starts.append((r.rid, b["id"], slot))

    def step(self, b):
        if b["key"] == self.fail_key:
            raise RuntimeError("injected batch failure")
        time.sleep(self.step_s)
        fin = []
        for slot in list(b["left"]):
            b["left"][slot] -= 1
            if b["left"][slot] == 0:
                fin.append(slot)
        return fin

    def result(self, b, slot, r):
        return {"particle_steps": r.ngrid ** 3 * (r.ngrid // 8), "batch": b["id"], "slot": slot}

    def release(self, b, slot):
        del b["left"][slot]


def test_batch_server_shares_batches_and_admits_while_others_run():
    be = FakeBatchBackend()
    arr = [0.0, 0.0, 0.0, 0.02]
    rs = [sv.Request(i, 64, arrival=t) for i, t in enumerate(arr)]
    with sv.BatchServer(be, 10 ** 9, slots=lambda k: 4) as s:
        out = s.serve(rs)
        assert all(o.status == "ok" for o in out)
        assert s.created == 1 and len({o.result["batch"] for o in out}) == 1
        # the late request joined while the first three were still running
        assert out[3].admitted < min(o.finished for o in out[:3])
        assert s.peak_running == 4
    assert not be.live


def test_batch_server_frees_slots_for_waiting_requests_and_keeps_order():
    be = FakeBatchBackend()
    rs = [sv.Request(i, g) for i, g in enumerate([32, 64, 32, 32, 64])]
    budget = 2 * 32 ** 3 + 2 * 64 ** 3  # one two-slot batch of each shape
    with sv.BatchServer(be, budget, slots=lambda k: 2) as s:
        out = s.serve(rs)
        assert all(o.status == "ok" for o in out)
        assert s.peak_running <= 4
    adm = [o.admitted for o in out]
    assert adm == sorted(adm)
    # the third 32^3 request took the slot the first one left
    slots = {rid: (b, sl) for rid, b, sl in be.starts}
    assert slots[3] == slots[0] or slots[3] == slots[2]


def test_batch_server_budget_eviction_and_rejection():
    be = FakeBatchBackend()
    budget = 2 * 32 ** 3  # one two-slot batch of 32^3, or one of 16^3
    with sv.BatchServer(be, budget, slots=lambda k: 2) as s:
        out = s.serve([sv.Request(0, 32), sv.Request(1, 16), sv.Request(2, 128), sv.Request(3, 32)])
        assert [o.status for o in out] == ["ok", "ok", "rejected", "ok"]
        assert s.evicted >= 1 and s.peak_in_use <= budget
    assert not be.live


def test_batch_server_failed_batch_fails_its_requests_only():
    be = FakeBatchBackend(fail_key=(16, 1, 0))
    with sv.BatchServer(be, 10 ** 9, slots=lambda k: 2) as s:
        out = s.serve([sv.Request(0, 16), sv.Request(1, 32), sv.Request(2, 16)])
    assert [o.status for o in out] == ["failed", "ok", "failed"]
    assert "injected" in out[0].error and not be.live
